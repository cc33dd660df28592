"""Learned tuner for the BCSR SpMM: an MLP cost model over matrix features.

For each matrix the model scores every candidate configuration (Jaccard
reorder threshold, block width, dense or 2:4 kernel, N-tile) and the lowest
score wins.  Features come from exact GPU statistics (smat_matrix_stats); the
network is trained on timings of this library's own exhaustive sweep over
seeded synthetic matrices, with the library's HIP kernels for the forward and
backward passes and Adam (smat_linear_forward / _backward, smat_mse_grad,
smat_adam_step), and inference runs in one HIP kernel (smat_tuner_predict).
The sweep itself (`sweep`) stays the fallback and the yardstick.

The reference has no learned part: it tunes by hand or by an exhaustive
search over the reorder threshold (preprocess/src/reorder_matrix.cpp:17-40,
preprocess/src/scripts/collect_taus.py:37-44).

Command line (run on a GPU):
    python -m smat.tuner collect --first 10000 --count 2000 --out sweep_train.npz
    python -m smat.tuner train --data sweep_train.npz --out mlp.npz
    python -m smat.tuner eval --data sweep_test.npz --weights mlp.npz
    python -m smat.tuner resweep --data sweep_train.npz --out sweep_train.npz   # after a kernel change
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

from . import (SMAT_24_SMFMAC, SMAT_KERNEL_DENSE, SMAT_KERNEL_MIXED24, SMAT_TAG_EXACT24, SMAT_TAG_REFERENCE,
               Bcsr, MatrixStats, SpmmPlan, check, lib)
from . import gen

NF = 26                      # SMAT_TUNER_NF
BKS = (1, 2, 4, 8, 16, 32)
NTILES = (16, 32, 64, 128)
NTILES_24 = (16, 32, 64)     # tile widths of the mixed dense / 2:4 kernel
# None: no reordering; else the Keeper threshold (16-row groups, 16-column
# blocks), the four values of the reference's R-MAT threshold sweep
TAUS = (None, 0.1, 0.3, 0.5, 0.7)
REORDER_MAX_ROWS = 16384     # the exact Keeper's host replay grows ~quadratically; larger matrices skip it
HIDDEN = 128
WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tuner", "mlp_weights.npz")


def _lib():
    return lib()


def _stream():
    import torch
    return torch.cuda.current_stream().cuda_stream


# ---- configuration space -------------------------------------------------------
def configs():
    """Candidate list: (tau, bk, n_tile, kernel)."""
    out = []
    for tau in TAUS:
        for bk in BKS:
            for nt in NTILES:
                out.append((tau, bk, nt, "dense"))
        for nt in NTILES_24:
            out.append((tau, 16, nt, "mixed24"))
    return out


CONFIGS = configs()
# encoding: one-hot block width, one-hot N-tile, 2:4 kernel, reordered, tau (0 without reordering)
ENC_REORDER = len(BKS) + len(NTILES) + 1
NE = ENC_REORDER + 2
PLAIN = CONFIGS.index((None, 16, 64, "dense"))  # the reference's configuration (16x16, dense, no reorder)


def encode(cfg) -> np.ndarray:
    tau, bk, nt, kernel = cfg
    e = np.zeros(NE, np.float32)
    e[BKS.index(bk)] = 1
    e[len(BKS) + NTILES.index(nt)] = 1
    e[len(BKS) + len(NTILES)] = kernel == "mixed24"
    e[ENC_REORDER] = tau is not None
    e[ENC_REORDER + 1] = tau or 0.0
    return e


ENC = np.stack([encode(c) for c in CONFIGS])


def available(rows: int) -> np.ndarray:
    return np.array([c[0] is None or rows <= REORDER_MAX_ROWS for c in CONFIGS], np.uint8)


# ---- features -------------------------------------------------------------------
def stats(M, K, d_rp, d_ci, d_v, nnz) -> MatrixStats:
    st = MatrixStats()
    check(_lib().smat_matrix_stats(M, K, nnz, d_rp.data_ptr(), d_ci.data_ptr(), d_v.data_ptr(), ctypes.byref(st),
                                   _stream()))
    return st


def features(st: MatrixStats, N: int) -> np.ndarray:
    out = np.zeros(NF, np.float32)
    check(_lib().smat_tuner_features(ctypes.byref(st), N, out.ctypes.data))
    return out


# ---- synthetic matrices (the batched workload's distribution) ------------------
def make_matrix(seed: int, large: bool = False):
    """Seeded matrix of the batched workload: n log-uniform in [2k, 200k], nnz
    log-uniform in [1e4, 5e6] (clamped by what the generator can place), the
    generator cycling over band, R-MAT, block-clustered and power-law.  With
    large=True: n in [100k, 600k], nnz in [1e6, 1.2e7] (single-matrix sizes
    such as R-MAT scale 17-19).  Returns (kind, (M, K, rowptr, colidx, FP16
    bits))."""
    rng = np.random.default_rng(seed)
    kind = ("band", "rmat", "clustered", "powerlaw")[seed % 4]
    n_lo, n_hi, z_lo, z_hi = (1e5, 6e5, 1e6, 1.2e7) if large else (2e3, 2e5, 1e4, 5e6)
    n = int(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi))))
    nnz = math.exp(rng.uniform(math.log(z_lo), math.log(z_hi)))
    nnz = min(nnz, 64.0 * n)
    gs = int(rng.integers(0, 2 ** 31))
    if kind == "band":
        p = float(rng.uniform(0.2, 0.9))
        hb = max(1, min(n // 4, int(nnz / n / p / 2)))
        return kind, gen.band(n=n, hb=hb, p=p, seed=gs)
    if kind == "rmat":
        scale = max(11, int(round(math.log2(n))))
        ef = max(1, int(round(nnz / 2 ** scale)))
        return kind, gen.rmat(scale=scale, edge_factor=min(ef, 48), seed=gs)
    if kind == "clustered":
        n16 = max(64, n // 16) * 16
        tiles = int(min(max(8, nnz / 128), (n16 // 16) ** 2 // 2))
        return kind, gen.block_clustered(n=n16, tiles=tiles, seed=gs)
    L = int(min(max(4, (nnz / (0.766 * n)) ** 2), n // 2, 4096))
    return kind, gen.power_law(n=n, max_len=L, seed=gs)


def make_tiled_matrix(seed: int):
    """Seeded tile-clustered matrix (nonzeros clustered in aligned bh x bw
    tiles, the structure merged tiles are for): n log-uniform in [4k, 131k],
    tile shape from {16x64, 32x32, 64x64, 64x128, 128x64, 128x128}, entry fill
    U[0.25, 1], pruned to 2:4 with probability 0.3, nnz log-uniform in
    [5e4, 8e6] (at most half the tile grid).  Returns ("tiled", CSR)."""
    rng = np.random.default_rng(seed)
    bh, bw = ((16, 64), (32, 32), (64, 64), (64, 128), (128, 64), (128, 128))[int(rng.integers(6))]
    n = int(math.exp(rng.uniform(math.log(4e3), math.log(1.31e5))))
    n = max(4, n // 128) * 128
    fill = float(rng.uniform(0.25, 1.0))
    prune = bool(rng.random() < 0.3)
    nnz = math.exp(rng.uniform(math.log(5e4), math.log(8e6)))
    per_tile = bh * bw * fill * (0.5 if prune else 1.0)
    tiles = int(min(max(8, nnz / per_tile), (n // bh) * (n // bw) // 2))
    gs = int(rng.integers(0, 2 ** 31))
    return "tiled", gen.block_clustered(n=n, tiles=tiles, seed=gs, prune=prune, bh=bh, bw=bw, fill=fill)


# ---- the sweep (fallback and yardstick) ----------------------------------------
def reorder_csr(rp, ci, v, tau):
    """Keeper grouping -> fixed-size groups -> row permutation, applied on the
    host arrays.  Returns (rp, ci, v, perm): new row r is old row perm[r]."""
    from . import fixed_size_grouping, grouping_to_permutation, reorder_grouping
    g, _ = reorder_grouping(rp, ci, np.float32(tau), 16, 16)
    perm = grouping_to_permutation(fixed_size_grouping(g, 16))
    lens = np.diff(rp)[perm]
    nrp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    idx = np.concatenate([np.arange(rp[r], rp[r + 1]) for r in perm]) if ci.size else np.zeros(0, np.int64)
    return nrp, ci[idx], v[idx], perm


def time_call(fn, reps=5, warmup=2):
    import torch
    for _ in range(warmup):
        fn()
    st = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for s, e in ev:
        s.record(st)
        fn()
        e.record(st)
    torch.cuda.synchronize()
    return float(np.median([s.elapsed_time(e) for s, e in ev]))


def sweep(M, K, rp, ci, v, N, B=None, reps=5, taus=TAUS, log=None):
    """Times every candidate configuration (ms per call, inf where not
    available).  Returns (times[len(CONFIGS)], extra info)."""
    import torch
    times = np.full(len(CONFIGS), np.inf)
    info = {"reorder_s": {}}
    if B is None:
        B = torch.from_numpy(gen.dense_b(K, N, seed=1).view(np.int16)).cuda()
    C = torch.empty((M, N), dtype=torch.float32, device="cuda")
    for tau in taus:
        if tau is not None and M > REORDER_MAX_ROWS:
            continue
        rowmap = None
        if tau is None:
            r_rp, r_ci, r_v = rp, ci, v
        else:
            t0 = time.perf_counter()
            r_rp, r_ci, r_v, perm = reorder_csr(rp, ci, v, tau)
            info["reorder_s"][str(tau)] = time.perf_counter() - t0
            rowmap = torch.from_numpy(perm.astype(np.int32)).cuda()
        d_rp = torch.from_numpy(r_rp).cuda()
        d_ci = torch.from_numpy(r_ci if r_ci.size else np.zeros(1, np.int32)).cuda()
        d_v = torch.from_numpy((r_v if r_v.size else np.zeros(1, np.uint16)).view(np.int16)).cuda()
        for bk in BKS:
            for kernel in ("dense", "mixed24"):
                if kernel == "mixed24" and bk != 16:
                    continue
                A = Bcsr(M, K, d_rp, d_ci, d_v, bk=bk,
                         tag_rule=SMAT_TAG_EXACT24 if kernel == "mixed24" else SMAT_TAG_REFERENCE)
                v24 = i24 = None
                if kernel == "mixed24":
                    v24, i24 = A.compress_24(SMAT_24_SMFMAC)
                for nt in NTILES_24 if kernel == "mixed24" else NTILES:
                    plan = SpmmPlan(A, N, n_tile=nt,
                                    kernel=SMAT_KERNEL_MIXED24 if kernel == "mixed24" else SMAT_KERNEL_DENSE)
                    t = time_call(lambda: plan.execute(B, C, values24=v24, index24=i24, c_rowmap=rowmap), reps=reps)
                    times[CONFIGS.index((tau, bk, nt, kernel))] = t
                    plan.close()
                A.close()
    return times, info


# ---- training (HIP kernels) -----------------------------------------------------
def _ptr(t):
    return t.data_ptr()


class Net:
    """Parameters of in -> HIDDEN -> HIDDEN -> 1 as device tensors."""

    def __init__(self, n_in, hidden=HIDDEN, seed=0, arrays=None):
        import torch
        self.n_in, self.h = n_in, hidden
        if arrays is None:
            rng = np.random.default_rng(seed)
            he = lambda fan_in, shape: (rng.standard_normal(shape) * math.sqrt(2.0 / fan_in)).astype(np.float32)  # noqa: E731
            arrays = [he(n_in, (hidden, n_in)), np.zeros(hidden, np.float32), he(hidden, (hidden, hidden)),
                      np.zeros(hidden, np.float32), he(hidden, (1, hidden)), np.zeros(1, np.float32)]
        self.p = [torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda() for a in arrays]

    def numpy(self):
        return [t.cpu().numpy() for t in self.p]


def forward_backward(net: Net, X, y, need_grad=True):
    """One pass on a device batch X [B, n_in], y [B] with the library kernels.
    Returns (loss tensor, predictions, grads list)."""
    import torch
    L = _lib()
    s = _stream()
    B = X.shape[0]
    W1, b1, W2, b2, W3, b3 = net.p
    h = net.h
    H1 = torch.empty((B, h), device="cuda")
    H2 = torch.empty((B, h), device="cuda")
    P = torch.empty((B, 1), device="cuda")
    check(L.smat_linear_forward(_ptr(X), B, net.n_in, _ptr(W1), _ptr(b1), h, 1, _ptr(H1), s))
    check(L.smat_linear_forward(_ptr(H1), B, h, _ptr(W2), _ptr(b2), h, 1, _ptr(H2), s))
    check(L.smat_linear_forward(_ptr(H2), B, h, _ptr(W3), _ptr(b3), 1, 0, _ptr(P), s))
    loss = torch.empty(1, device="cuda")
    dP = torch.empty((B, 1), device="cuda")
    check(L.smat_mse_grad(_ptr(P), _ptr(y), B, _ptr(dP), _ptr(loss), s))
    if not need_grad:
        return loss, P, None
    g = [torch.empty_like(t) for t in net.p]
    dH2 = torch.empty((B, h), device="cuda")
    dH1 = torch.empty((B, h), device="cuda")
    check(L.smat_linear_backward(_ptr(H2), _ptr(P), _ptr(dP), B, h, _ptr(W3), 1, 0, _ptr(dH2), _ptr(g[4]),
                                 _ptr(g[5]), s))
    check(L.smat_linear_backward(_ptr(H1), _ptr(H2), _ptr(dH2), B, h, _ptr(W2), h, 1, _ptr(dH1), _ptr(g[2]),
                                 _ptr(g[3]), s))
    check(L.smat_linear_backward(_ptr(X), _ptr(H1), _ptr(dH1), B, net.n_in, _ptr(W1), h, 1, None, _ptr(g[0]),
                                 _ptr(g[1]), s))
    return loss, P, g


def samples(feats, times):
    """(matrix, config) rows with finite times: inputs and per-matrix centred
    log times (the model ranks configurations of one matrix).  The centre is
    the mean log time of the configurations without reordering, which every
    matrix has, so the offset does not depend on which reordered candidates a
    matrix was swept with."""
    X, y, owner = [], [], []
    plain_order = np.array([c[0] is None for c in CONFIGS])
    for i in range(feats.shape[0]):
        ok = np.isfinite(times[i])
        if ok.sum() < 2:
            continue
        ref = ok & plain_order
        lt = np.log(times[i][ok])
        lt = lt - (np.log(times[i][ref]).mean() if ref.any() else lt.mean())
        for c, t in zip(np.nonzero(ok)[0], lt):
            X.append(np.concatenate([feats[i], ENC[c]]))
            y.append(t)
            owner.append(i)
    return np.asarray(X, np.float32), np.asarray(y, np.float32), np.asarray(owner)


def train(feats, times, epochs=100, batch=1024, lr=2e-3, seed=0, log=print):
    """Trains the cost model on the GPU with the library's kernels (Adam).
    Returns a dict of numpy arrays (weights, standardisation) and the loss curve."""
    import torch
    X, y, _ = samples(feats, times)
    mu = X.mean(0)
    sigma = X.std(0)
    # an input that never varies in the data carries no information: a huge
    # sigma makes its standardised value ~0 for any input at inference too
    sigma[sigma < 1e-6] = 1e30
    Xs = ((X - mu) / sigma).astype(np.float32)
    dX = torch.from_numpy(Xs).cuda()
    dy = torch.from_numpy(y).cuda()
    net = Net(X.shape[1], seed=seed)
    m = [torch.zeros_like(t) for t in net.p]
    v = [torch.zeros_like(t) for t in net.p]
    rng = np.random.default_rng(seed)
    L = _lib()
    s = _stream()
    step = 0
    curve = []
    for ep in range(epochs):
        order = torch.from_numpy(rng.permutation(len(y))).cuda()
        tot = 0.0
        nb = 0
        for b0 in range(0, len(y), batch):
            idx = order[b0:b0 + batch]
            xb, yb = dX[idx].contiguous(), dy[idx].contiguous()
            loss, _, g = forward_backward(net, xb, yb)
            step += 1
            for p_, g_, m_, v_ in zip(net.p, g, m, v):
                check(L.smat_adam_step(_ptr(p_), _ptr(g_), _ptr(m_), _ptr(v_), p_.numel(), lr, 0.9, 0.999, 1e-8,
                                       step, s))
            tot += float(loss.item())
            nb += 1
        curve.append(tot / max(nb, 1))
        if log and (ep % 50 == 0 or ep == epochs - 1):
            log(f"epoch {ep}: train mse {curve[-1]:.4f}")
    W = net.numpy()
    return {"W1": W[0], "b1": W[1], "W2": W[2], "b2": W[3], "W3": W[4], "b3": W[5], "mu": mu.astype(np.float32),
            "sigma": sigma.astype(np.float32), "hidden": np.int32(HIDDEN), "curve": np.asarray(curve, np.float32)}


def train_select(feats, times, rows, seeds=(0, 1), val_frac=1.0 / 7, epochs=100, log=print):
    """Trains one network per seed on a fixed split of the matrices (a
    seeded permutation; `val_frac` of them held out for validation) and keeps
    the one whose picks land within 10 % of the sweep best on the most
    validation matrices (ties: the higher geometric-mean gain over the plain
    configuration).  Returns (weights, report)."""
    perm = np.random.default_rng(0).permutation(len(feats))
    nv = int(len(perm) * val_frac)
    val, tr = perm[:nv], perm[nv:]
    best, best_key, report = None, None, []
    for seed in seeds:
        w = train(feats[tr], times[tr], epochs=epochs, seed=seed, log=log)
        r = evaluate(Model(w), feats[val], times[val], rows[val])
        key = (r["within_10pct"], r["geo_speedup_vs_plain"])
        report.append({"seed": int(seed), "val_within_10pct": r["within_10pct"],
                       "val_geo_speedup_vs_plain": r["geo_speedup_vs_plain"]})
        if log:
            log(f"seed {seed}: validation {key[0]:.3f} within 10 %, {key[1]:.3f}x plain")
        if best_key is None or key > best_key:
            best, best_key = w, key
    best = dict(best, val_frac=np.float32(val_frac), train_matrices=np.int32(len(tr)),
                selection=np.asarray(json.dumps(report)))
    return best, report


# ---- inference (HIP kernel) -----------------------------------------------------
class Model:
    def __init__(self, weights: dict | str | None = None):
        import torch
        if weights is None:
            weights = WEIGHTS_PATH
        if isinstance(weights, str):
            weights = dict(np.load(weights))
        self.w = weights
        self.h = int(weights["hidden"])
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()  # noqa: E731
        self.d = {k: t(weights[k]) for k in ("W1", "b1", "W2", "b2", "W3", "b3", "mu", "sigma")}
        self.enc = t(ENC)

    def scores(self, feats: np.ndarray, avail: np.ndarray):
        """(scores[len(CONFIGS)], best index) from the inference kernel."""
        import torch
        f = torch.from_numpy(np.ascontiguousarray(feats, np.float32)).cuda()
        a = torch.from_numpy(np.ascontiguousarray(avail, np.uint8)).cuda()
        sc = torch.empty(len(CONFIGS), device="cuda")
        best = torch.empty(1, dtype=torch.int32, device="cuda")
        d = self.d
        check(_lib().smat_tuner_predict(_ptr(f), NF, _ptr(self.enc), NE, _ptr(a), len(CONFIGS), _ptr(d["mu"]),
                                        _ptr(d["sigma"]), _ptr(d["W1"]), _ptr(d["b1"]), _ptr(d["W2"]),
                                        _ptr(d["b2"]), _ptr(d["W3"]), _ptr(d["b3"]), self.h, _ptr(sc),
                                        _ptr(best), _stream()))
        return sc.cpu().numpy(), int(best.item())

    def choose(self, feats, rows):
        return CONFIGS[self.scores(feats, available(rows))[1]]


def probe_locality(model: Model, M, K, d_rp, d_ci, d_v, nnz, N, B, C, pick=None, reps=10, margin=0.03):
    """Locality row order for a matrix above the exact Keeper's row limit
    (device CSR, B and C tensors): the model's pick on the matrix as given and
    on its locality-ordered copy (smat_locality_order + smat_permute_csr_rows;
    the copy also runs the as-given pick when the two differ), each timed over
    `reps` calls after 2 warm-up calls; the order is kept when it is faster by
    more than `margin` (a probe of a 0.1 ms call moves by a few per cent
    between runs).  Both picks skip the Keeper candidates (tau set).  Returns
    (ordered, pick, order, report): order is (rowptr, colidx, values, rowmap)
    of the ordered copy when kept, else None."""
    from . import Bcsr, SpmmPlan, SMAT_KERNEL_DENSE, SMAT_KERNEL_MIXED24, SMAT_TAG_EXACT24, SMAT_TAG_REFERENCE, \
        SMAT_24_SMFMAC, locality_order, permute_rows
    import torch
    avail = available(M) * (ENC[:, -1] == 0)
    if pick is None:
        pick = CONFIGS[model.scores(features(stats(M, K, d_rp, d_ci, d_v, nnz), N), avail)[1]]
    perm = locality_order(M, K, d_rp, d_ci)
    o_rp, o_ci, o_v = permute_rows(d_rp, d_ci, d_v, perm)
    rowmap = perm.to(torch.int32)
    o_pick = CONFIGS[model.scores(features(stats(M, K, o_rp, o_ci, o_v, nnz), N), avail)[1]]

    def probe(cfg, rp, ci, v, rm):
        _, bk, nt, kernel = cfg
        mixed = kernel == "mixed24"
        A = Bcsr(M, K, rp, ci, v, bk=bk, tag_rule=SMAT_TAG_EXACT24 if mixed else SMAT_TAG_REFERENCE)
        v24 = i24 = None
        if mixed:
            v24, i24 = A.compress_24(SMAT_24_SMFMAC)
        plan = SpmmPlan(A, N, n_tile=nt, kernel=SMAT_KERNEL_MIXED24 if mixed else SMAT_KERNEL_DENSE)
        ms = time_call(lambda: plan.execute(B, C, values24=v24, index24=i24, c_rowmap=rm), reps=reps)
        plan.close()
        A.close()
        return ms

    as_given = probe(pick, d_rp, d_ci, d_v, None)
    # the ordered copy runs the model's pick for it and, when that differs,
    # the pick for the matrix as given (the sweeps the model learned from
    # time matrices as given, so its pick on a reordered copy is the less
    # certain one); the faster of the two is compared
    ordered_ms = probe(o_pick, o_rp, o_ci, o_v, rowmap)
    if pick != o_pick:
        same_ms = probe(pick, o_rp, o_ci, o_v, rowmap)
        if same_ms < ordered_ms:
            o_pick, ordered_ms = pick, same_ms
    use = ordered_ms < as_given * (1.0 - margin)
    report = {"config": o_pick, "probe_event_ms": {"as_given": round(as_given, 4), "locality": round(ordered_ms, 4)},
              "used": bool(use)}
    return use, (o_pick if use else pick), ((o_rp, o_ci, o_v, rowmap) if use else None), report


def evaluate(model: Model, feats, times, rows):
    """Per held-out matrix: time of the model's pick / sweep-best time, and the
    speedups of the pick over the plain configuration and over the single
    configuration that is best on average (the method without a model)."""
    picks, ratio, vs_plain = [], [], []
    for i in range(feats.shape[0]):
        _, b = model.scores(feats[i], available(int(rows[i])))
        if not np.isfinite(times[i][b]):
            b = int(np.nanargmin(np.where(np.isfinite(times[i]), times[i], np.nan)))
        picks.append(b)
        best = np.min(times[i])
        ratio.append(times[i][b] / best)
        vs_plain.append(times[i][PLAIN] / times[i][b])
    ratio = np.asarray(ratio)
    return {"n": len(ratio), "within_10pct": float(np.mean(ratio <= 1.10)),
            "median_ratio_to_best": float(np.median(ratio)), "geo_speedup_vs_plain": float(np.exp(np.mean(np.log(vs_plain)))),
            "picks": picks, "ratio": ratio.tolist()}


def fixed_best(times):
    """The single configuration with the best geometric-mean time over matrices
    (only configurations available everywhere)."""
    ok = np.all(np.isfinite(times), axis=0)
    lg = np.where(ok, np.log(np.where(np.isfinite(times), times, 1.0)).mean(0), np.inf)
    return int(np.argmin(lg))


# ---- data collection ------------------------------------------------------------
def _gen_worker(seed):
    kind, (M, K, rp, ci, v) = make_matrix(seed)
    return seed, kind, M, K, rp, ci, v


def _gen_worker_large(seed):
    kind, (M, K, rp, ci, v) = make_matrix(seed, large=True)
    return seed, kind, M, K, rp, ci, v


def _gen_worker_tiled(seed):
    kind, (M, K, rp, ci, v) = make_tiled_matrix(seed)
    return seed, kind, M, K, rp, ci, v


def collect(first, count, out, N=128, workers=8, log=print, n_list=None, large=False, family="workload"):
    import multiprocessing as mp

    import torch
    rec = {"seed": [], "kind": [], "rows": [], "nnz": [], "N": [], "feats": [], "times": [], "stats": []}
    t0 = time.time()
    # the workers only run numpy generators; they are spawned, so no process
    # that has touched the GPU forks
    ctx = mp.get_context("spawn")
    with ctx.Pool(workers) as pool:
        torch.zeros(1).cuda()
        worker = _gen_worker_tiled if family == "tiled" else _gen_worker_large if large else _gen_worker
        for k, (seed, kind, M, K, rp, ci, v) in enumerate(pool.imap(worker, range(first, first + count))):
            n_cols = n_list[seed % len(n_list)] if n_list else N
            d_rp, d_ci = torch.from_numpy(rp).cuda(), torch.from_numpy(ci).cuda()
            d_v = torch.from_numpy(v.view(np.int16)).cuda()
            st = stats(M, K, d_rp, d_ci, d_v, int(ci.size))
            f = features(st, n_cols)
            times, _ = sweep(M, K, rp, ci, v, n_cols)
            for key, val in (("seed", seed), ("kind", kind), ("rows", M), ("nnz", int(ci.size)), ("N", n_cols),
                             ("feats", f), ("times", times), ("stats", json.dumps(st.as_dict()))):
                rec[key].append(val)
            if k % 20 == 0 or k == count - 1:
                log(f"[{k + 1}/{count}] seed {seed} {kind} rows {M} nnz {ci.size}: best "
                    f"{CONFIGS[int(np.argmin(times))]} {np.min(times):.3f} ms, plain {times[PLAIN]:.3f} ms "
                    f"({time.time() - t0:.0f}s)")
                _save(out, rec)
    _save(out, rec)
    return out


def _gen_worker_any(job):
    seed, family, large = job
    kind, (M, K, rp, ci, v) = make_tiled_matrix(seed) if family == "tiled" else make_matrix(seed, large=large)
    return seed, kind, M, K, rp, ci, v


def resweep(path_in, path_out, family="workload", large=False, count=None, workers=8, deadline_s=None, log=print):
    """Times every configuration again, with the current kernels, for the
    matrices of a sweep file (regenerated from their seeds and checked against
    the recorded sizes).  Resumable: the output keeps a `fresh` mask, and when
    `path_out` exists it is the starting point; at most `count` matrices (or
    until `deadline_s` seconds have passed) per call.  Returns the number of
    fresh matrices."""
    import multiprocessing as mp

    import torch
    base = np.load(path_out if os.path.exists(path_out) else path_in)
    old_cfgs = [tuple(json.loads(c)) for c in base["configs"]]
    col = {c: i for i, c in enumerate(old_cfgs)}
    n = len(base["seed"])
    times = np.full((n, len(CONFIGS)), np.inf)
    for j, c in enumerate(CONFIGS):
        if c in col:
            times[:, j] = base["times"][:, col[c]]
    fresh = base["fresh"].copy() if "fresh" in base else np.zeros(n, bool)
    rec = {k: list(base[k]) for k in ("seed", "kind", "rows", "nnz", "N", "stats")}
    rec["feats"] = list(base["feats"])
    todo = [i for i in range(n) if not fresh[i]][:count]
    t0 = time.time()
    # the workers only run numpy generators; they are spawned before the GPU
    # is touched, so no process that has used the GPU forks
    ctx = mp.get_context("spawn")
    with ctx.Pool(workers) as pool:
        torch.zeros(1).cuda()
        jobs = [(int(base["seed"][i]), family, large) for i in todo]
        for k, (i, (seed, kind, M, K, rp, ci, v)) in enumerate(zip(todo, pool.imap(_gen_worker_any, jobs))):
            if deadline_s is not None and time.time() - t0 > deadline_s:
                break
            assert seed == base["seed"][i] and M == base["rows"][i] and ci.size == base["nnz"][i], seed
            times[i], _ = sweep(M, K, rp, ci, v, int(base["N"][i]))
            fresh[i] = True
            if log:
                log(f"[{k + 1}/{len(todo)}] seed {seed} {kind} rows {M}: best {CONFIGS[int(np.argmin(times[i]))]} "
                    f"({time.time() - t0:.0f}s)")
            if k % 10 == 9:
                _save(path_out, dict(rec, times=list(times)), fresh)
        pool.terminate()
    _save(path_out, dict(rec, times=list(times)), fresh)
    return int(fresh.sum())


def _save(path, rec, fresh=None):
    extra = {} if fresh is None else {"fresh": np.asarray(fresh, bool)}
    np.savez_compressed(path, seed=np.asarray(rec["seed"]), kind=np.asarray(rec["kind"]),
                        rows=np.asarray(rec["rows"]), nnz=np.asarray(rec["nnz"]), N=np.asarray(rec["N"]),
                        feats=np.asarray(rec["feats"], np.float32).reshape(-1, NF),
                        times=np.asarray(rec["times"], np.float64).reshape(-1, len(CONFIGS)),
                        stats=np.asarray(rec["stats"]), configs=np.asarray([json.dumps(c) for c in CONFIGS]),
                        **extra)


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m smat.tuner")
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect")
    c.add_argument("--first", type=int, required=True)
    c.add_argument("--count", type=int, required=True)
    c.add_argument("--out", required=True)
    c.add_argument("--n", type=int, default=128)
    c.add_argument("--n-list", default="", help="comma-separated N values cycled over the seeds")
    c.add_argument("--workers", type=int, default=8)
    c.add_argument("--large", action="store_true", help="single-matrix sizes (n 100k-600k, nnz 1e6-1.2e7)")
    c.add_argument("--family", default="workload", choices=["workload", "tiled"],
                   help="workload: the batched workload's generators; tiled: tile-clustered matrices")
    t = sub.add_parser("train")
    t.add_argument("--data", nargs="+", required=True)
    t.add_argument("--out", required=True)
    t.add_argument("--epochs", type=int, default=100)
    t.add_argument("--seeds", default="0,1", help="one network per seed; the best on the validation split is kept")
    t.add_argument("--val-frac", type=float, default=1.0 / 7)
    f = sub.add_parser("resweep", help="time every configuration again with the current kernels")
    f.add_argument("--data", required=True)
    f.add_argument("--out", required=True)
    f.add_argument("--family", default="workload", choices=["workload", "tiled"])
    f.add_argument("--large", action="store_true")
    f.add_argument("--count", type=int, default=None)
    f.add_argument("--seconds", type=float, default=None, help="stop after this long (resume with the same --out)")
    f.add_argument("--workers", type=int, default=8)
    e = sub.add_parser("eval")
    e.add_argument("--data", nargs="+", required=True)
    e.add_argument("--weights", default=WEIGHTS_PATH)
    a = ap.parse_args(argv)
    log = lambda *x: print(*x, flush=True)  # noqa: E731
    if a.cmd == "collect":
        n_list = [int(x) for x in a.n_list.split(",")] if a.n_list else None
        collect(a.first, a.count, a.out, N=a.n, workers=a.workers, log=log, n_list=n_list, large=a.large,
                family=a.family)
    elif a.cmd == "resweep":
        resweep(a.data, a.out, family=a.family, large=a.large, count=a.count, workers=a.workers,
                deadline_s=a.seconds, log=log)
    elif a.cmd == "train":
        fs = [np.load(p) for p in a.data]
        feats = np.concatenate([f["feats"] for f in fs])
        times = np.concatenate([f["times"] for f in fs])
        rows = np.concatenate([f["rows"] for f in fs])
        w, _ = train_select(feats, times, rows, seeds=[int(x) for x in a.seeds.split(",")], val_frac=a.val_frac,
                            epochs=a.epochs, log=log)
        np.savez(a.out, **w)
        log(f"saved {a.out}")
    else:
        fs = [np.load(p) for p in a.data]
        feats = np.concatenate([f["feats"] for f in fs])
        times = np.concatenate([f["times"] for f in fs])
        rows = np.concatenate([f["rows"] for f in fs])
        r = evaluate(Model(a.weights), feats, times, rows)
        r.pop("picks")
        r.pop("ratio")
        log(json.dumps(r))


if __name__ == "__main__":
    sys.exit(main())
