"""Pins the oracle's tuner functions: statistics against a dense brute force,
MLP gradients against central finite differences of the FP64 loss, Adam
against its closed-form first two steps."""
import numpy as np
import pytest

import oracle_ffi as of
from smat import gen


def _brute_stats(M, K, rp, ci, v):
    D = np.zeros((M, K), bool)
    Z = np.zeros((M, K), bool)  # explicit entries with a nonzero value
    rows = np.repeat(np.arange(M), np.diff(rp))
    D[rows, ci] = True
    Z[rows, ci] = (v & 0x7FFF) != 0
    lens = np.diff(rp)
    hist = np.zeros(12, np.int64)
    for L in lens:
        b = 0 if L == 0 else min(11, int(L).bit_length())
        hist[b] += 1
    off = np.abs(rows - ci)
    out = {"rows": M, "cols": K, "nnz": int(ci.size), "max_row": int(lens.max()),
           "sum_abs_offset": int(off.sum()), "near_diag": int((off <= 64).sum())}
    out.update({f"row_hist{b}": int(hist[b]) for b in range(12)})
    for w in range(6):
        bk = 1 << w
        blocks = D.reshape(M // 16, 16, K // bk, bk).any(axis=(1, 3))
        out[f"nnzb{w}"] = int(blocks.sum())
    b16 = D.reshape(M // 16, 16, K // 16, 16).any(axis=(1, 3))
    viol = (Z.reshape(M // 16, 16, K // 16, 4, 4).sum(axis=4) > 2).any(axis=(1, 3))
    out["nnzb16"] = int(b16.sum())
    out["blocks24"] = int((b16 & ~viol).sum())
    return out


@pytest.mark.parametrize("which", ["band", "rmat", "clustered", "powerlaw"])
def test_matrix_stats_brute_force(which):
    m = {"band": lambda: gen.band(n=300, hb=20, p=0.5, seed=3),
         "rmat": lambda: gen.rmat(scale=9, edge_factor=6, seed=4),
         "clustered": lambda: gen.block_clustered(n=512, tiles=60, seed=5),
         "powerlaw": lambda: gen.power_law(n=700, max_len=300, seed=6)}[which]()
    M, K, rp, ci, v = m
    v = v.copy()
    v[::7] = 0  # explicit zeros do not count towards 2:4 groups
    assert of.matrix_stats(M, K, rp, ci, v) == _brute_stats(M, K, rp, ci, v)


def test_matrix_stats_hand_case():
    # one block row: row 0 has 3 nonzeros in group 0 (violates 2:4), row 1 an entry at column 70
    M, K = 16, 96
    rp = np.array([0, 3, 4] + [4] * 14, np.int32)
    ci = np.array([0, 1, 2, 70], np.int32)
    v = np.full(4, 0x3C00, np.uint16)
    s = of.matrix_stats(M, K, rp, ci, v)
    assert s["nnzb4"] == 2 and s["nnzb5"] == 2 and s["nnzb0"] == 4 and s["nnzb16"] == 2 and s["blocks24"] == 1
    assert s["row_hist0"] == 14 and s["row_hist1"] == 1 and s["row_hist2"] == 1 and s["max_row"] == 3
    assert s["sum_abs_offset"] == 0 + 1 + 2 + 69 and s["near_diag"] == 3


def _params(n_in, h, rng):
    return [rng.standard_normal((h, n_in)).astype(np.float32) * 0.5, rng.standard_normal(h).astype(np.float32) * 0.1,
            rng.standard_normal((h, h)).astype(np.float32) * 0.3, rng.standard_normal(h).astype(np.float32) * 0.1,
            rng.standard_normal((1, h)).astype(np.float32) * 0.3, rng.standard_normal(1).astype(np.float32)]


def test_mlp_gradients_finite_differences():
    rng = np.random.default_rng(0)
    n_in, h, B = 7, 5, 9
    X = rng.standard_normal((B, n_in)).astype(np.float32)
    y = rng.standard_normal(B).astype(np.float32)
    params = _params(n_in, h, rng)
    loss, grads = of.mlp_backward(X, y, params)
    pred, _, _ = of.mlp_forward(X, params)
    assert abs(loss - np.mean((pred - y) ** 2)) < 1e-12
    # central differences in FP64 through a float64 copy of the forward pass
    def f64_loss(ps):
        W1, b1, W2, b2, W3, b3 = ps
        H1 = np.maximum(X.astype(np.float64) @ W1.T + b1, 0)
        H2 = np.maximum(H1 @ W2.T + b2, 0)
        return np.mean(((H2 @ W3.T + b3).ravel() - y) ** 2)
    base = [p.astype(np.float64) for p in params]
    assert abs(f64_loss(base) - loss) < 1e-9
    eps = 1e-6
    for k, g in enumerate(grads):
        flat = base[k].ravel()
        for idx in range(flat.size):
            old = flat[idx]
            flat[idx] = old + eps
            lp = f64_loss(base)
            flat[idx] = old - eps
            lm = f64_loss(base)
            flat[idx] = old
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - g.ravel()[idx]) <= 1e-6 * max(1.0, abs(fd)), (k, idx, fd, g.ravel()[idx])


def test_adam_closed_form():
    lr, b1, b2, eps = 0.01, 0.9, 0.999, 1e-8
    p = np.array([1.0, -2.0, 0.5]); g1 = np.array([0.3, -4.0, 0.0]); g2 = np.array([-0.1, 1.0, 2.0])
    m, v = np.zeros(3), np.zeros(3)
    q = p.copy()
    of.adam_step(q, g1.copy(), m, v, lr, b1, b2, eps, 1)
    # first step: m_hat = g, v_hat = g^2 -> p - lr * g / (|g| + eps)
    assert np.allclose(q, p - lr * g1 / (np.abs(g1) + eps), rtol=0, atol=1e-15)
    q2 = q.copy()
    of.adam_step(q2, g2.copy(), m, v, lr, b1, b2, eps, 2)
    m2 = b1 * (1 - b1) * g1 + (1 - b1) * g2
    v2 = b2 * (1 - b2) * g1 ** 2 + (1 - b2) * g2 ** 2
    want = q - lr * (m2 / (1 - b1 ** 2)) / (np.sqrt(v2 / (1 - b2 ** 2)) + eps)
    assert np.allclose(q2, want, rtol=0, atol=1e-14)
