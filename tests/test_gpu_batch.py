"""GPU parity of the batched persistent launch: every matrix of a batch (mixed
block widths, N-tiles 16-128, split rows, reordered rows) equals its own SpMM call,
eagerly and through repeated HIP-graph replays, and the FP64 oracle."""
import numpy as np
import pytest

import oracle_ffi as of
from smat import gen

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _rel_fro(C, R):
    return float(np.linalg.norm(C - R) / max(np.linalg.norm(R), 1e-300))


def _mats(count, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        kind = i % 4
        if kind == 0:
            m = gen.rmat(scale=int(rng.integers(9, 12)), edge_factor=int(rng.integers(4, 16)), seed=seed * 100 + i)
        elif kind == 1:
            m = gen.band(n=int(rng.integers(300, 2000)), hb=int(rng.integers(4, 40)), p=0.4, seed=seed * 100 + i)
        elif kind == 2:
            n = int(rng.integers(40, 120)) * 16
            m = gen.block_clustered(n=n, tiles=int(rng.integers(20, 200)), seed=seed * 100 + i)
        else:
            m = gen.power_law(n=int(rng.integers(500, 3000)), max_len=300, seed=seed * 100 + i)
        out.append(m)
    return out


def _build(mats, N, seed=1, with_rowmap=False):
    import smat
    rng = np.random.default_rng(seed)
    entries, refs, keep = [], [], []
    for i, (M, K, rp, ci, v) in enumerate(mats):
        bk = [1, 2, 4, 8, 16, 32][i % 6]
        n_tile = [0, 16, 32, 64, 128][i % 5]
        split = [0, 0, 16][i % 3]
        perm = None
        if with_rowmap and i % 4 in (1, 2):  # dense entries and the mixed dense / 2:4 ones
            perm = rng.permutation(M).astype(np.int32)  # A's row r is C's row perm[r]
        rp_a, ci_a, v_a = rp, ci, v
        if perm is not None:
            inv = np.argsort(perm)  # A row r = original row perm[r]
            lens = np.diff(rp)[perm]
            rp_a = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
            ci_a = np.concatenate([ci[rp[r]:rp[r + 1]] for r in perm]).astype(np.int32)
            v_a = np.concatenate([v[rp[r]:rp[r + 1]] for r in perm])
            del inv
        mixed = i % 4 == 2  # block-clustered matrices with 2:4 tiles: the mixed dense / 2:4 kernel
        if mixed:
            A = smat.Bcsr(M, K, _dev(rp_a), _dev(ci_a), _dev(v_a.view(np.int16)), bk=16,
                          tag_rule=smat.SMAT_TAG_EXACT24)
            v24, i24 = A.compress_24(smat.SMAT_24_SMFMAC)
            plan = smat.SpmmPlan(A, N, n_tile=min(n_tile, 64), max_blocks_per_item=split,
                                 kernel=smat.SMAT_KERNEL_MIXED24)  # 2:4 tiles are at most 64 wide
        else:
            A = smat.Bcsr(M, K, _dev(rp_a), _dev(ci_a), _dev(v_a.view(np.int16)), bk=bk)
            v24 = i24 = None
            plan = smat.SpmmPlan(A, N, n_tile=n_tile, max_blocks_per_item=split)
        B = gen.dense_b(K, N, seed=seed + i)
        Bd = _dev(B.view(np.int16))
        C = torch.full((M, N), float("nan"), dtype=torch.float32, device="cuda")
        rm = _dev(perm) if perm is not None else None
        ops = plan.operands(Bd, C, c_rowmap=rm, values24=v24, index24=i24)
        entries.append((plan, ops))
        keep.append((A, Bd, C, rm, v24, i24))
        refs.append(of.spmm_csr_f64(M, rp, ci, v, B))
    return entries, refs, keep


@pytest.mark.parametrize("with_rowmap", [False, True])
@pytest.mark.parametrize("N", [64, 256])
@pytest.mark.parametrize("mode", [1, 2])
def test_batch_matches_each_call_and_oracle(with_rowmap, N, mode):
    _need_gpu()
    import smat
    mats = _mats(24, seed=3)
    entries, refs, keep = _build(mats, N, with_rowmap=with_rowmap)
    print(f"batch of {len(entries)} N={N} rowmap={with_rowmap} mode={mode}", flush=True)
    # one call each
    singles = []
    for (plan, ops), (A, Bd, C, rm, v24, i24) in zip(entries, keep):
        plan.execute(Bd, C, c_rowmap=rm, values24=v24, index24=i24)
        torch.cuda.synchronize()
        singles.append(C.cpu().numpy().copy())
        C.fill_(float("nan"))
    batch = smat.Batch(entries, mode=mode)
    info = batch.info()
    assert info["tasks"] > 0 and info["grid"] > 0
    batch()
    torch.cuda.synchronize()
    for i, (ref, single, (A, Bd, C, *_)) in enumerate(zip(refs, singles, keep)):
        got = C.cpu().numpy()
        assert np.array_equal(got, single), i  # same tiles, same order of accumulation
        assert _rel_fro(got.astype(np.float64), ref) < 1e-5, i
    batch.close()


@pytest.mark.parametrize("mode", [1, 2])
def test_batch_graph_replays(mode):
    _need_gpu()
    import smat
    N = 32
    mats = _mats(12, seed=5)
    entries, refs, keep = _build(mats, N, seed=2)
    print(f"graph replays mode={mode}", flush=True)
    batch = smat.Batch(entries, mode=mode)
    batch()
    torch.cuda.synchronize()
    eager = [C.cpu().numpy().copy() for (_, _, C, *_) in keep]
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        batch.capture(s)
    assert batch.info()["captured"]
    for rep in range(4):
        for (_, _, C, *_) in keep:
            C.fill_(float("nan"))  # eager work between replays
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            batch(s)
        s.synchronize()
        for i, (_, _, C, *_) in enumerate(keep):
            assert np.array_equal(C.cpu().numpy(), eager[i]), (rep, i)
    # new B values in the same buffers are picked up by a replay
    (_, Bd, C0, *_) = keep[0]
    Bd.view(torch.float16).mul_(2)  # B holds FP16 bits
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        batch(s)
    s.synchronize()
    assert np.allclose(C0.cpu().numpy(), 2 * eager[0], rtol=1e-6, atol=1e-6)
    batch.close()


def test_batch_rejects_bad_plans():
    _need_gpu()
    import smat
    M, K, rp, ci, v = gen.band(n=512, hb=8, p=0.5, seed=1)
    A = smat.Bcsr(M, K, _dev(rp), _dev(ci), _dev(v.view(np.int16)), bk=16)
    B = _dev(gen.dense_b(K, 64).view(np.int16))
    C = torch.empty((M, 64), dtype=torch.float32, device="cuda")
    good = smat.SpmmPlan(A, 64)
    for kw in [dict(out_dtype=smat.SMAT_F16), dict(b_staging=smat.SMAT_BSTAGE_DIRECT),
               dict(in_dtype=smat.SMAT_BF16)]:
        bad = smat.SpmmPlan(A, 64, **kw)
        with pytest.raises(smat.SmatError):
            smat.Batch([(good, good.operands(B, C)), (bad, bad.operands(B, C))], mode=smat.SMAT_BATCH_PERSISTENT)
    for mode in (smat.SMAT_BATCH_PERSISTENT, smat.SMAT_BATCH_BRANCHES):
        with pytest.raises(smat.SmatError):  # default stream cannot be captured
            smat.Batch([(good, good.operands(B, C))], mode=mode).capture(0)
        split = smat.SpmmPlan(A, 64, max_blocks_per_item=2)  # split rows share the plan's scratch
        with pytest.raises(smat.SmatError):
            smat.Batch([(split, split.operands(B, C)), (split, split.operands(B, C))], mode=mode)
    with pytest.raises(smat.SmatError):
        smat.Batch([(good, good.operands(B, C))], mode=7)


def test_batch_branches_take_any_plan():
    """Branches run each matrix's own call, so plans the persistent kernel
    cannot take (FP16 C, direct B gathers, BF16, reference-layout B) batch
    too, with results identical to the single calls."""
    _need_gpu()
    import smat
    M, K, rp, ci, v = gen.rmat(scale=10, edge_factor=8, seed=7)
    N = 64
    B = gen.dense_b(K, N, seed=3)
    Bd = _dev(B.view(np.int16))
    Bt = _dev(np.ascontiguousarray(B.T).view(np.int16))  # reference layout B[n*K + k]
    A = smat.Bcsr(M, K, _dev(rp), _dev(ci), _dev(v.view(np.int16)), bk=16)
    kws = [dict(out_dtype=smat.SMAT_F16), dict(b_staging=smat.SMAT_BSTAGE_DIRECT), dict(in_dtype=smat.SMAT_BF16),
           dict(b_layout=smat.SMAT_B_REFERENCE), dict()]
    entries, outs, singles = [], [], []
    for kw in kws:
        plan = smat.SpmmPlan(A, N, **kw)
        f16 = kw.get("out_dtype") == smat.SMAT_F16
        C = torch.full((M, N), float("nan"), dtype=torch.float16 if f16 else torch.float32, device="cuda")
        Bx = Bt if kw.get("b_layout") == smat.SMAT_B_REFERENCE else Bd
        ldb = K if kw.get("b_layout") == smat.SMAT_B_REFERENCE else N
        plan.execute(Bx, C, ldb=ldb)
        torch.cuda.synchronize()
        singles.append(C.cpu().numpy().copy())
        C.fill_(float("nan"))
        entries.append((plan, plan.operands(Bx, C, ldb=ldb)))
        outs.append(C)
    batch = smat.Batch(entries, mode=smat.SMAT_BATCH_BRANCHES)
    batch()
    torch.cuda.synchronize()
    for i, C in enumerate(outs):
        assert np.array_equal(C.cpu().numpy(), singles[i]), i
    batch.close()
