"""GPU parity of the BCSR build and the dense-block MFMA SpMM against the
CPU oracle (FP64 ground truth, bit-exact integer layout).  Calls go through
the C ABI (libsmat.so via ctypes)."""
import numpy as np
import pytest

import oracle_ffi as of
from smat import gen

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.view(dtype)
    return t.cuda()


def _bcsr(M, K, rp, ci, v, bk=16, tag_rule=None):
    import smat
    tr = smat.SMAT_TAG_REFERENCE if tag_rule is None else tag_rule
    return smat.Bcsr(M, K, _dev(rp), _dev(ci), _dev(v.view(np.int16)), bk=bk, tag_rule=tr)


def _rel_fro(C, R):
    return float(np.linalg.norm(C - R) / max(np.linalg.norm(R), 1e-300))


CASES = [
    ("band_small", lambda: gen.band(n=512, hb=32, p=0.5, seed=21)),
    ("rmat_small", lambda: gen.rmat(scale=10, edge_factor=8, seed=22)),
    ("clustered_small", lambda: gen.block_clustered(n=1024, tiles=300, seed=23)),
]


@pytest.mark.parametrize("name,make", CASES)
@pytest.mark.parametrize("bk", [1, 2, 4, 8, 16, 32])
def test_bcsr_build_bit_exact(name, make, bk):
    _need_gpu()
    M, K, rp, ci, v = make()
    print(f"case {name} bk={bk} M={M} nnz={ci.size}", flush=True)
    A = _bcsr(M, K, rp, ci, v, bk=bk)
    got = A.arrays()
    exp = of.csr_to_bcsr(M, K, bk, rp, ci, v)
    assert A.nnzb == exp[1].size
    for g, e, what in zip(got, exp, ("rowptr", "colidx", "values", "tags")):
        assert np.array_equal(g.numpy().view(e.dtype), e), what


def test_bcsr_build_unsorted_columns_and_tags():
    _need_gpu()
    # columns in file order (not sorted) change the reference's tag rule
    rng = np.random.default_rng(3)
    M, K = 64, 64
    rows, cols = [], []
    for r in range(M):
        c = rng.choice(K, size=rng.integers(0, 20), replace=False)
        rows += [r] * c.size
        cols += c.tolist()
    rows, cols = np.array(rows), np.array(cols, np.int32)
    rp = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=M))]).astype(np.int32)
    v = rng.uniform(-1, 1, cols.size).astype(np.float16)
    v[v == 0] = 1
    v = v.view(np.uint16)
    A = _bcsr(M, K, rp, cols, v, 16)
    got = A.arrays()
    exp = of.csr_to_bcsr(M, K, 16, rp, cols, v)
    for g, e in zip(got, exp):
        assert np.array_equal(g.numpy().view(e.dtype), e)
    import smat
    A2 = _bcsr(M, K, rp, cols, v, 16, tag_rule=smat.SMAT_TAG_EXACT24)
    assert np.array_equal(A2.arrays()[3].numpy(), of.tag_exact24(16, exp[2]))


SPMM_CASES = [
    ("band", lambda: gen.band(n=1024, hb=64, p=0.5, seed=31), 128),
    ("rmat", lambda: gen.rmat(scale=11, edge_factor=16, seed=32), 64),
    ("rmat_n256", lambda: gen.rmat(scale=10, edge_factor=16, seed=33), 256),
    ("clustered", lambda: gen.block_clustered(n=2048, tiles=600, seed=34), 48),
    ("band_n40", lambda: gen.band(n=512, hb=16, p=0.7, seed=35), 40),
    ("band_n36", lambda: gen.band(n=512, hb=16, p=0.7, seed=36), 36),
]


@pytest.mark.parametrize("name,make,N", SPMM_CASES)
@pytest.mark.parametrize("bk", [1, 2, 4, 8, 16, 32])
@pytest.mark.parametrize("n_tile", [0, 16, 32])
@pytest.mark.parametrize("stage", [1, 2])
def test_spmm_dense_vs_fp64(name, make, N, bk, n_tile, stage):
    _need_gpu()
    import smat
    if stage == smat.SMAT_BSTAGE_LDS and N % 8:
        pytest.skip("LDS staging needs N % 8 == 0 (AUTO falls back to direct gathers)")
    M, K, rp, ci, v = make()
    print(f"case {name} bk={bk} N={N} n_tile={n_tile} stage={stage} M={M} nnz={ci.size}", flush=True)
    A = _bcsr(M, K, rp, ci, v, bk=bk)
    B = gen.dense_b(K, N, seed=5)
    brp, bci, bv, _ = of.csr_to_bcsr(M, K, bk, rp, ci, v)
    R = of.spmm_bcsr_f64(bk, brp, bci, bv, B)
    dB = _dev(B.view(np.int16))
    C = torch.full((M, N), float("nan"), dtype=torch.float32, device="cuda")
    plan = smat.SpmmPlan(A, N, n_tile=n_tile, b_staging=stage)
    plan(dB, C)
    torch.cuda.synchronize()
    Cn = C.cpu().numpy().astype(np.float64)
    assert np.all(np.isfinite(Cn))
    # FP32 accumulation of FP16 products: relative Frobenius error well below 1e-5
    assert _rel_fro(Cn, R) < 1e-5


@pytest.mark.parametrize("S", [2, 4, 10])
@pytest.mark.parametrize("stage", [1, 2])
def test_spmm_split_rows_deterministic(S, stage):
    """Forces block-row splitting (slices of S blocks) and checks the partial
    tile reduction, and that two runs give identical bits."""
    _need_gpu()
    import smat
    M, K, rp, ci, v = gen.rmat(scale=10, edge_factor=16, seed=41)
    N = 64
    A = _bcsr(M, K, rp, ci, v)
    B = gen.dense_b(K, N, seed=6)
    brp, bci, bv, _ = of.csr_to_bcsr(M, K, 16, rp, ci, v)
    R = of.spmm_bcsr_f64(16, brp, bci, bv, B)
    plan = smat.SpmmPlan(A, N, max_blocks_per_item=S, b_staging=stage)
    info = plan.info()
    assert info.n_split_rows > 0
    dB = _dev(B.view(np.int16))
    C1 = torch.empty((M, N), dtype=torch.float32, device="cuda")
    C2 = torch.empty((M, N), dtype=torch.float32, device="cuda")
    plan(dB, C1)
    plan(dB, C2)
    torch.cuda.synchronize()
    assert torch.equal(C1, C2)
    assert _rel_fro(C1.cpu().numpy().astype(np.float64), R) < 1e-5


def test_spmm_fp16_out_and_reference_layout():
    """FP16 output and the reference's B[n*K+k] layout, compared with the
    FP64 oracle and the FP16-accumulate emulation of the reference kernel."""
    _need_gpu()
    import smat
    M, K, rp, ci, v = gen.band(n=1024, hb=64, p=0.5, seed=51)
    N = 64
    A = _bcsr(M, K, rp, ci, v)
    B = gen.dense_b(K, N, seed=7)
    brp, bci, bv, _ = of.csr_to_bcsr(M, K, 16, rp, ci, v)
    R = of.spmm_bcsr_f64(16, brp, bci, bv, B)
    Rsim = of.spmm_refsim_fp16acc(16, brp, bci, bv, B).astype(np.float64)
    Bt = np.ascontiguousarray(B.T)  # N x K: element (k, n) at n*K + k
    plan = smat.SpmmPlan(A, N, out_dtype=smat.SMAT_F16, b_layout=smat.SMAT_B_REFERENCE)
    C = torch.empty((M, N), dtype=torch.float16, device="cuda")
    plan(_dev(Bt.view(np.int16)), C, ldb=K)
    torch.cuda.synchronize()
    Cn = C.cpu().numpy().astype(np.float64)
    assert _rel_fro(Cn, R) < 1e-3
    # parity with the reference semantics (FP16 accumulate): normwise max error
    assert np.max(np.abs(Cn - Rsim)) / np.max(np.abs(Rsim)) < 1e-2


def test_spmm_all_ones_closed_form():
    """All-ones B (the only input the reference harness ever used,
    matrix.h:45): every output column is the row sum of A."""
    _need_gpu()
    import smat
    M, K, rp, ci, v = gen.rmat(scale=9, edge_factor=8, seed=61)
    N = 32
    A = _bcsr(M, K, rp, ci, v)
    B = np.full((K, N), 0x3C00, np.uint16)
    C = torch.empty((M, N), dtype=torch.float32, device="cuda")
    smat.SpmmPlan(A, N)(_dev(B.view(np.int16)), C)
    torch.cuda.synchronize()
    rows = gen.csr_to_dense(M, K, rp, ci, v).astype(np.float64).sum(1)
    Cn = C.cpu().numpy()
    assert np.allclose(Cn, rows[:, None] * np.ones((1, N)), rtol=1e-5, atol=1e-5)


def test_invalid_arguments_raise():
    _need_gpu()
    import smat
    M, K, rp, ci, v = gen.band(n=256, hb=8, p=0.5, seed=71)
    with pytest.raises(smat.SmatError):
        _bcsr(M, K, rp, ci, v, bk=12)
    A = _bcsr(M, K, rp, ci, v)
    with pytest.raises(smat.SmatError):
        smat.SpmmPlan(A, 0)
    with pytest.raises(smat.SmatError):
        smat.SpmmPlan(A, 64, n_tile=48)


@pytest.mark.parametrize("schedule", [0, 1, 2, 3])
@pytest.mark.parametrize("which", ["rmat", "band"])
def test_schedules_match_oracle(schedule, which):
    """Block-row order, longest first and the automatic choice all give the
    FP64 oracle's product (the slicing differs, so only the FP32 summation
    order changes)."""
    _need_gpu()
    import smat
    M, K, rp, ci, v = (gen.rmat(scale=12, edge_factor=12, seed=12) if which == "rmat"
                       else gen.band(n=1024, hb=24, p=0.5, seed=3))
    N = 64
    B = gen.dense_b(K, N, seed=8)
    R = of.spmm_csr_f64(M, rp, ci, v, B)
    for bk in (2, 16):
        for split in (0, 32):
            print(f"schedule {schedule} {which} bk {bk} split {split}", flush=True)
            A = smat.Bcsr(M, K, _dev(rp), _dev(ci), _dev(v.view(np.int16)), bk=bk)
            plan = smat.SpmmPlan(A, N, schedule=schedule, max_blocks_per_item=split)
            C = torch.full((M, N), float("nan"), dtype=torch.float32, device="cuda")
            plan(_dev(B.view(np.int16)), C)
            torch.cuda.synchronize()
            Cn = C.cpu().numpy().astype(np.float64)
            assert np.all(np.isfinite(Cn))
            assert np.linalg.norm(Cn - R) / np.linalg.norm(R) < 1e-5
            plan.close()
            A.close()


@pytest.mark.parametrize("N,n_tile", [(32, 32), (64, 16), (128, 16), (256, 16), (192, 64), (256, 32)])
@pytest.mark.parametrize("kernel", ["lds", "direct", "mixed24"])
@pytest.mark.parametrize("schedule", [3])
def test_tile_affinity_covers_every_tile(N, n_tile, kernel, schedule):
    """Tile-affinity order (each XCD a subset of the N tiles) for tile counts
    1, 4, 8, 16 and 3 (falls back to the plain order), on the LDS,
    direct-B and mixed dense/2:4 kernels: the FP64 oracle's product, every C
    element written."""
    _need_gpu()
    import smat
    M, K, rp, ci, v = gen.rmat(scale=11, edge_factor=10, seed=31)
    B = gen.dense_b(K, N, seed=9)
    R = of.spmm_csr_f64(M, rp, ci, v, B)
    print(f"schedule {schedule} N {N} n_tile {n_tile} {kernel}", flush=True)
    rule = smat.SMAT_TAG_EXACT24 if kernel == "mixed24" else smat.SMAT_TAG_REFERENCE
    A = smat.Bcsr(M, K, _dev(rp), _dev(ci), _dev(v.view(np.int16)), bk=16 if kernel == "mixed24" else 2,
                  tag_rule=rule)
    v24 = i24 = None
    if kernel == "mixed24":
        v24, i24 = A.compress_24(smat.SMAT_24_SMFMAC)
    stage = smat.SMAT_BSTAGE_DIRECT if kernel == "direct" else smat.SMAT_BSTAGE_AUTO
    plan = smat.SpmmPlan(A, N, n_tile=n_tile, b_staging=stage, schedule=schedule,
                         kernel=smat.SMAT_KERNEL_MIXED24 if kernel == "mixed24" else smat.SMAT_KERNEL_DENSE)
    C = torch.full((M, N), float("nan"), dtype=torch.float32, device="cuda")
    plan(_dev(B.view(np.int16)), C, values24=v24, index24=i24)
    torch.cuda.synchronize()
    Cn = C.cpu().numpy().astype(np.float64)
    assert np.all(np.isfinite(Cn))
    assert np.linalg.norm(Cn - R) / np.linalg.norm(R) < 1e-5
    plan.close()
    A.close()


@pytest.mark.parametrize("bk", [1, 2, 4, 8, 16, 32])
@pytest.mark.parametrize("N", [128, 200, 384])
@pytest.mark.parametrize("schedule", [1, 3])
def test_n_tile_128_bit_identical_to_64(bk, N, schedule):
    """128-column tiles run the same MFMA sequence per 16-column sub-tile as
    64-column tiles (same items, same step order), so C is bit-identical;
    and both are the FP64 oracle's product.  N = 200 leaves a partial tile;
    384 gives 3 tiles (tile affinity falls back to the plain order)."""
    _need_gpu()
    import smat
    M, K, rp, ci, v = gen.rmat(scale=11, edge_factor=10, seed=41)
    B = gen.dense_b(K, N, seed=10)
    R = of.spmm_csr_f64(M, rp, ci, v, B)
    print(f"n_tile 128 bk {bk} N {N} schedule {schedule}", flush=True)
    A = _bcsr(M, K, rp, ci, v, bk=bk)
    out = []
    for nt in (64, 128):
        plan = smat.SpmmPlan(A, N, n_tile=nt, schedule=schedule, max_blocks_per_item=8 * (32 // bk))
        C = torch.full((M, N), float("nan"), dtype=torch.float32, device="cuda")
        plan(_dev(B.view(np.int16)), C)
        torch.cuda.synchronize()
        out.append(C.cpu().numpy())
        plan.close()
    assert np.array_equal(out[0], out[1])
    assert _rel_fro(out[1].astype(np.float64), R) < 1e-5
    A.close()


def test_n_tile_128_unaligned_b_and_restrictions():
    """B that cannot be staged through LDS at call time (16-byte misaligned)
    runs 64-column direct tiles with the same result; plans that cannot use
    128-column tiles at all are refused."""
    _need_gpu()
    import smat
    M, K, rp, ci, v = gen.rmat(scale=10, edge_factor=8, seed=42)
    N = 256
    B = gen.dense_b(K, N, seed=11)
    R = of.spmm_csr_f64(M, rp, ci, v, B)
    A = _bcsr(M, K, rp, ci, v, bk=4)
    plan = smat.SpmmPlan(A, N, n_tile=128)
    Bd = torch.zeros(K * N + 1, dtype=torch.int16, device="cuda")
    Bd[1:] = _dev(B.view(np.int16)).reshape(-1)
    C = torch.full((M, N), float("nan"), dtype=torch.float32, device="cuda")
    plan(Bd[1:], C, ldb=N)
    torch.cuda.synchronize()
    assert _rel_fro(C.cpu().numpy().astype(np.float64), R) < 1e-5
    plan.close()
    for kw in (dict(b_staging=smat.SMAT_BSTAGE_DIRECT), dict(in_dtype=smat.SMAT_SPLIT_F16),
               dict(b_layout=smat.SMAT_B_REFERENCE)):
        with pytest.raises(smat.SmatError):
            smat.SpmmPlan(A, N, n_tile=128, **kw)
    A.close()
    A16 = _bcsr(M, K, rp, ci, v, bk=16, tag_rule=smat.SMAT_TAG_EXACT24)
    with pytest.raises(smat.SmatError):
        smat.SpmmPlan(A16, N, n_tile=128, kernel=smat.SMAT_KERNEL_MIXED24)
    A16.close()


def test_b_over_4gib_takes_the_64bit_path():
    """The LDS kernels address B with 32-bit byte offsets; a B of 4 GiB or
    more must take the 64-bit direct-gather kernel.  Entries in B rows past
    2^23 (byte offsets past 4 GiB at N = 256) give the oracle's product."""
    _need_gpu()
    import smat
    N = 256
    K = (1 << 23) + 4096  # B = K x 256 FP16 = 4.3 GB
    M = 64
    rng = np.random.default_rng(5)
    cols_hi = np.sort(rng.choice(np.arange(K - 4096, K), size=48, replace=False))
    cols_lo = np.sort(rng.choice(np.arange(0, 4096), size=16, replace=False))
    rows = np.repeat(np.arange(M), 1)
    ci = np.concatenate([cols_lo, cols_hi]).astype(np.int32)[:M]
    rp = np.arange(M + 1, dtype=np.int32)  # one entry per row
    v = np.float16(rng.uniform(-1, 1, M)).view(np.uint16)
    print("B over 4 GiB: allocating", flush=True)
    Bd = torch.empty((K, N), dtype=torch.float16, device="cuda")
    Bd[:4096].uniform_(-1, 1)
    Bd[4096:K - 4096].zero_()
    Bd[K - 4096:].uniform_(-1, 1)
    rows_used = ci.astype(np.int64)
    Bh = Bd[torch.from_numpy(rows_used).cuda()].float().cpu().numpy().astype(np.float64)
    R = v.view(np.float16).astype(np.float64)[:, None] * Bh  # row r = a_r * B[ci[r]]
    A = _bcsr(M, K, rp, ci, v, bk=2)
    plan = smat.SpmmPlan(A, N)
    C = torch.full((M, N), float("nan"), dtype=torch.float32, device="cuda")
    plan(Bd.view(torch.int16), C)
    torch.cuda.synchronize()
    assert _rel_fro(C.cpu().numpy().astype(np.float64), R) < 1e-6
    plan.close()
    A.close()
    del Bd
    torch.cuda.empty_cache()


def test_a_over_4gib_lds_matches_direct_and_oracle():
    """The LDS kernels address A from each item's first block with 32-bit
    offsets; with more than 4 GiB of A (power-law n = 131072 at bk = 16:
    5.3 GB of blocks) they agree with the 64-bit direct kernel and the FP64
    oracle."""
    _need_gpu()
    import smat
    M, K, rp, ci, v = gen.power_law(n=131072, seed=1004)
    N = 32
    B = gen.dense_b(K, N, seed=14)
    print("A over 4 GiB: building", flush=True)
    A = _bcsr(M, K, rp, ci, v, bk=16)
    assert A.nnzb * 16 * 16 * 2 > 4 * 2**30
    Bd = _dev(B.view(np.int16))
    out = []
    for stage in (smat.SMAT_BSTAGE_LDS, smat.SMAT_BSTAGE_DIRECT):
        plan = smat.SpmmPlan(A, N, b_staging=stage)
        C = torch.full((M, N), float("nan"), dtype=torch.float32, device="cuda")
        plan(Bd, C)
        torch.cuda.synchronize()
        out.append(C.cpu().numpy())
        plan.close()
    A.close()
    assert np.allclose(out[0], out[1], rtol=1e-5, atol=1e-5)
    R = of.spmm_csr_f64(M, rp, ci, v, B)
    assert _rel_fro(out[0].astype(np.float64), R) < 1e-5


@pytest.mark.parametrize("S", [32, 64, 96, 128, 160, 0])
@pytest.mark.parametrize("N", [64, 200])
@pytest.mark.parametrize("rowmap", [False, True])
def test_bk1_pipeline_paths(S, N, rowmap):
    """16x1 blocks run their own pipeline (A through LDS, loads predicated by
    buffer resources): items of 1, 2, 3, 4 or 5 steps of 32 blocks (S blocks
    per item) take the short path, the drain alone, and the steady loop plus
    drain; S = 0 is the planner's own slicing.  N = 200 leaves a partial
    column chunk in the last 64-column tile; the row map takes the row-map
    kernel.  FP64 oracle and bit-identical repeats."""
    _need_gpu()
    import smat
    M, K, rp, ci, v = gen.rmat(scale=11, edge_factor=24, seed=43)
    A = _bcsr(M, K, rp, ci, v, bk=1)
    B = gen.dense_b(K, N, seed=9)
    brp, bci, bv, _ = of.csr_to_bcsr(M, K, 1, rp, ci, v)
    R = of.spmm_bcsr_f64(1, brp, bci, bv, B)
    plan = smat.SpmmPlan(A, N, max_blocks_per_item=S, n_tile=64)
    perm = np.random.default_rng(S + N).permutation(M).astype(np.int32) if rowmap else None
    dB = _dev(B.view(np.int16))
    C1 = torch.full((M, N), float("nan"), dtype=torch.float32, device="cuda")
    C2 = torch.full((M, N), float("nan"), dtype=torch.float32, device="cuda")
    rm = _dev(perm) if rowmap else None
    plan.execute(dB, C1, c_rowmap=rm)
    plan.execute(dB, C2, c_rowmap=rm)
    torch.cuda.synchronize()
    assert torch.equal(C1, C2)
    Cn = C1.cpu().numpy().astype(np.float64)
    if rowmap:  # row r of A lands in C row perm[r]
        Cn = Cn[perm]
    assert _rel_fro(Cn, R) < 1e-5
    plan.close()
    A.close()


@pytest.mark.parametrize("scale,N,schedule", [(17, 256, 3), (15, 64, 2)])
def test_bk1_slice_length_rule(scale, N, schedule):
    """The planner's slicing of 16x1-block rows under the longest-first
    schedules: a block row of L blocks is one work item when its ceil(L/32)
    MFMA steps fit in S, else ceil(steps/S) slices, with S = max(ceil(total
    steps x N tiles / (16 waves x CUs)) + 32, 64) steps (a fair share plus
    32, at least 64; total steps = ceil(nnzb/32)).  Scale 17 at N = 256
    takes the fair-share term, scale 15 at N = 64 the floor.  Items, split
    rows and partial-tile scratch against that count on the host (from the
    oracle's BCSR), and the scale-15 product against the FP64 oracle."""
    _need_gpu()
    import smat
    M, K, rp, ci, v = gen.rmat(scale=scale, seed=47)
    print(f"slice rule R-MAT scale {scale} N={N} schedule {schedule}", flush=True)
    brp, bci, bv, _ = of.csr_to_bcsr(M, K, 1, rp, ci, v)
    steps = -(-np.diff(brp).astype(np.int64) // 32)
    total = -(-int(brp[-1]) // 32)
    ntl = -(-N // 64)
    waves = 16 * torch.cuda.get_device_properties(0).multi_processor_count
    s = max(-(-(total * ntl) // waves) + 32, 64)
    s = -(-min(max(s * 32, 12 * 32), 65536) // 32)
    if scale == 17:
        assert s > 64  # the fair-share term decides
    k = np.where(steps <= s, 1, -(-steps // s))
    A = _bcsr(M, K, rp, ci, v, bk=1)
    plan = smat.SpmmPlan(A, N, n_tile=64, schedule=schedule)
    info = plan.info()
    assert info.n_split_rows > 0
    assert (info.n_items, info.n_split_rows) == (int(k.sum()), int((k > 1).sum()))
    assert info.scratch_bytes == int(k[k > 1].sum()) * 16 * N * 4
    if scale == 15:
        B = gen.dense_b(K, N, seed=11)
        R = of.spmm_bcsr_f64(1, brp, bci, bv, B)
        C = torch.full((M, N), float("nan"), dtype=torch.float32, device="cuda")
        plan.execute(_dev(B.view(np.int16)), C)
        torch.cuda.synchronize()
        assert _rel_fro(C.cpu().numpy().astype(np.float64), R) < 1e-5
    plan.close()
    A.close()
