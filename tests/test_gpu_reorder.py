"""GPU Keeper reordering vs the compiled reference's golden vectors and the
oracle, bit-exact; GPU row permutation vs numpy."""
import os

import numpy as np
import pytest

import oracle_ffi as of
from smat import gen

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def test_keeper_golden_cases(golden_dir):
    _need_gpu()
    import smat
    d = np.load(os.path.join(golden_dir, "keeper_golden.npz"))
    cases = [str(c) for c in d["cases"]]
    for key in cases:
        name, params = key.split("__")
        t, b, B = params.split("_")
        tau, cb, mr = float(t[1:]), int(b[1:]), int(B[1:])
        rp, ci = d[f"{name}__rowptr"].astype(np.int32), d[f"{name}__colidx"].astype(np.int32)
        print(f"case {key}", flush=True)
        g, cnt = smat.reorder_grouping(rp, ci, np.float32(tau), cb, mr)
        assert np.array_equal(g, d[key + "__grouping"]), key
        assert np.array_equal(cnt, d[key + "__counters"]), key
        fsg = smat.fixed_size_grouping(g, mr)
        assert np.array_equal(fsg, d[key + "__fixed"]), key
        assert np.array_equal(smat.grouping_to_permutation(fsg), d[key + "__perm"]), key


@pytest.mark.parametrize("scale,tau,cb", [(12, 0.5, 16), (12, 0.1, 8), (11, 0.7, 16)])
def test_keeper_vs_oracle_rmat(scale, tau, cb):
    _need_gpu()
    import smat
    M, K, rp, ci, _ = gen.rmat(scale=scale, edge_factor=16, seed=200 + scale)
    rows = 1 << scale
    rp = rp[: rows + 1]
    print(f"rmat scale {scale} tau {tau} cb {cb}", flush=True)
    g, cnt = smat.reorder_grouping(rp, ci, np.float32(tau), cb, 16)
    eg, ecnt = of.keeper(rp.astype(np.int64), ci.astype(np.int64), np.float32(tau), cb, 16)
    assert np.array_equal(cnt, ecnt)
    assert np.array_equal(g, eg)


def test_keeper_band():
    _need_gpu()
    import smat
    M, K, rp, ci, _ = gen.band(n=2048, hb=40, p=0.4, seed=77)
    g, cnt = smat.reorder_grouping(rp, ci, np.float32(0.3), 16, 16)
    eg, ecnt = of.keeper(rp.astype(np.int64), ci.astype(np.int64), np.float32(0.3), 16, 16)
    assert np.array_equal(cnt, ecnt) and np.array_equal(g, eg)


def test_permute_rows_on_gpu():
    _need_gpu()
    import smat
    M, K, rp, ci, v = gen.rmat(scale=10, edge_factor=8, seed=5)
    rng = np.random.default_rng(1)
    perm = rng.permutation(M).astype(np.int64)
    d_rp, d_ci = torch.from_numpy(rp).cuda(), torch.from_numpy(ci).cuda()
    d_v = torch.from_numpy(v.view(np.int16)).cuda()
    d_perm = torch.from_numpy(perm).cuda()
    o_rp = torch.empty(M + 1, dtype=torch.int32, device="cuda")
    o_ci = torch.empty_like(d_ci)
    o_v = torch.empty_like(d_v)
    smat.check(smat.lib().smat_permute_csr_rows(M, d_rp.data_ptr(), d_ci.data_ptr(), d_v.data_ptr(),
                                                d_perm.data_ptr(), o_rp.data_ptr(), o_ci.data_ptr(),
                                                o_v.data_ptr(), smat.stream_ptr()))
    torch.cuda.synchronize()
    lens = np.diff(rp)[perm]
    erp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    eci = np.concatenate([ci[rp[r]:rp[r + 1]] for r in perm])
    ev = np.concatenate([v[rp[r]:rp[r + 1]] for r in perm])
    assert np.array_equal(o_rp.cpu().numpy(), erp)
    assert np.array_equal(o_ci.cpu().numpy(), eci)
    assert np.array_equal(o_v.cpu().numpy().view(np.uint16), ev)


def _transpose_sorted(M, K, rp, ci):
    rows = np.repeat(np.arange(M), np.diff(rp))
    order = np.lexsort((rows, ci))  # by column, then row: each transposed row ascending
    trp = np.zeros(K + 1, np.int64)
    np.add.at(trp, ci + 1, 1)
    return np.cumsum(trp), rows[order].astype(np.int64)


@pytest.mark.parametrize("scale,tau", [(10, 0.5), (11, 0.3)])
def test_column_pass_matches_oracle(scale, tau):
    """The column order equals the Keeper (oracle, itself pinned to the
    compiled reference) run on the sorted transpose (src/reorder.py:143-183)."""
    _need_gpu()
    import smat
    M, K, rp, ci, _ = gen.rmat(scale=scale, edge_factor=8, seed=300 + scale)
    print(f"column pass scale {scale} tau {tau}", flush=True)
    perm, cnt = smat.reorder_columns(rp, ci, K, np.float32(tau), 16, 16)
    trp, tci = _transpose_sorted(M, K, rp, ci)
    g, ecnt = of.keeper(trp, tci, np.float32(tau), 16, 16)
    eperm = of.get_permutation(of.fixed_size_grouping(g, 16))
    assert np.array_equal(cnt, ecnt)
    assert np.array_equal(perm, eperm)


def test_column_reordered_spmm_equals_original():
    _need_gpu()
    import smat
    M, K, rp, ci, v = gen.rmat(scale=11, edge_factor=8, seed=9)
    N = 64
    perm, _ = smat.reorder_columns(rp, ci, K, np.float32(0.5), 16, 16)
    inv = np.argsort(perm).astype(np.int64)
    d_ci = torch.from_numpy(ci).cuda()
    d_ci2 = torch.empty_like(d_ci)
    smat.check(smat.lib().smat_permute_csr_cols(ci.size, d_ci.data_ptr(), torch.from_numpy(inv).cuda().data_ptr(),
                                                d_ci2.data_ptr(), smat.stream_ptr()))
    B = gen.dense_b(K, N, seed=4)
    d_B = torch.from_numpy(B.view(np.int16)).cuda()
    d_B2 = torch.empty_like(d_B)
    smat.check(smat.lib().smat_permute_dense_rows(K, N, 2, d_B.data_ptr(), N, torch.from_numpy(perm).cuda().data_ptr(),
                                                  d_B2.data_ptr(), N, smat.stream_ptr()))
    torch.cuda.synchronize()
    assert np.array_equal(d_ci2.cpu().numpy(), inv[ci])
    assert np.array_equal(d_B2.cpu().numpy().view(np.uint16), B[perm])
    for bk in (2, 16):
        A = smat.Bcsr(M, K, torch.from_numpy(rp).cuda(), d_ci2, torch.from_numpy(v.view(np.int16)).cuda(), bk=bk)
        C = torch.empty((M, N), dtype=torch.float32, device="cuda")
        smat.SpmmPlan(A, N)(d_B2, C)
        torch.cuda.synchronize()
        R = of.spmm_csr_f64(M, rp, ci, v, B)
        err = np.linalg.norm(C.cpu().numpy() - R) / np.linalg.norm(R)
        assert err < 1e-5, (bk, err)


def test_threshold_search_matches_oracle_blocking_info():
    """Per-tau block counts equal the oracle's CollectBlockingInfo (pinned to
    the reference KAT) on the oracle's fixed-size Keeper grouping, and the
    winner follows reorder_matrix.cpp's `<=` rule."""
    _need_gpu()
    import smat
    M, K, rp, ci, _ = gen.rmat(scale=10, edge_factor=8, seed=41)
    taus = (0.9, 0.5, 0.3, 0.1)
    best_tau, best, counts = smat.threshold_search(rp, ci, K, 16, 16, taus)
    want_best_tau, want_best = -1.0, ci.size
    for tau in taus:
        g, _ = of.keeper(rp.astype(np.int64), ci.astype(np.int64), np.float32(tau), 16, 16)
        fixed = of.fixed_size_grouping(g, 16)
        info, _ = of.collect_blocking_info(rp.astype(np.int64), ci.astype(np.int64), fixed, K, 16)
        assert counts[tau] == int(info[1]), tau  # VBR_nzblocks_count
        if info[1] <= want_best:
            want_best_tau, want_best = tau, int(info[1])
    assert (best_tau, best) == (want_best_tau, want_best)


@pytest.mark.parametrize("tau,cb", [(0.5, 16), (0.3, 4), (0.9, 32)])
def test_keeper_long_rows_vs_oracle(tau, cb):
    """Power-law rows (up to 600 entries): distances of rows over 64 entries
    come from the wave-per-row kernel and must equal the per-thread ones."""
    _need_gpu()
    import smat
    M, K, rp, ci, _ = gen.power_law(n=3000, max_len=600, seed=17)
    print(f"power-law keeper tau {tau} cb {cb}, longest row {int(np.diff(rp).max())}", flush=True)
    g, cnt = smat.reorder_grouping(rp, ci, np.float32(tau), cb, 16)
    eg, ecnt = of.keeper(rp.astype(np.int64), ci.astype(np.int64), np.float32(tau), cb, 16)
    assert np.array_equal(cnt, ecnt)
    assert np.array_equal(g, eg)


def test_blocking_info_golden_and_kat(golden_dir):
    """GPU CollectBlockingInfo against the compiled reference's figures: the
    60 golden Keeper cases (fixed-size groupings) and the test-matrix KAT
    (raw algorithm-0 grouping, 3-column blocks over 10 columns)."""
    _need_gpu()
    import json
    import smat
    d = np.load(os.path.join(golden_dir, "keeper_golden.npz"))
    for key in [str(c) for c in d["cases"]]:
        name, params = key.split("__")
        cb = int(params.split("_")[1][1:])
        rp, ci = d[f"{name}__rowptr"].astype(np.int32), d[f"{name}__colidx"].astype(np.int32)
        cols = int(d[f"{name}__cols"][0])
        fixed = torch.from_numpy(d[key + "__fixed"].astype(np.int64)).cuda()
        info, avg = smat.blocking_info(rp.size - 1, cols, torch.from_numpy(rp).cuda(), torch.from_numpy(ci).cuda(),
                                       fixed, cb)
        exp = d[key + "__info"]
        assert info.tolist() == exp[:3].tolist(), key
        assert round(avg * 1e6) == exp[3], key
    kat = json.load(open(os.path.join(golden_dir, "kat_test_matrix.json")))
    e = np.array(kat["edges"])
    rows = int(e[:, 0].max()) + 1
    rp = np.zeros(rows + 1, np.int64)
    np.add.at(rp, e[:, 0].astype(np.int64) + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    ci = e[:, 1].astype(np.int32)
    g = torch.tensor(kat["compiled_reference"]["grouping"], dtype=torch.int64).cuda()
    info, avg = smat.blocking_info(rows, kat["cols"], torch.from_numpy(rp).cuda(), torch.from_numpy(ci).cuda(), g, 3)
    assert info.tolist() == [kat["VBR_nzcount"], kat["VBR_nzblocks_count"], kat["VBR_longest_row"]]
    assert avg == pytest.approx(kat["VBR_average_height"], abs=1e-6)


@pytest.mark.parametrize("which", ["rmat", "powerlaw", "band_empty_rows"])
@pytest.mark.parametrize("cb", [5, 8, 16])
def test_blocking_info_vs_oracle(which, cb):
    """Random, Keeper (values seed / seed + rows) and fixed-size groupings,
    column blocks that leave a short last block (cb = 5), empty rows: the
    GPU figures equal the oracle's exactly, the average height bit for bit."""
    _need_gpu()
    import smat
    rng = np.random.default_rng(7)
    if which == "rmat":
        M, K, rp, ci, _ = gen.rmat(scale=11, edge_factor=8, seed=51)
    elif which == "powerlaw":
        M, K, rp, ci, _ = gen.power_law(n=3000, max_len=600, seed=52)
    else:
        M, K, rp, ci, _ = gen.band(n=1000, hb=6, p=0.5, seed=53)
        keep = rng.random(M) < 0.7  # drop whole rows
        lens = np.diff(rp) * keep
        ci = np.concatenate([ci[rp[r]:rp[r + 1]] for r in range(M) if keep[r]] + [np.zeros(0, ci.dtype)])
        rp = np.concatenate([[0], np.cumsum(lens)]).astype(rp.dtype)
    K = int(K) + 3  # a few empty trailing columns
    g_keeper, _ = smat.reorder_grouping(rp.astype(np.int32), ci.astype(np.int32), np.float32(0.5), cb, 16)
    groupings = {"random": rng.integers(0, 2 * M, size=M, dtype=np.int64), "keeper": g_keeper,
                 "fixed": smat.fixed_size_grouping(g_keeper, 16)}
    d_rp = torch.from_numpy(rp.astype(np.int32)).cuda()
    d_ci = torch.from_numpy(ci.astype(np.int32)).cuda()
    for gname, g in groupings.items():
        print(f"blocking_info {which} cb {cb} {gname}", flush=True)
        info, avg = smat.blocking_info(M, K, d_rp, d_ci, torch.from_numpy(g).cuda(), cb)
        ref, ref_avg = of.collect_blocking_info(rp.astype(np.int64), ci.astype(np.int64), g, K, cb)
        assert info.tolist() == ref.tolist(), gname
        assert np.float32(avg).tobytes() == np.float32(ref_avg).tobytes(), gname


def test_blocking_info_errors_and_empty():
    _need_gpu()
    import smat
    rp = torch.tensor([0, 2, 3], dtype=torch.int32).cuda()
    ci = torch.tensor([0, 4, 1], dtype=torch.int32).cuda()
    with pytest.raises(smat.SmatError):
        smat.blocking_info(2, 5, rp, ci, torch.tensor([0, 4], dtype=torch.int64).cuda(), 2)   # 4 >= 2 * rows
    with pytest.raises(smat.SmatError):
        smat.blocking_info(2, 4, rp, ci, torch.tensor([0, 1], dtype=torch.int64).cuda(), 2)   # column 4 >= cols
    info, avg = smat.blocking_info(2, 5, rp, ci, torch.tensor([1, 1], dtype=torch.int64).cuda(), 2)
    # one block row of height 2, blocks {0, 2} (last, 1 wide) -> 2*2 + 2*2 - 2*1 = 6 entries, 2 blocks
    assert info.tolist() == [6, 2, 2] and avg == 2.0
    z = torch.zeros(3, dtype=torch.int32).cuda()
    info, avg = smat.blocking_info(2, 5, z, None, torch.tensor([0, 1], dtype=torch.int64).cuda(), 2)
    assert info.tolist() == [0, 0, 0] and np.isnan(avg)


def _rows_csr(rows, cols, seed, max_len=40, unsorted=False):
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, max_len, size=rows)
    lens[rng.random(rows) < 0.1] = 0
    parts = []
    for n in lens:
        c = rng.choice(cols, size=int(n), replace=False) if cols < 1 << 20 else \
            np.unique(rng.integers(0, cols, size=int(n)))
        c = np.sort(c)
        if unsorted:
            c = rng.permutation(c)
        parts.append(c)
    rp = np.concatenate([[0], np.cumsum([p.size for p in parts])]).astype(np.int32)
    ci = np.concatenate(parts).astype(np.int32) if rp[-1] else np.zeros(0, np.int32)
    return rp, ci


@pytest.mark.parametrize("case", ["rmat12", "powerlaw", "random", "unsorted", "wide_cols"])
def test_locality_order_matches_oracle(case):
    """GPU locality order (stable radix sort of the row keys) equals the
    oracle's lexsort, bit for bit, including empty rows, rows in file order
    and more than 2^29 columns (second key dropped)."""
    _need_gpu()
    import smat
    if case == "rmat12":
        M, K, rp, ci, _ = gen.rmat(scale=12, seed=5)
    elif case == "powerlaw":
        M, K, rp, ci, _ = gen.power_law(n=8192, seed=6)[:5]
    else:
        M = 3000
        K = (1 << 30) + 5 if case == "wide_cols" else 1000
        rp, ci = _rows_csr(M, K, seed=len(case), unsorted=case == "unsorted")
    print(f"locality order {case}: rows {M} cols {K} nnz {ci.size}", flush=True)
    d_rp = torch.from_numpy(rp).cuda()
    d_ci = torch.from_numpy(ci if ci.size else np.zeros(1, np.int32)).cuda()
    perm = smat.locality_order(M, K, d_rp, d_ci).cpu().numpy()
    assert np.array_equal(perm, of.locality_order(M, K, rp, ci))


@pytest.mark.parametrize("variant", ["f16", "split_rows", "f16_out", "bf16"])
def test_locality_order_spmm_writes_original_rows(variant):
    """Reordered SpMM end to end: locality order, GPU row permutation, BCSR,
    plan, and C written back through c_rowmap (the row-map LDS kernel for
    64-column 16-bit tiles) equals the FP64 product of the original matrix;
    also with every block row cut into 2-step slices (partial tiles reduced
    through the map), FP16 C, and BF16 inputs."""
    _need_gpu()
    import smat
    M, K, rp, ci, v = gen.rmat(scale=13, seed=9)
    N = 64
    vals = v
    if variant == "bf16":
        vals = gen.f32_to_bf16_bits(v.view(np.float16).astype(np.float32))
    d = [torch.from_numpy(x).cuda() for x in (rp, ci, vals.view(np.int16))]
    perm = smat.locality_order(M, K, d[0], d[1])
    prp, pci, pv = smat.permute_rows(d[0], d[1], d[2], perm)
    rowmap = perm.to(torch.int32)
    B = gen.dense_b(K, N, seed=3)
    if variant == "bf16":
        B = gen.f32_to_bf16_bits(B.view(np.float16).astype(np.float32)).reshape(K, N)
        A64 = gen.bf16_bits_to_f32(vals).astype(np.float64)
        B64 = gen.bf16_bits_to_f32(B).reshape(K, N).astype(np.float64)
        import scipy.sparse as sp
        R = sp.csr_matrix((A64, ci, rp), shape=(M, K)) @ B64
    else:
        R = of.spmm_csr_f64(M, rp, ci, v, B)
    Bd = torch.from_numpy(B.view(np.int16)).cuda()
    print(f"locality SpMM {variant}", flush=True)
    for bk in (1, 2, 16):
        A = smat.Bcsr(M, K, prp, pci, pv, bk=bk)
        kw = {}
        if variant == "split_rows":
            kw["max_blocks_per_item"] = 2 * (32 // bk)
        if variant == "f16_out":
            kw["out_dtype"] = smat.SMAT_F16
        if variant == "bf16":
            kw["in_dtype"] = smat.SMAT_BF16
        plan = smat.SpmmPlan(A, N, **kw)
        if variant == "split_rows":
            assert plan.info().n_split_rows > 0
        dt = torch.float16 if variant == "f16_out" else torch.float32
        C = torch.full((M, N), float("nan"), dtype=dt, device="cuda")
        plan.execute(Bd, C, c_rowmap=rowmap)
        torch.cuda.synchronize()
        Cn = C.float().cpu().numpy().astype(np.float64)
        tol = 1e-3 if variant == "f16_out" else 1e-5
        assert np.all(np.isfinite(Cn)), (variant, bk)
        assert np.linalg.norm(Cn - R) / np.linalg.norm(R) < tol, (variant, bk)
        plan.close()
        A.close()


def test_locality_order_mixed24_with_split_rows():
    """The mixed dense / 2:4 kernel on a locality-ordered matrix (as the
    locality probe and the batched workload may pick it): split rows, C
    written back through the row map, against the FP64 product of the
    original matrix."""
    _need_gpu()
    import smat
    M, K, rp, ci, v = gen.block_clustered(n=4096, tiles=900, seed=17)
    N = 64
    d = [torch.from_numpy(x).cuda() for x in (rp, ci, v.view(np.int16))]
    perm = smat.locality_order(M, K, d[0], d[1])
    prp, pci, pv = smat.permute_rows(d[0], d[1], d[2], perm)
    B = gen.dense_b(K, N, seed=4)
    R = of.spmm_csr_f64(M, rp, ci, v, B)
    A = smat.Bcsr(M, K, prp, pci, pv, bk=16, tag_rule=smat.SMAT_TAG_EXACT24)
    v24, i24 = A.compress_24(smat.SMAT_24_SMFMAC)
    plan = smat.SpmmPlan(A, N, kernel=smat.SMAT_KERNEL_MIXED24, max_blocks_per_item=2)
    assert plan.info().n_split_rows > 0
    C = torch.full((M, N), float("nan"), dtype=torch.float32, device="cuda")
    plan.execute(torch.from_numpy(B.view(np.int16)).cuda(), C, values24=v24, index24=i24,
                 c_rowmap=perm.to(torch.int32))
    torch.cuda.synchronize()
    Cn = C.cpu().numpy().astype(np.float64)
    assert np.all(np.isfinite(Cn))
    assert np.linalg.norm(Cn - R) / np.linalg.norm(R) < 1e-5
    plan.close()
    A.close()


def test_locality_order_rejects_bad_arguments():
    _need_gpu()
    import smat
    d = torch.zeros(4, dtype=torch.int32, device="cuda")
    out = torch.empty(4, dtype=torch.int64, device="cuda")
    assert smat.lib().smat_locality_order(0, 4, d.data_ptr(), d.data_ptr(), out.data_ptr(), None) == \
        smat.SMAT_ERR_INVALID
    assert smat.lib().smat_locality_order(3, 4, d.data_ptr(), None, out.data_ptr(), None) == smat.SMAT_ERR_INVALID
