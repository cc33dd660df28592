"""Merged tiles on the GPU against the oracle (or_bcsr_to_tiles,
or_spmm_tiles_f64; SparseMatrix::bcsrBlocking, matrix.h:794-1137): the tile
build bit-exact, the tile-grid maps, the 2:4 compression of info-1
sub-blocks, the tile SpMM (dense MFMA and mixed MFMA / SMFMAC, FP16 and BF16,
FP16 and FP32 output) against the FP64 product, and the reference drop-ins
mmaOBTKernel_tiled, preprocessing_mmaOBTSKernel_tiled_large and
mmaOBTSKernel_tiled_large on the reference's B layout, through the C ABI.

Tolerances: FP32 accumulation of FP16 / BF16 products is compared with the
FP64 product of the same operands at 1e-6 relative Frobenius error; FP16
output adds one rounding (2^-11 relative) per element, checked at 1e-3."""
import ctypes

import numpy as np
import pytest

import oracle_ffi as of
from smat import gen
from test_gpu_ref_api import _decompress_smfmac

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def tiled_matrix(M, K, bk, ntiles, seed, kinds=(0, 1, 2)):
    """ntiles random 64 x 4bk tiles; each 16 x bk sub-block empty (0), pruned
    to 2:4 (1) or dense (2), drawn from `kinds`.  Rows past M are dropped."""
    rng = np.random.default_rng(seed)
    tw = 4 * bk
    ntr, ntc = (M + 63) // 64, K // tw
    cells = rng.choice(ntr * ntc, size=min(ntiles, ntr * ntc), replace=False)
    d = np.zeros((ntr * 64, K), np.float16)
    for cell in cells:
        tr, tc = divmod(int(cell), ntc)
        for sr in range(4):
            for sc in range(4):
                kind = kinds[rng.integers(len(kinds))]
                if kind == 0:
                    continue
                blk = rng.uniform(-1, 1, (16, bk)).astype(np.float16)
                blk[blk == 0] = 0.5
                if kind == 1:
                    blk = gen.prune_2of4(blk)
                d[64 * tr + 16 * sr:64 * tr + 16 * sr + 16, tc * tw + bk * sc:tc * tw + bk * (sc + 1)] = blk
    d = d[:M]
    rows, cols = np.nonzero(d)
    rp = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=M))]).astype(np.int32)
    return M, K, rp, cols.astype(np.int32), d[rows, cols].view(np.uint16)


def _tiles(M, K, rp, ci, v, bk):
    import smat
    A = smat.Bcsr(M, K, _dev(rp), _dev(ci), _dev(v.view(np.int16)), bk=bk)
    return A, smat.Tiles(A)


def _oracle_tiles(M, K, bk, rp, ci, v):
    brp, bci, bv, _ = of.csr_to_bcsr(M, K, bk, rp, ci, v)
    return of.bcsr_to_tiles(M, K, bk, brp, bci, bv)


def _rel(C, R):
    return np.linalg.norm(C - R) / max(np.linalg.norm(R), 1e-30)


BUILD_CASES = [
    ("band", lambda bk: gen.band(n=1000, hb=40, p=0.4, seed=21)),
    ("rmat", lambda bk: gen.rmat(scale=11, edge_factor=8, seed=22)),
    ("clustered", lambda bk: gen.block_clustered(n=2048, tiles=300, seed=23)),
    ("tiled_partial", lambda bk: tiled_matrix(16 * 37, 4 * bk * 9, bk, 60, seed=24)),
]


@pytest.mark.parametrize("bk", [16, 32])
@pytest.mark.parametrize("name,make", BUILD_CASES, ids=[c[0] for c in BUILD_CASES])
def test_tiles_build_bit_exact(name, make, bk):
    _need_gpu()
    M, K, rp, ci, v = make(bk)
    K = gen.pad_to(K, bk)
    print(f"tiles build {name} bk={bk} M={M} K={K} nnz={ci.size}", flush=True)
    A, T = _tiles(M, K, rp, ci, v, bk)
    trp, tci, tv, tinfo = _oracle_tiles(M, K, bk, rp, ci, v)
    g_rp, g_ci, g_v, g_info = (x.numpy() for x in T.arrays())
    assert T.ntiles == tci.size and T.desc.ntrows == (M + 63) // 64 and T.desc.ntcols == -(-K // (4 * bk))
    assert np.array_equal(g_rp, trp)
    assert np.array_equal(g_ci, tci)
    assert np.array_equal(g_v.view(np.uint16), tv)
    assert np.array_equal(g_info, tinfo)
    # grid maps: mergedBlockInfo 1 / relative index at every tile's cell
    info, rel = T.grid_maps()
    torch.cuda.synchronize()
    ntc = int(T.desc.ntcols)
    ng = int(T.desc.ntrows) * ntc
    exp_info = np.zeros(ng, np.int32)
    exp_rel = np.full(ng, -1, np.int32)
    tr_of = np.repeat(np.arange(trp.size - 1), np.diff(trp))
    cells = tr_of * ntc + tci // (4 * bk)
    exp_info[cells] = 1
    exp_rel[cells] = np.arange(tci.size)
    assert np.array_equal(info.cpu().numpy()[:ng], exp_info) and np.array_equal(rel.cpu().numpy()[:ng], exp_rel)
    T.close()
    A.close()


@pytest.mark.parametrize("bk", [16, 32])
def test_tiles_compress_24_round_trip(bk):
    _need_gpu()
    M, K, rp, ci, v = tiled_matrix(512, 4 * bk * 8, bk, 40, seed=31)
    A, T = _tiles(M, K, rp, ci, v, bk)
    vals, idx = T.compress_24()
    torch.cuda.synchronize()
    _, _, tv, tinfo = _oracle_tiles(M, K, bk, rp, ci, v)
    n = tinfo.size
    assert set(np.unique(tinfo)) == {0, 1, 2}
    got = _decompress_smfmac(vals.cpu().numpy().view(np.uint16)[: n * 8 * bk], idx.cpu().numpy()[: n * bk], n,
                             bk).reshape(n, -1)
    exp = tv.reshape(n, -1)
    sel = tinfo == 1
    assert np.array_equal(got[sel], exp[sel])
    assert not got[~sel].any()


SPMM_CASES = [
    # (M, K, ntiles, N)
    (1024, 2048, 90, 64),
    (16 * 37, 2048, 70, 136),  # partial last tile row, N not a multiple of 64
    (256, 1024, 30, 8),
    (2048, 4096, 400, 256),
]


@pytest.mark.parametrize("bk", [16, 32])
@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("dtype", ["f16", "bf16"])
@pytest.mark.parametrize("case", SPMM_CASES, ids=[f"M{c[0]}_K{c[1]}_N{c[3]}" for c in SPMM_CASES])
def test_tiles_spmm_vs_fp64(case, dtype, mixed, bk):
    _need_gpu()
    import smat
    M0, K, nt, N = case
    M, K, rp, ci, v = tiled_matrix(M0, K, bk, nt, seed=M0 + N + bk)
    print(f"tiles spmm M={M} K={K} N={N} bk={bk} mixed={mixed} {dtype}", flush=True)
    A, T = _tiles(M, K, rp, ci, v, bk)
    trp, tci, tv, tinfo = _oracle_tiles(M, K, bk, rp, ci, v)
    B = gen.dense_b(K, N, seed=N)
    if dtype == "f16":
        vals_dev, B_dev, dt = None, _dev(B.view(np.int16)), smat.SMAT_F16
        R = of.spmm_tiles_f64(M, K, bk, trp, tci, tv, B)
    else:
        tvb = gen.f32_to_bf16_bits(tv.view(np.float16).astype(np.float32))
        Bb = gen.f32_to_bf16_bits(B.view(np.float16).astype(np.float32))
        vals_dev, B_dev, dt = _dev(tvb.view(np.int16)), _dev(Bb.view(np.int16)), smat.SMAT_BF16
        Bd = gen.bf16_bits_to_f32(Bb).astype(np.float64).reshape(K, N)
        # dense A from the bf16 tile values (layout: slot 4*sr+sc, 16 x bk row-major)
        vv = gen.bf16_bits_to_f32(tvb).astype(np.float64).reshape(-1, 4, 4, 16, bk)
        big = np.zeros((((M + 63) // 64) * 64, K + 4 * bk))
        for tr in range(trp.size - 1):
            for t in range(trp[tr], trp[tr + 1]):
                big[64 * tr:64 * tr + 64, tci[t]:tci[t] + 4 * bk] = vv[t].transpose(0, 2, 1, 3).reshape(64, 4 * bk)
        Ad = big[:M, :K]
        R = Ad @ Bd
    v24 = i24 = None
    if mixed:
        if dtype == "f16":
            v24, i24 = T.compress_24()
        else:
            d2 = smat.TilesDesc.from_buffer_copy(T.desc)
            d2.values = vals_dev.data_ptr()
            v24 = torch.empty(max(T.ntiles * 16, 1) * 8 * bk, dtype=torch.int16, device="cuda")
            i24 = torch.empty(max(T.ntiles * 16, 1) * bk, dtype=torch.int16, device="cuda")
            smat.check(smat.lib().smat_tiles_compress_24(ctypes.byref(d2), v24.data_ptr(), i24.data_ptr(), None))
    for out in (torch.float32, torch.float16):
        C = torch.full((M, N), float("nan"), dtype=out, device="cuda")
        T.spmm(B_dev, C, dtype=dt, values=vals_dev, values24=v24, index24=i24)
        torch.cuda.synchronize()
        Cn = C.float().cpu().numpy().astype(np.float64)
        assert np.all(np.isfinite(Cn)), "unwritten C elements"
        tol = 1e-6 if out == torch.float32 else 1e-3
        assert _rel(Cn, R) <= tol, (out, _rel(Cn, R))
    T.close()
    A.close()


def test_tiles_spmm_empty_matrix_writes_zeros():
    _need_gpu()
    M, K, N = 128, 256, 64
    rp = np.zeros(M + 1, np.int32)
    A, T = _tiles(M, K, rp, np.zeros(0, np.int32), np.zeros(0, np.uint16), 16)
    assert T.ntiles == 0
    C = torch.full((M, N), float("nan"), dtype=torch.float32, device="cuda")
    T.spmm(_dev(gen.dense_b(K, N).view(np.int16)), C)
    torch.cuda.synchronize()
    assert torch.count_nonzero(C).item() == 0


def test_tiles_spmm_rejects_bad_ldb():
    _need_gpu()
    import smat
    M, K, rp, ci, v = tiled_matrix(128, 256, 16, 4, seed=5)
    A, T = _tiles(M, K, rp, ci, v, 16)
    B = _dev(gen.dense_b(K, 24).view(np.int16))
    C = torch.zeros((M, 20), dtype=torch.float32, device="cuda")
    with pytest.raises(smat.SmatError) as e:
        T.spmm(B, C, N=20, ldb=20)  # ldb not a multiple of 8
    assert e.value.status == smat.SMAT_ERR_INVALID


def _ref_layout_b(B):
    """Row-major K x N -> the reference's B[n*K + k]."""
    return np.ascontiguousarray(B.T)


@pytest.mark.parametrize("N", [32, 100, 200])
def test_mmaOBTKernel_tiled_dropin(N):
    _need_gpu()
    import smat
    bk = 16
    M, K, rp, ci, v = tiled_matrix(1024, 2048, bk, 150, seed=41 + N)
    A, T = _tiles(M, K, rp, ci, v, bk)
    info, rel = T.grid_maps()
    B = gen.dense_b(K, N, seed=3)
    Bt = _dev(_ref_layout_b(B).view(np.int16))
    C = torch.full((M, N), float("nan"), dtype=torch.float16, device="cuda")
    d = T.desc
    smat.lib().smat_clear_error()
    smat.lib().mmaOBTKernel_tiled(d.values, d.rowptr, d.colidx, Bt.data_ptr(), C.data_ptr(), M, N, K, A.nnzb,
                                  info.data_ptr(), rel.data_ptr())
    torch.cuda.synchronize()
    assert smat.lib().smat_last_status() == 0, smat.lib().smat_last_error()
    trp, tci, tv, _ = _oracle_tiles(M, K, bk, rp, ci, v)
    R = of.spmm_tiles_f64(M, K, bk, trp, tci, tv, B)
    Cn = C.float().cpu().numpy().astype(np.float64)
    assert np.all(np.isfinite(Cn)) and _rel(Cn, R) <= 1e-3


@pytest.mark.parametrize("N", [64, 256])
def test_mmaOBTSKernel_tiled_large_dropin(N):
    _need_gpu()
    import smat
    bk = 32
    M, K, rp, ci, v = tiled_matrix(2048, 4096, bk, 300, seed=51 + N)
    A, T = _tiles(M, K, rp, ci, v, bk)
    info, rel = T.grid_maps()
    d = T.desc
    nsub = T.ntiles * 16
    # the reference's buffer sizes (tester.h:385-394), over the unmerged block count
    meta = torch.zeros(max(A.nnzb, 1) * 16 * 16 * 4, dtype=torch.uint8, device="cuda")
    sp = torch.zeros(max(A.nnzb, 1) * 16 * 4 * 16 * 4, dtype=torch.int16, device="cuda")
    smat.lib().smat_clear_error()
    smat.lib().preprocessing_mmaOBTSKernel_tiled_large(d.values, meta.data_ptr(), sp.data_ptr(), M, N, K, A.nnzb,
                                                       info.data_ptr(), rel.data_ptr(), d.info)
    torch.cuda.synchronize()
    assert smat.lib().smat_last_status() == 0, smat.lib().smat_last_error()
    # the pair is the SMFMAC layout of the info-1 sub-blocks
    v24, i24 = T.compress_24()
    torch.cuda.synchronize()
    assert torch.equal(sp[: nsub * 256], v24[: nsub * 256])
    assert torch.equal(meta[: nsub * 64].view(torch.int16), i24[: nsub * 32])
    B = gen.dense_b(K, N, seed=4)
    Bt = _dev(_ref_layout_b(B).view(np.int16))
    C = torch.full((M, N), float("nan"), dtype=torch.float16, device="cuda")
    smat.lib().mmaOBTSKernel_tiled_large(d.values, d.rowptr, d.colidx, meta.data_ptr(), sp.data_ptr(), Bt.data_ptr(),
                                         C.data_ptr(), M, N, K, A.nnzb, info.data_ptr(), rel.data_ptr(), d.info)
    torch.cuda.synchronize()
    assert smat.lib().smat_last_status() == 0, smat.lib().smat_last_error()
    trp, tci, tv, _ = _oracle_tiles(M, K, bk, rp, ci, v)
    R = of.spmm_tiles_f64(M, K, bk, trp, tci, tv, B)
    Cn = C.float().cpu().numpy().astype(np.float64)
    assert np.all(np.isfinite(Cn)) and _rel(Cn, R) <= 1e-3


def test_tiled_dropin_rejects_inconsistent_arrays():
    _need_gpu()
    import smat
    bk = 16
    M, K, rp, ci, v = tiled_matrix(256, 512, bk, 10, seed=61)
    A, T = _tiles(M, K, rp, ci, v, bk)
    info, rel = T.grid_maps()
    rel_bad = rel.clone()
    rel_bad[rel_bad >= 0] += T.ntiles  # every relative index past the last tile
    N = 64
    Bt = _dev(_ref_layout_b(gen.dense_b(K, N)).view(np.int16))
    C = torch.full((M, N), 7.0, dtype=torch.float16, device="cuda")
    d = T.desc
    smat.lib().smat_clear_error()
    smat.lib().mmaOBTKernel_tiled(d.values, d.rowptr, d.colidx, Bt.data_ptr(), C.data_ptr(), M, N, K, A.nnzb,
                                  info.data_ptr(), rel_bad.data_ptr())
    torch.cuda.synchronize()
    assert smat.lib().smat_last_status() == smat.SMAT_ERR_INVALID
    assert torch.all(C == 7.0)
