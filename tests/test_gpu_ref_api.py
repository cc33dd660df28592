"""GPU parity of the 2:4 compressor, the 2:4/dense split and the reference
drop-in launchers (mmaOBTKernel, mmaOBTSKernel, preprocessing_mmaSTKernel)
against the CPU oracle, through the C ABI."""
import ctypes

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


def mixed_matrix(n=1024, tiles=200, seed=81):
    """Half the tiles pruned to exact 2:4 (tag 1 under both rules), half dense
    (tag 2); sorted CSR columns."""
    rng = np.random.default_rng(seed)
    nb = n // 16
    pos = rng.choice(nb * nb, size=tiles, replace=False)
    dense = np.zeros((n, n), np.float16)
    for t, f in enumerate(pos):
        br, bc = divmod(int(f), nb)
        blk = rng.uniform(-1, 1, (16, 16)).astype(np.float16)
        blk[blk == 0] = 0.5
        if t % 2 == 0:
            blk = gen.prune_2of4(blk)
        dense[br * 16:(br + 1) * 16, bc * 16:(bc + 1) * 16] = blk
    rows, cols = np.nonzero(dense)
    rp = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=n))]).astype(np.int32)
    return n, n, rp, cols.astype(np.int32), dense[rows, cols].view(np.uint16)


def _bcsr(M, K, rp, ci, v, bk=16, rule=0):
    import smat
    return smat.Bcsr(M, K, _dev(rp), _dev(ci), _dev(v.view(np.int16)), bk=bk, tag_rule=rule)


def test_compress24_reference_layout_bit_exact():
    _need_gpu()
    import smat
    M, K, rp, ci, v = mixed_matrix()
    A = _bcsr(M, K, rp, ci, v)
    vals, meta = A.compress_24(smat.SMAT_24_REFERENCE_K16)
    torch.cuda.synchronize()
    brp, bci, bv, tags = of.csr_to_bcsr(M, K, 16, rp, ci, v)
    assert set(np.unique(tags)) == {1, 2}
    sp, me, ov = of.compress24_k16(bv, tags)
    assert ov == 0
    assert np.array_equal(vals.cpu().numpy().view(np.uint16)[: sp.size], sp)
    assert np.array_equal(meta.cpu().numpy()[: me.size], me)


def _decompress_smfmac(vals, idx, nnzb, bk):
    v = vals.reshape(nnzb, 16, bk // 2)
    ix = idx.view(np.uint16).reshape(nnzb, 16, bk // 16)
    out = np.zeros((nnzb, 16, bk), np.uint16)
    for m in range(bk // 2):
        pos = (ix[:, :, m // 8] >> (2 * (m % 8))) & 3
        col = 4 * (m // 2) + pos
        x = v[:, :, m]
        nzm = (x & 0x7FFF) != 0
        b, r = np.nonzero(nzm)
        out[b, r, col[b, r]] = x[b, r]
    return out.reshape(-1)


@pytest.mark.parametrize("bk", [16, 32])
def test_compress24_smfmac_layout_round_trip(bk):
    _need_gpu()
    import smat
    M, K, rp, ci, v = mixed_matrix(seed=82)
    A = _bcsr(M, K, rp, ci, v, bk=bk, rule=smat.SMAT_TAG_EXACT24)
    vals, idx = A.compress_24(smat.SMAT_24_SMFMAC)
    torch.cuda.synchronize()
    brp, bci, bv, _ = of.csr_to_bcsr(M, K, bk, rp, ci, v)
    tags = of.tag_exact24(bk, bv)
    n = tags.size
    got = _decompress_smfmac(vals.cpu().numpy().view(np.uint16)[: n * 8 * bk],
                             idx.cpu().numpy()[: n * bk], n, bk).reshape(n, -1)
    exp = bv.reshape(n, -1)
    for b in range(n):
        if tags[b] == 1:
            assert np.array_equal(got[b], exp[b]), b
    # the two positions of every group are distinct and ascending
    ix = idx.cpu().numpy().view(np.uint16)[: n * bk].reshape(n, 16, bk // 16)
    for m in range(0, bk // 2, 2):
        p0 = (ix[:, :, m // 8] >> (2 * (m % 8))) & 3
        p1 = (ix[:, :, m // 8] >> (2 * (m % 8 + 1))) & 3
        assert np.all(p0 < p1)


def test_compress24_rejects_non_24_blocks():
    _need_gpu()
    import smat
    # row 0 holds columns {0, 2, 3}: no three consecutive columns, so the
    # reference rule (matrix.h:689-709) tags the block 2:4, yet its first
    # 1x4 group has three nonzeros
    M, K = 32, 32
    rp = np.array([0, 3] + [3] * 31, np.int32)
    ci = np.array([0, 2, 3], np.int32)
    v = np.float16([1.0, 2.0, 3.0]).view(np.uint16)
    A = _bcsr(M, K, rp, ci, v)
    brp, bci, bv, tags = of.csr_to_bcsr(M, K, 16, rp, ci, v)
    assert tags.tolist() == [1]
    _, _, ov = of.compress24_k16(bv, tags)
    assert ov > 0  # the oracle overflows too
    with pytest.raises(smat.SmatError) as e:
        A.compress_24(smat.SMAT_24_REFERENCE_K16)
    assert e.value.status == smat.SMAT_ERR_NOT_24


def test_filter_matches_oracle():
    _need_gpu()
    M, K, rp, ci, v = mixed_matrix(seed=84)
    A = _bcsr(M, K, rp, ci, v)
    (srp, sci, sv), (drp, dci, dv) = A.filter()
    torch.cuda.synchronize()
    brp, bci, bv, tags = of.csr_to_bcsr(M, K, 16, rp, ci, v)
    (esrp, esci, esv), (edrp, edci, edv) = of.filter_blocks(16, brp, bci, bv, tags)
    for g, e in ((srp, esrp), (sci, esci), (sv, esv), (drp, edrp), (dci, edci), (dv, edv)):
        assert np.array_equal(g.cpu().numpy().view(e.dtype), e)


def test_grid_maps_match_reference_definition():
    _need_gpu()
    M, K, rp, ci, v = mixed_matrix(seed=85)
    A = _bcsr(M, K, rp, ci, v)
    info, rel = A.grid_maps()
    torch.cuda.synchronize()
    brp, bci, bv, tags = of.csr_to_bcsr(M, K, 16, rp, ci, v)
    ncb = K // 16
    e_info = np.zeros((M // 16) * ncb, np.int32)
    e_rel = np.full((M // 16) * ncb, -1, np.int32)
    for br in range(M // 16):
        for p in range(brp[br], brp[br + 1]):
            e_info[br * ncb + bci[p] // 16] = tags[p]
            e_rel[br * ncb + bci[p] // 16] = p
    assert np.array_equal(info.cpu().numpy(), e_info)
    assert np.array_equal(rel.cpu().numpy(), e_rel)


def _ref_case(seed, N):
    M, K, rp, ci, v = mixed_matrix(seed=seed)
    A = _bcsr(M, K, rp, ci, v)
    B = gen.dense_b(K, N, seed=seed)
    brp, bci, bv, tags = of.csr_to_bcsr(M, K, 16, rp, ci, v)
    R = of.spmm_bcsr_f64(16, brp, bci, bv, B)
    Rsim = of.spmm_refsim_fp16acc(16, brp, bci, bv, B).astype(np.float64)
    Bt = _dev(np.ascontiguousarray(B.T).view(np.int16))  # B[n*K + k]
    return M, K, A, Bt, R, Rsim


@pytest.mark.parametrize("N", [64, 40, 8])
def test_mmaOBTKernel_dropin(N):
    _need_gpu()
    import smat
    M, K, A, Bt, R, Rsim = _ref_case(91, N)
    info, rel = A.grid_maps()
    C = torch.full((M, N), float("nan"), dtype=torch.float16, device="cuda")
    d = A.desc
    smat.lib().smat_clear_error()
    smat.lib().mmaOBTKernel(d.values, d.rowptr, d.colidx, Bt.data_ptr(), C.data_ptr(), M, N, K, d.nnzb,
                            info.data_ptr(), rel.data_ptr())
    torch.cuda.synchronize()
    assert smat.lib().smat_last_status() == 0, smat.lib().smat_last_error()
    Cn = C.cpu().numpy().astype(np.float64)
    assert np.linalg.norm(Cn - R) / np.linalg.norm(R) < 1e-3
    assert np.max(np.abs(Cn - Rsim)) / np.max(np.abs(Rsim)) < 1e-2


def test_mmaOBTSKernel_dropin_mixed():
    _need_gpu()
    import smat
    N = 64
    M, K, A, Bt, R, Rsim = _ref_case(92, N)
    info, rel = A.grid_maps()
    d = A.desc
    sp = torch.zeros(d.nnzb * 128, dtype=torch.int16, device="cuda")
    meta = torch.zeros(d.nnzb * 32, dtype=torch.uint8, device="cuda")
    smat.lib().smat_clear_error()
    smat.lib().preprocessing_mmaSTKernel(d.values, meta.data_ptr(), sp.data_ptr(), M, K, K, d.nnzb,
                                         info.data_ptr(), rel.data_ptr())
    torch.cuda.synchronize()
    assert smat.lib().smat_last_status() == 0, smat.lib().smat_last_error()
    C = torch.full((M, N), float("nan"), dtype=torch.float16, device="cuda")
    # zero the dense copies of the 2:4 blocks so the kernel must use the compressed form
    vals = torch.empty(d.nnzb * 256, dtype=torch.int16, device="cuda")
    tags = torch.empty(d.nnzb, dtype=torch.int32)
    smat.copy_to_host(tags.data_ptr(), d.tags, d.nnzb * 4)
    host_vals = torch.empty(d.nnzb * 256, dtype=torch.int16)
    smat.copy_to_host(host_vals.data_ptr(), d.values, d.nnzb * 512)
    hv = host_vals.view(d.nnzb, 256).clone()
    hv[tags == 1] = 0
    vals.copy_(hv.view(-1))
    smat.lib().mmaOBTSKernel(vals.data_ptr(), d.rowptr, d.colidx, meta.data_ptr(), sp.data_ptr(), Bt.data_ptr(),
                             C.data_ptr(), M, N, K, d.nnzb, info.data_ptr(), rel.data_ptr())
    torch.cuda.synchronize()
    assert smat.lib().smat_last_status() == 0, smat.lib().smat_last_error()
    Cn = C.cpu().numpy().astype(np.float64)
    assert np.linalg.norm(Cn - R) / np.linalg.norm(R) < 1e-3
    assert np.max(np.abs(Cn - Rsim)) / np.max(np.abs(Rsim)) < 1e-2


def test_dropin_reports_bad_arguments():
    _need_gpu()
    import smat
    smat.lib().smat_clear_error()
    smat.lib().mmaOBTKernel(None, None, None, None, None, 16, 8, 16, 0, None, None)
    assert smat.lib().smat_last_status() == smat.SMAT_ERR_INVALID


def mixed_matrix_32(n=1024, tiles=200, seed=83):
    """16x32 tiles (aligned to the k=32 block grid): half pruned to exact
    2:4, half dense."""
    rng = np.random.default_rng(seed)
    nbr, nbc = n // 16, n // 32
    pos = rng.choice(nbr * nbc, size=tiles, replace=False)
    dense = np.zeros((n, n), np.float16)
    for t, f in enumerate(pos):
        br, bc = divmod(int(f), nbc)
        blk = rng.uniform(-1, 1, (16, 32)).astype(np.float16)
        blk[blk == 0] = 0.5
        if t % 2 == 0:
            blk = gen.prune_2of4(blk)
        dense[br * 16:(br + 1) * 16, bc * 32:(bc + 1) * 32] = blk
    rows, cols = np.nonzero(dense)
    rp = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=n))]).astype(np.int32)
    return n, n, rp, cols.astype(np.int32), dense[rows, cols].view(np.uint16)


@pytest.mark.parametrize("N", [64, 40])
def test_mmaOBTSKernel_large_dropin_mixed(N):
    """k=32 drop-ins: preprocessing_mmaSTKernel_large compresses the 2:4
    16x32 blocks, mmaOBTSKernel_large reads them from the compressed pair
    (their dense copies are zeroed first) and the dense ones from
    bcsrValuesA; FP16 C against the FP64 product."""
    _need_gpu()
    import smat
    M, K, rp, ci, v = mixed_matrix_32(seed=93 + N)
    A = _bcsr(M, K, rp, ci, v, bk=32, rule=smat.SMAT_TAG_EXACT24)
    info, rel = A.grid_maps()
    d = A.desc
    sp = torch.zeros(d.nnzb * 256, dtype=torch.int16, device="cuda")
    meta = torch.zeros(d.nnzb * 64, dtype=torch.uint8, device="cuda")
    smat.lib().smat_clear_error()
    smat.lib().preprocessing_mmaSTKernel_large(d.values, meta.data_ptr(), sp.data_ptr(), M, K, K, d.nnzb,
                                               info.data_ptr(), rel.data_ptr())
    torch.cuda.synchronize()
    assert smat.lib().smat_last_status() == 0, smat.lib().smat_last_error()
    tags = torch.empty(d.nnzb, dtype=torch.int32)
    smat.copy_to_host(tags.data_ptr(), d.tags, d.nnzb * 4)
    assert bool((tags == 1).any()) and bool((tags == 2).any())
    host_vals = torch.empty(d.nnzb * 512, dtype=torch.int16)
    smat.copy_to_host(host_vals.data_ptr(), d.values, d.nnzb * 1024)
    hv = host_vals.view(d.nnzb, 512).clone()
    hv[tags == 1] = 0
    vals = hv.view(-1).cuda()
    B = gen.dense_b(K, N, seed=N)
    Bt = _dev(np.ascontiguousarray(B.T).view(np.int16))  # B[n*K + k]
    C = torch.full((M, N), float("nan"), dtype=torch.float16, device="cuda")
    smat.lib().mmaOBTSKernel_large(vals.data_ptr(), d.rowptr, d.colidx, meta.data_ptr(), sp.data_ptr(),
                                   Bt.data_ptr(), C.data_ptr(), M, N, K, d.nnzb, info.data_ptr(), rel.data_ptr())
    torch.cuda.synchronize()
    assert smat.lib().smat_last_status() == 0, smat.lib().smat_last_error()
    R = of.spmm_csr_f64(M, rp, ci, v, B)
    Cn = C.cpu().numpy().astype(np.float64)
    assert np.all(np.isfinite(Cn))
    assert np.linalg.norm(Cn - R) / np.linalg.norm(R) < 1e-3


@pytest.mark.parametrize("N", [64, 40])
def test_mmaOBTSKernel_large_separate_dropin(N):
    """Split k=32 drop-in: the 2:4 list of smat_bcsr_filter read from the
    compressed pair of preprocessing_mmaSTKernel_large (their dense copies
    zeroed), the dense list from bcsrValuesA_sparse; bcsrValuesA_dense is NaN
    and must not be read.  The whole product (the reference returns only the
    dense blocks' share) against the FP64 product."""
    _need_gpu()
    import smat
    M, K, rp, ci, v = mixed_matrix_32(seed=95 + N)
    A = _bcsr(M, K, rp, ci, v, bk=32, rule=smat.SMAT_TAG_EXACT24)
    info, rel = A.grid_maps()
    d = A.desc
    sp = torch.zeros(d.nnzb * 256, dtype=torch.int16, device="cuda")
    meta = torch.zeros(d.nnzb * 64, dtype=torch.uint8, device="cuda")
    smat.lib().smat_clear_error()
    smat.lib().preprocessing_mmaSTKernel_large(d.values, meta.data_ptr(), sp.data_ptr(), M, K, K, d.nnzb,
                                               info.data_ptr(), rel.data_ptr())
    (srp, sci, _), (drp, dci, dvals) = A.filter()
    torch.cuda.synchronize()
    assert smat.lib().smat_last_status() == 0, smat.lib().smat_last_error()
    assert sci.numel() > 0 and dci.numel() > 0
    tags = torch.empty(d.nnzb, dtype=torch.int32)
    smat.copy_to_host(tags.data_ptr(), d.tags, d.nnzb * 4)
    host_vals = torch.empty(d.nnzb * 512, dtype=torch.int16)
    smat.copy_to_host(host_vals.data_ptr(), d.values, d.nnzb * 1024)
    hv = host_vals.view(d.nnzb, 512).clone()
    hv[tags == 1] = 0
    vals = hv.view(-1).cuda()
    nan_dense = torch.full_like(dvals, 0x7e00)  # FP16 NaN bits
    B = gen.dense_b(K, N, seed=N + 1)
    Bt = _dev(np.ascontiguousarray(B.T).view(np.int16))  # B[n*K + k]
    C = torch.full((M, N), float("nan"), dtype=torch.float16, device="cuda")
    smat.lib().mmaOBTSKernel_large_separate(vals.data_ptr(), srp.data_ptr(), sci.data_ptr(), nan_dense.data_ptr(),
                                            drp.data_ptr(), dci.data_ptr(), meta.data_ptr(), sp.data_ptr(),
                                            Bt.data_ptr(), C.data_ptr(), M, N, K, d.nnzb, info.data_ptr(),
                                            rel.data_ptr())
    torch.cuda.synchronize()
    assert smat.lib().smat_last_status() == 0, smat.lib().smat_last_error()
    R = of.spmm_csr_f64(M, rp, ci, v, B)
    Cn = C.cpu().numpy().astype(np.float64)
    assert np.all(np.isfinite(Cn))
    assert np.linalg.norm(Cn - R) / np.linalg.norm(R) < 1e-3
    # the 2:4 half contributes: without it the error would be of order one
    rows = np.repeat(np.arange(M), np.diff(rp))
    in24 = np.isin((rows // 16) * (K // 32) + ci // 32, _blocks_with_tag(A, 1))
    dense_only = of.spmm_csr_f64(M, rp, ci, np.where(in24, 0, v).astype(np.uint16), B)
    assert np.linalg.norm(dense_only - R) / np.linalg.norm(R) > 0.1
    smat.lib().smat_clear_error()
    smat.lib().mmaOBTSKernel_large_separate(vals.data_ptr(), srp.data_ptr(), sci.data_ptr(), nan_dense.data_ptr(),
                                            drp.data_ptr(), dci.data_ptr(), meta.data_ptr(), sp.data_ptr(),
                                            Bt.data_ptr(), C.data_ptr(), M, N, 48, d.nnzb, info.data_ptr(),
                                            rel.data_ptr())
    assert smat.lib().smat_last_status() == smat.SMAT_ERR_INVALID


def _blocks_with_tag(A, tag):
    """Grid cells (block row * K/bk + column block) of the BCSR blocks tagged `tag`."""
    import smat
    d = A.desc
    rowptr = torch.empty(d.nbrows + 1, dtype=torch.int32)
    colidx = torch.empty(max(d.nnzb, 1), dtype=torch.int32)
    tags = torch.empty(max(d.nnzb, 1), dtype=torch.int32)
    smat.copy_to_host(rowptr.data_ptr(), d.rowptr, (d.nbrows + 1) * 4)
    smat.copy_to_host(colidx.data_ptr(), d.colidx, d.nnzb * 4)
    smat.copy_to_host(tags.data_ptr(), d.tags, d.nnzb * 4)
    brow = np.repeat(np.arange(d.nbrows), np.diff(rowptr.numpy()))
    cells = brow * (d.K // d.bk) + colidx.numpy()[: d.nnzb] // d.bk
    return cells[tags.numpy()[: d.nnzb] == tag]


def test_large_dropins_alias_and_arguments():
    """mmaOBTKernel_large works on 16x16 blocks like the reference's copy of
    mmaOBTKernel; the k=32 entry points refuse K that is not a multiple of 32."""
    _need_gpu()
    import smat
    N = 64
    M, K, A, Bt, R, Rsim = _ref_case(94, N)
    info, rel = A.grid_maps()
    d = A.desc
    C = torch.full((M, N), float("nan"), dtype=torch.float16, device="cuda")
    smat.lib().smat_clear_error()
    smat.lib().mmaOBTKernel_large(d.values, d.rowptr, d.colidx, Bt.data_ptr(), C.data_ptr(), M, N, K, d.nnzb,
                                  info.data_ptr(), rel.data_ptr())
    torch.cuda.synchronize()
    assert smat.lib().smat_last_status() == 0, smat.lib().smat_last_error()
    Cn = C.cpu().numpy().astype(np.float64)
    assert np.linalg.norm(Cn - R) / np.linalg.norm(R) < 1e-3
    smat.lib().smat_clear_error()
    smat.lib().mmaOBTSKernel_large(d.values, d.rowptr, d.colidx, d.values, d.values, Bt.data_ptr(), C.data_ptr(),
                                   M, N, 48, d.nnzb, info.data_ptr(), rel.data_ptr())
    assert smat.lib().smat_last_status() == smat.SMAT_ERR_INVALID
    smat.lib().smat_clear_error()
    smat.lib().preprocessing_mmaSTKernel_large(d.values, d.values, d.values, M, N, 48, d.nnzb, info.data_ptr(),
                                               rel.data_ptr())
    assert smat.lib().smat_last_status() == smat.SMAT_ERR_INVALID


def test_dropin_planned_and_reference_paths():
    """The SpMM launchers run the planned kernel when their arrays are the
    BCSR csrToBcsr writes, and the per-block reference path otherwise: a grid
    map that stores the blocks at permuted slots (values moved with them)
    gives the same product through the reference path; the mixed launchers
    with intact dense copies of their 2:4 blocks take the planned path."""
    _need_gpu()
    import smat
    N = 72
    M, K, A, Bt, R, Rsim = _ref_case(93, N)
    info, rel = A.grid_maps()
    d = A.desc
    nb = int(d.nnzb)
    # blocks stored at permuted slots: relMap[g(p)] = sigma(p), values at sigma(p)
    sigma = np.random.default_rng(3).permutation(nb).astype(np.int32)
    hv = torch.empty(nb * 256, dtype=torch.int16)
    smat.copy_to_host(hv.data_ptr(), d.values, nb * 512)
    moved = torch.empty_like(hv).view(nb, 256)
    moved[torch.from_numpy(sigma).long()] = hv.view(nb, 256)
    vals2 = moved.view(-1).cuda()
    relh = rel.cpu().numpy()
    rel2 = np.where(relh >= 0, sigma[np.maximum(relh, 0)], -1).astype(np.int32)
    rel2d = torch.from_numpy(rel2).cuda()
    for vals, relmap, what in ((vals2, rel2d, "reference path"), (None, rel, "planned path")):
        C = torch.full((M, N), float("nan"), dtype=torch.float16, device="cuda")
        smat.lib().smat_clear_error()
        smat.lib().mmaOBTKernel(vals.data_ptr() if vals is not None else d.values, d.rowptr, d.colidx,
                                Bt.data_ptr(), C.data_ptr(), M, N, K, nb, info.data_ptr(), relmap.data_ptr())
        torch.cuda.synchronize()
        assert smat.lib().smat_last_status() == 0, (what, smat.lib().smat_last_error())
        Cn = C.cpu().numpy().astype(np.float64)
        assert np.linalg.norm(Cn - R) / np.linalg.norm(R) < 1e-3, what
        assert np.max(np.abs(Cn - Rsim)) / np.max(np.abs(Rsim)) < 1e-2, what
    # the mixed launcher with intact dense copies (planned path)
    sp = torch.zeros(nb * 128, dtype=torch.int16, device="cuda")
    meta = torch.zeros(nb * 32, dtype=torch.uint8, device="cuda")
    _, _, rp, ci, v = mixed_matrix(seed=93)
    A24 = _bcsr(M, K, rp, ci, v, rule=smat.SMAT_TAG_EXACT24)
    info24, rel24 = A24.grid_maps()
    d24 = A24.desc
    smat.lib().preprocessing_mmaSTKernel(d24.values, meta.data_ptr(), sp.data_ptr(), M, K, K, d24.nnzb,
                                         info24.data_ptr(), rel24.data_ptr())
    C = torch.full((M, N), float("nan"), dtype=torch.float16, device="cuda")
    smat.lib().mmaOBTSKernel(d24.values, d24.rowptr, d24.colidx, meta.data_ptr(), sp.data_ptr(), Bt.data_ptr(),
                             C.data_ptr(), M, N, K, d24.nnzb, info24.data_ptr(), rel24.data_ptr())
    torch.cuda.synchronize()
    assert smat.lib().smat_last_status() == 0, smat.lib().smat_last_error()
    Cn = C.cpu().numpy().astype(np.float64)
    assert np.linalg.norm(Cn - R) / np.linalg.norm(R) < 1e-3
