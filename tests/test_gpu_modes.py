"""GPU parity of the precision / kernel modes: the mixed dense + 2:4 SMFMAC
kernel and BF16 inputs, against FP64 references of the same rounded inputs."""
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


def mixed_matrix(n=2048, tiles=600, seed=81, frac24=0.5):
    rng = np.random.default_rng(seed)
    nb = n // 16
    pos = rng.choice(nb * nb, size=tiles, replace=False)
    dense = np.zeros((n, n), np.float16)
    for t, f in enumerate(pos):
        br, bc = divmod(int(f), nb)
        blk = rng.uniform(-1, 1, (16, 16)).astype(np.float16)
        blk[blk == 0] = 0.5
        if rng.random() < frac24:
            blk = gen.prune_2of4(blk)
            blk[rng.random((16, 16)) < 0.2] = 0  # some groups with fewer than 2
        dense[br * 16:(br + 1) * 16, bc * 16:(bc + 1) * 16] = blk
    rows, cols = np.nonzero(dense)
    rp = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=n))]).astype(np.int32)
    return n, n, rp, cols.astype(np.int32), dense[rows, cols].view(np.uint16)


@pytest.mark.parametrize("N,n_tile", [(64, 0), (128, 32), (48, 16), (256, 0)])
@pytest.mark.parametrize("frac24", [0.5, 1.0, 0.0])
def test_mixed_24_kernel_vs_fp64(N, n_tile, frac24):
    _need_gpu()
    import smat
    M, K, rp, ci, v = mixed_matrix(seed=100 + N, frac24=frac24)
    print(f"mixed N={N} n_tile={n_tile} frac24={frac24}", flush=True)
    A = smat.Bcsr(M, K, _dev(rp), _dev(ci), _dev(v.view(np.int16)), bk=16, tag_rule=smat.SMAT_TAG_EXACT24)
    vals24, idx24 = A.compress_24(smat.SMAT_24_SMFMAC)
    d = A.desc
    brp, bci, bv, _ = of.csr_to_bcsr(M, K, 16, rp, ci, v)
    tags = of.tag_exact24(16, bv)
    B = gen.dense_b(K, N, seed=9)
    R = of.spmm_bcsr_f64(16, brp, bci, bv, B)
    # wipe the dense copies of the 2:4 blocks: the kernel must use the compressed form
    hv = bv.copy().reshape(-1, 256)
    hv[tags == 1] = 0
    smat.copy_to_device(d.values, hv.ctypes.data, hv.nbytes)
    plan = smat.SpmmPlan(A, N, kernel=smat.SMAT_KERNEL_MIXED24, n_tile=n_tile)
    C = torch.full((M, N), float("nan"), dtype=torch.float32, device="cuda")
    plan(_dev(B.view(np.int16)), C, values24=vals24, index24=idx24)
    torch.cuda.synchronize()
    Cn = C.cpu().numpy().astype(np.float64)
    assert np.all(np.isfinite(Cn))
    assert _rel_fro(Cn, R) < 1e-5


def test_mixed_24_split_rows():
    _need_gpu()
    import smat
    M, K, rp, ci, v = mixed_matrix(n=512, tiles=700, seed=7)  # dense-ish block rows
    N = 64
    A = smat.Bcsr(M, K, _dev(rp), _dev(ci), _dev(v.view(np.int16)), bk=16, tag_rule=smat.SMAT_TAG_EXACT24)
    vals24, idx24 = A.compress_24(smat.SMAT_24_SMFMAC)
    brp, bci, bv, _ = of.csr_to_bcsr(M, K, 16, rp, ci, v)
    R = of.spmm_bcsr_f64(16, brp, bci, bv, gen.dense_b(K, N, seed=2))
    plan = smat.SpmmPlan(A, N, kernel=smat.SMAT_KERNEL_MIXED24, max_blocks_per_item=8)
    assert plan.info().n_split_rows > 0
    C = torch.empty((M, N), dtype=torch.float32, device="cuda")
    plan(_dev(gen.dense_b(K, N, seed=2).view(np.int16)), C, values24=vals24, index24=idx24)
    torch.cuda.synchronize()
    assert _rel_fro(C.cpu().numpy().astype(np.float64), R) < 1e-5


def _bf16_case(make, N, seed):
    M, K, rp, ci, v16 = make()
    rng = np.random.default_rng(seed)
    vals = rng.standard_normal(ci.size).astype(np.float32)
    vb = gen.f32_to_bf16_bits(vals)
    vb[vb == 0] = 0x3F80
    Bf = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    Bb = gen.f32_to_bf16_bits(Bf).reshape(K, N)
    A_d = np.zeros((M, K))
    rows = np.repeat(np.arange(M), np.diff(rp))
    A_d[rows, ci] = gen.bf16_bits_to_f32(vb)
    R = A_d @ gen.bf16_bits_to_f32(Bb).reshape(K, N).astype(np.float64)
    return M, K, rp, ci, vb, Bb, R


@pytest.mark.parametrize("bk", [1, 4, 16, 32])
@pytest.mark.parametrize("stage", [1, 2])
def test_bf16_dense_vs_fp64(bk, stage):
    _need_gpu()
    import smat
    M, K, rp, ci, vb, Bb, R = _bf16_case(lambda: gen.rmat(scale=10, edge_factor=8, seed=5), 64, 1)
    A = smat.Bcsr(M, K, _dev(rp), _dev(ci), _dev(vb.view(np.int16)), bk=bk)
    plan = smat.SpmmPlan(A, 64, in_dtype=smat.SMAT_BF16, b_staging=stage)
    C = torch.empty((M, 64), dtype=torch.float32, device="cuda")
    plan(_dev(Bb.view(np.int16)), C)
    torch.cuda.synchronize()
    assert _rel_fro(C.cpu().numpy().astype(np.float64), R) < 1e-5


def test_bf16_mixed_24():
    _need_gpu()
    import smat
    M, K, rp, ci, v = mixed_matrix(seed=31)
    rng = np.random.default_rng(2)
    vb = gen.f32_to_bf16_bits(v.view(np.float16).astype(np.float32))
    N = 64
    Bb = gen.f32_to_bf16_bits(rng.uniform(-1, 1, (K, N)).astype(np.float32)).reshape(K, N)
    A = smat.Bcsr(M, K, _dev(rp), _dev(ci), _dev(vb.view(np.int16)), bk=16, tag_rule=smat.SMAT_TAG_EXACT24)
    vals24, idx24 = A.compress_24(smat.SMAT_24_SMFMAC)
    A_d = np.zeros((M, K))
    A_d[np.repeat(np.arange(M), np.diff(rp)), ci] = gen.bf16_bits_to_f32(vb)
    R = A_d @ gen.bf16_bits_to_f32(Bb).reshape(K, N).astype(np.float64)
    plan = smat.SpmmPlan(A, N, kernel=smat.SMAT_KERNEL_MIXED24, in_dtype=smat.SMAT_BF16)
    C = torch.empty((M, N), dtype=torch.float32, device="cuda")
    plan(_dev(Bb.view(np.int16)), C, values24=vals24, index24=idx24)
    torch.cuda.synchronize()
    assert _rel_fro(C.cpu().numpy().astype(np.float64), R) < 1e-5


def test_mixed_requires_compressed_arrays():
    _need_gpu()
    import smat
    M, K, rp, ci, v = mixed_matrix(n=256, tiles=40, seed=3)
    A = smat.Bcsr(M, K, _dev(rp), _dev(ci), _dev(v.view(np.int16)), bk=16, tag_rule=smat.SMAT_TAG_EXACT24)
    plan = smat.SpmmPlan(A, 16, kernel=smat.SMAT_KERNEL_MIXED24)
    C = torch.empty((M, 16), dtype=torch.float32, device="cuda")
    with pytest.raises(smat.SmatError):
        plan(_dev(gen.dense_b(K, 16).view(np.int16)), C)
    A32 = smat.Bcsr(M, K, _dev(rp), _dev(ci), _dev(v.view(np.int16)), bk=32)
    with pytest.raises(smat.SmatError):
        smat.SpmmPlan(A32, 16, kernel=smat.SMAT_KERNEL_MIXED24)
