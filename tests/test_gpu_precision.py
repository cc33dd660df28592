"""GPU parity of the precision modes: value scatter, operand packing (bit-exact
with the oracle), the FP8 e4m3 and split-FP16 SpMM (vs the oracle's FP64
product of the same packed operands), accuracy of every mode against the
FP64 product of the FP32 inputs, and the FP64 error reduction."""
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


def _matrix(bk, n=1024, seed=31, max_len=200):
    """Power-law matrix (N(0,1) values, 1% x1e3) as a device BCSR with FP32 block values."""
    import smat
    M, K, rp, ci, v16, v32 = gen.power_law(n=n, seed=seed, max_len=max_len, f32=True)
    d_rp, d_ci = _dev(rp), _dev(ci)
    A = smat.Bcsr(M, K, d_rp, d_ci, _dev(v16.view(np.int16)), bk=bk)
    x = smat.scatter_values(A, d_rp, d_ci, _dev(v32))
    brp, bci, bv, _ = of.csr_to_bcsr(M, K, bk, rp, ci, v16)
    return A, M, K, rp, ci, v32, x, brp, bci, bv


@pytest.mark.parametrize("bk", [1, 2, 4, 8, 16, 32])
def test_scatter_values_matches_block_layout(bk):
    _need_gpu()
    print(f"scatter bk={bk}", flush=True)
    A, M, K, rp, ci, v32, x, brp, bci, bv = _matrix(bk)
    xs = x.cpu().numpy()[: bv.size]
    # the FP16 rounding of the scattered FP32 values is the BCSR's own value array
    assert np.array_equal(xs.astype(np.float16).view(np.uint16), bv)
    assert np.isclose(xs.astype(np.float64).sum(), v32.astype(np.float64).sum(), rtol=1e-9)


@pytest.mark.parametrize("dt", [of.F16, of.BF16, of.FP8, of.SPLIT])
@pytest.mark.parametrize("bk", [1, 16])
def test_pack_a_bit_exact(bk, dt):
    _need_gpu()
    import smat
    print(f"pack_a bk={bk} dt={dt}", flush=True)
    A, M, K, rp, ci, v32, x, brp, bci, bv = _matrix(bk, seed=40 + bk)
    q, s = smat.pack_a(A, x, dt)
    torch.cuda.synchronize()
    eq, es = of.pack_a(bk, brp, x.cpu().numpy()[: bv.size], dt)
    got = q.cpu().numpy().view(np.uint8 if dt == of.FP8 else np.uint16)[: eq.size]
    assert np.array_equal(got, eq)
    if s is not None:
        assert np.array_equal(s.cpu().numpy()[: es.size], es)


@pytest.mark.parametrize("dt", [of.F16, of.BF16, of.FP8, of.SPLIT])
def test_pack_b_bit_exact(dt):
    _need_gpu()
    import smat
    rng = np.random.default_rng(dt)
    K, N = 1100, 64
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    B[:, 3] *= 1e-4
    B[:, 9] = 0
    B[:, 10] *= 3e4
    # e4m3 midpoints and saturation cases in one column
    dec = np.array([of.lib().or_e4m3_decode(c) for c in range(0x7F)], np.float64)
    mids = ((dec[:-1] + dec[1:]) / 2).astype(np.float32)
    B[:, 11] = 0
    B[: mids.size, 11] = mids
    B[mids.size, 11] = 448.0  # column max 448 -> FP8 scale 1, the midpoints stay exact
    q, s = smat.pack_b(_dev(B), dt)
    torch.cuda.synchronize()
    eq, es = of.pack_b(B, dt)
    assert np.array_equal(q.cpu().numpy().view(np.uint8 if dt == of.FP8 else np.uint16), eq)
    if s is not None:
        assert np.array_equal(s.cpu().numpy(), es)


# Elementwise bound |C - R| <= TOL * (|A| |B|) (the product of the magnitudes of
# the packed operands).  FP16 / BF16 / split: FP32 accumulation.  FP8: the
# gfx950 FP8 dot product drops the low bits of products far below the largest
# one in the same instruction (measured on random full-range e4m3 operands:
# |D - exact| <= 2^-13.3 * sum |a b| over 8 chained MFMAs; exact when the
# products span a narrow range), so its bound is 2^-11.
_ELEM_TOL = {of.F16: 2.0 ** -18, of.BF16: 2.0 ** -18, of.SPLIT: 2.0 ** -18, of.FP8: 2.0 ** -11}


@pytest.mark.parametrize("dt", [of.FP8, of.SPLIT, of.BF16, of.F16])
@pytest.mark.parametrize("bk", [1, 2, 4, 8, 16, 32])
def test_spmm_packed_vs_oracle(bk, dt):
    """SpMM on packed operands vs the oracle's FP64 product of the same packed
    operands: only the matrix cores' accumulation separates them."""
    _need_gpu()
    import smat
    A, M, K, rp, ci, v32, x, brp, bci, bv = _matrix(bk, seed=60 + bk)
    N = 64
    rng = np.random.default_rng(bk)
    Bf = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    a, sa = smat.pack_a(A, x, dt)
    b, sb = smat.pack_b(_dev(Bf), dt)
    torch.cuda.synchronize()
    ha = a.cpu().numpy().view(np.uint8 if dt == of.FP8 else np.uint16)
    hb = b.cpu().numpy().view(np.uint8 if dt == of.FP8 else np.uint16)
    scaled = dt in (of.FP8, of.SPLIT)
    R = of.spmm_scaled_f64(bk, brp, bci, dt, ha, sa.cpu().numpy() if scaled else None, K, N, hb,
                           sb.cpu().numpy() if scaled else None)
    sign = 0x7F if dt == of.FP8 else 0x7FFF  # clear the sign bits: |A| |B|
    Rabs = of.spmm_scaled_f64(bk, brp, bci, dt, ha & sign, sa.cpu().numpy() if scaled else None, K, N, hb & sign,
                              sb.cpu().numpy() if scaled else None)
    cases = [(0, 0), (32, 0), (16, 64), (64, 32)]
    if dt != of.SPLIT:
        cases += [(128, 0), (128, 16)]  # 128-column tiles (half empty at N = 64)
    for n_tile, split in cases:
        print(f"spmm packed bk={bk} dt={dt} n_tile={n_tile} split={split}", flush=True)
        plan = smat.SpmmPlan(A, N, in_dtype=dt, n_tile=n_tile, max_blocks_per_item=split)
        C = torch.full((M, N), float("nan"), dtype=torch.float32, device="cuda")
        plan.execute(b, C, a_values=a, a_scale=sa, b_scale=sb)
        torch.cuda.synchronize()
        Cn = C.cpu().numpy().astype(np.float64)
        assert np.all(np.isfinite(Cn))
        excess = np.abs(Cn - R) - _ELEM_TOL[dt] * Rabs
        assert excess.max() <= 1e-30, (n_tile, split, float(excess.max()))
        assert _rel_fro(Cn, R) < (2e-4 if dt == of.FP8 else 1e-5), (n_tile, split)
        plan.close()


def test_precision_ladder_accuracy():
    """Every mode against the FP64 product of the FP32 inputs (the ladder's
    accuracy claims: split <= 1e-5, FP16 <= 2e-3, BF16 <= 1e-2, FP8 <= 8e-2)."""
    _need_gpu()
    import smat
    bk = 2
    A, M, K, rp, ci, v32, x, brp, bci, bv = _matrix(bk, n=2048, seed=77, max_len=400)
    N = 128
    rng = np.random.default_rng(5)
    Bf = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    # FP64 truth from the FP32 inputs (dense brute force over the CSR)
    R = np.zeros((M, N))
    rows = np.repeat(np.arange(M), np.diff(rp))
    np.add.at(R, rows, v32.astype(np.float64)[:, None] * Bf.astype(np.float64)[ci])
    Rd = _dev(R)
    errs = {}
    for dt, bound in [(of.SPLIT, 1e-5), (of.F16, 2e-3), (of.BF16, 1e-2), (of.FP8, 8e-2)]:
        a, sa = smat.pack_a(A, x, dt)
        b, sb = smat.pack_b(_dev(Bf), dt)
        plan = smat.SpmmPlan(A, N, in_dtype=dt)
        C = torch.empty((M, N), dtype=torch.float32, device="cuda")
        plan.execute(b, C, a_values=a, a_scale=sa, b_scale=sb)
        e_gpu = smat.rel_fro_error(C, Rd)
        e_cpu = _rel_fro(C.cpu().numpy().astype(np.float64), R)
        print(f"ladder dt={dt}: rel fro {e_gpu:.3e}", flush=True)
        assert abs(e_gpu - e_cpu) <= 1e-9 + 1e-6 * e_cpu
        assert e_gpu <= bound, (dt, e_gpu)
        errs[dt] = e_gpu
    assert errs[of.SPLIT] < errs[of.F16] < errs[of.BF16] < errs[of.FP8]


def test_error_sums_vs_numpy():
    _need_gpu()
    import smat
    rng = np.random.default_rng(2)
    R = rng.standard_normal((333, 70))
    C = (R + 1e-3 * rng.standard_normal(R.shape)).astype(np.float32)
    e = smat.rel_fro_error(_dev(C), _dev(R))
    assert abs(e - _rel_fro(C.astype(np.float64), R)) < 1e-12
    assert smat.rel_fro_error(_dev(R.astype(np.float32)), _dev(R.astype(np.float32).astype(np.float64))) == 0.0


def test_scaled_modes_reject_bad_options():
    _need_gpu()
    import smat
    A, M, K, *_ = _matrix(16, n=256, max_len=40)
    for kw in [dict(kernel=smat.SMAT_KERNEL_MIXED24), dict(out_dtype=smat.SMAT_F16),
               dict(b_layout=smat.SMAT_B_REFERENCE), dict(b_staging=smat.SMAT_BSTAGE_DIRECT)]:
        with pytest.raises(smat.SmatError):
            smat.SpmmPlan(A, 64, in_dtype=smat.SMAT_FP8_E4M3, **kw)
    plan = smat.SpmmPlan(A, 64, in_dtype=smat.SMAT_FP8_E4M3)
    C = torch.empty((M, 64), dtype=torch.float32, device="cuda")
    B = torch.zeros(K * 64, dtype=torch.uint8, device="cuda")
    with pytest.raises(smat.SmatError):  # no scales
        plan(B, C)
    with pytest.raises(smat.SmatError):
        plan.execute(B, C)


@pytest.mark.parametrize("bk", [1, 2, 16])
@pytest.mark.parametrize("N", [160, 256])
def test_fp8_default_128_column_tiles(bk, N):
    """FP8 plans default to 128-column tiles when N >= 128 (a whole 128-byte
    line per B row segment): same bound against the oracle's FP64 product of
    the packed operands as 64-column tiles, every column written."""
    _need_gpu()
    import smat
    A, M, K, rp, ci, v32, x, brp, bci, bv = _matrix(bk, seed=90 + bk)
    rng = np.random.default_rng(N + bk)
    Bf = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    a, sa = smat.pack_a(A, x, of.FP8)
    b, sb = smat.pack_b(_dev(Bf), of.FP8)
    torch.cuda.synchronize()
    ha = a.cpu().numpy().view(np.uint8)
    hb = b.cpu().numpy().view(np.uint8)
    R = of.spmm_scaled_f64(bk, brp, bci, of.FP8, ha, sa.cpu().numpy(), K, N, hb, sb.cpu().numpy())
    for n_tile in (0, 64):
        print(f"fp8 bk={bk} N={N} n_tile={n_tile}", flush=True)
        plan = smat.SpmmPlan(A, N, in_dtype=of.FP8, n_tile=n_tile)
        C = torch.full((M, N), float("nan"), dtype=torch.float32, device="cuda")
        plan.execute(b, C, a_values=a, a_scale=sa, b_scale=sb)
        torch.cuda.synchronize()
        Cn = C.cpu().numpy().astype(np.float64)
        assert np.all(np.isfinite(Cn))
        assert _rel_fro(Cn, R) < 2e-4, n_tile
        plan.close()
