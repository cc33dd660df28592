"""Pins the oracle's precision-mode functions (e4m3 / BF16 codecs, operand
packing, the scaled product) with closed forms, exhaustive code checks and
independent implementations (numpy FP16, torch BF16, dense brute force)."""
import numpy as np
import pytest

import oracle_ffi as of
from smat import gen

L = of.lib


def _finite_codes():
    return [c for c in range(256) if c not in (0x7F, 0xFF)]


def test_e4m3_known_values():
    enc = L().or_e4m3_encode
    assert enc(448.0) == 0x7E and enc(-448.0) == 0xFE
    assert enc(1.0) == 0x38 and enc(-1.5) == 0xBC and enc(0.5) == 0x30
    assert enc(2.0 ** -9) == 0x01 and enc(2.0 ** -6) == 0x08 and enc(7 * 2.0 ** -9) == 0x07
    assert enc(500.0) == 0x7E and enc(1e30) == 0x7E and enc(-1e30) == 0xFE  # saturation
    assert enc(0.0) == 0x00 and enc(-0.0) == 0x80
    assert enc(2.0 ** -11) == 0x00  # below half the smallest subnormal


def test_e4m3_every_code_round_trips():
    dec, enc = L().or_e4m3_decode, L().or_e4m3_encode
    for c in _finite_codes():
        assert enc(dec(c)) == c, hex(c)
    # value table: positive codes strictly increasing, max 448
    vals = [dec(c) for c in range(0x7F)]
    assert all(b > a for a, b in zip(vals, vals[1:])) and vals[-1] == 448.0


def test_e4m3_midpoints_tie_to_even():
    dec, enc = L().or_e4m3_decode, L().or_e4m3_encode
    for c in range(0x7E):
        lo, hi = dec(c), dec(c + 1)
        mid = np.float32((lo + hi) / 2)
        assert float(mid) == (lo + hi) / 2  # exactly representable
        assert enc(float(mid)) == (c if c % 2 == 0 else c + 1), hex(c)
        assert enc(float(np.nextafter(mid, np.float32(0)))) == c
        assert enc(float(np.nextafter(mid, np.float32(1e9)))) == c + 1


def test_bf16_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.standard_normal(4000).astype(np.float32) * 10.0 ** rng.integers(-20, 20, 4000),
                        np.float32([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -2.5, 0.0, 65504.0])]).astype(np.float32)
    ours = np.array([L().or_f2bf(float(v)) for v in x], np.uint16)
    theirs = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(ours, theirs)
    assert L().or_f2bf(1 + 2 ** -8) == 0x3F80 and L().or_f2bf(1 + 3 * 2 ** -8) == 0x3F82


def _small_bcsr(bk, seed=5):
    M, K, rp, ci, v = gen.power_law(n=512, seed=seed, max_len=96)
    B, bcol, bval, btag = of.csr_to_bcsr(M, K, bk, rp, ci, v)
    return M, K, rp, ci, B, bcol


@pytest.mark.parametrize("bk", [1, 4, 16])
def test_pack_a_scales_and_codes(bk):
    rng = np.random.default_rng(bk)
    M, K, rp, ci, brp, bcol = _small_bcsr(bk)
    n = int(brp[-1]) * 16 * bk
    x = (rng.standard_normal(n) * np.where(rng.random(n) < 0.01, 1e3, 1.0)).astype(np.float32)
    x[rng.random(n) < 0.5] = 0.0
    # F16 / BF16: plain round to nearest even (numpy FP16 is the independent check)
    q16, _ = of.pack_a(bk, brp, x, of.F16)
    assert np.array_equal(q16, x.astype(np.float16).view(np.uint16))
    # FP8: per-block-row scale amax / 448, the row maximum lands on +-448
    q8, s8 = of.pack_a(bk, brp, x, of.FP8)
    dec = np.array([L().or_e4m3_decode(c) for c in range(256)], np.float32)
    for r in range(brp.size - 1):
        seg = slice(int(brp[r]) * 16 * bk, int(brp[r + 1]) * 16 * bk)
        amax = np.abs(x[seg]).max() if x[seg].size else 0.0
        if amax == 0:
            assert s8[r] == 1.0
            continue
        assert s8[r] == np.float32(amax) / np.float32(448)
        assert np.abs(dec[q8[seg]]).max() == 448.0
        y = x[seg] / s8[r]
        rel = np.abs(dec[q8[seg]] - y) / np.maximum(np.abs(y), 2.0 ** -6)
        assert rel.max() <= 2.0 ** -4 + 1e-7  # half an e4m3 ulp (3 mantissa bits)
    # split: power-of-two scale with max in [2^14, 2^15); hi + lo reconstructs
    qs, ss = of.pack_a(bk, brp, x, of.SPLIT)
    hi, lo = qs[:n].view(np.float16).astype(np.float64), qs[n:].view(np.float16).astype(np.float64)
    for r in range(brp.size - 1):
        seg = slice(int(brp[r]) * 16 * bk, int(brp[r + 1]) * 16 * bk)
        amax = np.abs(x[seg]).max() if x[seg].size else 0.0
        if amax == 0:
            continue
        assert np.log2(ss[r]) == np.round(np.log2(ss[r]))
        assert 2 ** 14 <= amax / ss[r] < 2 ** 15
        y = x[seg].astype(np.float64) / float(ss[r])
        assert np.all(np.abs(hi[seg] + lo[seg] - y) <= 2.0 ** -22 * np.abs(y) + 2.0 ** -25)


def test_pack_b_columns():
    rng = np.random.default_rng(9)
    B = rng.uniform(-1, 1, (200, 48)).astype(np.float32)
    B[:, 5] *= 1e-3
    B[:, 7] = 0
    q, s = of.pack_b(B, of.FP8)
    assert s[7] == 1.0 and np.isclose(s[5], np.abs(B[:, 5]).max() / 448, rtol=0)
    dec = np.array([L().or_e4m3_decode(c) for c in range(256)], np.float32)
    D = dec[q.reshape(200, 48)] * s
    assert np.max(np.abs(D - B) / np.maximum(np.abs(B), 1e-30)[None] * (np.abs(B) > 1e-2 * np.abs(B).max(0))) <= 2 ** -4
    q16, _ = of.pack_b(B, of.F16)
    assert np.array_equal(q16, B.astype(np.float16).view(np.uint16).ravel())


@pytest.mark.parametrize("dt", [of.F16, of.FP8, of.SPLIT])
def test_scaled_product_brute_force(dt):
    bk = 4
    rng = np.random.default_rng(dt)
    M, K, rp, ci, brp, bcol = _small_bcsr(bk, seed=11)
    n = int(brp[-1]) * 16 * bk
    x = rng.standard_normal(n).astype(np.float32)
    Bf = rng.uniform(-1, 1, (K, 32)).astype(np.float32)
    a, sa = of.pack_a(bk, brp, x, dt)
    b, sb = of.pack_b(Bf, dt)
    unscaled = dt == of.F16
    C = of.spmm_scaled_f64(bk, brp, bcol, dt, a, None if unscaled else sa, K, 32, b, None if unscaled else sb)
    # dense brute force from decoded operands
    dec8 = np.array([L().or_e4m3_decode(c) for c in range(256)], np.float64)
    if dt == of.FP8:
        av, bv = dec8[a], dec8[b].reshape(K, 32)
    elif dt == of.SPLIT:
        av = a[:n].view(np.float16).astype(np.float64) + a[n:].view(np.float16).astype(np.float64)
        bh = b[: K * 32].view(np.float16).astype(np.float64).reshape(K, 32)
        bl = b[K * 32:].view(np.float16).astype(np.float64).reshape(K, 32)
        bv = bh + bl
        alo = a[n:].view(np.float16).astype(np.float64)
    else:
        av, bv = a.view(np.float16).astype(np.float64), b.view(np.float16).astype(np.float64).reshape(K, 32)
    Ad = np.zeros((M, K))
    Alo = np.zeros((M, K))
    for r in range(brp.size - 1):
        for p in range(brp[r], brp[r + 1]):
            blk = av[p * 16 * bk:(p + 1) * 16 * bk].reshape(16, bk)
            sc = 1.0 if unscaled else float(sa[r])
            Ad[r * 16:(r + 1) * 16, bcol[p]:bcol[p] + bk] = blk * sc
            if dt == of.SPLIT:
                Alo[r * 16:(r + 1) * 16, bcol[p]:bcol[p] + bk] = alo[p * 16 * bk:(p + 1) * 16 * bk].reshape(16, bk) * sc
    sbv = np.ones(32) if unscaled else sb.astype(np.float64)
    R = Ad @ (bv * sbv)
    if dt == of.SPLIT:  # lo*lo is not formed
        R -= Alo @ (bl * sbv)
    np.testing.assert_allclose(C, R, rtol=1e-12, atol=1e-12 * np.abs(R).max())
    if dt == of.F16:  # unit scales reduce to the FP16 oracle
        C16 = of.spmm_bcsr_f64(bk, brp, bcol, a, b.reshape(K, 32))
        np.testing.assert_array_equal(C, C16)
