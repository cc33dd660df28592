"""CPU tests that pin the oracle: the reference's own KAT, golden vectors
written by the compiled reference (tests/golden/make_golden.py), closed forms
and brute force.  No GPU needed."""
import json
import os

import numpy as np
import pytest

import oracle_ffi as of
from smat import gen

# ---------------------------------------------------------------- Jaccard / merge


def test_jaccard_closed_form_test_similarities():
    # preprocess/test/general/TEST_similarities.cpp:14-22 inputs, block 3
    # (CLI default, input.h:31).  Column blocks: A -> {0,1,3,4,6}, B -> {0,1,3,5}
    # count = |A\B| + |B\A| = 2 + 1 = 3; distance = 2*3/(5 + 4 + 3) = 0.5.
    a = [1, 2, 5, 10, 12, 20]
    b = [0, 2, 4, 10, 16]
    assert of.jaccard(a, 1, b, 1, 3) == pytest.approx(0.5)


def test_jaccard_special_cases():
    assert of.jaccard([], 1, [], 1, 4) == 0.0
    assert of.jaccard([1], 1, [], 1, 4) == 1.0
    assert of.jaccard([], 1, [3], 1, 4) == 1.0
    # identical block patterns -> 0, disjoint -> 2n/(2n+n..)
    assert of.jaccard([0, 5, 9], 2, [1, 6, 10], 1, 4) == 0.0
    # group weighting: A (group 3) has block 0 only, B has block 1 only:
    # count = 1*1 + 1*3 = 4, den = 1*3 + 1*1 + 4 = 8 -> 1.0
    assert of.jaccard([0], 3, [4], 1, 4) == 1.0
    # A blocks {0,1}, B {0}: count = 1 (|A\B| * g_B = 1), den = 2*2 + 1 + 1 = 6
    assert of.jaccard([0, 4], 2, [1], 1, 4) == pytest.approx(np.float32(2.0 / 6.0))


def test_merge_rows_truncation_quirk():
    # utilities.cpp:145-173 is not a union (documented quirk)
    assert of.merge_rows([1, 5], [3, 7]).tolist() == [1, 3, 7]
    assert of.merge_rows([1, 5, 9], [5]).tolist() == [1, 5]
    assert of.merge_rows([], [2, 4]).tolist() == [2, 4]
    assert of.merge_rows([2, 4], []).tolist() == []
    assert of.merge_rows([1, 2, 3], [2, 3]).tolist() == [1, 2, 3]


def test_jaccard_merge_golden(golden_dir):
    d = np.load(os.path.join(golden_dir, "jaccard_merge_golden.npz"))
    ao = np.concatenate([[0], np.cumsum(d["a_lens"])])
    bo = np.concatenate([[0], np.cumsum(d["b_lens"])])
    mo = np.concatenate([[0], np.cumsum(d["m_lens"])])
    for i in range(d["jaccard"].size):
        a = d["a_vals"][ao[i]:ao[i + 1]]
        b = d["b_vals"][bo[i]:bo[i + 1]]
        ga, blk = d["params"][i]
        assert of.jaccard(a, int(ga), b, 1, int(blk)) == float(d["jaccard"][i]), i
        assert of.merge_rows(a, b).tolist() == d["m_vals"][mo[i]:mo[i + 1]].tolist(), i


# ---------------------------------------------------------------- KAT


def test_reference_kat_algo0(golden_dir):
    kat = json.load(open(os.path.join(golden_dir, "kat_test_matrix.json")))
    e = np.array(kat["edges"])
    rows = int(e[:, 0].max()) + 1
    rp = np.zeros(rows + 1, np.int64)
    np.add.at(rp, e[:, 0].astype(np.int64) + 1, 1)
    rp = np.cumsum(rp)
    ci = e[:, 1].astype(np.int64)
    assert (rows, int(ci.max()) + 1, ci.size) == (kat["rows"], kat["cols"], kat["nnz"])
    g, cnt = of.algo0(rp, ci, 0.5, 3)
    assert cnt.tolist() == [kat["comparison_counter"], kat["merge_counter"]]
    assert g.tolist() == kat["compiled_reference"]["grouping"]
    info, avg = of.collect_blocking_info(rp, ci, g, kat["cols"], 3)
    assert info.tolist() == [kat["VBR_nzcount"], kat["VBR_nzblocks_count"], kat["VBR_longest_row"]]
    assert avg == pytest.approx(kat["VBR_average_height"], abs=1e-6)


# ---------------------------------------------------------------- Keeper


def _keeper_cases(golden_dir):
    d = np.load(os.path.join(golden_dir, "keeper_golden.npz"))
    return d, [str(c) for c in d["cases"]]


def test_keeper_golden(golden_dir):
    d, cases = _keeper_cases(golden_dir)
    assert len(cases) == 60
    for key in cases:
        name, params = key.split("__")
        t, b, B = params.split("_")
        tau, cb, mr = float(t[1:]), int(b[1:]), int(B[1:])
        rp, ci = d[f"{name}__rowptr"], d[f"{name}__colidx"]
        cols = int(d[f"{name}__cols"][0])
        g, cnt = of.keeper(rp, ci, np.float32(tau), cb, mr)
        assert np.array_equal(g, d[key + "__grouping"]), key
        assert np.array_equal(cnt, d[key + "__counters"]), key
        fsg = of.fixed_size_grouping(g, mr)
        assert np.array_equal(fsg, d[key + "__fixed"]), key
        perm = of.get_permutation(fsg)
        assert np.array_equal(perm, d[key + "__perm"]), key
        info, avg = of.collect_blocking_info(rp, ci, fsg, cols, cb)
        exp = d[key + "__info"]
        assert info.tolist() == exp[:3].tolist(), key
        assert round(avg * 1e6) == exp[3], key


def test_keeper_invariants():
    # every row grouped; groups of <= max_rows; complete groups numbered by seed
    M, K, rp, ci, _ = gen.rmat(scale=9, edge_factor=6, seed=3)
    g, cnt = of.keeper(rp.astype(np.int64), ci.astype(np.int64), np.float32(0.5), 8, 16)
    rows = rp.size - 1
    assert np.all(g >= 0)
    vals, counts = np.unique(g, return_counts=True)
    assert counts.max() <= 16
    for v, c in zip(vals, counts):
        if v < rows:
            assert c == 16 and g[v] == v  # complete group keyed by its seed
        else:
            assert c < 16 and g[v - rows] == v
    perm = of.get_permutation(of.fixed_size_grouping(g, 16))
    assert np.array_equal(np.sort(perm), np.arange(rows))


def test_keeper_catches_planted_mistake():
    # the golden comparison must fail if the prune test used `>=` instead of `>`
    # (simulated by shifting tau by one ulp), showing the goldens are sensitive
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "keeper_golden.npz"))
    key = "rmat11__t0.5_b8_B16"
    rp, ci = d["rmat11__rowptr"], d["rmat11__colidx"]
    g, cnt = of.keeper(rp, ci, np.nextafter(np.float32(0.5), np.float32(0)), 8, 16)
    assert not (np.array_equal(g, d[key + "__grouping"]) and np.array_equal(cnt, d[key + "__counters"]))


# ---------------------------------------------------------------- CSR -> BCSR


def _rand_csr(rng, M, K, density, shuffle_cols=False):
    dense = (rng.random((M, K)) < density)
    vals = rng.uniform(-1, 1, (M, K)).astype(np.float16)
    vals[vals == 0] = np.float16(0.25)
    rows, cols = np.nonzero(dense)
    v = vals[rows, cols]
    rp = np.concatenate([[0], np.cumsum(dense.sum(1))]).astype(np.int32)
    ci = cols.astype(np.int32)
    if shuffle_cols:
        for r in range(M):
            seg = slice(rp[r], rp[r + 1])
            p = rng.permutation(rp[r + 1] - rp[r])
            ci[seg] = ci[seg][p]
            v[seg] = v[seg][p]
    return rp, ci, v.view(np.uint16)


@pytest.mark.parametrize("bk", [16, 32])
def test_csr_to_bcsr_invariants(bk):
    rng = np.random.default_rng(1)
    M, K = 64, 128
    rp, ci, v = _rand_csr(rng, M, K, 0.05, shuffle_cols=True)
    brp, bci, bv, tags = of.csr_to_bcsr(M, K, bk, rp, ci, v)
    nnzb = bci.size
    assert brp[0] == 0 and brp[-1] == nnzb and np.all(np.diff(brp) >= 0)
    assert np.all(bci % bk == 0)
    # every nonzero at rel*16*bk + (r%16)*bk + c%bk
    dense = gen.csr_to_dense(M, K, rp, ci, v)
    for br in range(M // 16):
        cols_in_row = bci[brp[br]:brp[br + 1]]
        assert np.all(np.diff(cols_in_row) > 0)
        for p in range(brp[br], brp[br + 1]):
            blk = bv[p * 16 * bk:(p + 1) * 16 * bk].view(np.float16).reshape(16, bk).astype(np.float32)
            assert np.array_equal(blk, dense[br * 16:(br + 1) * 16, bci[p]:bci[p] + bk])
            assert np.any(blk != 0)
    # blocks not stored are all zero
    stored = np.zeros((M // 16, K // bk), bool)
    for br in range(M // 16):
        stored[br, bci[brp[br]:brp[br + 1]] // bk] = True
    blocksum = np.abs(dense).reshape(M // 16, 16, K // bk, bk).sum(axis=(1, 3))
    assert np.all((blocksum > 0) == stored)
    assert set(np.unique(tags)) <= {1, 2}


def test_reference_tag_rule_hand_cases():
    # one row, cols 0..15 in block 0: entries (0,1,2) -> c=2 has c-1, c-2 -> dense
    def tag_of(cols):
        rp = np.array([0, len(cols)] + [len(cols)] * 15, np.int32)
        ci = np.array(cols, np.int32)
        v = np.full(len(cols), 0x3C00, np.uint16)  # 1.0
        return of.csr_to_bcsr(16, 16, 16, rp, ci, v)[3].tolist()
    assert tag_of([0, 1, 2]) == [2]
    assert tag_of([1, 2, 3]) == [2]
    assert tag_of([0, 2, 3]) == [1]          # 3 nonzeros in a group, still "2:4"
    assert tag_of([0, 1, 3]) == [1]
    assert tag_of([2, 1, 0]) == [1]          # order-sensitive: descending file order
    assert tag_of([4, 5, 6, 9]) == [2]
    assert tag_of([3, 4, 5]) == [1]          # c=5 has c%4 == 1


def test_exact24_and_filter():
    rng = np.random.default_rng(2)
    M, K = 48, 64
    rp, ci, v = _rand_csr(rng, M, K, 0.3)
    brp, bci, bv, tags = of.csr_to_bcsr(M, K, 16, rp, ci, v)
    ex = of.tag_exact24(16, bv)
    # exact rule by brute force
    for p in range(bci.size):
        blk = bv[p * 256:(p + 1) * 256].reshape(16, 4, 4)
        ok = np.all(((blk & 0x7FFF) != 0).sum(-1) <= 2)
        assert ex[p] == (1 if ok else 2)
    (srp, sci, sv), (drp, dci, dv) = of.filter_blocks(16, brp, bci, bv, ex)
    assert srp[-1] + drp[-1] == bci.size
    assert np.array_equal(np.sort(np.concatenate([sci + 0, dci])), np.sort(bci))
    for br in range(M // 16):
        s = set(bci[brp[br]:brp[br + 1]][ex[brp[br]:brp[br + 1]] == 1].tolist())
        assert s == set(sci[srp[br]:srp[br + 1]].tolist())


# ---------------------------------------------------------------- 2:4


def test_compress24_roundtrip_and_hand_bytes():
    # hand case: row 0, cols 0..7 = [0, a, 0, b | c, 0, 0, 0]
    blk = np.zeros(256, np.uint16)
    a, b, c = 0x3C00, 0x4000, 0x4200  # 1, 2, 3
    blk[1], blk[3], blk[4] = a, b, c
    sp, me, ov = of.compress24_k16(blk, np.array([1], np.int32))
    assert ov == 0
    # lane 0: group 0 -> slots (a@1, b@3), group 1 -> (c@0, 0@0)
    assert sp[:4].tolist() == [a, b, c, 0]
    # meta = idx00<<4 | idx01<<6 | idx10<<0 | idx11<<2 = 1<<4 | 3<<6 | 0 | 0
    assert int(me[0]) == (1 << 4) | (3 << 6)
    # round trip on random exact 2:4 blocks
    rng = np.random.default_rng(4)
    dense = rng.uniform(-1, 1, (20, 16, 16)).astype(np.float16)
    dense = gen.prune_2of4(dense)
    dense[rng.random(dense.shape) < 0.3] = 0
    u = dense.view(np.uint16).reshape(-1)
    sp, me, ov = of.compress24_k16(u, np.ones(20, np.int32))
    assert ov == 0
    assert np.array_equal(of.decompress24_k16(sp, me), u)


def test_compress24_counts_overflow():
    blk = np.zeros(256, np.uint16)
    blk[0:3] = 0x3C00
    _, _, ov = of.compress24_k16(blk, np.array([1], np.int32))
    assert ov == 1


# ---------------------------------------------------------------- SpMM oracles


def test_spmm_f64_vs_dense_and_closed_forms():
    rng = np.random.default_rng(6)
    M, K, N = 64, 96, 24
    rp, ci, v = _rand_csr(rng, M, K, 0.1)
    B = gen.dense_b(K, N, seed=3)
    Ad = gen.csr_to_dense(M, K, rp, ci, v).astype(np.float64)
    Bd = B.view(np.float16).astype(np.float64)
    ref = Ad @ Bd
    assert np.allclose(of.spmm_csr_f64(M, rp, ci, v, B), ref, rtol=0, atol=1e-12)
    for bk in (16, 32):
        brp, bci, bv, _ = of.csr_to_bcsr(M, K, bk, rp, ci, v)
        assert np.allclose(of.spmm_bcsr_f64(bk, brp, bci, bv, B), ref, rtol=0, atol=1e-12)
    # all-ones B: row sums (the only configuration the reference ever ran)
    ones = np.full((K, N), 0x3C00, np.uint16)
    brp, bci, bv, _ = of.csr_to_bcsr(M, K, 16, rp, ci, v)
    C1 = of.spmm_bcsr_f64(16, brp, bci, bv, ones)
    assert np.allclose(C1, Ad.sum(1, keepdims=True) * np.ones((1, N)))


def test_refsim_fp16acc_close_to_f64():
    rng = np.random.default_rng(8)
    M, K, N = 32, 64, 16
    rp, ci, v = _rand_csr(rng, M, K, 0.2)
    B = gen.dense_b(K, N, seed=1)
    brp, bci, bv, _ = of.csr_to_bcsr(M, K, 16, rp, ci, v)
    C = of.spmm_refsim_fp16acc(16, brp, bci, bv, B).astype(np.float64)
    R = of.spmm_bcsr_f64(16, brp, bci, bv, B)
    assert np.max(np.abs(C - R)) / np.max(np.abs(R)) < 1e-2


def _refsim_case(bk):
    """One 16-row block row, three blocks of 16 x bk (columns 0, bk, 2 bk), B
    all ones, N = 2.  Rows hold closed-form cases of FP16 accumulation:
      row 0: 2048, then +1, then +1 in three k-steps: each 2049 ties to the
             even 2048, so C = 2048 (an FP32 accumulator ends at 2050);
      row 1: 2048, 1, 1 inside one k-step: summed before the one rounding,
             2050 is exact in FP16, so C = 2050 (rounding after every
             product gives 2048);
      row 2: 1, 1, then 2048 in block order: 1, 2, 2050, so C = 2050 (the
             blocks visited in reverse give 2048);
      row 3: 2048, +1, -2048: the 1 is lost, C = 0 (FP32 accumulation: 1).
    Returns (rowptr, colidx, values, B, K, expected C[:4, 0])."""
    K = 3 * bk
    A = np.zeros((16, K), np.float32)
    ks = min(bk, 16)  # the second k-step's first column (bk = 32 has two per block)
    A[0, 0], A[0, bk], A[0, 2 * bk] = 2048, 1, 1
    A[1, 0], A[1, 1], A[1, 2] = 2048, 1, 1
    A[2, 0], A[2, bk], A[2, 2 * bk] = 1, 1, 2048
    A[3, 0], A[3, bk], A[3, 2 * bk] = 2048, 1, -2048
    if bk == 32:  # k-steps inside one 16 x 32 block: the same rounding between them
        A[4, 0], A[4, ks], A[4, 2 * bk] = 2048, 1, 1
    rowptr = np.array([0, 3], np.int32)
    colidx = np.array([0, bk, 2 * bk], np.int32)
    vals = np.concatenate([A[:, b * bk:(b + 1) * bk].reshape(-1) for b in range(3)]).astype(np.float16)
    B = np.full((K, 2), 0x3C00, np.uint16)
    exp = [2048.0, 2050.0, 2050.0, 0.0] + ([2048.0] if bk == 32 else [])
    return rowptr, colidx, vals.view(np.uint16), B, K, np.array(exp)


def _refsim_mutant(kind, bk, rowptr, colidx, vals, B):
    """Plausible mistakes in the emulation (tiny cases only): an FP32
    accumulator, rounding after every product, blocks in reverse order."""
    v = vals.view(np.float16).astype(np.float32).reshape(-1, 16, bk)
    Bf = B.view(np.float16).astype(np.float32)
    out = np.zeros(16, np.float64)
    blocks = list(range(rowptr[0], rowptr[1]))
    if kind == "reverse":
        blocks = blocks[::-1]
    for r in range(16):
        acc = np.float32(0) if kind == "fp32acc" else np.float16(0)
        for p in blocks:
            for k0 in range(0, bk, 16):
                if kind == "per_product":
                    for c in range(k0, min(k0 + 16, bk)):
                        acc = np.float16(np.float32(acc) + v[p, r, c] * Bf[colidx[p] + c, 0])
                    continue
                s = np.float32(0)
                for c in range(k0, min(k0 + 16, bk)):
                    s = np.float32(s + v[p, r, c] * Bf[colidx[p] + c, 0])
                acc = np.float32(acc + s) if kind == "fp32acc" else np.float16(np.float32(acc) + s)
        out[r] = float(np.float16(acc))
    return out


@pytest.mark.parametrize("bk", [16, 32])
def test_refsim_fp16acc_closed_forms(bk):
    """or_spmm_refsim_fp16acc against hand-derived FP16-accumulation results
    (HMMA16816 with an FP16 accumulator rounds once per 16-wide k-step,
    ptx.h:20-25; blocks in BCSR order, mmaOBT.cu:87-147); each planted mistake
    misses at least one of them."""
    rowptr, colidx, vals, B, K, exp = _refsim_case(bk)
    C = of.spmm_refsim_fp16acc(bk, rowptr, colidx, vals, B).astype(np.float64)
    n = exp.size
    assert np.array_equal(C[:n, 0], exp) and np.array_equal(C[:n, 1], exp)
    assert np.all(C[n:] == 0)
    for kind in ("fp32acc", "per_product", "reverse"):
        got = _refsim_mutant(kind, bk, rowptr, colidx, vals, B)
        assert not np.array_equal(got[:n], exp), kind


def test_csr_multiply_golden(golden_dir):
    d = np.load(os.path.join(golden_dir, "csr_multiply_golden.npz"))
    n, bc = d["n"]
    C = of.csr_multiply_f32(int(n), d["rowptr"], d["colidx"], d["vals"], d["B"], int(bc))
    assert np.array_equal(C, d["C"])
    C4, used = of.csr_multiply_f32_omp(int(n), d["rowptr"], d["colidx"], d["vals"], d["B"], int(bc), 4)
    assert used >= 1 and np.array_equal(C4, d["C"])


def test_csr_multiply_omp_bit_identical_on_skewed_rows():
    """The all-cores CPU baseline sums each C element in the serial order, so
    it is bit-identical to the serial product on a skewed R-MAT sample (long
    rows spread over chunks and threads), and both agree with the FP64
    product to FP32 accuracy."""
    from smat import gen
    M, K, rp, ci, v = gen.rmat(scale=11, edge_factor=16, seed=4)
    rp64, ci64 = rp.astype(np.int64), ci.astype(np.int64)
    vals = v.view(np.float16).astype(np.float32)
    N = 8
    B = np.random.default_rng(2).uniform(-1, 1, M * N).astype(np.float32)
    C1 = of.csr_multiply_f32(M, rp64, ci64, vals, B, N)
    C8, used = of.csr_multiply_f32_omp(M, rp64, ci64, vals, B, N, 8)
    assert used >= 1 and np.array_equal(C1, C8)
    rows = np.repeat(np.arange(M), np.diff(rp64))
    R = np.zeros((M, N))
    np.add.at(R, rows, vals[:, None].astype(np.float64) * B.reshape(N, M).T[ci64].astype(np.float64))
    assert np.allclose(C8.reshape(N, M).T, R, rtol=1e-4, atol=1e-4)


def test_fp16_conversions():
    xs = np.array([0.0, -0.0, 1.0, -2.5, 65504.0, 6.1e-5, 5.96e-8, 1e-9, 70000.0, 0.1, 1 / 3], np.float32)
    for x in xs:
        assert of.lib().or_f2h(float(x)) == int(np.float16(x).view(np.uint16))
    for h in range(0, 0x7C00, 37):
        assert of.lib().or_h2f(h) == float(np.uint16(h).view(np.float16))


# ---------------------------------------------------------------- locality order
def _locality_brute(rows, cols, rp, ci, with_second):
    """The order's definition as a plain Python sort of per-row tuples."""
    def key(r):
        row = list(ci[rp[r]:rp[r + 1]])
        n = len(row)
        lb = n.bit_length() - 1 if n > 1 else 0
        c0 = row[0] if n > 0 else cols
        c1 = row[1] if n > 1 else cols
        return (-lb, c0, c1, r) if with_second else (-lb, c0, r)
    return sorted(range(rows), key=key)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_locality_order_brute_force(seed):
    rng = np.random.default_rng(seed)
    rows, cols = 200, 97
    lens = rng.integers(0, 12, size=rows)
    lens[rng.random(rows) < 0.1] = 0            # empty rows
    lens[:3] = [33, 64, 65]                      # bucket edges 32 / 64
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = np.concatenate([np.sort(rng.choice(cols, size=min(n, cols), replace=False)) for n in lens]).astype(np.int32)
    rp = np.concatenate([[0], np.cumsum(np.minimum(lens, cols))]).astype(np.int32)
    perm = of.locality_order(rows, cols, rp, ci)
    assert perm.tolist() == _locality_brute(rows, cols, rp, ci, True)
    # a planted mistake (length buckets ascending instead of descending) is caught
    lens_k = np.diff(rp)
    planted = sorted(range(rows), key=lambda r: (int(lens_k[r]).bit_length() - 1 if lens_k[r] > 1 else 0,
                                                 ci[rp[r]] if lens_k[r] else cols, r))
    assert perm.tolist() != planted


def test_locality_order_wide_columns_drop_second_key():
    rows, cols = 6, 1 << 30
    rp = np.array([0, 2, 4, 5, 6, 7, 9], np.int32)
    ci = np.array([5, 9, 5, 7, 5, 3, 1, 5, 6], np.int32)
    perm = of.locality_order(rows, cols, rp, ci)
    # rows 0, 1, 5 have two entries (bucket 1) and tie on the first column:
    # row order decides, where the second column would give 5, 1, 0
    assert perm.tolist() == _locality_brute(rows, cols, rp, ci, False) == [0, 1, 5, 4, 3, 2]
    assert of.locality_order(rows, 1 << 20, rp, ci).tolist() == [5, 1, 0, 4, 3, 2]
