"""Full-size GPU parity at every workload configuration of the benchmark.

  C1  band n = 4096, half-bandwidth 64, p = 0.5 (seed 1001), N = 128
  C2  R-MAT scale 18 (262144 rows, 16 edge draws per row, seed 1002), N = 64 and 256,
      as given and in the locality row order (C written back through the row map)
  C3  block-clustered n = 65536, 12000 tiles pruned to 2:4 (seed 1003), N = 256,
      FP16 and BF16, the 2:4 SMFMAC kernel against dense-block MFMA
  C4  power-law n = 131072 (seed 1004), N = 128, the four precision modes
  C5  64 seeded matrices (seeds 0-63 of the batched workload), N = 128, the model's
      picks as one graph-captured batch (branches and persistent queues)

Ground truth is the FP64 product of the rounded inputs the kernels consume
(host scipy, FP64), compared on the GPU with the FP64 error reduction
(smat_error_sums).  FP16 / BF16 inputs with FP32 accumulation stay below a
relative Frobenius error of 1e-5.  C1 and C2 are also compared with the
oracle's emulation of the reference kernel's FP16 accumulation
(or_spmm_refsim_fp16acc, mmaOBT.cu:87-147 with ptx.h:20-25), normwise max
error max|C - R_ref| / max|R_ref|.  B values are U[-1, 1] from numpy seed 7, as
the benchmark draws them."""
import numpy as np
import pytest

import oracle_ffi as of
from smat import gen

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
sp = pytest.importorskip("scipy.sparse")


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _b_f32(K, N):
    return np.random.default_rng(7).uniform(-1.0, 1.0, size=(K, N)).astype(np.float32)


def _fp64(M, K, rp, ci, vals64, B64):
    """FP64 product on the host, returned as a device tensor."""
    A = sp.csr_matrix((vals64, ci, rp), shape=(M, K))
    return _dev(np.ascontiguousarray(A @ B64))


def _run(A, N, B, C, **kw):
    import smat
    ex = {k: kw.pop(k) for k in ("values24", "index24", "c_rowmap", "a_values", "a_scale", "b_scale") if k in kw}
    plan = smat.SpmmPlan(A, N, **kw)
    C.fill_(float("nan"))
    plan.execute(B, C, **ex)
    torch.cuda.synchronize()
    plan.close()
    return C


def _normwise_max(C, R):
    return float(np.max(np.abs(C - R)) / max(float(np.max(np.abs(R))), 1e-300))


def _refsim_rows(brp, bci, bv, B_u16, lo, hi, ncols):
    """The reference-semantics emulation on block rows [lo, hi) and the first
    `ncols` columns (each C column depends only on its B column)."""
    sub_rp = (brp[lo:hi + 1] - brp[lo]).astype(np.int32)
    sci = bci[brp[lo]:brp[hi]]
    sv = bv[brp[lo] * 256:brp[hi] * 256]
    Bs = np.ascontiguousarray(B_u16[:, :ncols])
    return (of.spmm_refsim_fp16acc(16, sub_rp, sci, sv, Bs).astype(np.float64),
            of.spmm_bcsr_f64(16, sub_rp, sci, sv, Bs))


# ---------------------------------------------------------------------------- C1
def test_c1_band4096_n128():
    _need_gpu()
    import smat
    M, K, rp, ci, v = gen.band(n=4096, hb=64, p=0.5, seed=1001)
    N = 128
    Bh = _b_f32(K, N).astype(np.float16)
    B = _dev(Bh.view(np.int16))
    R = _fp64(M, K, rp, ci, v.view(np.float16).astype(np.float64), Bh.astype(np.float64))
    brp, bci, bv, _ = of.csr_to_bcsr(M, K, 16, rp, ci, v)
    Rsim = of.spmm_refsim_fp16acc(16, brp, bci, bv, Bh.view(np.uint16)).astype(np.float64)
    C = torch.empty((M, N), dtype=torch.float32, device="cuda")
    d = (_dev(rp), _dev(ci), _dev(v.view(np.int16)))
    for bk in (1, 2, 4, 8, 16, 32):
        A = smat.Bcsr(M, K, *d, bk=bk)
        for nt in (0, 32, 128):
            print(f"C1 bk={bk} n_tile={nt}", flush=True)
            _run(A, N, B, C, n_tile=nt)
            assert smat.rel_fro_error(C, R) < 1e-5, (bk, nt)
            if bk == 16 and nt == 0:  # the plain configuration against the reference semantics
                e = _normwise_max(C.cpu().numpy().astype(np.float64), Rsim)
                print(f"C1 plain vs FP16-accumulate emulation: {e:.2e}", flush=True)
                assert e < 1e-2
        A.close()
    # the 2:4 kernel on the band (few exact 2:4 blocks: most run as dense pairs)
    A = smat.Bcsr(M, K, *d, bk=16, tag_rule=smat.SMAT_TAG_EXACT24)
    v24, i24 = A.compress_24(smat.SMAT_24_SMFMAC)
    _run(A, N, B, C, kernel=smat.SMAT_KERNEL_MIXED24, values24=v24, index24=i24)
    assert smat.rel_fro_error(C, R) < 1e-5
    # FP16 C in the reference's own B layout (B[n*K + k]) against the emulation
    Cf = torch.empty((M, N), dtype=torch.float16, device="cuda")
    plan = smat.SpmmPlan(A, N, out_dtype=smat.SMAT_F16, b_layout=smat.SMAT_B_REFERENCE)
    plan(_dev(np.ascontiguousarray(Bh.T).view(np.int16)), Cf, ldb=K)
    torch.cuda.synchronize()
    Cn = Cf.cpu().numpy().astype(np.float64)
    assert _normwise_max(Cn, Rsim) < 1e-2
    assert smat.rel_fro_error(Cf.float(), R) < 1e-3
    plan.close()
    A.close()


# ---------------------------------------------------------------------------- C2
@pytest.fixture(scope="module")
def rmat18():
    M, K, rp, ci, v = gen.rmat(scale=18, seed=1002)
    return M, K, rp, ci, v


@pytest.mark.timeout(600)
@pytest.mark.parametrize("N", [64, 256])
def test_c2_rmat18(rmat18, N):
    _need_gpu()
    import smat
    M, K, rp, ci, v = rmat18
    assert M == 262144 and ci.size > 3_900_000
    Bh = _b_f32(K, N).astype(np.float16)
    B = _dev(Bh.view(np.int16))
    R = _fp64(M, K, rp, ci, v.view(np.float16).astype(np.float64), Bh.astype(np.float64))
    C = torch.empty((M, N), dtype=torch.float32, device="cuda")
    d = (_dev(rp), _dev(ci), _dev(v.view(np.int16)))
    # as given: the plain 16x16 call, the learned picks' block widths, every schedule
    for bk, nt, sched in ((16, 0, 0), (1, 64, 0), (2, 64, 0), (2, 64, 3), (2, 64, 2), (1, 128, 0), (4, 32, 1)):
        print(f"C2 N={N} bk={bk} n_tile={nt} schedule={sched}", flush=True)
        A = smat.Bcsr(M, K, *d, bk=bk)
        _run(A, N, B, C, n_tile=nt, schedule=sched)
        e = smat.rel_fro_error(C, R)
        assert e < 1e-5, (bk, nt, sched, e)
        if bk == 16:
            Cplain = C.cpu().numpy()
        A.close()
    # locality row order, C written back through the row map
    perm = smat.locality_order(M, K, d[0], d[1])
    o = smat.permute_rows(*d, perm)
    rowmap = perm.to(torch.int32)
    for bk in (1, 2, 16):
        print(f"C2 N={N} locality bk={bk}", flush=True)
        A = smat.Bcsr(M, K, *o, bk=bk)
        _run(A, N, B, C, c_rowmap=rowmap)
        assert smat.rel_fro_error(C, R) < 1e-5, bk
        A.close()
    # against the emulation of the reference's FP16 accumulation: hub block rows,
    # a middle range and the tail, first 16 columns
    brp, bci, bv, _ = of.csr_to_bcsr(M, K, 16, rp, ci, v)
    for lo, hi in ((0, 32), (8000, 8064), (16320, 16384)):
        Rsim, R64 = _refsim_rows(brp, bci, bv, Bh.view(np.uint16), lo, hi, 16)
        got = Cplain[lo * 16:hi * 16, :16].astype(np.float64)
        e_ours = _normwise_max(got, Rsim)
        e_ref = _normwise_max(R64, Rsim)  # the reference's own distance from the exact product
        print(f"C2 N={N} block rows {lo}-{hi}: vs emulation {e_ours:.2e}, emulation vs FP64 {e_ref:.2e}", flush=True)
        # never further from the reference's result than the exact product is
        assert np.max(np.abs(got - Rsim)) <= np.max(np.abs(R64 - Rsim)) + 1e-5 * np.max(np.abs(R64))
        if e_ref < 5e-3:  # where FP16 accumulation itself stays accurate: the stated 1e-2 bound
            assert e_ours < 1e-2


# ---------------------------------------------------------------------------- C3
@pytest.mark.parametrize("dtype", ["f16", "bf16"])
def test_c3_clustered_24(dtype):
    _need_gpu()
    import smat
    M, K, rp, ci, v = gen.block_clustered(n=65536, tiles=12000, seed=1003)
    N = 256
    Bf = _b_f32(K, N)
    if dtype == "bf16":
        v = gen.f32_to_bf16_bits(v.view(np.float16).astype(np.float32))
        vals64 = gen.bf16_bits_to_f32(v).astype(np.float64)
        Bbits = gen.f32_to_bf16_bits(Bf).reshape(K, N)
        B64 = gen.bf16_bits_to_f32(Bbits).reshape(K, N).astype(np.float64)
        in_dt = smat.SMAT_BF16
    else:
        vals64 = v.view(np.float16).astype(np.float64)
        Bbits = Bf.astype(np.float16).view(np.uint16)
        B64 = Bbits.view(np.float16).astype(np.float64)
        in_dt = smat.SMAT_F16
    B = _dev(Bbits.view(np.int16))
    R = _fp64(M, K, rp, ci, vals64, B64)
    d = (_dev(rp), _dev(ci), _dev(v.view(np.int16)))
    C = torch.empty((M, N), dtype=torch.float32, device="cuda")
    A = smat.Bcsr(M, K, *d, bk=16, tag_rule=smat.SMAT_TAG_EXACT24)
    tags = A.arrays()[3].numpy()
    assert A.nnzb == 12000 and np.all(tags == 1)  # every tile is 2:4: all of them run on SMFMAC
    v24, i24 = A.compress_24(smat.SMAT_24_SMFMAC)
    out = {}
    for kernel, nt in ((smat.SMAT_KERNEL_MIXED24, 0), (smat.SMAT_KERNEL_MIXED24, 32),
                       (smat.SMAT_KERNEL_DENSE, 0), (smat.SMAT_KERNEL_DENSE, 32)):
        print(f"C3 {dtype} kernel={kernel} n_tile={nt}", flush=True)
        _run(A, N, B, C, kernel=kernel, n_tile=nt, in_dtype=in_dt, values24=v24, index24=i24)
        assert smat.rel_fro_error(C, R) < 1e-5, (kernel, nt)
        out[(kernel, nt)] = C.clone()
    # SMFMAC and dense MFMA agree (each sums exact products in FP32)
    a, b = out[(smat.SMAT_KERNEL_MIXED24, 0)], out[(smat.SMAT_KERNEL_DENSE, 0)]
    assert float((a - b).abs().max()) <= 1e-5 * float(b.abs().max())
    A.close()
    for bk in (4, 8):  # the tuned block widths
        A = smat.Bcsr(M, K, *d, bk=bk)
        _run(A, N, B, C, in_dtype=in_dt)
        assert smat.rel_fro_error(C, R) < 1e-5, bk
        A.close()


# ---------------------------------------------------------------------------- C4
_LADDER = {"split_f16": 1e-5, "f16": 2e-3, "bf16": 1e-2, "fp8_e4m3": 8e-2}


@pytest.mark.timeout(600)
def test_c4_powerlaw_ladder():
    _need_gpu()
    import smat
    M, K, rp, ci, v16, v32 = gen.power_law(n=131072, seed=1004, f32=True)
    N = 128
    Bf = _b_f32(K, N)
    R32 = _fp64(M, K, rp, ci, v32.astype(np.float64), Bf.astype(np.float64))  # truth of the FP32 inputs
    d = (_dev(rp), _dev(ci), _dev(v16.view(np.int16)))
    Bd = _dev(Bf)
    C = torch.empty((M, N), dtype=torch.float32, device="cuda")
    errs = {}
    for bk in (1, 2):
        A = smat.Bcsr(M, K, *d, bk=bk)
        x = smat.scatter_values(A, d[0], d[1], _dev(v32))
        brp, bci, _, _ = of.csr_to_bcsr(M, K, bk, rp, ci, v16)
        nb = M // 16
        ranges = ((0, 64), (nb // 2, nb // 2 + 64), (nb - 64, nb))  # block rows checked against the oracle
        for name, dt in (("f16", smat.SMAT_F16), ("bf16", smat.SMAT_BF16), ("fp8_e4m3", smat.SMAT_FP8_E4M3),
                         ("split_f16", smat.SMAT_SPLIT_F16)):
            print(f"C4 bk={bk} {name}", flush=True)
            a, sa = smat.pack_a(A, x, dt)
            b, sb = smat.pack_b(Bd, dt)
            _run(A, N, b, C, in_dtype=dt, a_values=a, a_scale=sa, b_scale=sb)
            e = smat.rel_fro_error(C, R32)
            assert e <= _LADDER[name], (bk, name, e)
            errs[(bk, name)] = e
            # the oracle's FP64 product of the same packed operands, on sampled block rows
            ha = a.cpu().numpy().view(np.uint8 if dt == of.FP8 else np.uint16)
            hb = b.cpu().numpy().view(np.uint8 if dt == of.FP8 else np.uint16)
            scaled = dt in (of.FP8, of.SPLIT)
            hsa = sa.cpu().numpy() if scaled else None
            hsb = sb.cpu().numpy() if scaled else None
            per = 16 * bk
            nnzb = bci.size
            for lo, hi in ranges:
                p0, p1 = brp[lo], brp[hi]
                sub_rp = (brp[lo:hi + 1] - p0).astype(np.int32)
                sub_a = ha[p0 * per:p1 * per]
                if dt == of.SPLIT:  # hi parts, then lo parts
                    sub_a = np.concatenate([ha[p0 * per:p1 * per], ha[nnzb * per + p0 * per:nnzb * per + p1 * per]])
                sub_sa = hsa[lo:hi] if scaled else None
                Rp = of.spmm_scaled_f64(bk, sub_rp, bci[p0:p1], dt, sub_a, sub_sa, K, N, hb, hsb)
                sign = 0x7F if dt == of.FP8 else 0x7FFF
                Rabs = of.spmm_scaled_f64(bk, sub_rp, bci[p0:p1], dt, sub_a & sign, sub_sa, K, N, hb & sign, hsb)
                got = C[lo * 16:hi * 16].cpu().numpy().astype(np.float64)
                tol = 2.0 ** -11 if dt == of.FP8 else 2.0 ** -18
                assert (np.abs(got - Rp) - tol * Rabs).max() <= 1e-30, (bk, name, lo)
        A.close()
    for bk in (1, 2):
        assert errs[(bk, "split_f16")] < errs[(bk, "f16")] < errs[(bk, "bf16")] < errs[(bk, "fp8_e4m3")]


# ---------------------------------------------------------------------------- C5
def _gen(seed):
    from smat import tuner
    return tuner.make_matrix(seed)


@pytest.mark.timeout(900)
def test_c5_batch64_graph():
    _need_gpu()
    import multiprocessing as mp
    import smat
    from smat import tuner
    N = 128
    with mp.get_context("spawn").Pool(8) as pool:  # numpy generators only; spawned, not forked
        mats = pool.map(_gen, range(64))
    model = tuner.Model()
    entries, keep, refs = [], [], []
    n_mixed = n_rowmap = 0
    for i, (kind, (M, K, rp, ci, v)) in enumerate(mats):
        d = (_dev(rp), _dev(ci), _dev(v.view(np.int16)))
        st = tuner.stats(M, K, *d, int(ci.size))
        tau, bk, nt, kernel = model.choose(tuner.features(st, N), M)
        rowmap = None
        if tau is not None:
            rp2, ci2, v2, perm = tuner.reorder_csr(rp, ci, v, tau)
            d = (_dev(rp2), _dev(ci2), _dev(v2.view(np.int16)))
            rowmap = _dev(perm.astype(np.int32))
        elif M > tuner.REORDER_MAX_ROWS and i % 3 == 0:  # the locality order on a third of the large ones
            perm = smat.locality_order(M, K, d[0], d[1])
            d = smat.permute_rows(*d, perm)
            rowmap = perm.to(torch.int32)
        mixed = kernel == "mixed24"
        A = smat.Bcsr(M, K, *d, bk=bk, tag_rule=smat.SMAT_TAG_EXACT24 if mixed else smat.SMAT_TAG_REFERENCE)
        v24 = i24 = None
        if mixed:
            v24, i24 = A.compress_24(smat.SMAT_24_SMFMAC)
        plan = smat.SpmmPlan(A, N, n_tile=nt, max_blocks_per_item=64 * (32 // bk),
                             kernel=smat.SMAT_KERNEL_MIXED24 if mixed else smat.SMAT_KERNEL_DENSE)
        Bh = gen.dense_b(K, N, seed=100 + i)
        Bd = _dev(Bh.view(np.int16))
        C = torch.full((M, N), float("nan"), dtype=torch.float32, device="cuda")
        entries.append((plan, plan.operands(Bd, C, values24=v24, index24=i24, c_rowmap=rowmap)))
        keep.append((A, d, Bd, C, v24, i24, rowmap))
        refs.append(_fp64(M, K, rp, ci, v.view(np.float16).astype(np.float64),
                          Bh.view(np.float16).astype(np.float64)))
        n_mixed += mixed
        n_rowmap += rowmap is not None
    print(f"C5: 64 matrices, {sum(m[1][3].size for m in mats)} nnz, {n_mixed} on the 2:4 kernel, "
          f"{n_rowmap} with a row map", flush=True)
    assert n_rowmap > 0
    s = torch.cuda.Stream()
    for mode in (smat.SMAT_BATCH_BRANCHES, smat.SMAT_BATCH_PERSISTENT):
        batch = smat.Batch(entries, mode=mode)
        with torch.cuda.stream(s):
            batch.capture(s)
        first = None
        for rep in range(3):
            for k in keep:
                k[3].fill_(float("nan"))  # eager work between replays
            torch.cuda.synchronize()
            with torch.cuda.stream(s):
                batch(s)
            s.synchronize()
            errs = [smat.rel_fro_error(k[3], r) for k, r in zip(keep, refs)]
            print(f"C5 mode={mode} replay {rep}: max rel Frobenius error {max(errs):.2e}", flush=True)
            assert max(errs) < 1e-5
            snap = [k[3].cpu().numpy() for k in keep[:8]]
            if first is None:
                first = snap
            else:
                assert all(np.array_equal(a, b) for a, b in zip(first, snap))
        batch.close()
