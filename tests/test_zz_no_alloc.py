"""GPU tests of smat_spmm_noalloc, the SpMM entry point that allocates
nothing: through the C ABI, its C equals that of smat_spmm_plan_create +
smat_spmm_execute_ex with the same options (same work decomposition, so the
same bits), on calls in a row whose second is smaller than the first (the
static scratch keeps the previous call's contents), and a call that needs
more than the scratch is refused before any SpMM kernel runs."""
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


def _pair(A, N, B, kw, ex, out_dtype=torch.float32):
    """C from plan + execute, and from two no-alloc calls' worth of inputs."""
    import smat
    M = int(A.desc.M)
    C_ref = torch.full((M, N), float("nan"), dtype=out_dtype, device="cuda")
    plan = smat.SpmmPlan(A, N, **kw)
    plan.execute(B, C_ref, **ex)
    C = torch.full((M, N), float("nan"), dtype=out_dtype, device="cuda")
    smat.spmm_noalloc(A, N, B, C, **kw, **ex)
    torch.cuda.synchronize()
    info = plan.info()
    plan.close()
    return C_ref, C, info


CASES = [
    # (name, first matrix, second (smaller) matrix, N first, N second, options)
    ("rmat_split_rows", lambda: gen.rmat(scale=13, edge_factor=16, seed=81),
     lambda: gen.rmat(scale=10, edge_factor=8, seed=82), 256, 64, dict(bk=2)),
    ("rmat_tile_affinity", lambda: gen.rmat(scale=12, edge_factor=12, seed=83),
     lambda: gen.band(n=512, hb=16, p=0.5, seed=84), 256, 128, dict(bk=2, schedule=3, n_tile=64)),
    ("band_block_rows_forced_split", lambda: gen.band(n=2048, hb=64, p=0.5, seed=85),
     lambda: gen.band(n=1024, hb=8, p=0.5, seed=86), 128, 32, dict(bk=16, max_blocks_per_item=4)),
    ("mixed24", lambda: gen.block_clustered(n=2048, tiles=500, seed=87),
     lambda: gen.block_clustered(n=1024, tiles=60, seed=88), 128, 64, dict(bk=16, mixed=True)),
    ("n_tile_128_f16_out", lambda: gen.rmat(scale=12, edge_factor=8, seed=89),
     lambda: gen.rmat(scale=9, edge_factor=8, seed=90), 256, 128, dict(bk=1, n_tile=128)),
    ("reference_layout_f16_out", lambda: gen.band(n=1024, hb=32, p=0.5, seed=91),
     lambda: gen.band(n=256, hb=8, p=0.5, seed=92), 64, 32, dict(bk=16, ref_layout=True)),
]


@pytest.mark.parametrize("name,first,second,N1,N2,opt", CASES)
def test_noalloc_matches_plan_execute(name, first, second, N1, N2, opt):
    _need_gpu()
    import smat
    for call, (make, N) in enumerate(((first, N1), (second, N2))):
        M, K, rp, ci, v = make()
        print(f"noalloc {name} call {call}: M={M} nnz={ci.size} N={N}", flush=True)
        mixed = opt.get("mixed", False)
        A = smat.Bcsr(M, K, _dev(rp), _dev(ci), _dev(v.view(np.int16)), bk=opt["bk"],
                      tag_rule=smat.SMAT_TAG_EXACT24 if mixed else smat.SMAT_TAG_REFERENCE)
        Bh = gen.dense_b(K, N, seed=call + 3)
        kw = {k: opt[k] for k in ("schedule", "n_tile", "max_blocks_per_item") if k in opt}
        ex = {}
        out_dtype = torch.float32
        B = _dev(Bh.view(np.int16))
        if mixed:
            kw["kernel"] = smat.SMAT_KERNEL_MIXED24
            ex["values24"], ex["index24"] = A.compress_24(smat.SMAT_24_SMFMAC)
        if opt.get("ref_layout"):
            kw.update(out_dtype=smat.SMAT_F16, b_layout=smat.SMAT_B_REFERENCE)
            B = _dev(np.ascontiguousarray(Bh.T).view(np.int16))
            ex["ldb"] = K
            out_dtype = torch.float16
        if name == "n_tile_128_f16_out":
            kw["out_dtype"] = smat.SMAT_F16
            out_dtype = torch.float16
        C_ref, C, info = _pair(A, N, B, kw, ex, out_dtype)
        if call == 0 and name in ("rmat_split_rows", "band_block_rows_forced_split"):
            assert info.n_split_rows > 0  # the partial-tile scratch is exercised
        assert torch.equal(C.view(torch.int16 if out_dtype == torch.float16 else torch.int32),
                           C_ref.view(torch.int16 if out_dtype == torch.float16 else torch.int32)), (name, call)
        R = of.spmm_csr_f64(M, rp, ci, v, Bh)
        Cn = C.float().cpu().numpy().astype(np.float64)
        tol = 1e-3 if out_dtype == torch.float16 else 1e-5
        assert np.linalg.norm(Cn - R) / np.linalg.norm(R) < tol, (name, call)
        A.close()


def test_noalloc_row_map_and_fp8():
    """A C row map (reordered rows) and FP8 operands with their scales."""
    _need_gpu()
    import smat
    M, K, rp, ci, v16, v32 = gen.power_law(n=4096, seed=93, max_len=300, f32=True)
    d = (_dev(rp), _dev(ci), _dev(v16.view(np.int16)))
    perm = smat.locality_order(M, K, d[0], d[1])
    o = smat.permute_rows(*d, perm)
    rowmap = perm.to(torch.int32)
    N = 128
    Bh = gen.dense_b(K, N, seed=4)
    A = smat.Bcsr(M, K, *o, bk=1)
    C_ref, C, _ = _pair(A, N, _dev(Bh.view(np.int16)), {}, {"c_rowmap": rowmap})
    assert torch.equal(C.view(torch.int32), C_ref.view(torch.int32))
    A.close()
    A = smat.Bcsr(M, K, *d, bk=2)
    x = smat.scatter_values(A, d[0], d[1], _dev(v32))
    a, sa = smat.pack_a(A, x, smat.SMAT_FP8_E4M3)
    b, sb = smat.pack_b(_dev(np.random.default_rng(1).uniform(-1, 1, (K, N)).astype(np.float32)),
                        smat.SMAT_FP8_E4M3)
    C_ref, C, _ = _pair(A, N, b, {"in_dtype": smat.SMAT_FP8_E4M3},
                        {"a_values": a, "a_scale": sa, "b_scale": sb})
    assert torch.equal(C.view(torch.int32), C_ref.view(torch.int32))
    A.close()


def test_noalloc_refuses_oversized_call_then_works():
    """A call whose partial tiles would exceed the static scratch is refused
    with a clear error; the next call on the same GPU still gives the right C."""
    _need_gpu()
    import smat
    M, K, rp, ci, v = gen.rmat(scale=12, edge_factor=16, seed=94)
    A = smat.Bcsr(M, K, _dev(rp), _dev(ci), _dev(v.view(np.int16)), bk=16)
    N = 2048  # 16 x 2048 FP32 partial tiles of 128 KiB each, one per slice of 2 blocks
    plan = smat.SpmmPlan(A, N, max_blocks_per_item=2)
    need = plan.info().scratch_bytes
    plan.close()
    assert need > smat.SMAT_NOALLOC_SCRATCH_BYTES
    B = torch.zeros((K, N), dtype=torch.int16, device="cuda")
    C = torch.zeros((M, N), dtype=torch.float32, device="cuda")
    with pytest.raises(smat.SmatError) as e:
        smat.spmm_noalloc(A, N, B, C, max_blocks_per_item=2)
    assert "static device scratch" in str(e.value)
    torch.cuda.synchronize()
    N = 64
    Bh = gen.dense_b(K, N, seed=5)
    C_ref, C2, _ = _pair(A, N, _dev(Bh.view(np.int16)), {"max_blocks_per_item": 8}, {})
    assert torch.equal(C2.view(torch.int32), C_ref.view(torch.int32))
    A.close()


def test_noalloc_symbol_in_header():
    import smat
    assert "smat_spmm_noalloc" in smat.SIGNATURES
    assert isinstance(smat.lib().smat_spmm_noalloc, ctypes._CFuncPtr)
