"""The reference-named C++ API (include/smat_reference.hpp: SparseMatrix,
CSR, BlockingEngine, get_permutation, get_fixed_size_grouping, the half*
launchers) driven from C++ (tests/cpp/reference_api.cpp, built by
__graft_entry__.build() with the host compiler only) the way the reference's
code calls it; what it writes is checked here:
  * SparseMatrix(name, file.mtx, 16): the BCSR, the grid maps and the
    filtered 2:4 / dense lists against the oracle's csrToBcsr /
    filterBcsrBlocks restatement on the same padded CSR;
  * mmaOBTKernel on its arrays: the FP64 oracle product (FP16 C);
  * bcsrBlocking(): the merged tiles, tile types and tile-grid maps against
    the oracle's restatement (or_bcsr_to_tiles), and mmaOBTKernel_tiled /
    preprocessing_mmaOBTSKernel_tiled_large + mmaOBTSKernel_tiled_large
    (SparseMatrix with k = 32) on them against the FP64 product;
  * CSR(ifstream) + BlockingEngine(Keeper) + get_fixed_size_grouping +
    CollectBlockingInfo + CSR::reorder on the reference's cs_department.el
    fixture: the compiled reference's groupings, counters, blocking info
    and permutation (tests/golden/keeper_golden.npz)."""
import os
import subprocess

import numpy as np
import pytest

import oracle_ffi as of

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "reference_api")


def _read(d, name, dt):
    return np.fromfile(os.path.join(d, name), dtype=dt)


@pytest.mark.parametrize("tau,cb,rb", [(0.5, 8, 16), (0.3, 16, 16), (0.7, 3, 3)])
def test_reference_api_from_cpp(tmp_path, golden_dir, tau, cb, rb):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    assert os.path.exists(BIN), "tests/cpp/reference_api is not built (run __graft_entry__.build())"
    import smat
    rng = np.random.default_rng(int(tau * 10) + cb)
    n, m = 200, 150
    # one full 16 x 16 tile written first, row by row in column order (a dense block by the
    # reference's tag rule, which looks at consecutive entries in file order), then scattered
    # entries in random order (2:4 blocks)
    tile = (np.arange(16)[:, None] * m + np.arange(32, 48)[None, :]).ravel()
    rest = np.setdiff1d(rng.choice(n * m, size=1800, replace=False), tile)
    rng.shuffle(rest)
    cells = np.concatenate([tile, rest])
    r, c = cells // m, cells % m
    lines = "".join(f"{i + 1} {j + 1} {rng.uniform(-1, 1):.4f}\n" for i, j in zip(r, c))
    mtx = tmp_path / "a.mtx"
    mtx.write_text(f"%%MatrixMarket matrix coordinate real general\n{n} {m} {cells.size}\n" + lines)
    el = os.path.join(golden_dir, "cs_department.el")
    out = tmp_path / "out"
    out.mkdir()
    print(f"reference_api tau={tau} b={cb} B={rb}", flush=True)
    res = subprocess.run([BIN, str(mtx), el, str(tau), str(cb), str(rb), str(out)], capture_output=True, text=True,
                         timeout=120)
    print(res.stdout, res.stderr, flush=True)
    assert res.returncode == 0, res.stderr

    # ---- SparseMatrix against the oracle on the same padded CSR
    h = smat.read_mtx(str(mtx), k=16)
    M, K = h["rows"], h["cols"]
    brp, bci, bv, tags = of.csr_to_bcsr(M, K, 16, h["rowptr"], h["colidx"], h["values_f16"])
    assert np.array_equal(_read(out, "bcsr_rowptr.bin", np.int32), brp)
    assert np.array_equal(_read(out, "bcsr_colidx.bin", np.int32), bci)
    assert np.array_equal(_read(out, "bcsr_values.bin", np.uint16), bv)
    grid = (M // 16) * (K // 16)
    info = _read(out, "block_info.bin", np.int32)
    rel = _read(out, "rel_map.bin", np.int32)
    assert info.size == rel.size == grid
    exp_info = np.zeros(grid, np.int32)
    exp_rel = np.full(grid, -1, np.int32)
    brow = np.repeat(np.arange(M // 16), np.diff(brp))
    gidx = brow * (K // 16) + bci // 16
    exp_info[gidx] = tags
    exp_rel[gidx] = np.arange(bci.size)
    assert np.array_equal(info, exp_info) and np.array_equal(rel, exp_rel)
    (srp, sci, _), (drp, dci, _) = of.filter_blocks(16, brp, bci, bv, tags)
    assert np.array_equal(_read(out, "sparse_rowptr.bin", np.int32), srp)
    assert np.array_equal(_read(out, "sparse_colidx.bin", np.int32), sci)
    assert np.array_equal(_read(out, "dense_rowptr.bin", np.int32), drp)
    assert np.array_equal(_read(out, "dense_colidx.bin", np.int32), dci)
    srel = _read(out, "sparse_rel_map.bin", np.int32)
    exp_srel = np.full(grid, -1, np.int32)
    exp_srel[exp_info == 1] = np.arange(int((exp_info == 1).sum()))
    assert np.array_equal(srel, exp_srel)
    assert sci.size > 0 and dci.size > 0  # both lists exercised

    # ---- mmaOBTKernel: B[n*K + k] FP16, C row-major FP16
    N = 64
    B = _read(out, "B.bin", np.uint16).reshape(N, K).T.copy()  # K x N
    C = _read(out, "C.bin", np.uint16).view(np.float16).astype(np.float64).reshape(M, N)
    R = of.spmm_bcsr_f64(16, brp, bci, bv, B)
    assert np.linalg.norm(C - R) / np.linalg.norm(R) < 1e-3

    # ---- bcsrBlocking + the tiled launchers
    trp, tci, tv, tinfo = of.bcsr_to_tiles(M, K, 16, brp, bci, bv)
    assert np.array_equal(_read(out, "tiles_rowptr.bin", np.int32), trp)
    assert np.array_equal(_read(out, "tiles_colidx.bin", np.int32), tci)
    assert np.array_equal(_read(out, "tiles_values.bin", np.uint16), tv)
    assert np.array_equal(_read(out, "tiles_info.bin", np.int32), tinfo)
    ntr, ntc = (M + 63) // 64, (K + 63) // 64
    exp_tinfo = np.zeros(ntr * ntc, np.int32)
    exp_trel = np.full(ntr * ntc, -1, np.int32)
    cells = np.repeat(np.arange(ntr), np.diff(trp)) * ntc + tci // 64
    exp_tinfo[cells], exp_trel[cells] = 1, np.arange(tci.size)
    assert np.array_equal(_read(out, "tiles_block_info.bin", np.int32), exp_tinfo)
    assert np.array_equal(_read(out, "tiles_rel_map.bin", np.int32), exp_trel)
    Ct = _read(out, "C_tiled.bin", np.uint16).view(np.float16).astype(np.float64).reshape(M, N)
    assert np.linalg.norm(Ct - R) / np.linalg.norm(R) < 1e-3
    h32 = smat.read_mtx(str(mtx), k=32)
    M32, K32 = h32["rows"], h32["cols"]
    B32 = _read(out, "B32.bin", np.uint16).reshape(N, K32).T.copy()
    R32 = of.spmm_csr_f64(M32, h32["rowptr"], h32["colidx"], h32["values_f16"], B32)
    C32 = _read(out, "C_tiled_large.bin", np.uint16).view(np.float16).astype(np.float64).reshape(M32, N)
    assert np.linalg.norm(C32 - R32) / np.linalg.norm(R32) < 1e-3

    # ---- CSR + BlockingEngine + reorder against the compiled reference's goldens
    g = np.load(os.path.join(golden_dir, "keeper_golden.npz"))
    key = f"cs_department__t{tau}_b{cb}_B{rb}"
    cnt = _read(out, "counters.bin", np.int64)
    assert np.array_equal(_read(out, "grouping.bin", np.int64), g[key + "__grouping"])
    assert np.array_equal(cnt[:2], g[key + "__counters"])
    assert np.array_equal(_read(out, "fixed.bin", np.int64), g[key + "__fixed"])
    assert np.array_equal(cnt[2:], g[key + "__info"])
    perm = _read(out, "perm.bin", np.int64)
    assert np.array_equal(perm, g[key + "__perm"])
    rp0 = g["cs_department__rowptr"]
    ci0 = g["cs_department__colidx"]
    rrp = _read(out, "reordered_rowptr.bin", np.int64)
    rci = _read(out, "reordered_colidx.bin", np.int64)
    exp = np.concatenate([ci0[rp0[p]:rp0[p + 1]] for p in perm])
    assert np.array_equal(np.diff(rrp), np.diff(rp0)[perm]) and np.array_equal(rci, exp)
