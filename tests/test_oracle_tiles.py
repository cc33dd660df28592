"""Pins of the oracle's merged-tile restatement (or_bcsr_to_tiles,
or_spmm_tiles_f64; SparseMatrix::bcsrBlocking, matrix.h:794-1137) by checks
that do not repeat its loops: the tiles expand to the same dense matrix as
the CSR they come from (a transposed slot, a wrong tile column or a lost
block fails this), the tile set is the set of distinct (row/64, col/4bk)
cells of the stored entries, tiles are sorted inside every tile row, each
slot's type agrees with a classification of the dense matrix, and hand cases
fix -0, a 3-of-4 group and an explicitly stored zero block.  The tile SpMM
is checked against a dense GEMM."""
import numpy as np
import pytest

import oracle_ffi as of
from smat import gen


def _expand_tiles(M, K, bk, rp, ci, tv):
    """Dense FP32 matrix of a tile set, from the layout definition: slot
    sb = 4*sr + sc of tile t holds rows 16*sr.. and columns bk*sc.. of the
    tile, 16 x bk row-major at (16*t + sb)*16*bk."""
    tw = 4 * bk
    d = np.zeros((pad64(M), K + tw), np.float32)
    v = tv.view(np.float16).astype(np.float32).reshape(-1, 4, 4, 16, bk)  # tile, sr, sc, r, c
    for tr in range(rp.size - 1):
        for t in range(rp[tr], rp[tr + 1]):
            blk = v[t].transpose(0, 2, 1, 3).reshape(64, tw)  # (sr, r) x (sc, c)
            d[64 * tr:64 * tr + 64, ci[t]:ci[t] + tw] = blk
    assert not d[M:].any() and not d[:, K:].any()
    return d[:M, :K]


def pad64(M):
    return (M + 63) // 64 * 64


def _classify(d_u16, M, K, bk, tr, col):
    """Slot types of the tile at (tile row tr, first column col) from the
    dense FP16 bit matrix."""
    tw = 4 * bk
    sub = np.zeros((64, tw), np.uint16)
    rows = d_u16[64 * tr:min(64 * tr + 64, M), col:min(col + tw, K)]
    sub[:rows.shape[0], :rows.shape[1]] = rows
    nz = (sub & 0x7FFF) != 0
    out = []
    for sr in range(4):
        for sc in range(4):
            g = nz[16 * sr:16 * sr + 16, bk * sc:bk * sc + bk].reshape(16, bk // 4, 4).sum(-1)
            out.append(0 if g.max() == 0 else (1 if g.max() <= 2 else 2))
    return out


def _dense_bits(M, K, rp, ci, v):
    d = np.zeros((M, K), np.uint16)
    rows = np.repeat(np.arange(M), np.diff(rp))
    d[rows, ci] = v
    return d


CASES = [
    ("band", lambda: gen.band(n=400, hb=20, p=0.5, seed=3)),
    ("clustered", lambda: gen.block_clustered(n=512, tiles=120, seed=5)),
    ("rmat", lambda: gen.rmat(scale=9, edge_factor=8, seed=7)),
]


@pytest.mark.parametrize("bk", [16, 32])
@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_tiles_expand_to_the_csr_matrix(name, make, bk):
    M, K, rp, ci, v = make()
    K = gen.pad_to(K, bk)
    brp, bci, bv, _ = of.csr_to_bcsr(M, K, bk, rp, ci, v)
    trp, tci, tv, tinfo = of.bcsr_to_tiles(M, K, bk, brp, bci, bv)
    dense = gen.csr_to_dense(M, K, rp, ci, v)
    assert np.array_equal(_expand_tiles(M, K, bk, trp, tci, tv), dense)
    # tile set = distinct (row/64, col/4bk) cells of the stored entries, sorted per tile row
    rows = np.repeat(np.arange(M), np.diff(rp))
    cells = np.unique((rows // 64) * (K // (4 * bk) + 1) + ci // (4 * bk))
    tr_of = np.repeat(np.arange(trp.size - 1), np.diff(trp))
    got = tr_of * (K // (4 * bk) + 1) + tci // (4 * bk)
    assert np.array_equal(got, cells)
    assert trp.size == (M + 63) // 64 + 1 and trp[0] == 0 and trp[-1] == tci.size
    # slot types from the dense matrix
    d16 = _dense_bits(M, K, rp, ci, v)
    exp = np.concatenate([_classify(d16, M, K, bk, tr_of[t], tci[t]) for t in range(tci.size)])
    assert np.array_equal(tinfo, exp)
    assert set(np.unique(tinfo)) <= {0, 1, 2}


def test_tiles_partial_last_tile_row():
    # 80 rows: the second tile row has one 16-row band
    M, K = 80, 64
    rp = np.zeros(M + 1, np.int32)
    rows = np.array([0, 17, 70, 79])
    cols = np.array([3, 60, 5, 33])
    for r in rows:
        rp[r + 1:] += 1
    v = np.float16([1, 2, 3, 4]).view(np.uint16)
    brp, bci, bv, _ = of.csr_to_bcsr(M, K, 16, rp, cols.astype(np.int32), v)
    trp, tci, tv, tinfo = of.bcsr_to_tiles(M, K, 16, brp, bci, bv)
    assert trp.tolist() == [0, 1, 2] and tci.tolist() == [0, 0]
    assert np.array_equal(_expand_tiles(M, K, 16, trp, tci, tv), gen.csr_to_dense(M, K, rp, cols, v))
    # slots: tile 0 rows 0 (col 3 -> sc 0) and 17 (col 60 -> sc 3); tile 1 row 70 (sr 0), row 79 (sr 0, sc 2)
    nz0 = [sb for sb in range(16) if tinfo[sb]]
    nz1 = [sb for sb in range(16) if tinfo[16 + sb]]
    assert nz0 == [0, 7] and nz1 == [0, 2]


def test_tiles_slot_type_hand_cases():
    M, K, bk = 64, 64, 16
    h = lambda x: np.float16(x).view(np.uint16)  # noqa: E731
    neg0 = np.uint16(0x8000)
    # slot 0: row 0 holds 1, -0, 2 in group 0 -> two nonzeros: 2:4
    # slot 1: row 5 holds 1, 2, 3 in columns 16..18 -> three: dense
    # slot 2: an explicitly stored 0 at (0, 32): the block is stored, its type is empty
    entries = [(0, 0, h(1)), (0, 1, neg0), (0, 2, h(2)), (5, 16, h(1)), (5, 17, h(2)), (5, 18, h(3)),
               (0, 32, h(0))]
    entries.sort()
    rp = np.zeros(M + 1, np.int32)
    for r, _, _ in entries:
        rp[r + 1:] += 1
    ci = np.array([c for _, c, _ in entries], np.int32)
    v = np.array([x for _, _, x in entries], np.uint16)
    brp, bci, bv, _ = of.csr_to_bcsr(M, K, bk, rp, ci, v)
    assert bci.tolist() == [0, 16, 32]
    trp, tci, tv, tinfo = of.bcsr_to_tiles(M, K, bk, brp, bci, bv)
    assert tci.tolist() == [0]
    assert tinfo[:4].tolist() == [1, 2, 0, 0] and not tinfo[4:].any()


@pytest.mark.parametrize("bk", [16, 32])
def test_tile_spmm_f64_matches_dense_gemm(bk):
    M, K, rp, ci, v = gen.band(n=300, hb=24, p=0.4, seed=11)
    K = gen.pad_to(K, bk)
    brp, bci, bv, _ = of.csr_to_bcsr(M, K, bk, rp, ci, v)
    trp, tci, tv, _ = of.bcsr_to_tiles(M, K, bk, brp, bci, bv)
    B = gen.dense_b(K, 40, seed=2)
    C = of.spmm_tiles_f64(M, K, bk, trp, tci, tv, B)
    R = gen.csr_to_dense(M, K, rp, ci, v).astype(np.float64) @ B.view(np.float16).astype(np.float64)
    assert np.allclose(C, R, rtol=0, atol=1e-12)
    assert np.allclose(C, of.spmm_csr_f64(M, rp, ci, v, B), rtol=0, atol=1e-12)
