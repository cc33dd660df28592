"""Seeded synthetic sparse matrices for the SMaT workloads.

Every generator returns a host CSR: (M, K, rowptr int32[M+1], colidx int32[nnz],
values uint16[nnz] = FP16 bits), rows padded to a multiple of 16 and columns to
a multiple of 32 (the reference pads rows to 16 and columns to k,
mmio_highlevel.h:514-520; padding to 32 keeps one matrix valid for bk = 16
and bk = 32).  Columns are sorted inside each row.  The reference ships only
unseeded generators (src/generate_matrices.py uses `random`), so these follow
the workload definitions instead:

  band           n, half-bandwidth hb, each in-band entry kept with prob p,
                 values U[-1, 1]
  rmat           scale s, 16 * 2^s edge draws, quadrant probs (a, b, c, d),
                 duplicates merged, values U[-1, 1]
  block_clustered  n, `tiles` dense 16x16 tiles at distinct random block
                 coordinates, every 1x4 group pruned to its 2 largest |v|
  power_law      n, row lengths Zipf(s) truncated at max_len, columns uniform
                 without repetition, values N(0, 1) with 1% scaled by 1e3
"""
from __future__ import annotations

import numpy as np


def pad_to(x: int, m: int) -> int:
    return (x + m - 1) // m * m


def _to_csr(n_rows: int, n_cols: int, rows: np.ndarray, cols: np.ndarray, vals: np.ndarray, f32: bool = False):
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    M, K = pad_to(n_rows, 16), pad_to(n_cols, 32)
    counts = np.bincount(rows, minlength=M)
    rowptr = np.zeros(M + 1, dtype=np.int64)
    np.cumsum(counts, out=rowptr[1:])
    out = (M, K, rowptr.astype(np.int32), cols.astype(np.int32), vals.astype(np.float16).view(np.uint16))
    return out + (vals.astype(np.float32),) if f32 else out


def _uniform_vals(rng: np.random.Generator, n: int) -> np.ndarray:
    v = rng.uniform(-1.0, 1.0, size=n).astype(np.float16)
    # keep explicit zeros out of the pattern (they would change the 2:4 tags)
    v[v == 0] = np.float16(0.5)
    return v


def band(n: int = 4096, hb: int = 64, p: float = 0.5, seed: int = 1001):
    rng = np.random.default_rng(seed)
    i = np.repeat(np.arange(n, dtype=np.int64), 2 * hb + 1)
    off = np.tile(np.arange(-hb, hb + 1, dtype=np.int64), n)
    j = i + off
    ok = (j >= 0) & (j < n)
    i, j = i[ok], j[ok]
    keep = rng.random(i.size) < p
    i, j = i[keep], j[keep]
    return _to_csr(n, n, i, j, _uniform_vals(rng, i.size))


def rmat(scale: int = 18, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
         seed: int = 1002):
    rng = np.random.default_rng(seed)
    n = 1 << scale
    m = edge_factor * n
    rows = np.zeros(m, dtype=np.int64)
    cols = np.zeros(m, dtype=np.int64)
    ab, abc = a + b, a + b + c
    for lvl in range(scale):
        u = rng.random(m)
        bit = 1 << (scale - 1 - lvl)
        right = ((u >= a) & (u < ab)) | (u >= abc)   # quadrants b, d
        down = u >= ab                                # quadrants c, d
        cols += right * bit
        rows += down * bit
    key = np.unique(rows * n + cols)
    rows, cols = key // n, key % n
    return _to_csr(n, n, rows, cols, _uniform_vals(rng, rows.size))


def prune_2of4(dense: np.ndarray) -> np.ndarray:
    """Keeps the two largest-magnitude entries of every aligned 1x4 group
    (ties: lower column first)."""
    g = dense.reshape(*dense.shape[:-1], dense.shape[-1] // 4, 4)
    order = np.argsort(-np.abs(g), axis=-1, kind="stable")
    mask = np.zeros_like(g, dtype=bool)
    np.put_along_axis(mask, order[..., :2], True, axis=-1)
    return np.where(mask, g, 0).reshape(dense.shape)


def block_clustered(n: int = 65536, tiles: int = 12000, seed: int = 1003, prune: bool = True, bh: int = 16,
                    bw: int = 16, fill: float = 1.0):
    """`tiles` distinct aligned bh x bw blocks of an n x n matrix, U[-1, 1]
    values, each entry kept with probability `fill`, optionally pruned to 2
    of every aligned 1x4 group."""
    rng = np.random.default_rng(seed)
    nbr, nbc = n // bh, n // bw
    flat = rng.choice(nbr * nbc, size=tiles, replace=False)
    br, bc = flat // nbc, flat % nbc
    vals = rng.uniform(-1.0, 1.0, size=(tiles, bh, bw)).astype(np.float16)
    vals[vals == 0] = np.float16(0.5)
    if fill < 1.0:
        vals[rng.random(vals.shape) >= fill] = 0
    if prune:
        vals = prune_2of4(vals)
    rows = (br[:, None, None] * bh + np.arange(bh)[None, :, None]) * np.ones((1, 1, bw), dtype=np.int64)
    cols = (bc[:, None, None] * bw + np.arange(bw)[None, None, :]) * np.ones((1, bh, 1), dtype=np.int64)
    keep = vals != 0
    return _to_csr(n, n, rows[keep], cols[keep], vals[keep])


def power_law(n: int = 131072, s: float = 1.5, max_len: int = 4096, seed: int = 1004,
              outlier_frac: float = 0.01, outlier_scale: float = 1e3, f32: bool = False):
    """Power-law rows (Zipf(s) lengths, at most max_len, distinct uniform
    columns), values N(0, 1) with a fraction scaled by outlier_scale.  Returns
    the CSR with FP16 values, and with f32=True also the FP32 values (the
    exact inputs of the precision modes)."""
    rng = np.random.default_rng(seed)
    lens = rng.zipf(s, size=n)
    lens = np.minimum(lens, max_len).astype(np.int64)
    rows = np.repeat(np.arange(n, dtype=np.int64), lens)
    cols = np.empty(rows.size, dtype=np.int64)
    pos = 0
    for i, L in enumerate(lens):
        cols[pos:pos + L] = rng.choice(n, size=L, replace=False) if L > 64 else _distinct_small(rng, n, L)
        pos += L
    vals = rng.standard_normal(rows.size).astype(np.float32)
    out = rng.random(rows.size) < outlier_frac
    vals[out] *= outlier_scale
    vals[vals == 0] = 1.0
    return _to_csr(n, n, rows, cols, vals, f32)


def _distinct_small(rng: np.random.Generator, n: int, L: int) -> np.ndarray:
    out = set()
    while len(out) < L:
        out.update(rng.integers(0, n, size=L - len(out)).tolist())
    return np.fromiter(out, dtype=np.int64, count=L)


def dense_b(K: int, N: int, seed: int = 7) -> np.ndarray:
    """Row-major K x N FP16 matrix, U[-1, 1], as uint16 bits."""
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.0, 1.0, size=(K, N)).astype(np.float16).view(np.uint16)


def f32_to_bf16_bits(x: np.ndarray) -> np.ndarray:
    """FP32 -> BF16 bits, round to nearest even (finite inputs)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return r.astype(np.uint16)


def bf16_bits_to_f32(b: np.ndarray) -> np.ndarray:
    return (np.ascontiguousarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32)


def csr_to_dense(M, K, rowptr, colidx, values_u16) -> np.ndarray:
    d = np.zeros((M, K), dtype=np.float32)
    rows = np.repeat(np.arange(M), np.diff(rowptr))
    d[rows, colidx] = values_u16.view(np.float16).astype(np.float32)
    return d
