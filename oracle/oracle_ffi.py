"""ctypes binding of the CPU oracle (liboracle.so) and, where it was built,
of the reference's own preprocessing code (_ref/libsmat_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py.  The product never imports this module.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libsmat_ref.so")

_p = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_f32 = ctypes.c_float

_SIG = {
    "or_h2f": (_f32, [ctypes.c_uint16]),
    "or_f2h": (ctypes.c_uint16, [_f32]),
    "or_csr_to_bcsr": (_i64, [_i64, _i64, ctypes.c_int, _p, _p, _p, _p, _p, _p, _p, _i64]),
    "or_tag_exact24": (None, [_i64, ctypes.c_int, _p, _p]),
    "or_filter_blocks": (None, [_i64, ctypes.c_int, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
    "or_compress24_k16": (_i64, [_i64, _p, _p, _p, _p]),
    "or_decompress24_k16": (None, [_i64, _p, _p, _p]),
    "or_spmm_bcsr_f64": (None, [_i64, ctypes.c_int, _p, _p, _p, _i64, _p, _i64, _p, _i64]),
    "or_spmm_csr_f64": (None, [_i64, _p, _p, _p, _i64, _p, _i64, _p, _i64]),
    "or_bcsr_to_tiles": (_i64, [_i64, _i64, ctypes.c_int, _p, _p, _p, _p, _p, _p, _p, _i64]),
    "or_spmm_tiles_f64": (None, [_i64, _i64, ctypes.c_int, _p, _p, _p, _i64, _p, _i64, _p, _i64]),
    "or_spmm_refsim_fp16acc": (None, [_i64, ctypes.c_int, _p, _p, _p, _i64, _p, _i64, _p, _i64]),
    "or_csr_multiply_f32": (None, [_i64, _p, _p, _p, _p, _i64, _p]),
    "or_csr_multiply_f32_omp": (ctypes.c_int32, [_i64, _p, _p, _p, _p, _i64, _p, ctypes.c_int32]),
    "or_jaccard_distance_group": (_f32, [_p, _i64, _i64, _p, _i64, _i64, _i64]),
    "or_merge_rows": (_i64, [_p, _i64, _p, _i64, _p]),
    "or_blocking_algo0": (None, [_i64, _p, _p, _f32, _i64, ctypes.c_int, _p, _p]),
    "or_keeper": (None, [_i64, _p, _p, _f32, _i64, _i64, _p, _p]),
    "or_get_permutation": (None, [_i64, _p, _p]),
    "or_fixed_size_grouping": (None, [_i64, _p, _i64, _p]),
    "or_collect_blocking_info": (None, [_i64, _i64, _p, _p, _p, _i64, _p, ctypes.POINTER(_f32)]),
    "or_e4m3_decode": (_f32, [ctypes.c_uint8]),
    "or_e4m3_encode": (ctypes.c_uint8, [_f32]),
    "or_f2bf": (ctypes.c_uint16, [_f32]),
    "or_bf2f": (_f32, [ctypes.c_uint16]),
    "or_scale_of": (_f32, [_f32, ctypes.c_int]),
    "or_pack_a": (None, [_i64, ctypes.c_int, _p, _p, ctypes.c_int, _p, _p]),
    "or_pack_b": (None, [_i64, _i64, _i64, _p, ctypes.c_int, _p, _p]),
    "or_spmm_bcsr_scaled_f64": (None, [_i64, ctypes.c_int, _p, _p, ctypes.c_int, _p, _p, _i64, _i64, _p, _p,
                                       _p, _i64]),
    "or_matrix_stats": (None, [_i64, _i64, _p, _p, _p, _p]),
    "or_mlp_forward": (None, [_i64, ctypes.c_int, ctypes.c_int, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
    "or_mlp_backward": (ctypes.c_double, [_i64, ctypes.c_int, ctypes.c_int, _p, _p, _p, _p, _p, _p, _p, _p, _p,
                                          _p, _p, _p, _p, _p]),
    "or_adam_step": (None, [_p, _p, _p, _p, _i64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                            ctypes.c_double, ctypes.c_int]),
}

_REF_SIG = {
    "ref_jaccard_distance_group": (_f32, [_p, _i64, _i64, _p, _i64, _i64, _i64]),
    "ref_merge_rows": (_i64, [_p, _i64, _p, _i64, _p]),
    "ref_keeper": (ctypes.c_double, [_i64, _i64, _p, _p, _f32, _i64, _i64, _p, _p]),
    "ref_algo0": (None, [_i64, _i64, _p, _p, _f32, _i64, ctypes.c_int, _p, _p]),
    "ref_get_permutation": (None, [_i64, _p, _p]),
    "ref_fixed_size_grouping": (None, [_i64, _p, _i64, _p]),
    "ref_collect_blocking_info": (None, [_i64, _i64, _p, _p, _p, _i64, _p, ctypes.POINTER(_f32)]),
    "ref_csr_multiply": (ctypes.c_double, [_i64, _i64, _p, _p, _p, _p, _i64, _p]),
    "ref_reorder_rows": (None, [_i64, _i64, _p, _p, _p, _p, _p]),
}

_lib = None
_ref = None


def _load(path, sig):
    lib = ctypes.CDLL(path)
    for k, (r, a) in sig.items():
        f = getattr(lib, k)
        f.restype = r
        f.argtypes = a
    return lib


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _lib = _load(ORACLE_SO, _SIG)
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        _ref = _load(REF_SO, _REF_SIG)
    return _ref


def P(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"], "arrays passed to the oracle must be contiguous"
    return a.ctypes.data


# ----------------------------------------------------------------- wrappers
def csr_to_bcsr(M, K, bk, rowptr, colidx, values):
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    colidx = np.ascontiguousarray(colidx, np.int32)
    values = np.ascontiguousarray(values, np.uint16)
    cap = max(int(colidx.size), 1)
    o_rp = np.zeros(M // 16 + 1, np.int32)
    o_ci = np.zeros(cap, np.int32)
    o_v = np.zeros(cap * 16 * bk, np.uint16)
    o_t = np.zeros(cap, np.int32)
    nnzb = lib().or_csr_to_bcsr(M, K, bk, P(rowptr), P(colidx), P(values), P(o_rp), P(o_ci), P(o_v),
                                P(o_t), cap)
    assert nnzb >= 0
    return o_rp, o_ci[:nnzb].copy(), o_v[: nnzb * 16 * bk].copy(), o_t[:nnzb].copy()


def tag_exact24(bk, values):
    values = np.ascontiguousarray(values, np.uint16)
    nnzb = values.size // (16 * bk)
    t = np.zeros(nnzb, np.int32)
    lib().or_tag_exact24(nnzb, bk, P(values), P(t))
    return t


def filter_blocks(bk, rowptr, colidx, values, tags):
    nbrows = rowptr.size - 1
    ns, nd = int((tags == 1).sum()), int((tags == 2).sum())
    s_rp, d_rp = np.zeros(nbrows + 1, np.int32), np.zeros(nbrows + 1, np.int32)
    s_ci, d_ci = np.zeros(max(ns, 1), np.int32), np.zeros(max(nd, 1), np.int32)
    s_v, d_v = np.zeros(max(ns, 1) * 16 * bk, np.uint16), np.zeros(max(nd, 1) * 16 * bk, np.uint16)
    args = [np.ascontiguousarray(x) for x in (rowptr, colidx, values, tags)]
    lib().or_filter_blocks(nbrows, bk, *(P(a) for a in args), P(s_rp), P(s_ci), P(s_v), P(d_rp), P(d_ci),
                           P(d_v))
    return (s_rp, s_ci[:ns], s_v[: ns * 256 * bk // 16]), (d_rp, d_ci[:nd], d_v[: nd * 256 * bk // 16])


def compress24_k16(values, tags):
    nnzb = tags.size
    sp = np.zeros(max(nnzb, 1) * 128, np.uint16)
    me = np.zeros(max(nnzb, 1) * 32, np.uint8)
    ov = lib().or_compress24_k16(nnzb, P(np.ascontiguousarray(values, np.uint16)),
                                 P(np.ascontiguousarray(tags, np.int32)), P(sp), P(me))
    return sp[: nnzb * 128], me[: nnzb * 32], int(ov)


def decompress24_k16(sparse, meta):
    nnzb = meta.size // 32
    v = np.zeros(max(nnzb, 1) * 256, np.uint16)
    lib().or_decompress24_k16(nnzb, P(sparse), P(meta), P(v))
    return v[: nnzb * 256]


def spmm_bcsr_f64(bk, rowptr, colidx, values, B_u16):
    nbrows = rowptr.size - 1
    K, N = B_u16.shape
    C = np.zeros((nbrows * 16, N), np.float64)
    B_u16 = np.ascontiguousarray(B_u16, np.uint16)
    lib().or_spmm_bcsr_f64(nbrows, bk, P(np.ascontiguousarray(rowptr, np.int32)),
                           P(np.ascontiguousarray(colidx, np.int32)), P(np.ascontiguousarray(values, np.uint16)),
                           N, P(B_u16), N, P(C), N)
    return C


def bcsr_to_tiles(M, K, bk, rowptr, colidx, values):
    """Merged 64 x 4bk tiles of a BCSR: (rowptr, colidx, values[u16], info)."""
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    colidx = np.ascontiguousarray(colidx, np.int32)
    values = np.ascontiguousarray(values, np.uint16)
    cap = max(int(colidx.size), 1)
    ntr = (M + 63) // 64
    o_rp = np.zeros(ntr + 1, np.int32)
    o_ci = np.zeros(cap, np.int32)
    o_v = np.zeros(cap * 256 * bk, np.uint16)
    o_i = np.zeros(cap * 16, np.int32)
    nt = lib().or_bcsr_to_tiles(M, K, bk, P(rowptr), P(colidx), P(values), P(o_rp), P(o_ci), P(o_v), P(o_i), cap)
    assert nt >= 0
    return o_rp, o_ci[:nt].copy(), o_v[: nt * 256 * bk].copy(), o_i[: nt * 16].copy()


def spmm_tiles_f64(M, K, bk, rowptr, colidx, values, B_u16):
    Kb, N = B_u16.shape
    C = np.zeros((M, N), np.float64)
    B_u16 = np.ascontiguousarray(B_u16, np.uint16)
    lib().or_spmm_tiles_f64(M, K, bk, P(np.ascontiguousarray(rowptr, np.int32)),
                            P(np.ascontiguousarray(colidx, np.int32)), P(np.ascontiguousarray(values, np.uint16)),
                            N, P(B_u16), N, P(C), N)
    return C


def spmm_csr_f64(M, rowptr, colidx, values, B_u16):
    K, N = B_u16.shape
    C = np.zeros((M, N), np.float64)
    B_u16 = np.ascontiguousarray(B_u16, np.uint16)
    lib().or_spmm_csr_f64(M, P(np.ascontiguousarray(rowptr, np.int32)), P(np.ascontiguousarray(colidx, np.int32)),
                          P(np.ascontiguousarray(values, np.uint16)), N, P(B_u16), N, P(C), N)
    return C


def spmm_refsim_fp16acc(bk, rowptr, colidx, values, B_u16):
    nbrows = rowptr.size - 1
    K, N = B_u16.shape
    C = np.zeros((nbrows * 16, N), np.uint16)
    B_u16 = np.ascontiguousarray(B_u16, np.uint16)
    lib().or_spmm_refsim_fp16acc(nbrows, bk, P(np.ascontiguousarray(rowptr, np.int32)),
                                 P(np.ascontiguousarray(colidx, np.int32)),
                                 P(np.ascontiguousarray(values, np.uint16)), N, P(B_u16), N, P(C), N)
    return C.view(np.float16)


def csr_multiply_f32(rows, rowptr64, colidx64, vals32, B_colmajor, B_cols):
    C = np.zeros(rows * B_cols, np.float32)
    lib().or_csr_multiply_f32(rows, P(rowptr64), P(colidx64), P(vals32), P(B_colmajor), B_cols, P(C))
    return C


def csr_multiply_f32_omp(rows, rowptr64, colidx64, vals32, B_colmajor, B_cols, threads):
    """Row-parallel (OpenMP) CSR::multiply; returns (C, threads used)."""
    C = np.zeros(rows * B_cols, np.float32)
    used = lib().or_csr_multiply_f32_omp(rows, P(rowptr64), P(colidx64), P(vals32), P(B_colmajor), B_cols, P(C),
                                         threads)
    return C, int(used)


def jaccard(a, ga, b, gb, block, use_ref=False):
    a = np.ascontiguousarray(a, np.int64)
    b = np.ascontiguousarray(b, np.int64)
    f = ref().ref_jaccard_distance_group if use_ref else lib().or_jaccard_distance_group
    return float(np.float32(f(P(a) if a.size else None, a.size, ga, P(b) if b.size else None, b.size, gb, block)))


def merge_rows(a, b, use_ref=False):
    a = np.ascontiguousarray(a, np.int64)
    b = np.ascontiguousarray(b, np.int64)
    out = np.zeros(a.size + b.size + 1, np.int64)
    f = ref().ref_merge_rows if use_ref else lib().or_merge_rows
    n = f(P(a) if a.size else None, a.size, P(b) if b.size else None, b.size, P(out))
    return out[:n].copy()


def keeper(rowptr64, colidx64, tau, col_block, max_rows, cols=None, use_ref=False):
    rows = rowptr64.size - 1
    rowptr64 = np.ascontiguousarray(rowptr64, np.int64)
    colidx64 = np.ascontiguousarray(colidx64, np.int64) if colidx64.size else np.zeros(1, np.int64)
    g = np.zeros(rows, np.int64)
    cnt = np.zeros(2, np.int64)
    if use_ref:
        cols = cols if cols is not None else int(colidx64.max()) + 1
        t = ref().ref_keeper(rows, cols, P(rowptr64), P(colidx64), tau, col_block, max_rows, P(g), P(cnt))
        return g, cnt, t
    lib().or_keeper(rows, P(rowptr64), P(colidx64), tau, col_block, max_rows, P(g), P(cnt))
    return g, cnt


def algo0(rowptr64, colidx64, tau, block, use_size=0, cols=None, use_ref=False):
    rows = rowptr64.size - 1
    g = np.zeros(rows, np.int64)
    cnt = np.zeros(2, np.int64)
    rowptr64 = np.ascontiguousarray(rowptr64, np.int64)
    colidx64 = np.ascontiguousarray(colidx64, np.int64)
    if use_ref:
        cols = cols if cols is not None else int(colidx64.max()) + 1
        ref().ref_algo0(rows, cols, P(rowptr64), P(colidx64), tau, block, use_size, P(g), P(cnt))
    else:
        lib().or_blocking_algo0(rows, P(rowptr64), P(colidx64), tau, block, use_size, P(g), P(cnt))
    return g, cnt


def get_permutation(grouping, use_ref=False):
    g = np.ascontiguousarray(grouping, np.int64)
    out = np.zeros(g.size, np.int64)
    (ref().ref_get_permutation if use_ref else lib().or_get_permutation)(g.size, P(g), P(out))
    return out


def fixed_size_grouping(grouping, row_block, use_ref=False):
    g = np.ascontiguousarray(grouping, np.int64)
    out = np.zeros(g.size, np.int64)
    (ref().ref_fixed_size_grouping if use_ref else lib().or_fixed_size_grouping)(g.size, P(g), row_block, P(out))
    return out


def collect_blocking_info(rowptr64, colidx64, grouping, cols, col_block, use_ref=False):
    rows = rowptr64.size - 1
    out = np.zeros(3, np.int64)
    avg = _f32(0)
    args = [np.ascontiguousarray(x, np.int64) for x in (rowptr64, colidx64, grouping)]
    f = ref().ref_collect_blocking_info if use_ref else lib().or_collect_blocking_info
    f(rows, cols, P(args[0]), P(args[1]), P(args[2]), col_block, P(out), ctypes.byref(avg))
    return out, float(avg.value)


def locality_order(rows, cols, rowptr, colidx):
    """Locality row order (no reference counterpart; the library's stand-in
    for the Keeper on matrices too large for it): rows by descending
    floor(log2 length), then first and second stored column (`cols` when
    missing; the second dropped when cols >= 2^29), ties in row order.
    Returns perm (new row r = old row perm[r])."""
    rp = np.asarray(rowptr, np.int64)
    ci = np.asarray(colidx, np.int64)
    lens = np.diff(rp)
    lb = np.where(lens > 1, np.floor(np.log2(np.maximum(lens, 1))), 0).astype(np.int64)

    def kth(k):
        out = np.full(rows, cols, np.int64)
        ok = lens > k
        out[ok] = ci[rp[:-1][ok] + k]
        return out

    if cols < (1 << 29):
        return np.lexsort((kth(1), kth(0), -lb))
    return np.lexsort((kth(0), -lb))


def read_edgelist(text: str, weighted: bool = True):
    """Edge-list reader with CSR::read_from_edgelist conventions
    (preprocess/src/general/csr.cpp:166-290, el format): rows must ascend,
    columns kept in file order, rows = max row + 1, cols = max col + 1."""
    rows, cols, vals = [], [], []
    for line in text.splitlines():
        if not line.strip() or line[0] in "#%":
            continue
        parts = line.split(" ")
        rows.append(int(parts[0]))
        cols.append(int(parts[1]))
        vals.append(float(parts[2]) if weighted and len(parts) > 2 else 1.0)
    r = np.array(rows, np.int64)
    assert np.all(np.diff(r) >= 0), "rows must be in ascending order"
    nrows = int(r.max()) + 1
    rowptr = np.zeros(nrows + 1, np.int64)
    np.add.at(rowptr, r + 1, 1)
    rowptr = np.cumsum(rowptr)
    return nrows, int(max(cols)) + 1, rowptr, np.array(cols, np.int64), np.array(vals, np.float32)


# ---- precision modes ---------------------------------------------------------
F16, BF16, FP8, SPLIT = 1, 2, 4, 5


def pack_a(bk, rowptr, values_f32, dt):
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    x = np.ascontiguousarray(values_f32, np.float32)
    nbrows = rowptr.size - 1
    n = int(rowptr[-1]) * 16 * bk
    out = np.zeros(max(2 * n if dt == SPLIT else n, 1), np.uint8 if dt == FP8 else np.uint16)
    scale = np.ones(max(nbrows, 1), np.float32)
    lib().or_pack_a(nbrows, bk, P(rowptr), P(x) if x.size else None, dt, P(out), P(scale))
    return out[: (2 * n if dt == SPLIT else n)], scale[:nbrows]


def pack_b(B_f32, dt):
    B = np.ascontiguousarray(B_f32, np.float32)
    K, N = B.shape
    out = np.zeros(2 * K * N if dt == SPLIT else K * N, np.uint8 if dt == FP8 else np.uint16)
    scale = np.ones(N, np.float32)
    lib().or_pack_b(K, N, N, P(B), dt, P(out), P(scale))
    return out, scale


def spmm_scaled_f64(bk, rowptr, colidx, dt, a, a_scale, K, N, b, b_scale):
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    colidx = np.ascontiguousarray(colidx, np.int32)
    nbrows = rowptr.size - 1
    C = np.zeros((nbrows * 16, N), np.float64)
    lib().or_spmm_bcsr_scaled_f64(nbrows, bk, P(rowptr), P(colidx) if colidx.size else None, dt,
                                  P(np.ascontiguousarray(a)), None if a_scale is None else P(a_scale), K, N,
                                  P(np.ascontiguousarray(b)), None if b_scale is None else P(b_scale), P(C), N)
    return C


# ---- learned tuner ------------------------------------------------------------
STAT_FIELDS = ["rows", "cols", "nnz", "max_row"] + [f"row_hist{b}" for b in range(12)] + [
    "sum_abs_offset", "near_diag"] + [f"nnzb{w}" for w in range(6)] + ["nnzb16", "blocks24"]


def matrix_stats(M, K, rowptr, colidx, values_u16):
    out = np.zeros(26, np.int64)
    rp = np.ascontiguousarray(rowptr, np.int32)
    ci = np.ascontiguousarray(colidx, np.int32) if colidx.size else np.zeros(1, np.int32)
    v = np.ascontiguousarray(values_u16, np.uint16) if values_u16.size else np.zeros(1, np.uint16)
    lib().or_matrix_stats(M, K, P(rp), P(ci), P(v), P(out))
    return dict(zip(STAT_FIELDS, out.tolist()))


def _as_f32(a):
    return np.ascontiguousarray(a, np.float32)


def mlp_forward(X, params):
    W1, b1, W2, b2, W3, b3 = [_as_f32(p) for p in params]
    X = _as_f32(X)
    B, n_in = X.shape
    h = W1.shape[0]
    pred = np.zeros(B)
    H1, H2 = np.zeros((B, h)), np.zeros((B, h))
    lib().or_mlp_forward(B, n_in, h, P(X), P(W1), P(b1), P(W2), P(b2), P(W3), P(b3), P(pred), P(H1), P(H2))
    return pred, H1, H2


def mlp_backward(X, y, params):
    W1, b1, W2, b2, W3, b3 = [_as_f32(p) for p in params]
    X, y = _as_f32(X), _as_f32(y)
    B, n_in = X.shape
    h = W1.shape[0]
    g = [np.zeros(p.shape) for p in (W1, b1, W2, b2, W3, b3)]
    loss = lib().or_mlp_backward(B, n_in, h, P(X), P(y), P(W1), P(b1), P(W2), P(b2), P(W3), P(b3),
                                 *[P(x) for x in g])
    return loss, g


def adam_step(p, g, m, v, lr, b1, b2, eps, t):
    for a in (p, g, m, v):
        assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    lib().or_adam_step(P(p), P(g), P(m), P(v), p.size, lr, b1, b2, eps, t)
