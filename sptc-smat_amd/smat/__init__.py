"""Python binding of libsmat (the MI355X-native SMaT C ABI, include/smat.h).

The binding is a thin ctypes layer: every call goes straight to the C entry
point.  Device memory is passed as raw pointers; torch tensors are accepted
for convenience (PyTorch is used only as an allocator and stream provider).
The library must be built (``make -C sptc-smat_amd``); loading fails loudly
otherwise — there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libsmat.so")
HEADER_PATH = os.path.join(os.path.dirname(os.path.dirname(_HERE)), "include", "smat.h")

SMAT_OK = 0
SMAT_ERR_INVALID = -1
SMAT_ERR_HIP = -2
SMAT_ERR_ALLOC = -3
SMAT_ERR_NOT_24 = -4
SMAT_ERR_UNSUPPORTED = -5

SMAT_F16, SMAT_BF16, SMAT_F32 = 1, 2, 3
SMAT_FP8_E4M3, SMAT_SPLIT_F16 = 4, 5
SMAT_B_ROW_MAJOR, SMAT_B_REFERENCE = 0, 1
SMAT_TAG_REFERENCE, SMAT_TAG_EXACT24 = 0, 1
SMAT_24_REFERENCE_K16, SMAT_24_SMFMAC = 0, 1
SMAT_KERNEL_AUTO, SMAT_KERNEL_DENSE, SMAT_KERNEL_MIXED24 = 0, 1, 2

_c = ctypes
_p = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_sz = ctypes.c_size_t


class BcsrDesc(ctypes.Structure):
    _fields_ = [
        ("M", _i64), ("K", _i64), ("bm", _i32), ("bk", _i32), ("nbrows", _i64),
        ("nnzb", _i64), ("rowptr", _p), ("colidx", _p), ("values", _p), ("tags", _p),
    ]


class TilesDesc(ctypes.Structure):
    _fields_ = [
        ("M", _i64), ("K", _i64), ("bk", _i32), ("tw", _i32), ("ntrows", _i64), ("ntcols", _i64),
        ("ntiles", _i64), ("rowptr", _p), ("colidx", _p), ("values", _p), ("info", _p),
    ]


class TilesOperands(ctypes.Structure):
    _fields_ = [
        ("dtype", _i32), ("out_dtype", _i32), ("values", _p), ("values24", _p), ("index24", _p), ("b", _p),
        ("ldb", _i64), ("N", _i64), ("c", _p), ("ldc", _i64),
    ]


class SpmmOperands(ctypes.Structure):
    _fields_ = [
        ("a_values", _p), ("a_scale", _p), ("b", _p), ("b_scale", _p), ("ldb", _i64), ("c", _p),
        ("ldc", _i64), ("values24", _p), ("index24", _p), ("c_rowmap", _p),
    ]


class SpmmOpts(ctypes.Structure):
    _fields_ = [
        ("kernel", _i32), ("out_dtype", _i32), ("b_layout", _i32), ("n_tile", _i32),
        ("max_blocks_per_item", _i32), ("b_staging", _i32), ("in_dtype", _i32), ("schedule", _i32),
    ]


SMAT_BSTAGE_AUTO, SMAT_BSTAGE_DIRECT, SMAT_BSTAGE_LDS = 0, 1, 2
SMAT_SCHED_AUTO, SMAT_SCHED_BLOCK_ROWS, SMAT_SCHED_LONGEST_FIRST, SMAT_SCHED_TILE_AFFINITY = 0, 1, 2, 3
SMAT_BATCH_AUTO, SMAT_BATCH_PERSISTENT, SMAT_BATCH_BRANCHES = 0, 1, 2


class MatrixStats(ctypes.Structure):
    _fields_ = [("rows", ctypes.c_int64), ("cols", ctypes.c_int64), ("nnz", ctypes.c_int64),
                ("max_row", ctypes.c_int64), ("row_hist", ctypes.c_int64 * 12),
                ("sum_abs_offset", ctypes.c_int64), ("near_diag", ctypes.c_int64),
                ("nnzb", ctypes.c_int64 * 6), ("nnzb16", ctypes.c_int64), ("blocks24", ctypes.c_int64)]

    def as_dict(self):
        d = {"rows": self.rows, "cols": self.cols, "nnz": self.nnz, "max_row": self.max_row}
        d.update({f"row_hist{b}": self.row_hist[b] for b in range(12)})
        d.update({"sum_abs_offset": self.sum_abs_offset, "near_diag": self.near_diag})
        d.update({f"nnzb{w}": self.nnzb[w] for w in range(6)})
        d.update({"nnzb16": self.nnzb16, "blocks24": self.blocks24})
        return d


class CsrHost(ctypes.Structure):
    _fields_ = [("rows", _i64), ("cols", _i64), ("nnz", _i64), ("rowptr", ctypes.POINTER(ctypes.c_int32)),
                ("colidx", ctypes.POINTER(ctypes.c_int32)), ("values_f16", ctypes.POINTER(ctypes.c_uint16)),
                ("values_f32", ctypes.POINTER(ctypes.c_float)), ("symmetric", _i32)]


# name -> (restype, argtypes); must cover every function declared in smat.h
SIGNATURES = {
    "smat_last_error": (ctypes.c_char_p, []),
    "smat_last_status": (_i32, []),
    "smat_clear_error": (None, []),
    "smat_version": (ctypes.c_char_p, []),
    "smat_bcsr_from_csr": (_i32, [_i64, _i64, _i64, _p, _p, _p, _i32, _i32, _p, ctypes.POINTER(_p)]),
    "smat_bcsr_describe": (_i32, [_p, ctypes.POINTER(BcsrDesc)]),
    "smat_bcsr_destroy": (_i32, [_p]),
    "smat_bcsr_retag_exact24": (_i32, [_p, _p]),
    "smat_bcsr_grid_maps": (_i32, [ctypes.POINTER(BcsrDesc), _p, _p, _p]),
    "smat_bcsr_filter": (_i32, [ctypes.POINTER(BcsrDesc), ctypes.POINTER(_i64), ctypes.POINTER(_i64),
                                _p, _p, _p, _p, _p, _p, _p]),
    "smat_compress_24": (_i32, [ctypes.POINTER(BcsrDesc), _i32, _p, _p, _p]),
    "smat_spmm_plan_create": (_i32, [ctypes.POINTER(BcsrDesc), _i64, ctypes.POINTER(SpmmOpts), _p,
                                     ctypes.POINTER(_p)]),
    "smat_spmm_execute": (_i32, [_p, _p, _p, _p, _i64, _p, _i64, _p]),
    "smat_spmm_execute_ex": (_i32, [_p, ctypes.POINTER(SpmmOperands), _p]),
    "smat_spmm_plan_destroy": (_i32, [_p]),
    "smat_bcsr_scatter_values": (_i32, [ctypes.POINTER(BcsrDesc), _p, _p, _p, _p, _p]),
    "smat_reorder_columns": (_i32, [_i64, _i64, _p, _p, ctypes.c_float, _i64, _i64, _p, _p, _p]),
    "smat_permute_csr_cols": (_i32, [_i64, _p, _p, _p, _p]),
    "smat_permute_dense_rows": (_i32, [_i64, _i64, _i32, _p, _i64, _p, _p, _i64, _p]),
    "smat_read_mtx": (_i32, [ctypes.c_char_p, _i32, ctypes.POINTER(CsrHost)]),
    "smat_read_edgelist": (_i32, [ctypes.c_char_p, ctypes.c_char_p, _i32, ctypes.POINTER(CsrHost)]),
    "smat_read_edgelist_ex": (_i32, [ctypes.c_char_p, ctypes.c_char_p, _i32, _i32, _i32, ctypes.POINTER(CsrHost)]),
    "smat_parse_edgelist": (_i32, [ctypes.c_char_p, _i64, ctypes.c_char_p, _i32, _i32, _i32, ctypes.POINTER(CsrHost)]),
    "smat_csr_host_free": (None, [ctypes.POINTER(CsrHost)]),
    "smat_batch_create": (_i32, [_i32, _p, _p, _p, ctypes.POINTER(_p)]),
    "smat_batch_create_ex": (_i32, [_i32, _p, _p, _i32, _p, ctypes.POINTER(_p)]),
    "smat_batch_capture": (_i32, [_p, _p]),
    "smat_batch_execute": (_i32, [_p, _p]),
    "smat_batch_info": (_i32, [_p, ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_i32),
                               ctypes.POINTER(_i32)]),
    "smat_batch_destroy": (_i32, [_p]),
    "smat_matrix_stats": (ctypes.c_int32, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(MatrixStats),
                                           ctypes.c_void_p]),
    "smat_tuner_features": (ctypes.c_int32, [ctypes.POINTER(MatrixStats), ctypes.c_int64, ctypes.c_void_p]),
    "smat_linear_forward": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p,
                                             ctypes.c_void_p]),
    "smat_linear_backward": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                              ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "smat_mse_grad": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p]),
    "smat_adam_step": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_int64, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                        ctypes.c_float, ctypes.c_int32, ctypes.c_void_p]),
    "smat_tuner_predict": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32,
                                            ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p]),
    "smat_pack_a": (_i32, [ctypes.POINTER(BcsrDesc), _p, _i32, _p, _p, _p]),
    "smat_pack_b": (_i32, [_p, _i64, _i64, _i64, _i32, _p, _p, _p]),
    "smat_error_sums": (_i32, [_p, _i64, _p, _i64, _i64, _i64, _p, _p, _p]),
    "smat_spmm_noalloc": (_i32, [ctypes.POINTER(BcsrDesc), _i64, ctypes.POINTER(SpmmOpts),
                                 ctypes.POINTER(SpmmOperands), _p]),
    "smat_spmm_plan_info": (_i32, [_p, ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    "smat_reorder_grouping": (_i32, [_i64, _p, _p, _p, _p, ctypes.c_float, _i64, _i64, _p, _p, _p]),
    "smat_fixed_size_grouping": (_i32, [_i64, _p, _i64, _p]),
    "smat_grouping_to_permutation": (_i32, [_i64, _p, _p]),
    "smat_blocking_info": (_i32, [_i64, _i64, _i64, _p, _p, _p, _i64, _p, _p, _p]),
    "smat_permute_csr_rows": (_i32, [_i64, _p, _p, _p, _p, _p, _p, _p, _p]),
    "smat_locality_order": (_i32, [_i64, _i64, _p, _p, _p, _p]),
    "smat_tiles_from_bcsr": (_i32, [ctypes.POINTER(BcsrDesc), _p, ctypes.POINTER(_p)]),
    "smat_tiles_describe": (_i32, [_p, ctypes.POINTER(TilesDesc)]),
    "smat_tiles_destroy": (_i32, [_p]),
    "smat_tiles_grid_maps": (_i32, [ctypes.POINTER(TilesDesc), _p, _p, _p]),
    "smat_tiles_compress_24": (_i32, [ctypes.POINTER(TilesDesc), _p, _p, _p]),
    "smat_tiles_spmm": (_i32, [ctypes.POINTER(TilesDesc), ctypes.POINTER(TilesOperands), _p]),
    "mmaOBTKernel_tiled": (None, [_p, _p, _p, _p, _p, _sz, _sz, _sz, _sz, _p, _p]),
    "preprocessing_mmaOBTSKernel_tiled_large": (None, [_p, _p, _p, _sz, _sz, _sz, _sz, _p, _p, _p]),
    "mmaOBTSKernel_tiled_large": (None, [_p, _p, _p, _p, _p, _p, _p, _sz, _sz, _sz, _sz, _p, _p, _p]),
    "mmaOBTKernel": (None, [_p, _p, _p, _p, _p, _sz, _sz, _sz, _sz, _p, _p]),
    "mmaOBTSKernel": (None, [_p, _p, _p, _p, _p, _p, _p, _sz, _sz, _sz, _sz, _p, _p]),
    "preprocessing_mmaSTKernel": (None, [_p, _p, _p, _sz, _sz, _sz, _sz, _p, _p]),
    "mmaOBTKernel_large": (None, [_p, _p, _p, _p, _p, _sz, _sz, _sz, _sz, _p, _p]),
    "mmaOBTSKernel_large": (None, [_p, _p, _p, _p, _p, _p, _p, _sz, _sz, _sz, _sz, _p, _p]),
    "mmaOBTSKernel_large_separate": (None, [_p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _sz, _sz, _sz, _sz, _p, _p]),
    "preprocessing_mmaSTKernel_large": (None, [_p, _p, _p, _sz, _sz, _sz, _sz, _p, _p]),
}

_lib = None


def lib() -> ctypes.CDLL:
    """Loads libsmat.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libsmat.so not built at {LIB_PATH}; run `make -C sptc-smat_amd`")
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


class SmatError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"smat status {status}: {msg}")
        self.status = status


def check(status: int) -> None:
    if status != SMAT_OK:
        raise SmatError(status, lib().smat_last_error().decode(errors="replace"))


def ptr(x) -> int | None:
    """Device pointer of a torch tensor (or an int / None passthrough)."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


def stream_ptr(stream=None) -> int | None:
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    return stream if isinstance(stream, int) else stream.cuda_stream


class Bcsr:
    """Device-resident BCSR (16 x bk blocks) built on the GPU from a device CSR."""

    def __init__(self, M: int, K: int, rowptr, colidx, values, bk: int = 16,
                 tag_rule: int = SMAT_TAG_REFERENCE, stream=None):
        h = _p()
        nnz = int(colidx.numel()) if hasattr(colidx, "numel") else 0
        check(lib().smat_bcsr_from_csr(M, K, nnz, ptr(rowptr), ptr(colidx), ptr(values), bk, tag_rule,
                                       stream_ptr(stream), ctypes.byref(h)))
        self.handle = h
        self.desc = BcsrDesc()
        check(lib().smat_bcsr_describe(h, ctypes.byref(self.desc)))

    @property
    def nnzb(self) -> int:
        return int(self.desc.nnzb)

    def arrays(self):
        """Copies (rowptr, colidx, values[u16], tags) to torch CPU tensors."""
        import torch
        d = self.desc
        out = []
        for p, n, dt in ((d.rowptr, d.nbrows + 1, torch.int32), (d.colidx, d.nnzb, torch.int32),
                         (d.values, d.nnzb * 16 * d.bk, torch.int16), (d.tags, d.nnzb, torch.int32)):
            t = torch.empty(int(n), dtype=dt)
            if n:
                torch.cuda.synchronize()
                copy_to_host(t.data_ptr(), p, int(n) * t.element_size())
            out.append(t)
        return out

    def retag_exact24(self, stream=None):
        check(lib().smat_bcsr_retag_exact24(self.handle, stream_ptr(stream)))

    def grid_maps(self, stream=None):
        """(blockInfo, relativeBlockIndexMapping) as device int32 tensors over
        the dense block grid, as the reference kernels take them."""
        import torch
        d = self.desc
        ng = int(d.nbrows * (d.K // d.bk))
        info = torch.empty(ng, dtype=torch.int32, device="cuda")
        rel = torch.empty(ng, dtype=torch.int32, device="cuda")
        check(lib().smat_bcsr_grid_maps(ctypes.byref(d), ptr(info), ptr(rel), stream_ptr(stream)))
        return info, rel

    def compress_24(self, layout: int = SMAT_24_SMFMAC, stream=None):
        """Compressed 2:4 arrays (values, metadata) as device tensors."""
        import torch
        d = self.desc
        n = int(d.nnzb)
        if layout == SMAT_24_REFERENCE_K16:
            vals = torch.empty(max(n, 1) * 128, dtype=torch.int16, device="cuda")
            meta = torch.empty(max(n, 1) * 32, dtype=torch.uint8, device="cuda")
        else:
            vals = torch.empty(max(n, 1) * 16 * (d.bk // 2), dtype=torch.int16, device="cuda")
            meta = torch.empty(max(n, 1) * 16 * (d.bk // 16), dtype=torch.int16, device="cuda")
        check(lib().smat_compress_24(ctypes.byref(d), layout, ptr(vals), ptr(meta), stream_ptr(stream)))
        return vals, meta

    def filter(self, stream=None):
        """(sparse list, dense list), each (rowptr, colidx, values) device tensors."""
        import torch
        d = self.desc
        ns, nd = _i64(), _i64()
        check(lib().smat_bcsr_filter(ctypes.byref(d), ctypes.byref(ns), ctypes.byref(nd), None, None, None,
                                     None, None, None, stream_ptr(stream)))
        mk = lambda n, dt: torch.empty(max(int(n), 1), dtype=dt, device="cuda")  # noqa: E731
        bs = 16 * d.bk
        s = (mk(d.nbrows + 1, torch.int32), mk(ns.value, torch.int32), mk(ns.value * bs, torch.int16))
        dn = (mk(d.nbrows + 1, torch.int32), mk(nd.value, torch.int32), mk(nd.value * bs, torch.int16))
        check(lib().smat_bcsr_filter(ctypes.byref(d), None, None, *(ptr(x) for x in s), *(ptr(x) for x in dn),
                                     stream_ptr(stream)))
        return (s[0], s[1][: ns.value], s[2][: ns.value * bs]), (dn[0], dn[1][: nd.value], dn[2][: nd.value * bs])

    def close(self):
        if self.handle:
            lib().smat_bcsr_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Tiles:
    """Merged 64 x 4bk tiles (4 x 4 sub-blocks of 16 x bk, bk 16 or 32) built
    on the GPU from a Bcsr (SparseMatrix::bcsrBlocking, matrix.h:794-1137)."""

    def __init__(self, A: Bcsr | BcsrDesc, stream=None):
        desc = A.desc if isinstance(A, Bcsr) else A
        h = _p()
        check(lib().smat_tiles_from_bcsr(ctypes.byref(desc), stream_ptr(stream), ctypes.byref(h)))
        self.handle = h
        self.desc = TilesDesc()
        check(lib().smat_tiles_describe(h, ctypes.byref(self.desc)))

    @property
    def ntiles(self) -> int:
        return int(self.desc.ntiles)

    def arrays(self):
        """Copies (rowptr, colidx, values[u16 as int16], info) to torch CPU tensors."""
        import torch
        d = self.desc
        out = []
        for p, n, dt in ((d.rowptr, d.ntrows + 1, torch.int32), (d.colidx, d.ntiles, torch.int32),
                         (d.values, d.ntiles * 256 * d.bk, torch.int16), (d.info, d.ntiles * 16, torch.int32)):
            t = torch.empty(int(n), dtype=dt)
            if n:
                torch.cuda.synchronize()
                copy_to_host(t.data_ptr(), p, int(n) * t.element_size())
            out.append(t)
        return out

    def grid_maps(self, stream=None):
        """(mergedBlockInfo, mergedRelativeBlockIndexMapping) over the tile grid, device int32."""
        import torch
        d = self.desc
        ng = max(int(d.ntrows * d.ntcols), 1)
        info = torch.empty(ng, dtype=torch.int32, device="cuda")
        rel = torch.empty(ng, dtype=torch.int32, device="cuda")
        check(lib().smat_tiles_grid_maps(ctypes.byref(d), ptr(info), ptr(rel), stream_ptr(stream)))
        return info, rel

    def compress_24(self, stream=None):
        """(values24, index24) of the info-1 sub-blocks, SMFMAC layout, device tensors."""
        import torch
        d = self.desc
        n = max(int(d.ntiles) * 16, 1)
        vals = torch.empty(n * 16 * (d.bk // 2), dtype=torch.int16, device="cuda")
        idx = torch.empty(n * 16 * (d.bk // 16), dtype=torch.int16, device="cuda")
        check(lib().smat_tiles_compress_24(ctypes.byref(d), ptr(vals), ptr(idx), stream_ptr(stream)))
        return vals, idx

    def spmm(self, B, C, *, N: int | None = None, dtype: int = SMAT_F16, values=None, values24=None, index24=None,
             ldb: int | None = None, ldc: int | None = None, stream=None):
        """C = A * B (B row-major K x N, C row-major M x N, FP16 or FP32 by C's dtype)."""
        import torch
        n = int(B.shape[-1]) if N is None else N
        out = SMAT_F32 if C.dtype == torch.float32 else SMAT_F16
        ops = TilesOperands(dtype, out, ptr(values), ptr(values24), ptr(index24), ptr(B),
                            int(B.shape[-1]) if ldb is None else ldb, n, ptr(C),
                            int(C.shape[-1]) if ldc is None else ldc)
        check(lib().smat_tiles_spmm(ctypes.byref(self.desc), ctypes.byref(ops), stream_ptr(stream)))
        return C

    def close(self):
        if self.handle:
            lib().smat_tiles_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_hip = None


def copy_to_host(dst_host: int, src_dev: int, nbytes: int) -> None:
    """Synchronous device-to-host copy of library-owned memory (hipMemcpy)."""
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.restype = ctypes.c_int
        _hip.hipMemcpy.argtypes = [_p, _p, _sz, ctypes.c_int]
    ret = _hip.hipMemcpy(dst_host, src_dev, nbytes, 2)  # hipMemcpyDeviceToHost
    if ret != 0:
        raise RuntimeError(f"hipMemcpy D2H failed: {ret}")


def copy_to_device(dst_dev: int, src_host: int, nbytes: int) -> None:
    """Synchronous host-to-device copy into library-owned memory (hipMemcpy)."""
    copy_to_host(0, 0, 0)  # loads the runtime binding
    ret = _hip.hipMemcpy(dst_dev, src_host, nbytes, 1)  # hipMemcpyHostToDevice
    if ret != 0:
        raise RuntimeError(f"hipMemcpy H2D failed: {ret}")


@dataclass
class PlanInfo:
    n_items: int
    n_split_rows: int
    scratch_bytes: int


class SpmmPlan:
    def __init__(self, A: Bcsr | BcsrDesc, N: int, *, kernel: int = SMAT_KERNEL_AUTO,
                 out_dtype: int = SMAT_F32, b_layout: int = SMAT_B_ROW_MAJOR, n_tile: int = 0,
                 max_blocks_per_item: int = 0, b_staging: int = 0, in_dtype: int = SMAT_F16, schedule: int = 0,
                 stream=None):
        desc = A.desc if isinstance(A, Bcsr) else A
        self._keep = A
        o = SpmmOpts(kernel, out_dtype, b_layout, n_tile, max_blocks_per_item, b_staging, in_dtype, schedule)
        h = _p()
        check(lib().smat_spmm_plan_create(ctypes.byref(desc), N, ctypes.byref(o), stream_ptr(stream),
                                          ctypes.byref(h)))
        self.handle = h
        self.N = N

    def info(self) -> PlanInfo:
        a, b, c = _i64(), _i64(), _i64()
        check(lib().smat_spmm_plan_info(self.handle, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return PlanInfo(a.value, b.value, c.value)

    def __call__(self, B, C, ldb: int | None = None, ldc: int | None = None, values24=None,
                 index24=None, stream=None):
        if ldb is None:
            ldb = B.shape[-1]
        if ldc is None:
            ldc = C.shape[-1]
        check(lib().smat_spmm_execute(self.handle, ptr(values24), ptr(index24), ptr(B), ldb, ptr(C),
                                      ldc, stream_ptr(stream)))
        return C

    def operands(self, B, C, *, a_values=None, a_scale=None, b_scale=None, ldb: int | None = None,
                 ldc: int | None = None, values24=None, index24=None, c_rowmap=None) -> SpmmOperands:
        if ldb is None:
            ldb = self.N
        if ldc is None:
            ldc = C.shape[-1]
        return SpmmOperands(ptr(a_values), ptr(a_scale), ptr(B), ptr(b_scale), ldb, ptr(C), ldc, ptr(values24),
                            ptr(index24), ptr(c_rowmap))

    def execute(self, B, C, *, a_values=None, a_scale=None, b_scale=None, ldb: int | None = None,
                ldc: int | None = None, values24=None, index24=None, c_rowmap=None, stream=None):
        """smat_spmm_execute_ex: every operand explicit (FP8 / split plans, reordered rows)."""
        ops = self.operands(B, C, a_values=a_values, a_scale=a_scale, b_scale=b_scale, ldb=ldb, ldc=ldc,
                            values24=values24, index24=index24, c_rowmap=c_rowmap)
        check(lib().smat_spmm_execute_ex(self.handle, ctypes.byref(ops), stream_ptr(stream)))
        return C

    def close(self):
        if self.handle:
            lib().smat_spmm_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


SMAT_NOALLOC_SCRATCH_BYTES = 64 << 20


def spmm_noalloc(A: Bcsr | BcsrDesc, N: int, B, C, *, kernel: int = SMAT_KERNEL_AUTO, out_dtype: int = SMAT_F32,
                 b_layout: int = SMAT_B_ROW_MAJOR, n_tile: int = 0, max_blocks_per_item: int = 0,
                 b_staging: int = 0, in_dtype: int = SMAT_F16, schedule: int = 0, a_values=None, a_scale=None,
                 b_scale=None, ldb: int | None = None, ldc: int | None = None, values24=None, index24=None,
                 c_rowmap=None, stream=None):
    """smat_spmm_noalloc: plan + execute in one call, with the plan's device
    arrays in the library's static scratch (no device allocation)."""
    desc = A.desc if isinstance(A, Bcsr) else A
    o = SpmmOpts(kernel, out_dtype, b_layout, n_tile, max_blocks_per_item, b_staging, in_dtype, schedule)
    ops = SpmmOperands(ptr(a_values), ptr(a_scale), ptr(B), ptr(b_scale), N if ldb is None else ldb, ptr(C),
                       C.shape[-1] if ldc is None else ldc, ptr(values24), ptr(index24), ptr(c_rowmap))
    check(lib().smat_spmm_noalloc(ctypes.byref(desc), N, ctypes.byref(o), ctypes.byref(ops), stream_ptr(stream)))
    return C


def reorder_grouping(rowptr, colidx, tau: float, col_block: int, max_rows: int, stream=None):
    """Keeper (algorithm 5) grouping of a host CSR pattern (int32 numpy
    arrays, columns sorted per row); returns (grouping int64[rows],
    [comparisons, merges])."""
    import numpy as np
    import torch
    rp = np.ascontiguousarray(rowptr, np.int32)
    ci = np.ascontiguousarray(colidx, np.int32)
    rows = rp.size - 1
    d_rp = torch.from_numpy(rp).cuda()
    d_ci = torch.from_numpy(ci if ci.size else np.zeros(1, np.int32)).cuda()
    g = np.zeros(rows, np.int64)
    cnt = np.zeros(2, np.int64)
    check(lib().smat_reorder_grouping(rows, d_rp.data_ptr(), d_ci.data_ptr(), rp.ctypes.data,
                                      (ci if ci.size else np.zeros(1, np.int32)).ctypes.data, float(tau),
                                      col_block, max_rows, g.ctypes.data, cnt.ctypes.data, stream_ptr(stream)))
    return g, cnt


def reorder_columns(rowptr, colidx, cols: int, tau: float, col_block: int, max_rows: int, stream=None):
    """Column order from the Keeper on the transposed pattern (host int32
    arrays); returns (col_perm int64[cols], [comparisons, merges])."""
    import numpy as np
    rp = np.ascontiguousarray(rowptr, np.int32)
    ci = np.ascontiguousarray(colidx, np.int32) if len(colidx) else np.zeros(1, np.int32)
    perm = np.zeros(cols, np.int64)
    cnt = np.zeros(2, np.int64)
    check(lib().smat_reorder_columns(rp.size - 1, cols, rp.ctypes.data, ci.ctypes.data, float(tau), col_block,
                                     max_rows, perm.ctypes.data, cnt.ctypes.data, stream_ptr(stream)))
    return perm, cnt


REFERENCE_TAUS = (0.9, 0.8, 0.6, 0.5, 0.4, 0.3, 0.2, 0.1, 0.01)


def blocking_info(rows: int, cols: int, d_rowptr, d_colidx, d_grouping, col_block: int, stream=None):
    """VBR statistics of a row grouping on the GPU (BlockingEngine::
    CollectBlockingInfo, preprocess/src/general/blocking.cpp:562-614): device
    int32 rowptr / colidx and int64 grouping (torch tensors); returns
    ([VBR_nzcount, VBR_nzblocks_count, VBR_longest_row], VBR_average_height)."""
    import numpy as np
    out = np.zeros(3, np.int64)
    avg = ctypes.c_float(0.0)
    nnz = int(d_colidx.numel()) if d_colidx is not None else 0
    check(lib().smat_blocking_info(rows, cols, nnz, ptr(d_rowptr), ptr(d_colidx), ptr(d_grouping), col_block,
                                   out.ctypes.data, ctypes.byref(avg), stream_ptr(stream)))
    return out, float(avg.value)


def threshold_search(rowptr, colidx, cols: int, col_block: int = 16, row_block: int = 16, taus=REFERENCE_TAUS):
    """The reference's exhaustive threshold search
    (preprocess/test/general/reorder_matrix.cpp:17-33): for each tau, Keeper
    grouping forced to fixed-size groups (BlockingEngine::GetGrouping with
    -F 1, blocking.cpp:651-654), then the number of nonzero (group, column
    block) pairs (VBR_nzblocks_count of CollectBlockingInfo,
    blocking.cpp:562-614); the best count wins, ties going to the later tau
    (`<=`), starting from the nonzero count.  Returns (best_tau, best_blocks,
    {tau: blocks})."""
    import numpy as np
    import torch
    rp = np.ascontiguousarray(rowptr, np.int32)
    ci = np.ascontiguousarray(colidx, np.int32)
    rows = rp.size - 1
    d_rp = torch.from_numpy(rp).cuda()
    d_ci = torch.from_numpy(ci).cuda() if ci.size else None
    best_tau, best = -1.0, int(ci.size)
    counts = {}
    for tau in taus:
        g, _ = reorder_grouping(rp, ci, np.float32(tau), col_block, row_block)
        fixed = torch.from_numpy(fixed_size_grouping(g, row_block)).cuda()
        blocks = int(blocking_info(rows, cols, d_rp, d_ci, fixed, col_block)[0][1])
        counts[tau] = blocks
        if blocks <= best:
            best_tau, best = tau, blocks
    return best_tau, best, counts


def fixed_size_grouping(grouping, row_block: int):
    import numpy as np
    g = np.ascontiguousarray(grouping, np.int64)
    out = np.zeros_like(g)
    check(lib().smat_fixed_size_grouping(g.size, g.ctypes.data, row_block, out.ctypes.data))
    return out


def locality_order(rows: int, cols: int, rowptr, colidx, stream=None):
    """Locality row order of a device CSR (smat_locality_order): a device
    int64 tensor, new row r = old row perm[r]."""
    import torch
    perm = torch.empty(rows, dtype=torch.int64, device="cuda")
    check(lib().smat_locality_order(rows, cols, ptr(rowptr), ptr(colidx), ptr(perm), stream_ptr(stream)))
    return perm


def permute_rows(rowptr, colidx, values, perm, stream=None):
    """Rows of a device CSR permuted on the GPU (smat_permute_csr_rows):
    (rowptr, colidx, values) device tensors, new row r = old row perm[r]."""
    import torch
    rows = perm.numel()
    out_rp = torch.empty(rows + 1, dtype=torch.int32, device="cuda")
    out_ci = torch.empty_like(colidx)
    out_v = torch.empty_like(values) if values is not None else None
    check(lib().smat_permute_csr_rows(rows, ptr(rowptr), ptr(colidx), ptr(values), ptr(perm), ptr(out_rp),
                                      ptr(out_ci), ptr(out_v), stream_ptr(stream)))
    return out_rp, out_ci, out_v


def grouping_to_permutation(grouping):
    import numpy as np
    g = np.ascontiguousarray(grouping, np.int64)
    out = np.zeros_like(g)
    check(lib().smat_grouping_to_permutation(g.size, g.ctypes.data, out.ctypes.data))
    return out


# ---- precision modes --------------------------------------------------------
def elem_bytes(dtype: int) -> int:
    return 1 if dtype == SMAT_FP8_E4M3 else 2


def scatter_values(A: Bcsr, csr_rowptr, csr_colidx, csr_values_f32, stream=None):
    """FP32 values of a device CSR -> A's block layout (torch float32 tensor)."""
    import torch
    d = A.desc
    out = torch.empty(max(d.nnzb * 16 * d.bk, 1), dtype=torch.float32, device="cuda")
    check(lib().smat_bcsr_scatter_values(ctypes.byref(d), ptr(csr_rowptr), ptr(csr_colidx), ptr(csr_values_f32),
                                         ptr(out), stream_ptr(stream)))
    return out


def pack_a(A: Bcsr, values_f32, dtype: int, stream=None):
    """Operand A of `dtype` (uint8 / int16 torch tensor) and its per-block-row
    scales (float32, FP8 and split only; None otherwise)."""
    import torch
    d = A.desc
    n = d.nnzb * 16 * d.bk
    cnt = 2 * n if dtype == SMAT_SPLIT_F16 else n
    out = torch.empty(max(cnt, 8), dtype=torch.uint8 if dtype == SMAT_FP8_E4M3 else torch.int16, device="cuda")
    scale = None
    if dtype in (SMAT_FP8_E4M3, SMAT_SPLIT_F16):
        scale = torch.empty(max(d.nbrows, 1), dtype=torch.float32, device="cuda")
    check(lib().smat_pack_a(ctypes.byref(d), ptr(values_f32), dtype, ptr(out), ptr(scale), stream_ptr(stream)))
    return out, scale


def pack_b(B_f32, dtype: int, stream=None):
    """FP32 B (K x N torch tensor) -> operand B of `dtype` and per-column scales."""
    import torch
    K, N = B_f32.shape
    cnt = 2 * K * N if dtype == SMAT_SPLIT_F16 else K * N
    out = torch.empty(cnt, dtype=torch.uint8 if dtype == SMAT_FP8_E4M3 else torch.int16, device="cuda")
    scale = None
    if dtype in (SMAT_FP8_E4M3, SMAT_SPLIT_F16):
        scale = torch.empty(N, dtype=torch.float32, device="cuda")
    check(lib().smat_pack_b(ptr(B_f32), K, N, B_f32.stride(0), dtype, ptr(out), ptr(scale), stream_ptr(stream)))
    return out, scale


def rel_fro_error(C, R, stream=None) -> float:
    """||C - R||_F / ||R||_F with the sums formed in FP64 on the GPU
    (C float32, R float64, both rows x N device tensors)."""
    import torch
    rows, N = C.shape
    out = torch.zeros(2, dtype=torch.float64, device="cuda")
    work = torch.empty(2 * 1024, dtype=torch.float64, device="cuda")
    check(lib().smat_error_sums(ptr(C), C.stride(0), ptr(R), R.stride(0), rows, N, ptr(out), ptr(work),
                                stream_ptr(stream)))
    e2, r2 = out.cpu().tolist()
    return (e2 / r2) ** 0.5 if r2 > 0 else (0.0 if e2 == 0 else float("inf"))


class Batch:
    """Many planned SpMMs in one launch (smat_batch_*): `mode` SMAT_BATCH_AUTO
    / SMAT_BATCH_BRANCHES (each matrix's own kernels on 4 parallel branches)
    or SMAT_BATCH_PERSISTENT (persistent queue kernels).  `entries` is a list
    of (SpmmPlan, SpmmOperands); plans and buffers must outlive the batch."""

    def __init__(self, entries, stream=None, mode: int = 0):
        self._keep = entries
        n = len(entries)
        plans = (_p * n)(*[e[0].handle for e in entries])
        ops = (SpmmOperands * n)(*[e[1] for e in entries])
        h = _p()
        check(lib().smat_batch_create_ex(n, ctypes.cast(plans, _p), ctypes.cast(ops, _p), mode, stream_ptr(stream),
                                         ctypes.byref(h)))
        self.handle = h

    def capture(self, stream):
        check(lib().smat_batch_capture(self.handle, stream_ptr(stream)))

    def __call__(self, stream=None):
        check(lib().smat_batch_execute(self.handle, stream_ptr(stream)))

    def info(self):
        t, r, g, c = _i64(), _i64(), _i32(), _i32()
        check(lib().smat_batch_info(self.handle, ctypes.byref(t), ctypes.byref(r), ctypes.byref(g), ctypes.byref(c)))
        return {"tasks": t.value, "split_rows": r.value, "grid": g.value, "captured": bool(c.value)}

    def close(self):
        if self.handle:
            lib().smat_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- readers ------------------------------------------------------------------
def _csr_from_host(c: CsrHost):
    import numpy as np
    try:
        rows, nnz = int(c.rows), int(c.nnz)
        rp = np.ctypeslib.as_array(c.rowptr, shape=(rows + 1,)).copy()
        ci = np.ctypeslib.as_array(c.colidx, shape=(max(nnz, 1),))[:nnz].copy()
        v16 = np.ctypeslib.as_array(c.values_f16, shape=(max(nnz, 1),))[:nnz].copy()
        v32 = np.ctypeslib.as_array(c.values_f32, shape=(max(nnz, 1),))[:nnz].copy()
        return {"rows": rows, "cols": int(c.cols), "nnz": nnz, "rowptr": rp, "colidx": ci, "values_f16": v16,
                "values_f32": v32, "symmetric": bool(c.symmetric)}
    finally:
        lib().smat_csr_host_free(ctypes.byref(c))


def read_mtx(path: str, k: int = 16):
    """Matrix Market file -> dict of host CSR arrays (smat_read_mtx)."""
    c = CsrHost()
    check(lib().smat_read_mtx(path.encode(), k, ctypes.byref(c)))
    return _csr_from_host(c)


SMAT_FMT_EL, SMAT_FMT_MTX = 0, 1


def read_edgelist(path: str, delimiter: str = " ", pattern_only: bool = False, mat_fmt: int = SMAT_FMT_EL,
                  symmetrize: bool = False):
    """Edge-list (or, with mat_fmt SMAT_FMT_MTX, the reference's MTX-as-edge-list)
    file -> dict of host CSR arrays (smat_read_edgelist_ex)."""
    c = CsrHost()
    check(lib().smat_read_edgelist_ex(path.encode(), delimiter.encode(), int(pattern_only), mat_fmt, int(symmetrize),
                                      ctypes.byref(c)))
    return _csr_from_host(c)


def parse_edgelist(text: str, delimiter: str = " ", pattern_only: bool = False, mat_fmt: int = SMAT_FMT_EL,
                   symmetrize: bool = False):
    """The same reader on text in memory (smat_parse_edgelist)."""
    c = CsrHost()
    b = text.encode()
    check(lib().smat_parse_edgelist(b, len(b), delimiter.encode(), int(pattern_only), mat_fmt, int(symmetrize),
                                    ctypes.byref(c)))
    return _csr_from_host(c)
