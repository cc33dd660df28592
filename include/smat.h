/*
 * smat.h — C ABI of the MI355X-native SMaT library (libsmat.so).
 *
 * Two layers:
 *
 *  1. Reference drop-in launchers.  Same names, parameter order and meaning
 *     as the reference's kernel launchers (src/cuda_hgemm/src/main.cu:10-77),
 *     with `half*` spelled `uint16_t*` (IEEE binary16 bits) so the signatures
 *     stay plain C.  Like the reference they return void, launch on the
 *     default stream and are asynchronous; unlike the reference they never
 *     exit the process: a failure is recorded and readable through
 *     smat_last_error() / smat_last_status().
 *
 *  2. The native API (smat_*): explicit streams, status codes, a device
 *     resident BCSR object built on the GPU from CSR, and SpMM plans that
 *     hold the load-balanced work decomposition.
 *
 * All pointers are device pointers unless the parameter name ends in `_host`.
 * Objects are allocated by *_create / smat_bcsr_from_csr and freed by the
 * matching *_destroy.  The SpMM execute calls allocate nothing (a plan owns
 * its scratch).  The preprocessing and statistics calls (conversion, filter,
 * compression, reorder, locality order, blocking info, matrix statistics,
 * the reference drop-in packers) allocate temporary device memory and free
 * it before they return.  Calls on distinct objects and streams may run
 * concurrently; one object must not be used from two threads at once.
 */
#ifndef SMAT_H_
#define SMAT_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status and errors                                                        */
/* ------------------------------------------------------------------------ */
typedef int32_t smat_status;
#define SMAT_OK 0
#define SMAT_ERR_INVALID (-1)     /* bad argument: null pointer, size, option */
#define SMAT_ERR_HIP (-2)         /* HIP runtime error; message has the detail */
#define SMAT_ERR_ALLOC (-3)       /* device allocation failed */
#define SMAT_ERR_NOT_24 (-4)      /* a block tagged 2:4 has a 1x4 group with >2 nonzeros */
#define SMAT_ERR_UNSUPPORTED (-5) /* valid request the library does not implement */

/* Thread-local message of the last failing call on this thread ("" if none). */
const char* smat_last_error(void);
/* Status of the last failing call on this thread (SMAT_OK if none). */
smat_status smat_last_status(void);
/* Clears the thread-local error state. */
void smat_clear_error(void);
const char* smat_version(void);

/* hipStream_t, kept opaque so this header needs no HIP include. NULL = default stream. */
typedef void* smat_stream;

/* ------------------------------------------------------------------------ */
/* Element types and layouts                                                */
/* ------------------------------------------------------------------------ */
#define SMAT_F16 1
#define SMAT_BF16 2
#define SMAT_F32 3
/* OCP FP8 e4m3 (finite, max 448), one byte per element.  A carries one FP32
   scale per block row, B one per column: C = a_scale[row/16] * b_scale[n] *
   sum_k qA * qB.  See smat_pack_a / smat_pack_b. */
#define SMAT_FP8_E4M3 4
/* Split FP16: every operand x (pre-scaled by a power of two, per block row of
   A and per column of B) is stored as hi = fp16(x), lo = fp16(x - hi); the
   kernel accumulates Ahi*Bhi + Ahi*Blo + Alo*Bhi in FP32. */
#define SMAT_SPLIT_F16 5

/* Layout of the dense operand B (K x N). */
#define SMAT_B_ROW_MAJOR 0   /* B[k*ldb + n] */
#define SMAT_B_REFERENCE 1   /* B[n*ldb + k] — the reference kernels' layout (mmaOBT.cu:75) */

/* Block tagging rules for CSR -> BCSR. */
#define SMAT_TAG_REFERENCE 0 /* csrToBcsr heuristic, matrix.h:689-709 (CSR-order dependent) */
#define SMAT_TAG_EXACT24 1   /* 1 iff every aligned 1x4 group of every row has <= 2 nonzeros */

/* ------------------------------------------------------------------------ */
/* Host readers                                                             */
/* ------------------------------------------------------------------------ */
typedef struct {
  int64_t rows, cols, nnz;
  int32_t* rowptr;       /* [rows + 1] */
  int32_t* colidx;       /* [nnz], each row in file order */
  uint16_t* values_f16;  /* [nnz] FP16 bits (one rounding from the file's value) */
  float* values_f32;     /* [nnz] */
  int32_t symmetric;     /* MTX: the file was symmetric / hermitian and was mirrored */
} smat_csr_host;

/*
 * Matrix Market coordinate file -> host CSR, as mmio_allinone
 * (src/cuda_hgemm/src/common/mmio_highlevel.h:523-680): rows padded to a
 * multiple of 16 and columns to a multiple of k (padSparseMatrix, :514-520),
 * pattern entries 1, complex entries their real part, symmetric / hermitian
 * files mirrored, file order kept in each row, duplicates kept.
 */
smat_status smat_read_mtx(const char* path, int32_t k, smat_csr_host* out);
/*
 * Edge list (`row<d>col[<d>value]` per line, rows not descending) -> host
 * CSR, as CSR::read_from_edgelist (preprocess/src/general/csr.cpp:166-290)
 * in edge-list mode: rows = last row + 1, cols = largest column + 1, values 1
 * when pattern_only.  delimiter NULL or "" means " ".
 */
smat_status smat_read_edgelist(const char* path, const char* delimiter, int32_t pattern_only,
                               smat_csr_host* out);
/* Formats of CSR::read_from_edgelist (preprocess/include/definitions.h MatrixFormat). */
#define SMAT_FMT_EL 0
#define SMAT_FMT_MTX 1
/*
 * CSR::read_from_edgelist with every option (csr.cpp:166-290, matrices.h:
 * 61-78): mat_fmt SMAT_FMT_MTX also skips the size line after the comments
 * and swaps the two index fields (indices kept as in the file); symmetrize
 * != 0 mirrors every entry when none lies below the diagonal (pattern-only
 * input; the reference throws for weighted input, and indexes past its rows
 * when a column has no row: both are SMAT_ERR_INVALID here).
 */
smat_status smat_read_edgelist_ex(const char* path, const char* delimiter, int32_t pattern_only,
                                  int32_t mat_fmt, int32_t symmetrize, smat_csr_host* out);
/* Same, on text in memory (the reference's CSR(std::ifstream&) constructor). */
smat_status smat_parse_edgelist(const char* text, int64_t len, const char* delimiter, int32_t pattern_only,
                                int32_t mat_fmt, int32_t symmetrize, smat_csr_host* out);
void smat_csr_host_free(smat_csr_host* csr);

/* ------------------------------------------------------------------------ */
/* BCSR object: 16 x bk blocks, built on the GPU                            */
/* ------------------------------------------------------------------------ */
typedef struct smat_bcsr_t* smat_bcsr;

typedef struct {
  int64_t M;          /* rows, multiple of 16 */
  int64_t K;          /* columns, multiple of bk */
  int32_t bm;         /* block rows (16) */
  int32_t bk;         /* block columns: 1, 2, 4, 8, 16 or 32 */
  int64_t nbrows;     /* M / 16 */
  int64_t nnzb;       /* stored blocks */
  int32_t* rowptr;    /* [nbrows + 1] */
  int32_t* colidx;    /* [nnzb] first column of the block (not the block index), matrix.h:743 */
  uint16_t* values;   /* [nnzb * 16 * bk] FP16, row-major inside each block */
  int32_t* tags;      /* [nnzb] 1 = 2:4 block, 2 = dense block */
} smat_bcsr_desc;

/*
 * CSR (int32 row pointers and columns, FP16 values, columns of a row in any
 * order — mmio_allinone keeps file order, mmio_highlevel.h:655-661) -> BCSR.
 * Bit-identical to SparseMatrix::csrToBcsr (matrix.h:666-792) for rowptr,
 * colidx, values and, with SMAT_TAG_REFERENCE, the block tags.  M and K must
 * already be padded (padSparseMatrix, mmio_highlevel.h:514-520).  Runs on
 * `stream`; returns after the block count is known (one device->host copy).
 * Duplicate (row, col) entries are not supported (the reference keeps the
 * last one in CSR order; here which one is kept is unspecified).
 */
smat_status smat_bcsr_from_csr(int64_t M, int64_t K, int64_t nnz,
                               const int32_t* csr_rowptr, const int32_t* csr_colidx,
                               const uint16_t* csr_values, int32_t bk, int32_t tag_rule,
                               smat_stream stream, smat_bcsr* out);
smat_status smat_bcsr_describe(smat_bcsr A, smat_bcsr_desc* out);
smat_status smat_bcsr_destroy(smat_bcsr A);

/*
 * Writes FP32 values of a CSR with the pattern A was built from into A's block
 * layout: out_values[nnzb * 16 * bk] (zeros where no entry falls).  The
 * reference keeps half values only (matrix.h:757-775); this is what the FP8
 * and split precision modes quantise from (smat_pack_a).
 */
smat_status smat_bcsr_scatter_values(const smat_bcsr_desc* A, const int32_t* csr_rowptr,
                                     const int32_t* csr_colidx, const float* csr_values,
                                     float* out_values, smat_stream stream);

/*
 * Re-tag every block with the exact 2:4 rule (SMAT_TAG_EXACT24) in place.
 */
smat_status smat_bcsr_retag_exact24(smat_bcsr A, smat_stream stream);

/*
 * Dense-grid views the reference kernels take (matrix.h:672-720): blockInfo
 * (0 empty, 1 2:4, 2 dense) and relativeBlockIndexMapping (-1 or the BCSR
 * block index), each [nbrows * (K / bk)] int32, caller-allocated.
 */
smat_status smat_bcsr_grid_maps(const smat_bcsr_desc* A, int32_t* block_info,
                                int32_t* relative_block_index, smat_stream stream);

/*
 * Split into the 2:4 list and the dense list (SparseMatrix::filterBcsrBlocks,
 * matrix.h:1140-1288).  Counts first: s_nnzb_host / d_nnzb_host receive the
 * list sizes when the list pointers are NULL; otherwise the caller passes
 * buffers of at least those sizes and the lists are written.
 */
smat_status smat_bcsr_filter(const smat_bcsr_desc* A, int64_t* s_nnzb_host,
                             int64_t* d_nnzb_host, int32_t* s_rowptr, int32_t* s_colidx,
                             uint16_t* s_values, int32_t* d_rowptr, int32_t* d_colidx,
                             uint16_t* d_values, smat_stream stream);

/* ------------------------------------------------------------------------ */
/* Row reordering (Jaccard "Keeper" grouping, the SMaT reorder pipeline)    */
/* ------------------------------------------------------------------------ */
/*
 * Replaces BlockingEngine::GetGrouping with blocking algorithm 5
 * (IterativeBlockingKeeper, preprocess/src/general/blocking.cpp:431-547,
 * selected by `Matrix_Blocking -a 5`, src/reorder.py:131): groups rows of a
 * pattern (columns sorted inside each row) into groups of at most max_rows
 * by block-level Jaccard distance over column blocks of width col_block.
 * Bit-exact with the reference, including its counters
 * (counters_host = [comparison_counter, merge_counter]).  The distance
 * evaluation runs on the GPU (rowptr/colidx, device); the order-dependent
 * bookkeeping runs on the host, which also needs the pattern
 * (rowptr_host/colidx_host).  grouping_host receives `rows` group ids: the
 * seed row of a complete group, seed + rows for an incomplete one.
 */
smat_status smat_reorder_grouping(int64_t rows, const int32_t* rowptr, const int32_t* colidx,
                                  const int32_t* rowptr_host, const int32_t* colidx_host, float tau,
                                  int64_t col_block, int64_t max_rows, int64_t* grouping_host,
                                  int64_t* counters_host, smat_stream stream);

/* get_fixed_size_grouping (preprocess/src/general/utilities.cpp:45-54). Host arrays. */
smat_status smat_fixed_size_grouping(int64_t rows, const int64_t* grouping_host, int64_t row_block,
                                     int64_t* out_host);

/* get_permutation (utilities.cpp:8-20): the row order CSR::reorder applies. Host arrays. */
smat_status smat_grouping_to_permutation(int64_t rows, const int64_t* grouping_host, int64_t* perm_host);

/*
 * VBR blocking statistics of a row grouping on the GPU: replaces
 * BlockingEngine::CollectBlockingInfo (preprocess/src/general/blocking.cpp:
 * 562-614) with get_partition (utilities.cpp:23-43).  Rows with equal
 * grouping values form one block row; out_host[0] = VBR_nzcount,
 * out_host[1] = VBR_nzblocks_count, out_host[2] = VBR_longest_row;
 * *avg_height_host = VBR_average_height (NaN when there are no blocks, as
 * the reference's 0/0).  rowptr/colidx/grouping are device arrays; grouping
 * values must lie in [0, 2 * rows) (as GetGrouping and
 * smat_fixed_size_grouping produce).  Synchronous.
 */
smat_status smat_blocking_info(int64_t rows, int64_t cols, int64_t nnz, const int32_t* rowptr, const int32_t* colidx,
                               const int64_t* grouping, int64_t col_block, int64_t* out_host, float* avg_height_host,
                               smat_stream stream);

/*
 * CSR::reorder / permute_rows (csr.cpp:65-74, 97-105) on the GPU: row r of
 * the output is row perm[r] of the input (perm: device int64[rows]); values
 * may be NULL (pattern only).  Unlike the reference pipeline, which writes
 * the reordered matrix as a pattern and sets every value to 1
 * (Matrix_Blocking.cpp:46, reorder.py:39-40), values travel with their rows.
 */
smat_status smat_permute_csr_rows(int64_t rows, const int32_t* rowptr, const int32_t* colidx,
                                  const uint16_t* values, const int64_t* perm, int32_t* out_rowptr,
                                  int32_t* out_colidx, uint16_t* out_values, smat_stream stream);

/*
 * Locality row order — no counterpart in the reference; it stands in for the
 * Jaccard Keeper (smat_reorder_grouping) on matrices too large for the
 * Keeper's host replay.  Rows sorted by descending floor(log2 length), then
 * by their first stored column, then by their second (a missing column
 * counts as `cols`; the second is dropped when cols >= 2^29), ties in row
 * order.  perm (device int64[rows]) receives the old row of each new row, as
 * smat_permute_csr_rows takes it; the SpMM writes C back to the original
 * rows through smat_spmm_operands.c_rowmap.
 */
smat_status smat_locality_order(int64_t rows, int64_t cols, const int32_t* rowptr, const int32_t* colidx,
                                int64_t* perm, smat_stream stream);

/*
 * Column pass of the reference's reorder driver (src/reorder.py:143-183):
 * the pattern (host arrays) is transposed with each transposed row sorted, as
 * transpose_el_matrix does, grouped by the same Keeper (GPU) and turned into
 * a column order: new column c is old column col_perm_host[c].  The product
 * is unchanged when A's columns are relabelled (smat_permute_csr_cols with
 * the inverse order) and B's rows permuted the same way
 * (smat_permute_dense_rows with col_perm).
 */
smat_status smat_reorder_columns(int64_t rows, int64_t cols, const int32_t* rowptr_host,
                                 const int32_t* colidx_host, float tau, int64_t col_block,
                                 int64_t max_rows, int64_t* col_perm_host, int64_t* counters_host,
                                 smat_stream stream);
/* out_colidx[q] = inv_perm[colidx[q]] (device arrays; inv_perm: int64[cols]). */
smat_status smat_permute_csr_cols(int64_t nnz, const int32_t* colidx, const int64_t* inv_perm,
                                  int32_t* out_colidx, smat_stream stream);
/* Row r of out (K x N, ld_out) = row perm[r] of in (ld_in); elem_bytes 2 or 4;
 * perm: device int64[K], a permutation of 0..K-1. */
smat_status smat_permute_dense_rows(int64_t K, int64_t N, int32_t elem_bytes, const void* in, int64_t ld_in,
                                    const int64_t* perm, void* out, int64_t ld_out, smat_stream stream);

/* ------------------------------------------------------------------------ */
/* 2:4 compression                                                          */
/* ------------------------------------------------------------------------ */
/* Byte layouts of compressed 2:4 blocks. */
#define SMAT_24_REFERENCE_K16 0 /* preprocessing_mmaSTKernelSparse, mmaST.cu:22-140:
                                   values [blk][32 lanes][4] halves, metadata [blk][32] bytes */
#define SMAT_24_SMFMAC 1        /* this library's SMFMAC operand layout: values
                                   [blk][16 rows][bk/2] halves, index [blk][16 rows][bk/8] uint16 */

/*
 * Compress every block with tag 1 of a 16x16 (or 16x32 for SMAT_24_SMFMAC)
 * BCSR.  Values and metadata arrays are indexed by BCSR block; blocks with
 * tag 2 are written as zeros.  Returns SMAT_ERR_NOT_24 if a tag-1 block
 * violates 2:4 (the reference overflows its slots in that case,
 * mmaST.cu:103-107); the output is then still written with the extra
 * nonzeros dropped.
 */
smat_status smat_compress_24(const smat_bcsr_desc* A, int32_t layout, uint16_t* values24,
                             void* metadata, smat_stream stream);

/* ------------------------------------------------------------------------ */
/* Merged tiles: 64 x 4bk tiles of 4 x 4 sub-blocks of 16 x bk              */
/* ------------------------------------------------------------------------ */
/*
 * Replaces SparseMatrix::bcsrBlocking (src/cuda_hgemm/src/common/matrix.h:
 * 794-1137, BLOCK = 4; bk = 16 gives the 64 x 64 tiles of mmaOBT_tiled.cu,
 * bk = 32 the 64 x 128 tiles of mmaOBTS_tiled_large.cu).  As in the
 * reference, sub-block sb = 4*sr + sc of tile t (sr its 16-row band, sc its
 * bk-column band) is stored 16 x bk row-major at (16*t + sb)*16*bk, colidx
 * holds the tile's first column, and info holds one entry per sub-block
 * (0 empty, 1 every aligned 1x4 group of every row has <= 2 nonzeros, 2
 * otherwise).  Unlike bcsrBlocking, tiles are sorted by (tile row, tile
 * column), rowptr has an entry for every tile row, and info is computed
 * from each sub-block's own values (matrix.h:1042-1050 reads other
 * sub-blocks' rows).
 */
typedef struct smat_tiles_t* smat_tiles;

typedef struct {
  int64_t M;        /* rows of A (a multiple of 16) */
  int64_t K;        /* columns of A (a multiple of bk) */
  int32_t bk;       /* sub-block width: 16 or 32 */
  int32_t tw;       /* tile width, 4 * bk */
  int64_t ntrows;   /* tile rows, ceil(M / 64) */
  int64_t ntcols;   /* tile columns, ceil(K / tw) */
  int64_t ntiles;   /* stored tiles */
  int32_t* rowptr;  /* [ntrows + 1] */
  int32_t* colidx;  /* [ntiles] first column of the tile, a multiple of tw */
  uint16_t* values; /* [ntiles * 16 * 16 * bk] the BCSR's 16-bit values */
  int32_t* info;    /* [ntiles * 16] sub-block types */
} smat_tiles_desc;

/* Groups the blocks of a BCSR with 16 x bk blocks (bk 16 or 32) into tiles
 * on the GPU (sort of block keys by tile, O(nnzb log nnzb)). */
smat_status smat_tiles_from_bcsr(const smat_bcsr_desc* A, smat_stream stream, smat_tiles* out);
smat_status smat_tiles_describe(smat_tiles T, smat_tiles_desc* out);
smat_status smat_tiles_destroy(smat_tiles T);

/* mergedBlockInfo / mergedRelativeBlockIndexMapping (matrix.h:919-927,
 * 1083-1089): [ntrows * ntcols] int32 each, caller-allocated: 1 / the stored
 * tile index where a tile exists, 0 / -1 elsewhere. */
smat_status smat_tiles_grid_maps(const smat_tiles_desc* T, int32_t* tile_info_grid,
                                 int32_t* relative_tile_index, smat_stream stream);

/* Compresses every info-1 sub-block into the SMAT_24_SMFMAC layout, indexed
 * by sub-block (16*t + sb): values24 [ntiles*16 * 16 * bk/2] halves,
 * index24 [ntiles*16 * 16 * bk/16] uint16. */
smat_status smat_tiles_compress_24(const smat_tiles_desc* T, uint16_t* values24, uint16_t* index24,
                                   smat_stream stream);

typedef struct {
  int32_t dtype;             /* SMAT_F16 or SMAT_BF16: element type of values, values24 and B */
  int32_t out_dtype;         /* SMAT_F16 or SMAT_F32 */
  const void* values;        /* NULL: the tiles' own values, read as dtype; else [ntiles*16*16*bk] */
  const uint16_t* values24;  /* with index24: info-1 sub-blocks on SMFMAC (smat_tiles_compress_24 of */
  const uint16_t* index24;   /* the same values); both NULL: every sub-block on dense MFMA */
  const void* b;             /* row-major B: K rows of ldb elements, 16-byte aligned */
  int64_t ldb;               /* multiple of 8, >= N */
  int64_t N;
  void* c;                   /* row-major C: M rows of ldc elements, every row written */
  int64_t ldc;
} smat_tiles_operands;

/* C = A * B over the tiles, FP32 accumulation.  Empty sub-blocks are
 * skipped; with values24/index24 a 64-deep window of a sub-row that holds no
 * dense sub-block and would need both of its 32-deep MFMA steps runs on one
 * SMFMAC, every other window on MFMA. */
smat_status smat_tiles_spmm(const smat_tiles_desc* T, const smat_tiles_operands* ops, smat_stream stream);

/* ------------------------------------------------------------------------ */
/* SpMM: C = A * B, A BCSR (FP16), B dense K x N, FP32 accumulation         */
/* ------------------------------------------------------------------------ */
#define SMAT_KERNEL_AUTO 0
#define SMAT_KERNEL_DENSE 1   /* every stored block on dense MFMA */
#define SMAT_KERNEL_MIXED24 2 /* tag-1 blocks on SMFMAC (2:4), tag-2 blocks on MFMA */

/* How row-major B tiles reach the matrix cores. */
#define SMAT_BSTAGE_AUTO 0   /* LDS when B rows are 16-byte aligned, else direct */
#define SMAT_BSTAGE_DIRECT 1 /* per-lane 16-bit gathers straight into MFMA operands */
#define SMAT_BSTAGE_LDS 2    /* 128-bit row loads into LDS, ds_read_b64_tr_b16 operand reads */

typedef struct {
  int32_t kernel;          /* SMAT_KERNEL_* */
  int32_t out_dtype;       /* SMAT_F32 or SMAT_F16 */
  int32_t b_layout;        /* SMAT_B_ROW_MAJOR or SMAT_B_REFERENCE */
  int32_t n_tile;          /* columns per wave: 16, 32, 64 or 128 (128: dense kernel, FP16 / BF16 /
                              FP8, row-major B staged through LDS); 0 = auto (FP8: 128 when N >= 128) */
  int32_t max_blocks_per_item; /* split block rows longer than this (at most 65536); 0 = auto */
  int32_t b_staging;       /* SMAT_BSTAGE_* (row-major B only) */
  int32_t in_dtype;        /* SMAT_F16 (0 = F16), SMAT_BF16, SMAT_FP8_E4M3 or SMAT_SPLIT_F16 */
  int32_t schedule;        /* SMAT_SCHED_* */
} smat_spmm_opts;

/* Order in which work items reach the GPU. */
#define SMAT_SCHED_AUTO 0          /* block-row lengths skewed (longest over 8x the mean): tile
                                      affinity with 4 N tiles (16x1 blocks) or 2 or 4 (wider
                                      blocks), else longest first; block-row order otherwise */
#define SMAT_SCHED_BLOCK_ROWS 1    /* block-row order; each XCD takes a contiguous range (L2 locality) */
#define SMAT_SCHED_LONGEST_FIRST 2 /* items sorted by length, longest dispatched first: hub rows start
                                      early, so only rows over half a wave's fair share are split */
#define SMAT_SCHED_TILE_AFFINITY 3 /* longest first, and each XCD runs only some of the N tiles, so
                                      its L2 holds one column slice of B; needs 8 % tiles == 0 or
                                      tiles % 8 == 0 (tiles = ceil(N / n_tile)), else as 2 */

typedef struct smat_plan_t* smat_plan;

/*
 * Operands of one SpMM call (smat_spmm_execute_ex).  a_values NULL means the
 * BCSR's own values.  For SMAT_FP8_E4M3 and SMAT_SPLIT_F16 plans, a_values /
 * a_scale come from smat_pack_a and b / b_scale from smat_pack_b (split
 * arrays hold hi then lo: A's lo starts nnzb*16*bk elements in, B's K*ldb).
 */
typedef struct {
  const void* a_values;
  const float* a_scale;        /* [nbrows]; FP8 and split only */
  const void* b;               /* K x N */
  const float* b_scale;        /* [N]; FP8 and split only */
  int64_t ldb;                 /* elements */
  void* c;                     /* M x N row-major, overwritten */
  int64_t ldc;
  const uint16_t* values24;    /* SMAT_KERNEL_MIXED24 only */
  const void* index24;
  const int32_t* c_rowmap;     /* NULL, or [M]: row r of A is row c_rowmap[r] of C (A's rows
                                  reordered, e.g. by smat_reorder_grouping; C keeps the
                                  original order) */
} smat_spmm_operands;

/*
 * Builds the work decomposition (on the GPU) and the scratch for split rows.
 * SMAT_KERNEL_MIXED24 needs bk = 16 and row-major B; its plan also holds the
 * per-block-row lists of tag-1 (2:4) and tag-2 (dense) blocks.
 */
smat_status smat_spmm_plan_create(const smat_bcsr_desc* A, int64_t N,
                                  const smat_spmm_opts* opts, smat_stream stream,
                                  smat_plan* out);
/*
 * Runs the SpMM.  B: K x N (layout from the plan), leading dimension ldb;
 * C: M x N row-major, leading dimension ldc, overwritten (beta = 0).
 * For SMAT_KERNEL_MIXED24, values24/index24 are the SMAT_24_SMFMAC arrays of
 * smat_compress_24 (NULL otherwise).
 */
smat_status smat_spmm_execute(smat_plan plan, const uint16_t* values24, const void* index24,
                              const uint16_t* B, int64_t ldb, void* C, int64_t ldc,
                              smat_stream stream);
/* Same as smat_spmm_execute with every operand explicit; the only entry for
 * FP8 and split plans. */
smat_status smat_spmm_execute_ex(smat_plan plan, const smat_spmm_operands* ops, smat_stream stream);
smat_status smat_spmm_plan_destroy(smat_plan plan);

/* ------------------------------------------------------------------------ */
/* Precision modes: operand packing and the FP64 error reduction           */
/* ------------------------------------------------------------------------ */
/*
 * FP32 block values (smat_bcsr_scatter_values) -> operand A of `dtype`:
 *  SMAT_F16 / SMAT_BF16: uint16[nnzb*16*bk], round to nearest even;
 *  SMAT_FP8_E4M3: uint8[nnzb*16*bk] = e4m3(x / s), s = a_scale[r] =
 *    max|x| over block row r / 448 (1 for an all-zero row), round to nearest
 *    even, clamped to +-448;
 *  SMAT_SPLIT_F16: uint16[2*nnzb*16*bk] (hi, then lo) of x / s with
 *    s = a_scale[r] = 2^(e-14), e the exponent of max|x| (max|x|/s in
 *    [2^14, 2^15)), 1 for an all-zero row.
 */
smat_status smat_pack_a(const smat_bcsr_desc* A, const float* values_f32, int32_t dtype,
                        void* a_out, float* a_scale, smat_stream stream);
/*
 * FP32 B (K x N, row-major, leading dimension ldb) -> operand B of `dtype`,
 * row-major with leading dimension N; per-column scales as smat_pack_a uses
 * per block row (b_scale[N], FP8 and split only).  Split output: hi [K*N]
 * then lo [K*N].
 */
smat_status smat_pack_b(const float* B, int64_t K, int64_t N, int64_t ldb, int32_t dtype,
                        void* b_out, float* b_scale, smat_stream stream);
/*
 * Sums in FP64 on the GPU: out[0] = sum (C - R)^2, out[1] = sum R^2 over a
 * rows x N window; C FP32 (ldc), R FP64 (ldr).  out: device double[2];
 * work: device double[2 * 1024].  rel. Frobenius error = sqrt(out0 / out1).
 */
smat_status smat_error_sums(const float* C, int64_t ldc, const double* R, int64_t ldr, int64_t rows,
                            int64_t N, double* out, double* work, smat_stream stream);

/* ------------------------------------------------------------------------ */
/* Batches: many SpMMs in one launch (one HIP graph when captured)           */
/* ------------------------------------------------------------------------ */
/*
 * A batch runs `count` planned SpMMs (each with its own block width, N-tile,
 * work schedule and operands).  Two ways to run it:
 *   SMAT_BATCH_PERSISTENT: one persistent queue kernel per N-tile width
 *     (concurrent), drawing (matrix, work item, N-tile) tasks from per-XCD
 *     device counters, then one kernel reducing the split rows of every
 *     matrix.  Plans (dense or mixed 2:4) must share one 16-bit input type
 *     (F16 or BF16), use row-major B staged through LDS and FP32 C.
 *   SMAT_BATCH_BRANCHES: each matrix's own call (its specialised kernels and
 *     work order), the matrices spread over 4 streams balanced by estimated
 *     work, longest first: parallel branches of the captured graph.  Any plan.
 * SMAT_BATCH_AUTO is SMAT_BATCH_BRANCHES.  Entries may share a plan only if
 * it has no split rows (split rows use the plan's scratch).  The plans and
 * every operand buffer must outlive the batch (their addresses are baked
 * in).  No reference counterpart: the reference launches one kernel per
 * matrix (tester.h:944-1011).
 */
typedef struct smat_batch_t* smat_batch;
#define SMAT_BATCH_AUTO 0
#define SMAT_BATCH_PERSISTENT 1
#define SMAT_BATCH_BRANCHES 2
smat_status smat_batch_create_ex(int32_t count, const smat_plan* plans, const smat_spmm_operands* ops,
                                 int32_t mode, smat_stream stream, smat_batch* out);
/* smat_batch_create_ex with SMAT_BATCH_AUTO. */
smat_status smat_batch_create(int32_t count, const smat_plan* plans, const smat_spmm_operands* ops,
                              smat_stream stream, smat_batch* out);
/* Captures the batch's launches (no memset nodes) into a HIP graph on
 * `stream` (not the default stream); later smat_batch_execute calls replay
 * the graph.  smat_batch_info: tasks = wave tasks of all plans, split_rows =
 * rows reduced from partial tiles, grid = resident workgroups of the largest
 * persistent group (persistent) or branches (branches). */
smat_status smat_batch_capture(smat_batch batch, smat_stream stream);
smat_status smat_batch_execute(smat_batch batch, smat_stream stream);
smat_status smat_batch_info(smat_batch batch, int64_t* tasks, int64_t* split_rows, int32_t* grid,
                            int32_t* captured);
smat_status smat_batch_destroy(smat_batch batch);

/*
 * SpMM that allocates nothing.  No counterpart in the reference (its
 * launchers take caller-allocated arrays and keep no work decomposition); it
 * is for callers that cannot allocate device memory: serving loops and
 * harnesses that forbid hipMalloc.  One call does what smat_spmm_plan_create
 * (A, N, opts) followed by smat_spmm_execute_ex(ops) and
 * smat_spmm_plan_destroy do, with the same work decomposition and therefore
 * the same C bits, but keeps every device array of the plan (the planner's
 * counters and scan space, work items, split-row records, partial tiles of
 * split block rows) in a static device scratch of SMAT_NOALLOC_SCRATCH_BYTES
 * that each GPU holds from the library's first use on it until the process
 * ends.  A call that needs more returns SMAT_ERR_INVALID before any SpMM
 * kernel is launched.  The scratch keeps the last call's contents; nothing
 * in it has to start at zero (every array is written before it is read).
 * The call reads plan sizes back, so it synchronizes `stream` once, and then
 * returns with the SpMM kernels enqueued on `stream`.  Calls on one GPU share
 * its scratch: calls that overlap in time (different streams or threads
 * without synchronization between them) race; calls in order on one stream
 * do not.
 */
#define SMAT_NOALLOC_SCRATCH_BYTES (64ull << 20)
smat_status smat_spmm_noalloc(const smat_bcsr_desc* A, int64_t N, const smat_spmm_opts* opts,
                              const smat_spmm_operands* ops, smat_stream stream);

/* Number of work items and split rows of a plan (for reports and tests). */
smat_status smat_spmm_plan_info(smat_plan plan, int64_t* n_items, int64_t* n_split_rows,
                                int64_t* scratch_bytes);

/* ------------------------------------------------------------------------ */
/* Learned tuner: matrix statistics, features and the MLP cost model         */
/* ------------------------------------------------------------------------ */
/*
 * No reference counterpart: the reference tunes by hand or by an exhaustive
 * threshold search (reorder_matrix.cpp:17-40).  The model scores candidate
 * configurations (reorder threshold, block width, dense / 2:4 kernel, N-tile)
 * of a matrix from its features; the exhaustive sweep stays the fallback.
 */
typedef struct {
  int64_t rows, cols, nnz;
  int64_t max_row;          /* longest row */
  int64_t row_hist[12];     /* rows of length 0, 1, 2-3, 4-7, ..., 512-1023, >= 1024 */
  int64_t sum_abs_offset;   /* sum over entries of |row - col| */
  int64_t near_diag;        /* entries with |row - col| <= 64 */
  int64_t nnzb[6];          /* stored blocks of the 16 x bk BCSR, bk = 1, 2, 4, 8, 16, 32 */
  int64_t nnzb16;           /* stored 16 x 16 blocks */
  int64_t blocks24;         /* of those, blocks that are 2:4 in every aligned 1x4 group */
} smat_matrix_stats_t;

/* Exact integer statistics of a device CSR (GPU; one host sync). */
smat_status smat_matrix_stats(int64_t M, int64_t K, int64_t nnz, const int32_t* rowptr,
                              const int32_t* colidx, const uint16_t* values,
                              smat_matrix_stats_t* out_host, smat_stream stream);

#define SMAT_TUNER_NF 26
/*
 * Feature vector (host): log2 rows, log2 nnz, log2 mean row length,
 * log2 longest / mean row, the 12 row-length bucket fractions, mean |row - col|
 * / rows, fraction of entries within 64 of the diagonal, log2 fill of the
 * 16 x bk blocks for bk = 1..32, fraction of 2:4 blocks, log2 N.
 */
smat_status smat_tuner_features(const smat_matrix_stats_t* stats, int64_t N, float* out_host);

/* Dense layer: Y[b][o] = act(sum_i X[b][i] W[o][i] + bias[o]), act = ReLU if relu. */
smat_status smat_linear_forward(const float* X, int64_t B, int32_t n_in, const float* W,
                                const float* bias, int32_t n_out, int32_t relu, float* Y,
                                smat_stream stream);
/* Its gradients from dY (Y: the layer's output, for the ReLU mask): dW, db and,
 * when dX is not NULL, dX.  Deterministic (fixed summation order). */
smat_status smat_linear_backward(const float* X, const float* Y, const float* dY, int64_t B,
                                 int32_t n_in, const float* W, int32_t n_out, int32_t relu,
                                 float* dX, float* dW, float* db, smat_stream stream);
/* dpred = 2 (pred - y) / B and loss[0] = mean (pred - y)^2 (device scalars). */
smat_status smat_mse_grad(const float* pred, const float* y, int64_t B, float* dpred,
                          float* loss, smat_stream stream);
/* One bias-corrected Adam step t >= 1 on n parameters. */
smat_status smat_adam_step(float* p, const float* g, float* m, float* v, int64_t n, float lr,
                           float beta1, float beta2, float eps, int32_t t, smat_stream stream);
/*
 * Inference: scores the n_cfg candidates (cfg_enc: [n_cfg][n_enc]) of one
 * matrix (feats: [n_feat]) with the network in -> hidden -> hidden -> 1
 * (ReLU; inputs standardised with mu / sigma) and writes scores[n_cfg] and
 * best[0], the lowest-scoring candidate with avail != 0 (-1 if none).
 * Device arrays; n_feat + n_enc <= 64, hidden <= 128, n_cfg <= 256.
 */
smat_status smat_tuner_predict(const float* feats, int32_t n_feat, const float* cfg_enc, int32_t n_enc,
                               const uint8_t* avail, int32_t n_cfg, const float* mu, const float* sigma,
                               const float* W1, const float* b1, const float* W2, const float* b2,
                               const float* W3, const float* b3, int32_t hidden, float* scores,
                               int32_t* best, smat_stream stream);

/* ------------------------------------------------------------------------ */
/* Reference drop-in launchers (src/cuda_hgemm/src/main.cu:10-77)           */
/* ------------------------------------------------------------------------ */
/*
 * HGEMM_FUNC_SPARSE2 (main.cu:25-28) — replaces mmaOBTKernel
 * (src/cuda_hgemm/src/mma/mmaOBT.cu:160-170): plain dense-block BCSR SpMM,
 * 16x16 blocks.  B[n*K + k] (the reference layout), C row-major M x N FP16.
 * Block values are read through relativeBlockIndexMapping exactly as the
 * reference does; blockInfo is ignored (as in mmaOBT.cu).  Accumulates in
 * FP32 (the reference accumulates in FP16, ptx.h:20-25).
 */
void mmaOBTKernel(uint16_t* bcsrValuesA, int* bcsrRowPtrA, int* bcsrColIdxA, uint16_t* B,
                  uint16_t* C, size_t M, size_t N, size_t K, size_t nonzeroBlocks,
                  int* blockInfo, int* relativeBlockIndexMapping);

/*
 * HGEMM_FUNC_SPARSE24_2 (main.cu:20-24) — replaces mmaOBTSKernel
 * (src/cuda_hgemm/src/mma/mmaOBTS.cu:18-215): mixed 16x16 BCSR SpMM; blocks
 * with blockInfo 2 are read from bcsrValuesA, blocks with blockInfo 1 from
 * the compressed (sparseMatrixA, metadata) pair in SMAT_24_REFERENCE_K16
 * layout.
 */
void mmaOBTSKernel(uint16_t* bcsrValuesA, int* bcsrRowPtrA, int* bcsrColIdxA, char* metadata,
                   uint16_t* sparseMatrixA, uint16_t* B, uint16_t* C, size_t M, size_t N,
                   size_t K, size_t nonzeroBlocks, int* blockInfo,
                   int* relativeBlockIndexMapping);

/*
 * Replaces preprocessing_mmaSTKernel (main.cu:61-64, mmaST.cu:316-326):
 * compresses every blockInfo==1 block of a 16x16 BCSR into the reference
 * k=16 layout.  K here is the column count of A (the reference passes the
 * harness --K, tester.h:360-364, which is wrong whenever A is not K wide).
 */
void preprocessing_mmaSTKernel(uint16_t* bcsrValuesA, char* metadata, uint16_t* sparseMatrixA,
                               size_t M, size_t N, size_t K, size_t nonzeroBlocks,
                               int* blockInfo, int* relativeBlockIndexMapping);

/*
 * HGEMM_FUNC_SPARSE2 (main.cu:25-28) — replaces mmaOBTKernel_large
 * (src/cuda_hgemm/src/mma/mmaOBT_large.cu:160-170), which is a copy of
 * mmaOBTKernel still on 16x16 blocks: same behaviour as mmaOBTKernel.
 */
void mmaOBTKernel_large(uint16_t* bcsrValuesA, int* bcsrRowPtrA, int* bcsrColIdxA, uint16_t* B,
                        uint16_t* C, size_t M, size_t N, size_t K, size_t nonzeroBlocks,
                        int* blockInfo, int* relativeBlockIndexMapping);

/*
 * HGEMM_FUNC_SPARSE24_2 (main.cu:20-24) — replaces mmaOBTSKernel_large
 * (src/cuda_hgemm/src/mma/mmaOBTS_large.cu:19-266), the reference's evaluated
 * kernel (main.cu:306-308): mixed BCSR SpMM on 16x32 blocks (K a multiple of
 * 32; blockInfo / relativeBlockIndexMapping over the k=32 grid).  blockInfo 2
 * blocks are read from bcsrValuesA (16 x 32 row-major), blockInfo 1 blocks
 * from the (sparseMatrixA, metadata) pair of preprocessing_mmaSTKernel_large.
 */
void mmaOBTSKernel_large(uint16_t* bcsrValuesA, int* bcsrRowPtrA, int* bcsrColIdxA, char* metadata,
                         uint16_t* sparseMatrixA, uint16_t* B, uint16_t* C, size_t M, size_t N,
                         size_t K, size_t nonzeroBlocks, int* blockInfo,
                         int* relativeBlockIndexMapping);

/*
 * HGEMM_FUNC_SPLIT (main.cu:35-41) — replaces mmaOBTSKernel_large_separate
 * (src/cuda_hgemm/src/mma/mmaOBTS_large_separate.cu:18-275, evaluated by
 * tester.h:552-606 from main.cu:309-311): 16x32 BCSR SpMM over two block
 * lists, the 2:4 list (bcsrRowPtrA_sparse / bcsrColIdxA_sparse, as written by
 * smat_bcsr_filter) read from the (sparseMatrixA, metadata) pair of
 * preprocessing_mmaSTKernel_large, and the dense list (bcsrRowPtrA_dense /
 * bcsrColIdxA_dense) read from bcsrValuesA_sparse at each block's
 * relativeBlockIndexMapping slot, as the reference kernel reads its first
 * values argument.  The reference's 2:4 loop is commented out
 * (mmaOBTS_large_separate.cu:53-154), so it returns only the dense blocks'
 * share; this computes the whole product.  bcsrValuesA_dense, nonzeroBlocks
 * and blockInfo are not read.
 */
void mmaOBTSKernel_large_separate(uint16_t* bcsrValuesA_sparse, int* bcsrRowPtrA_sparse,
                                  int* bcsrColIdxA_sparse, uint16_t* bcsrValuesA_dense,
                                  int* bcsrRowPtrA_dense, int* bcsrColIdxA_dense, char* metadata,
                                  uint16_t* sparseMatrixA, uint16_t* B, uint16_t* C, size_t M, size_t N,
                                  size_t K, size_t nonzeroBlocks, int* blockInfo,
                                  int* relativeBlockIndexMapping);

/*
 * Replaces preprocessing_mmaSTKernel_large (main.cu:64-68, mmaST_large.cu:
 * 324-335): compresses every blockInfo==1 16x32 block.  The reference's k=32
 * producer and consumer disagree on metadata byte offsets (mmaST_large.cu:
 * 143-146 vs mmaOBTS_large.cu:111-115), so no reference byte layout exists;
 * each block is stored as its two 16x16 halves (columns 0-15, 16-31) in the
 * SMAT_24_REFERENCE_K16 layout: 256 halves of sparseMatrixA and 64 bytes of
 * metadata per block (the sizes the reference allocates, tester.h:318-320).
 */
void preprocessing_mmaSTKernel_large(uint16_t* bcsrValuesA, char* metadata, uint16_t* sparseMatrixA,
                                     size_t M, size_t N, size_t K, size_t nonzeroBlocks,
                                     int* blockInfo, int* relativeBlockIndexMapping);

/*
 * HGEMM_FUNC_SPARSE2 (main.cu:25-28) on merged 64 x 64 tiles — replaces
 * mmaOBTKernel_tiled (src/cuda_hgemm/src/mma/mmaOBT_tiled.cu:23-218), fed by
 * tester.h:279-286 with the merged arrays of bcsrBlocking: bcsrValuesA,
 * bcsrRowPtrA (one entry per tile row), bcsrColIdxA (tile first columns) and
 * relativeBlockIndexMapping over the tile grid, as written by
 * smat_tiles_from_bcsr (bk 16) and smat_tiles_grid_maps.  Tile values are
 * read through relativeBlockIndexMapping like the reference; every
 * sub-block is multiplied as dense (there is no tileInfo argument).  B[n*K +
 * k], C row-major M x N FP16.  nonzeroBlocks and blockInfo are not read.
 * The arrays are checked on the GPU first (rowptr monotone from 0, column
 * starts inside K, relative indices inside [0, rowptr[last])); a failure is
 * recorded for smat_last_error() and C is left alone.
 */
void mmaOBTKernel_tiled(uint16_t* bcsrValuesA, int* bcsrRowPtrA, int* bcsrColIdxA, uint16_t* B,
                        uint16_t* C, size_t M, size_t N, size_t K, size_t nonzeroBlocks,
                        int* blockInfo, int* relativeBlockIndexMapping);

/*
 * Replaces preprocessing_mmaOBTSKernel_tiled_large (main.cu:74-77,
 * mmaOBTS_tiled_large.cu:21-150, 416-429): compresses every tileInfo 1 sub-block of
 * the merged 64 x 128 tiles in bcsrValuesA into (sparseMatrixA, metadata) in
 * the SMAT_24_SMFMAC layout of 16 x 32 blocks, indexed by sub-block
 * (16*t + sb): 256 halves and 64 bytes per sub-block, within the sizes the
 * reference allocates (tester.h:385-394).  The reference copies dense
 * sub-blocks into sparseMatrixA as well; mmaOBTSKernel_tiled_large reads
 * those from bcsrValuesA.  tileInfo is indexed by stored tile, as
 * bcsrBlocking writes it (matrix.h:1081); the tile count is taken from
 * relativeBlockIndexMapping (largest index + 1).  blockInfo is not read.
 */
void preprocessing_mmaOBTSKernel_tiled_large(uint16_t* bcsrValuesA, char* metadata,
                                             uint16_t* sparseMatrixA, size_t M, size_t N, size_t K,
                                             size_t nonzeroBlocks, int* blockInfo,
                                             int* relativeBlockIndexMapping, int* tileInfo);

/*
 * HGEMM_FUNC_SPARSE24_2_tiled (main.cu:29-33) — replaces
 * mmaOBTSKernel_tiled_large (src/cuda_hgemm/src/mma/mmaOBTS_tiled_large.cu:
 * 152-449): mixed SpMM over merged 64 x 128 tiles (bk 32).  Per 16-row
 * sub-row of a tile and 64-column window: two tileInfo 1 sub-blocks run on
 * one SMFMAC from the pair of preprocessing_mmaOBTSKernel_tiled_large; a
 * window with a tileInfo 2 sub-block, or with one non-empty sub-block, runs
 * on MFMA from bcsrValuesA; tileInfo 0 sub-blocks are skipped.  Arrays,
 * layouts and checks as mmaOBTKernel_tiled.
 */
void mmaOBTSKernel_tiled_large(uint16_t* bcsrValuesA, int* bcsrRowPtrA, int* bcsrColIdxA, char* metadata,
                               uint16_t* sparseMatrixA, uint16_t* B, uint16_t* C, size_t M, size_t N,
                               size_t K, size_t nonzeroBlocks, int* blockInfo,
                               int* relativeBlockIndexMapping, int* tileInfo);

#ifdef __cplusplus
}
#endif

#endif /* SMAT_H_ */
