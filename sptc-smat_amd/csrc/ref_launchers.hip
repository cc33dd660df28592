// Reference drop-in launchers: the exact names, parameter order and meaning of
// the reference's kernel entry points (src/cuda_hgemm/src/main.cu:10-77), on
// MI355X kernels.  Differences, all deliberate:
//  * accumulation is FP32 (the reference's HMMA accumulates in FP16,
//    src/cuda_hgemm/src/common/ptx.h:20-41);
//  * a failure is recorded for smat_last_error() instead of exit();
//  * the SpMM launchers first check on the GPU that their arrays are the
//    BCSR csrToBcsr writes (relativeBlockIndexMapping points every block at
//    its own position, and for the mixed ones every blockInfo 1 block's
//    compressed pair holds exactly its values in bcsrValuesA); then they run
//    the planned dense-block kernel of smat_spmm_execute on the reference's
//    B layout, planned per call in the library's static scratch
//    (smat_spmm_noalloc; one host sync per call for the check and two for
//    the plan), with B first transposed into a library-owned row-major copy
//    (grown on demand, kept per device until the process ends).  Otherwise (or when the plan does not fit
//    the scratch) they run k_ref_spmm: one wave per (block row, 64-column
//    tile), blocks read through relativeBlockIndexMapping and blockInfo like
//    the reference kernels (mmaOBT.cu:56-61).
#include <algorithm>
#include <mutex>

#include "smat_internal.h"

namespace smat {
smat_status compress_ref16(int64_t nnzb, const uint16_t* val, const int32_t* tags, uint16_t* sp,
                           uint8_t* meta, hipStream_t s, bool* overflowed);
smat_status tiles_spmm_rel(int bk, bool mixed, int64_t M, int64_t K, int64_t N, const int32_t* rowptr,
                           const int32_t* colidx, const int32_t* relmap, const uint16_t* values,
                           const int32_t* info, const uint16_t* values24, const uint16_t* index24,
                           const uint16_t* B, int64_t ldb, uint16_t* C, int64_t ldc, hipStream_t s);
}

namespace {

constexpr int kWaves = 4;
constexpr int kNsub = 4;  // 64 columns per wave

// Expands one reference-layout 2:4 "lane" (4 halves + 1 metadata byte,
// mmaST.cu:98-137) into the 8 dense halves it stands for.
__device__ inline smat_h8 expand_ref24(const uint16_t* __restrict__ sp4, uint8_t m) {
  uint16_t v[4];
  *reinterpret_cast<uint2*>(v) = *reinterpret_cast<const uint2*>(sp4);
  uint16_t out[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int j = 0; j < 2; j++)
#pragma unroll
    for (int cur = 0; cur < 2; cur++) {
      const int pos = (m >> (6 - (2 * (1 - cur) + 4 * j))) & 3;
      const uint16_t x = v[2 * j + cur];
      if ((x & 0x7fff) != 0) out[4 * j + pos] = x;
    }
  return __builtin_bit_cast(smat_h8, *reinterpret_cast<uint4*>(out));
}

// How a list's blocks are read: every block dense (bcsrValuesA), every block
// 2:4 (the compressed arrays), or per block as blockInfo says.
enum RefInfo { kAllDense = 0, kAll24 = 1, kPerBlock = 2 };

// Accumulates the blocks of block row `brow` listed in rowptr/colidx into the
// wave's 16 x (16*kNsub) tile.  BKW = 16: 16x16 blocks, two per MFMA step;
// BKW = 32: 16x32 blocks (the reference's k=32 "large" BCSR), one per step,
// whose 2:4 form is two k=16 halves (see preprocessing_mmaSTKernel_large).
// Blocks are found through relativeBlockIndexMapping like the reference
// kernels (mmaOBT.cu:56-61).
template <int BKW, RefInfo INFO>
__device__ inline void ref_list(smat_f4 (&acc)[kNsub], const uint16_t* __restrict__ vals,
                                const int* __restrict__ rowptr, const int* __restrict__ colidx,
                                const uint8_t* __restrict__ meta, const uint16_t* __restrict__ sparseA,
                                const uint16_t* __restrict__ B, int64_t brow, int64_t n0, int64_t N, int64_t K,
                                const int* __restrict__ blockInfo, const int* __restrict__ relMap, int lane) {
  const int64_t ncr = K / BKW;
  const int r16 = lane & 15, grp = lane >> 4, sub = lane >> 5, half = grp & 1;
  constexpr int PER = 32 / BKW;  // blocks per MFMA step
  const int pb = rowptr[brow], pe = rowptr[brow + 1];
  for (int p = pb; p < pe; p += PER) {
    const int blk = BKW == 16 ? p + sub : p;
    const bool valid = blk < pe;
    smat_h8 a = smat_h8{0, 0, 0, 0, 0, 0, 0, 0};
    int64_t c0 = 0;
    if (valid) {
      const int col = colidx[blk];
      const int64_t g = brow * ncr + col / BKW;
      const int64_t rel = relMap[g];
      const int info = INFO == kPerBlock ? blockInfo[g] : INFO == kAll24 ? 1 : 2;
      // k-slots of this lane: columns 8*half .. of block `blk` (BKW = 16), or
      // columns 8*grp .. of the 16x32 block, i.e. of its 16-column half grp/2
      const int64_t sb = BKW == 16 ? rel : rel * 2 + (grp >> 1);  // 16x16 (sub-)block of the 2:4 arrays
      if (info == 1) {
        const int q = r16 * 2 + half;
        a = expand_ref24(sparseA + sb * 128 + q * 4, meta[sb * 32 + q]);
      } else if (BKW == 16) {
        a = __builtin_bit_cast(smat_h8, *reinterpret_cast<const uint4*>(vals + rel * 256 + r16 * 16 + half * 8));
      } else {
        a = __builtin_bit_cast(smat_h8, *reinterpret_cast<const uint4*>(vals + rel * 512 + r16 * 32 + grp * 8));
      }
      c0 = col + (BKW == 16 ? half * 8 : grp * 8);
    }
#pragma unroll
    for (int j = 0; j < kNsub; j++) {
      const int64_t n = n0 + 16 * j + r16;
      smat_h8 b = smat_h8{0, 0, 0, 0, 0, 0, 0, 0};
      if (valid && n < N)
        b = __builtin_bit_cast(smat_h8, *reinterpret_cast<const uint4*>(B + n * K + c0));
      acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, acc[j], 0, 0, 0);
    }
  }
}

// C[M x N] (FP16, row-major, ldc = N) = A * B with B[n*K + k]; one wave per
// (block row, 64-column tile).  SPLIT: the block row's 2:4 list
// (rowptr2/colidx2, read from the compressed arrays) and then its dense list
// (rowptr/colidx) go into one FP32 tile; otherwise the one list rowptr/colidx
// is read as INFO says.
template <RefInfo INFO, int BKW, bool SPLIT>
__global__ __launch_bounds__(256) void k_ref_spmm(const uint16_t* __restrict__ vals,
                                                  const int* __restrict__ rowptr,
                                                  const int* __restrict__ colidx,
                                                  const int* __restrict__ rowptr2,
                                                  const int* __restrict__ colidx2,
                                                  const uint8_t* __restrict__ meta,
                                                  const uint16_t* __restrict__ sparseA,
                                                  const uint16_t* __restrict__ B,
                                                  _Float16* __restrict__ C, int64_t M, int64_t N,
                                                  int64_t K, const int* __restrict__ blockInfo,
                                                  const int* __restrict__ relMap) {
  const int lane = threadIdx.x & 63;
  const int64_t task = (int64_t)blockIdx.x * kWaves + (threadIdx.x >> 6);
  const int64_t ntiles = (N + 16 * kNsub - 1) / (16 * kNsub);
  const int64_t brow = task / ntiles;
  if (brow >= M / 16) return;
  const int64_t n0 = (task % ntiles) * 16 * kNsub;
  const int r16 = lane & 15, grp = lane >> 4;

  smat_f4 acc[kNsub];
#pragma unroll
  for (int j = 0; j < kNsub; j++) acc[j] = smat_f4{0.f, 0.f, 0.f, 0.f};
  if (SPLIT)
    ref_list<BKW, kAll24>(acc, vals, rowptr2, colidx2, meta, sparseA, B, brow, n0, N, K, blockInfo, relMap, lane);
  ref_list<BKW, INFO>(acc, vals, rowptr, colidx, meta, sparseA, B, brow, n0, N, K, blockInfo, relMap, lane);
#pragma unroll
  for (int j = 0; j < kNsub; j++) {
    const int64_t n = n0 + 16 * j + r16;
    if (n >= N) continue;
#pragma unroll
    for (int i = 0; i < 4; i++) C[(brow * 16 + 4 * grp + i) * N + n] = (_Float16)acc[j][i];
  }
}

// preprocessing_mmaSTKernel over the dense block grid: one thread per grid
// cell with blockInfo == 1 gathers its block into a compact list.
__global__ void k_gather_tags(int64_t ncells, const int* __restrict__ info, const int* __restrict__ rel,
                              int32_t* __restrict__ tags_by_block, int64_t nnzb) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ncells) return;
  const int r = rel[g];
  if (r >= 0 && r < nnzb) tags_by_block[r] = info[g];
}

// 16x32 blocks -> pairs of 16x16 blocks (columns 0-15, 16-31), tags copied
// to both halves: the k=32 compressor runs the k=16 one on the halves.
__global__ void k_split32(int64_t nnzb, const uint16_t* __restrict__ vals, const int32_t* __restrict__ tags,
                          uint16_t* __restrict__ halves, int32_t* __restrict__ tags2) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // (block, row)
  if (t >= nnzb * 16) return;
  const int64_t b = t / 16;
  const int r = (int)(t % 16);
  const uint4* src = reinterpret_cast<const uint4*>(vals + b * 512 + r * 32);
  uint4* d0 = reinterpret_cast<uint4*>(halves + (2 * b) * 256 + r * 16);
  uint4* d1 = reinterpret_cast<uint4*>(halves + (2 * b + 1) * 256 + r * 16);
  d0[0] = src[0];
  d0[1] = src[1];
  d1[0] = src[2];
  d1[1] = src[3];
  if (r == 0) tags2[2 * b] = tags2[2 * b + 1] = tags[b];
}

// Fast-path check, one thread per block p: rowptr bounds, the block's
// column start (a multiple of BKW inside K), relativeBlockIndexMapping of its
// grid cell == p and, with MIXED, a blockInfo 1 block's compressed 2:4 pair
// (reference k=16 layout, two halves per 16x32 block) expanding to its values
// in `vals` (zeros of either sign count as equal).  Any mismatch sets *flag.
template <int BKW, bool MIXED>
__global__ void k_check_planned(int64_t nbrows, int64_t nnzb, int64_t K, const int* __restrict__ rowptr,
                                const int* __restrict__ colidx, const uint16_t* __restrict__ vals,
                                const int* __restrict__ blockInfo, const int* __restrict__ relMap,
                                const uint8_t* __restrict__ meta, const uint16_t* __restrict__ sparseA,
                                int* __restrict__ flag) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0 && (rowptr[0] != 0 || rowptr[nbrows] != nnzb)) atomicOr(flag, 1);
  if (p >= nnzb) return;
  // block row of p: the last r with rowptr[r] <= p
  int64_t lo = 0, hi = nbrows;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (rowptr[mid] <= p) lo = mid;
    else hi = mid;
  }
  const int col = colidx[p];
  const int64_t ncr = K / BKW;
  if (col < 0 || col % BKW != 0 || col / BKW >= ncr) {
    atomicOr(flag, 1);
    return;
  }
  const int64_t g = lo * ncr + col / BKW;
  if (relMap[g] != p) {
    atomicOr(flag, 1);
    return;
  }
  if (!MIXED || blockInfo[g] != 1) return;
  for (int h = 0; h < BKW / 16; h++) {
    const int64_t sb = BKW == 16 ? p : 2 * p + h;  // 16x16 (half-)block of the compressed arrays
    for (int q = 0; q < 32; q++) {
      const smat_h8 e = expand_ref24(sparseA + sb * 128 + q * 4, meta[sb * 32 + q]);
      const uint4 eb = __builtin_bit_cast(uint4, e);
      const uint16_t* ev = reinterpret_cast<const uint16_t*>(&eb);
      const int r = q / 2, c0 = (q % 2) * 8;
      for (int j = 0; j < 8; j++) {
        const uint16_t d = vals[p * 16 * BKW + r * BKW + h * 16 + c0 + j];
        if (d != ev[j] && ((d | ev[j]) & 0x7fff) != 0) {
          atomicOr(flag, 1);
          return;
        }
      }
    }
  }
}

// B in the reference's layout (element (k, n) at n*K + k) -> row-major K x N
// with row pitch ldo (k*ldo + n), 64 x 64 tiles through LDS: 16-byte reads along k, 2-byte
// writes along n coalesced per wave row.
__global__ __launch_bounds__(256) void k_transpose_b(const uint16_t* __restrict__ bt, uint16_t* __restrict__ b,
                                                     int64_t K, int64_t N, int64_t ldo) {
  __shared__ uint16_t tile[64][65];
  const int64_t k0 = (int64_t)blockIdx.x * 64, n0 = (int64_t)blockIdx.y * 64;
  const int t = threadIdx.x;
  // read: 64 rows n of 64 halves (k), 8 halves per thread
  for (int i = t; i < 64 * 8; i += 256) {
    const int nn = i / 8, kk = (i % 8) * 8;
    const int64_t n = n0 + nn, k = k0 + kk;
    if (n < N && k + 8 <= K) {
      const uint4 v = *reinterpret_cast<const uint4*>(bt + n * K + k);
      const uint16_t* e = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
      for (int j = 0; j < 8; j++) tile[kk + j][nn] = e[j];
    } else if (n < N) {
      for (int j = 0; j < 8 && k + j < K; j++) tile[kk + j][nn] = bt[n * K + k + j];
    }
  }
  __syncthreads();
  for (int i = t; i < 64 * 64; i += 256) {
    const int kk = i / 64, nn = i % 64;
    const int64_t k = k0 + kk, n = n0 + nn;
    if (k < K && n < N) b[k * ldo + n] = tile[kk][nn];
  }
}

// Library-owned row-major copy of B for the launchers, per device, grown on
// demand and kept until the process ends.
constexpr int kMaxDev = 64;
std::mutex g_bstage_mu;
void* g_bstage[kMaxDev] = {};
size_t g_bstage_bytes[kMaxDev] = {};

uint16_t* staged_b(size_t bytes) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDev) {
    (void)hipGetLastError();
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_bstage_mu);
  if (g_bstage_bytes[dev] < bytes) {
    if (g_bstage[dev]) (void)hipFree(g_bstage[dev]);
    g_bstage[dev] = nullptr;
    g_bstage_bytes[dev] = 0;
    if (hipMalloc(&g_bstage[dev], bytes) != hipSuccess) {
      (void)hipGetLastError();
      g_bstage[dev] = nullptr;
      return nullptr;
    }
    g_bstage_bytes[dev] = bytes;
  }
  return static_cast<uint16_t*>(g_bstage[dev]);
}

__global__ void k_zero_flag(int* flag) {
  if (threadIdx.x == 0) *flag = 0;
}

// The planned path of the SpMM launchers (see the file comment); false when
// it does not apply and the caller runs k_ref_spmm.
template <int BKW, bool MIXED>
bool planned_spmm(const uint16_t* vals, const int* rowptr, const int* colidx, const uint8_t* meta,
                  const uint16_t* sparseA, const uint16_t* B, uint16_t* C, size_t M, size_t N, size_t K,
                  size_t nnzb, const int* blockInfo, const int* relMap) {
  if (nnzb == 0 || nnzb >= (size_t)INT32_MAX) return false;
  void* base = nullptr;
  if (smat::noalloc_scratch(&base) != SMAT_OK) {
    smat_clear_error();
    return false;
  }
  int* flag = static_cast<int*>(base);
  const int64_t nbrows = (int64_t)M / 16;
  hipLaunchKernelGGL(k_zero_flag, 1, 64, 0, 0, flag);
  hipLaunchKernelGGL((k_check_planned<BKW, MIXED>), (unsigned)((nnzb + 255) / 256), 256, 0, 0, nbrows,
                     (int64_t)nnzb, (int64_t)K, rowptr, colidx, vals, blockInfo, relMap, meta, sparseA, flag);
  int h = 1;
  if (hipGetLastError() != hipSuccess || hipMemcpy(&h, flag, 4, hipMemcpyDeviceToHost) != hipSuccess || h != 0) {
    (void)hipGetLastError();
    return false;
  }
  smat_bcsr_desc d{};
  d.M = (int64_t)M;
  d.K = (int64_t)K;
  d.bm = 16;
  d.bk = BKW;
  d.nbrows = nbrows;
  d.nnzb = (int64_t)nnzb;
  d.rowptr = const_cast<int32_t*>(rowptr);
  d.colidx = const_cast<int32_t*>(colidx);
  d.values = const_cast<uint16_t*>(vals);
  smat_spmm_opts o{};
  o.kernel = SMAT_KERNEL_DENSE;
  o.out_dtype = SMAT_F16;
  o.in_dtype = SMAT_F16;
  smat_spmm_operands ops{};
  ops.c = C;
  ops.ldc = (int64_t)N;
  // B[n*K + k] gathers touch 32 bytes of a 128-byte line per block row
  // segment; a row-major copy (one transpose, 2 x K x N x 2 bytes of traffic)
  // lets the kernel stage whole B lines through LDS
  uint16_t* Brm = (reinterpret_cast<uintptr_t>(B) & 15) == 0 ? staged_b(K * N * sizeof(uint16_t)) : nullptr;
  if (Brm) {
    hipLaunchKernelGGL(k_transpose_b, dim3((unsigned)((K + 63) / 64), (unsigned)((N + 63) / 64)), 256, 0, 0, B, Brm,
                       (int64_t)K, (int64_t)N, (int64_t)N);
    o.b_layout = SMAT_B_ROW_MAJOR;
    ops.b = Brm;
    ops.ldb = (int64_t)N;
  } else {
    o.b_layout = SMAT_B_REFERENCE;
    ops.b = B;
    ops.ldb = (int64_t)K;
  }
  if (smat_spmm_noalloc(&d, (int64_t)N, &o, &ops, nullptr) != SMAT_OK) {
    smat_clear_error();
    return false;
  }
  return true;
}

// Array check of the tiled launchers, one thread per row pointer and per
// tile position: rowptr starts at 0 and never decreases, every column start
// is a multiple of tw inside K, and the relative index of every tile's grid
// cell lies in [0, ntiles), ntiles = rowptr[ntrows].  Any mismatch sets *flag.
__global__ void k_check_tiled(int64_t ntrows, int64_t ntiles, int64_t ntc, int tw, int64_t K,
                              const int* __restrict__ rowptr, const int* __restrict__ colidx,
                              const int* __restrict__ relMap, int* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0 && rowptr[0] != 0) atomicOr(flag, 1);
  if (i < ntrows && rowptr[i] > rowptr[i + 1]) atomicOr(flag, 1);
  if (i >= ntiles) return;
  int64_t lo = 0, hi = ntrows;  // tile row of position i
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (rowptr[mid] <= i) lo = mid;
    else hi = mid;
  }
  const int col = colidx[i];
  if (col < 0 || col % tw != 0 || col >= K) {
    atomicOr(flag, 1);
    return;
  }
  const int r = relMap[lo * ntc + col / tw];
  if (r < 0 || r >= ntiles) atomicOr(flag, 1);
}

// Largest entry of a relative index map (-1 when every cell is empty).
__global__ __launch_bounds__(256) void k_max_rel(int64_t n, const int* __restrict__ rel, int* __restrict__ out) {
  int m = -1;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    m = max(m, rel[i]);
  for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

__global__ void k_set_flag(int* flag, int v) {
  if (threadIdx.x == 0) *flag = v;
}

// The tiled SpMM launchers: checks the arrays (two host syncs), stages B
// row-major and runs the tile kernel through relativeBlockIndexMapping.
void tiled_spmm(const char* who, int bk, bool mixed, const uint16_t* vals, const int* rowptr, const int* colidx,
                const int* relMap, const int* tileInfo, const uint8_t* meta, const uint16_t* sparseA,
                const uint16_t* B, uint16_t* C, size_t M, size_t N, size_t K) {
  const int tw = 4 * bk;
  if (!vals || !rowptr || !colidx || !relMap || !B || !C || (mixed && (!tileInfo || !meta || !sparseA))) {
    smat::set_error(SMAT_ERR_INVALID, "%s: null pointer", who);
    return;
  }
  if (M % 16 || K % bk || N == 0 || (reinterpret_cast<uintptr_t>(B) & 15) != 0) {
    smat::set_error(SMAT_ERR_INVALID, "%s: M must be a multiple of 16, K of %d, N > 0 and B 16-byte aligned "
                    "(M=%zu N=%zu K=%zu)", who, bk, M, N, K);
    return;
  }
  const int64_t ntrows = ((int64_t)M + 63) / 64, ntc = ((int64_t)K + tw - 1) / tw;
  int ntiles = -1;
  if (hipMemcpy(&ntiles, rowptr + ntrows, 4, hipMemcpyDeviceToHost) != hipSuccess) {
    smat::set_error(SMAT_ERR_HIP, "%s: reading bcsrRowPtrA failed", who);
    return;
  }
  smat::DevBuf flag;
  if (flag.alloc(4) != SMAT_OK) return;
  hipLaunchKernelGGL(k_set_flag, 1, 64, 0, 0, flag.as<int>(), ntiles < 0 ? 1 : 0);
  const int64_t nchk = std::max<int64_t>(ntrows, ntiles);
  if (ntiles >= 0)
    hipLaunchKernelGGL(k_check_tiled, (unsigned)((nchk + 255) / 256), 256, 0, 0, ntrows, (int64_t)ntiles, ntc, tw,
                       (int64_t)K, rowptr, colidx, relMap, flag.as<int>());
  int h = 1;
  if (hipGetLastError() != hipSuccess || hipMemcpy(&h, flag.p, 4, hipMemcpyDeviceToHost) != hipSuccess) {
    smat::set_error(SMAT_ERR_HIP, "%s: array check failed to run", who);
    return;
  }
  if (h != 0) {
    smat::set_error(SMAT_ERR_INVALID,
                    "%s: the tile arrays are inconsistent (rowptr must start at 0 and have an entry per tile row, "
                    "column starts must be multiples of %d inside K, relative indices inside [0, %d))",
                    who, tw, ntiles);
    return;
  }
  const int64_t ldo = ((int64_t)N + 7) / 8 * 8;
  uint16_t* Brm = staged_b((size_t)K * (size_t)ldo * sizeof(uint16_t));
  if (!Brm) {
    smat::set_error(SMAT_ERR_ALLOC, "%s: staging B failed", who);
    return;
  }
  hipLaunchKernelGGL(k_transpose_b, dim3((unsigned)((K + 63) / 64), (unsigned)((N + 63) / 64)), 256, 0, 0, B, Brm,
                     (int64_t)K, (int64_t)N, ldo);
  if (hipGetLastError() != hipSuccess) {
    smat::set_error(SMAT_ERR_HIP, "%s: B transpose launch failed", who);
    return;
  }
  (void)smat::tiles_spmm_rel(bk, mixed, (int64_t)M, (int64_t)K, (int64_t)N, rowptr, colidx, relMap, vals,
                             mixed ? tileInfo : nullptr, mixed ? sparseA : nullptr,
                             mixed ? reinterpret_cast<const uint16_t*>(meta) : nullptr, Brm, ldo, C, (int64_t)N,
                             nullptr);
}

bool check_ref_args(const void* a, const void* b, const void* c, size_t M, size_t N, size_t K, const char* who,
                    size_t kw = 16) {
  if (!a || !b || !c) {
    smat::set_error(SMAT_ERR_INVALID, "%s: null pointer", who);
    return false;
  }
  if (M % 16 || K % kw || N == 0) {
    smat::set_error(SMAT_ERR_INVALID, "%s: M must be a multiple of 16, K of %zu, N > 0 (M=%zu N=%zu K=%zu)", who,
                    kw, M, N, K);
    return false;
  }
  return true;
}

}  // namespace

extern "C" {

void mmaOBTKernel(uint16_t* bcsrValuesA, int* bcsrRowPtrA, int* bcsrColIdxA, uint16_t* B, uint16_t* C,
                  size_t M, size_t N, size_t K, size_t nonzeroBlocks, int* blockInfo,
                  int* relativeBlockIndexMapping) {
  (void)blockInfo;
  if (!check_ref_args(bcsrRowPtrA, B, C, M, N, K, "mmaOBTKernel")) return;
  if (!relativeBlockIndexMapping || !bcsrColIdxA || !bcsrValuesA) {
    smat::set_error(SMAT_ERR_INVALID, "mmaOBTKernel: null BCSR array");
    return;
  }
  if (planned_spmm<16, false>(bcsrValuesA, bcsrRowPtrA, bcsrColIdxA, nullptr, nullptr, B, C, M, N, K, nonzeroBlocks,
                              nullptr, relativeBlockIndexMapping))
    return;
  const int64_t ntiles = ((int64_t)N + 63) / 64;
  const int64_t tasks = (int64_t)(M / 16) * ntiles;
  hipLaunchKernelGGL((k_ref_spmm<kAllDense, 16, false>), (unsigned)((tasks + kWaves - 1) / kWaves), 64 * kWaves,
                     0, 0, bcsrValuesA, bcsrRowPtrA, bcsrColIdxA, nullptr, nullptr, nullptr, nullptr, B,
                     reinterpret_cast<_Float16*>(C), (int64_t)M, (int64_t)N, (int64_t)K, nullptr,
                     relativeBlockIndexMapping);
  if (hipGetLastError() != hipSuccess) smat::set_error(SMAT_ERR_HIP, "mmaOBTKernel: launch failed");
}

void mmaOBTSKernel(uint16_t* bcsrValuesA, int* bcsrRowPtrA, int* bcsrColIdxA, char* metadata,
                   uint16_t* sparseMatrixA, uint16_t* B, uint16_t* C, size_t M, size_t N, size_t K,
                   size_t nonzeroBlocks, int* blockInfo, int* relativeBlockIndexMapping) {
  if (!check_ref_args(bcsrRowPtrA, B, C, M, N, K, "mmaOBTSKernel")) return;
  if (!relativeBlockIndexMapping || !blockInfo || !bcsrColIdxA || !bcsrValuesA || !metadata ||
      !sparseMatrixA) {
    smat::set_error(SMAT_ERR_INVALID, "mmaOBTSKernel: null BCSR or 2:4 array");
    return;
  }
  if (planned_spmm<16, true>(bcsrValuesA, bcsrRowPtrA, bcsrColIdxA, reinterpret_cast<const uint8_t*>(metadata),
                             sparseMatrixA, B, C, M, N, K, nonzeroBlocks, blockInfo, relativeBlockIndexMapping))
    return;
  const int64_t ntiles = ((int64_t)N + 63) / 64;
  const int64_t tasks = (int64_t)(M / 16) * ntiles;
  hipLaunchKernelGGL((k_ref_spmm<kPerBlock, 16, false>), (unsigned)((tasks + kWaves - 1) / kWaves), 64 * kWaves,
                     0, 0, bcsrValuesA, bcsrRowPtrA, bcsrColIdxA, nullptr, nullptr, reinterpret_cast<const uint8_t*>(metadata),
                     sparseMatrixA, B, reinterpret_cast<_Float16*>(C), (int64_t)M, (int64_t)N, (int64_t)K,
                     blockInfo, relativeBlockIndexMapping);
  if (hipGetLastError() != hipSuccess) smat::set_error(SMAT_ERR_HIP, "mmaOBTSKernel: launch failed");
}

void preprocessing_mmaSTKernel(uint16_t* bcsrValuesA, char* metadata, uint16_t* sparseMatrixA, size_t M,
                               size_t N, size_t K, size_t nonzeroBlocks, int* blockInfo,
                               int* relativeBlockIndexMapping) {
  (void)N;
  if (!bcsrValuesA || !metadata || !sparseMatrixA || !blockInfo || !relativeBlockIndexMapping) {
    smat::set_error(SMAT_ERR_INVALID, "preprocessing_mmaSTKernel: null pointer");
    return;
  }
  if (M % 16 || K % 16) {
    smat::set_error(SMAT_ERR_INVALID, "preprocessing_mmaSTKernel: M and K must be multiples of 16");
    return;
  }
  const int64_t nnzb = (int64_t)nonzeroBlocks;
  const int64_t ncells = (int64_t)(M / 16) * (int64_t)(K / 16);
  smat::DevBuf tags;
  if (tags.alloc((nnzb + 1) * 4) != SMAT_OK) return;
  if (hipMemsetAsync(tags.p, 0, (nnzb + 1) * 4, 0) != hipSuccess) {
    smat::set_error(SMAT_ERR_HIP, "preprocessing_mmaSTKernel: memset failed");
    return;
  }
  if (ncells > 0)
    hipLaunchKernelGGL(k_gather_tags, (unsigned)((ncells + 255) / 256), 256, 0, 0, ncells, blockInfo,
                       relativeBlockIndexMapping, tags.as<int32_t>(), nnzb);
  bool over = false;
  if (smat::compress_ref16(nnzb, bcsrValuesA, tags.as<int32_t>(), sparseMatrixA,
                           reinterpret_cast<uint8_t*>(metadata), 0, &over) != SMAT_OK)
    return;
  if (over)
    smat::set_error(SMAT_ERR_NOT_24,
                    "preprocessing_mmaSTKernel: a blockInfo==1 block holds a 1x4 group with more than two "
                    "nonzeros; the extra values were dropped");
}

// The reference's mmaOBTKernel_large (mmaOBT_large.cu:160-170) is a copy of
// mmaOBTKernel that still works on 16x16 blocks (MMA_K 16), so this is too.
void mmaOBTKernel_large(uint16_t* bcsrValuesA, int* bcsrRowPtrA, int* bcsrColIdxA, uint16_t* B, uint16_t* C,
                        size_t M, size_t N, size_t K, size_t nonzeroBlocks, int* blockInfo,
                        int* relativeBlockIndexMapping) {
  mmaOBTKernel(bcsrValuesA, bcsrRowPtrA, bcsrColIdxA, B, C, M, N, K, nonzeroBlocks, blockInfo,
               relativeBlockIndexMapping);
}

// Mixed 16x32 BCSR SpMM (mmaOBTS_large.cu:256-266): blockInfo 2 blocks from
// bcsrValuesA (16 x 32 row-major), blockInfo 1 blocks from the compressed
// pair written by preprocessing_mmaSTKernel_large.
void mmaOBTSKernel_large(uint16_t* bcsrValuesA, int* bcsrRowPtrA, int* bcsrColIdxA, char* metadata,
                         uint16_t* sparseMatrixA, uint16_t* B, uint16_t* C, size_t M, size_t N, size_t K,
                         size_t nonzeroBlocks, int* blockInfo, int* relativeBlockIndexMapping) {
  if (!check_ref_args(bcsrRowPtrA, B, C, M, N, K, "mmaOBTSKernel_large", 32)) return;
  if (!relativeBlockIndexMapping || !blockInfo || !bcsrColIdxA || !bcsrValuesA || !metadata ||
      !sparseMatrixA) {
    smat::set_error(SMAT_ERR_INVALID, "mmaOBTSKernel_large: null BCSR or 2:4 array");
    return;
  }
  if (planned_spmm<32, true>(bcsrValuesA, bcsrRowPtrA, bcsrColIdxA, reinterpret_cast<const uint8_t*>(metadata),
                             sparseMatrixA, B, C, M, N, K, nonzeroBlocks, blockInfo, relativeBlockIndexMapping))
    return;
  const int64_t ntiles = ((int64_t)N + 63) / 64;
  const int64_t tasks = (int64_t)(M / 16) * ntiles;
  hipLaunchKernelGGL((k_ref_spmm<kPerBlock, 32, false>), (unsigned)((tasks + kWaves - 1) / kWaves), 64 * kWaves,
                     0, 0, bcsrValuesA, bcsrRowPtrA, bcsrColIdxA, nullptr, nullptr, reinterpret_cast<const uint8_t*>(metadata),
                     sparseMatrixA, B, reinterpret_cast<_Float16*>(C), (int64_t)M, (int64_t)N, (int64_t)K,
                     blockInfo, relativeBlockIndexMapping);
  if (hipGetLastError() != hipSuccess) smat::set_error(SMAT_ERR_HIP, "mmaOBTSKernel_large: launch failed");
}

// Split 16x32 SpMM (mmaOBTS_large_separate.cu:262-275, evaluated by
// tester.h:552-606): the 2:4 blocks come from their own BCSR list
// (bcsrRowPtrA_sparse/bcsrColIdxA_sparse, filterBcsrBlocks) and are read
// from the compressed pair written by preprocessing_mmaSTKernel_large; the
// dense blocks come from the dense list and are read from bcsrValuesA_sparse
// at their relativeBlockIndexMapping slot, as the reference kernel reads its
// first values argument (its only caller passes the same BCSR values for
// both).  The reference has its 2:4 loop commented out
// (mmaOBTS_large_separate.cu:53-154) and so returns the dense blocks' share
// only; this computes the whole product.  bcsrValuesA_dense and blockInfo
// are not read, as in the reference.
void mmaOBTSKernel_large_separate(uint16_t* bcsrValuesA_sparse, int* bcsrRowPtrA_sparse, int* bcsrColIdxA_sparse,
                                  uint16_t* bcsrValuesA_dense, int* bcsrRowPtrA_dense, int* bcsrColIdxA_dense,
                                  char* metadata, uint16_t* sparseMatrixA, uint16_t* B, uint16_t* C, size_t M,
                                  size_t N, size_t K, size_t nonzeroBlocks, int* blockInfo,
                                  int* relativeBlockIndexMapping) {
  (void)bcsrValuesA_dense;
  (void)nonzeroBlocks;
  (void)blockInfo;
  if (!check_ref_args(bcsrRowPtrA_dense, B, C, M, N, K, "mmaOBTSKernel_large_separate", 32)) return;
  if (!relativeBlockIndexMapping || !bcsrRowPtrA_sparse || !bcsrColIdxA_sparse || !bcsrColIdxA_dense ||
      !bcsrValuesA_sparse || !metadata || !sparseMatrixA) {
    smat::set_error(SMAT_ERR_INVALID, "mmaOBTSKernel_large_separate: null BCSR or 2:4 array");
    return;
  }
  const int64_t ntiles = ((int64_t)N + 63) / 64;
  const int64_t tasks = (int64_t)(M / 16) * ntiles;
  hipLaunchKernelGGL((k_ref_spmm<kAllDense, 32, true>), (unsigned)((tasks + kWaves - 1) / kWaves), 64 * kWaves,
                     0, 0, bcsrValuesA_sparse, bcsrRowPtrA_dense, bcsrColIdxA_dense, bcsrRowPtrA_sparse,
                     bcsrColIdxA_sparse, reinterpret_cast<const uint8_t*>(metadata), sparseMatrixA, B,
                     reinterpret_cast<_Float16*>(C), (int64_t)M, (int64_t)N, (int64_t)K, nullptr,
                     relativeBlockIndexMapping);
  if (hipGetLastError() != hipSuccess)
    smat::set_error(SMAT_ERR_HIP, "mmaOBTSKernel_large_separate: launch failed");
}

// k=32 compressor (replaces mmaST_large.cu:324-335).  The reference's k=32
// producer and consumer disagree on the metadata byte offsets
// (mmaST_large.cu:143-146 writes bytes lane, lane+1; mmaOBTS_large.cu:111-115
// reads 2*lane, 2*lane+1), so there is no byte layout to match; here each
// 16x32 block is two consecutive 16x16 halves (columns 0-15, 16-31) in the
// k=16 layout: sparseMatrixA holds 256 halves and metadata 64 bytes per
// block, the same sizes the reference allocates (tester.h:318-320).
void preprocessing_mmaSTKernel_large(uint16_t* bcsrValuesA, char* metadata, uint16_t* sparseMatrixA, size_t M,
                                     size_t N, size_t K, size_t nonzeroBlocks, int* blockInfo,
                                     int* relativeBlockIndexMapping) {
  (void)N;
  if (!bcsrValuesA || !metadata || !sparseMatrixA || !blockInfo || !relativeBlockIndexMapping) {
    smat::set_error(SMAT_ERR_INVALID, "preprocessing_mmaSTKernel_large: null pointer");
    return;
  }
  if (M % 16 || K % 32) {
    smat::set_error(SMAT_ERR_INVALID, "preprocessing_mmaSTKernel_large: M must be a multiple of 16 and K of 32");
    return;
  }
  const int64_t nnzb = (int64_t)nonzeroBlocks;
  const int64_t ncells = (int64_t)(M / 16) * (int64_t)(K / 32);
  smat::DevBuf tags, tags2, halves;
  if (tags.alloc((nnzb + 1) * 4) != SMAT_OK || tags2.alloc((2 * nnzb + 1) * 4) != SMAT_OK ||
      halves.alloc((2 * nnzb + 1) * 512) != SMAT_OK)
    return;
  if (hipMemsetAsync(tags.p, 0, (nnzb + 1) * 4, 0) != hipSuccess) {
    smat::set_error(SMAT_ERR_HIP, "preprocessing_mmaSTKernel_large: memset failed");
    return;
  }
  if (ncells > 0)
    hipLaunchKernelGGL(k_gather_tags, (unsigned)((ncells + 255) / 256), 256, 0, 0, ncells, blockInfo,
                       relativeBlockIndexMapping, tags.as<int32_t>(), nnzb);
  if (nnzb > 0)
    hipLaunchKernelGGL(k_split32, (unsigned)((nnzb * 16 + 255) / 256), 256, 0, 0, nnzb, bcsrValuesA,
                       tags.as<int32_t>(), halves.as<uint16_t>(), tags2.as<int32_t>());
  if (hipGetLastError() != hipSuccess) {
    smat::set_error(SMAT_ERR_HIP, "preprocessing_mmaSTKernel_large: launch failed");
    return;
  }
  bool over = false;
  if (smat::compress_ref16(2 * nnzb, halves.as<uint16_t>(), tags2.as<int32_t>(), sparseMatrixA,
                           reinterpret_cast<uint8_t*>(metadata), 0, &over) != SMAT_OK)
    return;
  if (over)
    smat::set_error(SMAT_ERR_NOT_24,
                    "preprocessing_mmaSTKernel_large: a blockInfo==1 block holds a 1x4 group with more than two "
                    "nonzeros; the extra values were dropped");
}

// Dense SpMM over merged 64 x 64 tiles (mmaOBT_tiled.cu:208-218, called
// with bcsrBlocking's arrays by tester.h:279-286): every sub-block of a tile
// is multiplied as dense, tiles found through relativeBlockIndexMapping.
void mmaOBTKernel_tiled(uint16_t* bcsrValuesA, int* bcsrRowPtrA, int* bcsrColIdxA, uint16_t* B, uint16_t* C,
                        size_t M, size_t N, size_t K, size_t nonzeroBlocks, int* blockInfo,
                        int* relativeBlockIndexMapping) {
  (void)nonzeroBlocks;
  (void)blockInfo;
  tiled_spmm("mmaOBTKernel_tiled", 16, false, bcsrValuesA, bcsrRowPtrA, bcsrColIdxA, relativeBlockIndexMapping,
             nullptr, nullptr, nullptr, B, C, M, N, K);
}

// Compressor of the 64 x 128 tiles (mmaOBTS_tiled_large.cu:21-150, declared at
// main.cu:74-77): the SMFMAC compressor over the tiles' 16 x 32 sub-blocks,
// tagged by tileInfo (indexed by stored tile).
void preprocessing_mmaOBTSKernel_tiled_large(uint16_t* bcsrValuesA, char* metadata, uint16_t* sparseMatrixA,
                                             size_t M, size_t N, size_t K, size_t nonzeroBlocks, int* blockInfo,
                                             int* relativeBlockIndexMapping, int* tileInfo) {
  (void)N;
  (void)nonzeroBlocks;
  (void)blockInfo;
  const char* who = "preprocessing_mmaOBTSKernel_tiled_large";
  if (!bcsrValuesA || !metadata || !sparseMatrixA || !relativeBlockIndexMapping || !tileInfo) {
    smat::set_error(SMAT_ERR_INVALID, "%s: null pointer", who);
    return;
  }
  if (M % 16 || K % 32) {
    smat::set_error(SMAT_ERR_INVALID, "%s: M must be a multiple of 16 and K of 32", who);
    return;
  }
  const int64_t ng = (((int64_t)M + 63) / 64) * (((int64_t)K + 127) / 128);
  smat::DevBuf mx;
  if (mx.alloc(4) != SMAT_OK) return;
  hipLaunchKernelGGL(k_set_flag, 1, 64, 0, 0, mx.as<int>(), -1);
  if (ng > 0)
    hipLaunchKernelGGL(k_max_rel, (unsigned)std::min<int64_t>((ng + 255) / 256, 4096), 256, 0, 0, ng,
                       relativeBlockIndexMapping, mx.as<int>());
  int last = -1;
  if (hipGetLastError() != hipSuccess || hipMemcpy(&last, mx.p, 4, hipMemcpyDeviceToHost) != hipSuccess) {
    smat::set_error(SMAT_ERR_HIP, "%s: tile count failed", who);
    return;
  }
  smat_bcsr_desc d{};
  d.M = 16;
  d.K = 32;
  d.bm = 16;
  d.bk = 32;
  d.nnzb = (int64_t)(last + 1) * 16;
  d.values = bcsrValuesA;
  d.tags = tileInfo;
  if (smat_compress_24(&d, SMAT_24_SMFMAC, sparseMatrixA, metadata, nullptr) == SMAT_ERR_NOT_24)
    smat::set_error(SMAT_ERR_NOT_24,
                    "%s: a tileInfo==1 sub-block holds a 1x4 group with more than two nonzeros; the extra values "
                    "were dropped", who);
}

// Mixed SpMM over merged 64 x 128 tiles (mmaOBTS_tiled_large.cu:431-449).
void mmaOBTSKernel_tiled_large(uint16_t* bcsrValuesA, int* bcsrRowPtrA, int* bcsrColIdxA, char* metadata,
                               uint16_t* sparseMatrixA, uint16_t* B, uint16_t* C, size_t M, size_t N, size_t K,
                               size_t nonzeroBlocks, int* blockInfo, int* relativeBlockIndexMapping, int* tileInfo) {
  (void)nonzeroBlocks;
  (void)blockInfo;
  tiled_spmm("mmaOBTSKernel_tiled_large", 32, true, bcsrValuesA, bcsrRowPtrA, bcsrColIdxA,
             relativeBlockIndexMapping, tileInfo, reinterpret_cast<const uint8_t*>(metadata), sparseMatrixA, B, C, M,
             N, K);
}

}  // extern "C"
