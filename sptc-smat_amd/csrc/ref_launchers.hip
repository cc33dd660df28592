// Reference drop-in launchers: the exact names, parameter order and meaning of
// the reference's kernel entry points (src/cuda_hgemm/src/main.cu:10-77), on
// MI355X kernels.  Differences, all deliberate:
//  * accumulation is FP32 (the reference's HMMA accumulates in FP16,
//    src/cuda_hgemm/src/common/ptx.h:20-41);
//  * a failure is recorded for smat_last_error() instead of exit();
//  * no plan: one wave per (block row, 64-column tile), blocks read through
//    relativeBlockIndexMapping like the reference kernels (mmaOBT.cu:56-61).
#include "smat_internal.h"

namespace smat {
smat_status compress_ref16(int64_t nnzb, const uint16_t* val, const int32_t* tags, uint16_t* sp,
                           uint8_t* meta, hipStream_t s, bool* overflowed);
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

// C[M x N] (FP16, row-major, ldc = N) = A * B with B[n*K + k].
template <bool MIXED>
__global__ __launch_bounds__(256) void k_ref_spmm(const uint16_t* __restrict__ vals,
                                                  const int* __restrict__ rowptr,
                                                  const int* __restrict__ colidx,
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
  const int64_t ncr = K / 16;
  const int r16 = lane & 15, grp = lane >> 4, sub = lane >> 5, half = grp & 1;

  smat_f4 acc[kNsub];
#pragma unroll
  for (int j = 0; j < kNsub; j++) acc[j] = smat_f4{0.f, 0.f, 0.f, 0.f};

  const int pb = rowptr[brow], pe = rowptr[brow + 1];
  for (int p = pb; p < pe; p += 2) {
    const int blk = p + sub;
    const bool valid = blk < pe;
    smat_h8 a = smat_h8{0, 0, 0, 0, 0, 0, 0, 0};
    int64_t c0 = 0;
    if (valid) {
      const int col = colidx[blk];
      const int64_t g = brow * ncr + col / 16;
      const int64_t rel = relMap[g];
      const int info = MIXED ? blockInfo[g] : 2;
      if (info == 1) {
        const int q = r16 * 2 + half;
        a = expand_ref24(sparseA + rel * 128 + q * 4, meta[rel * 32 + q]);
      } else {
        a = __builtin_bit_cast(smat_h8, *reinterpret_cast<const uint4*>(vals + rel * 256 + r16 * 16 + half * 8));
      }
      c0 = col + half * 8;
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

bool check_ref_args(const void* a, const void* b, const void* c, size_t M, size_t N, size_t K, const char* who) {
  if (!a || !b || !c) {
    smat::set_error(SMAT_ERR_INVALID, "%s: null pointer", who);
    return false;
  }
  if (M % 16 || K % 16 || N == 0) {
    smat::set_error(SMAT_ERR_INVALID, "%s: M and K must be multiples of 16, N > 0 (M=%zu N=%zu K=%zu)", who,
                    M, N, K);
    return false;
  }
  return true;
}

}  // namespace

extern "C" {

void mmaOBTKernel(uint16_t* bcsrValuesA, int* bcsrRowPtrA, int* bcsrColIdxA, uint16_t* B, uint16_t* C,
                  size_t M, size_t N, size_t K, size_t nonzeroBlocks, int* blockInfo,
                  int* relativeBlockIndexMapping) {
  (void)nonzeroBlocks;
  (void)blockInfo;
  if (!check_ref_args(bcsrRowPtrA, B, C, M, N, K, "mmaOBTKernel")) return;
  if (!relativeBlockIndexMapping || !bcsrColIdxA || !bcsrValuesA) {
    smat::set_error(SMAT_ERR_INVALID, "mmaOBTKernel: null BCSR array");
    return;
  }
  const int64_t ntiles = ((int64_t)N + 63) / 64;
  const int64_t tasks = (int64_t)(M / 16) * ntiles;
  hipLaunchKernelGGL(k_ref_spmm<false>, (unsigned)((tasks + kWaves - 1) / kWaves), 64 * kWaves, 0, 0,
                     bcsrValuesA, bcsrRowPtrA, bcsrColIdxA, nullptr, nullptr, B,
                     reinterpret_cast<_Float16*>(C), (int64_t)M, (int64_t)N, (int64_t)K, nullptr,
                     relativeBlockIndexMapping);
  if (hipGetLastError() != hipSuccess) smat::set_error(SMAT_ERR_HIP, "mmaOBTKernel: launch failed");
}

void mmaOBTSKernel(uint16_t* bcsrValuesA, int* bcsrRowPtrA, int* bcsrColIdxA, char* metadata,
                   uint16_t* sparseMatrixA, uint16_t* B, uint16_t* C, size_t M, size_t N, size_t K,
                   size_t nonzeroBlocks, int* blockInfo, int* relativeBlockIndexMapping) {
  (void)nonzeroBlocks;
  if (!check_ref_args(bcsrRowPtrA, B, C, M, N, K, "mmaOBTSKernel")) return;
  if (!relativeBlockIndexMapping || !blockInfo || !bcsrColIdxA || !bcsrValuesA || !metadata ||
      !sparseMatrixA) {
    smat::set_error(SMAT_ERR_INVALID, "mmaOBTSKernel: null BCSR or 2:4 array");
    return;
  }
  const int64_t ntiles = ((int64_t)N + 63) / 64;
  const int64_t tasks = (int64_t)(M / 16) * ntiles;
  hipLaunchKernelGGL(k_ref_spmm<true>, (unsigned)((tasks + kWaves - 1) / kWaves), 64 * kWaves, 0, 0,
                     bcsrValuesA, bcsrRowPtrA, bcsrColIdxA, reinterpret_cast<const uint8_t*>(metadata),
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

}  // extern "C"
