// 2:4 compression of BCSR blocks on the GPU.
//
// Two byte layouts:
//  * SMAT_24_REFERENCE_K16 — the reference's k=16 packing
//    (preprocessing_mmaSTKernelSparse, src/cuda_hgemm/src/mma/mmaST.cu:22-140):
//    "lane" q = 2*row + half covers columns 8*half .. 8*half+7 of one row as two
//    4-groups; the first two nonzeros of a group go to slots 0 and 1 (missing
//    slots are zero) and their positions to one metadata byte:
//    group 0 slot 0 -> bits 4-5, slot 1 -> bits 6-7, group 1 slot 0 -> bits
//    0-1, slot 1 -> bits 2-3 (mmaST.cu:106).  values[b*128 + 4q + s],
//    meta[b*32 + q].
//  * SMAT_24_SMFMAC — the operand layout of this library's SMFMAC kernel:
//    per block row r, bk/2 compressed halves values[(b*16 + r)*(bk/2) + m]
//    and bk/16 uint16 index words; compressed value m belongs to 4-group m/2
//    and its 2-bit position sits at bits 2*(m%8)+1 .. 2*(m%8) of word m/8.
//    The two positions of a group are distinct and ascending (zero padding
//    takes the free positions), as the sparse matrix core expects.
//
// One thread per (block, row, 8-column chunk) — a 16-byte load — so a block is
// handled by 16*bk/8 threads and the whole pass is one coalesced sweep.
#include "smat_internal.h"

namespace {

__device__ inline bool nz16(uint16_t v) { return (v & 0x7fff) != 0; }

// Reference k=16 layout.  Thread t of block b: q = t (0..31), row q/2, half q%2.
__global__ void k_compress_ref16(int64_t nnzb, const uint16_t* __restrict__ val,
                                 const int32_t* __restrict__ tags, uint16_t* __restrict__ sp,
                                 uint8_t* __restrict__ meta, int32_t* __restrict__ overflow) {
  int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t b = gid >> 5;
  int q = (int)(gid & 31);
  if (b >= nnzb) return;
  uint16_t out[4] = {0, 0, 0, 0};
  uint8_t m = 0;
  if (tags[b] == 1) {
    const uint4 raw = *reinterpret_cast<const uint4*>(val + b * 256 + (q >> 1) * 16 + (q & 1) * 8);
    uint16_t src[8];
    __builtin_memcpy(src, &raw, 16);
    int over = 0;
#pragma unroll
    for (int j = 0; j < 2; j++) {
      int cur = 0;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        if (nz16(src[4 * j + i])) {
          if (cur < 2) {
            out[cur + 2 * j] = src[4 * j + i];
            m |= (uint8_t)(i << (6 - (2 * (1 - cur) + 4 * j)));
          } else {
            over = 1;
          }
          cur++;
        }
      }
    }
    if (over) atomicOr(overflow, 1);
  }
  *reinterpret_cast<uint2*>(sp + b * 128 + q * 4) = *reinterpret_cast<uint2*>(out);
  meta[b * 32 + q] = m;
}

// Native SMFMAC layout.  Thread t of block b: row r = t / (bk/8), chunk c of 8 columns.
template <int BK>
__global__ void k_compress_smfmac(int64_t nnzb, const uint16_t* __restrict__ val,
                                  const int32_t* __restrict__ tags, uint16_t* __restrict__ sp,
                                  uint16_t* __restrict__ idx, int32_t* __restrict__ overflow) {
  constexpr int CH = BK / 8;  // 8-column chunks per row
  int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t b = gid / (16 * CH);
  int t = (int)(gid % (16 * CH));
  if (b >= nnzb) return;
  int r = t / CH, c = t % CH;
  uint16_t out[4] = {0, 0, 0, 0};
  uint32_t bits = 0;  // 4 positions x 2 bits for this chunk
  if (tags[b] == 1) {
    const uint4 raw = *reinterpret_cast<const uint4*>(val + (b * 16 + r) * BK + c * 8);
    uint16_t src[8];
    __builtin_memcpy(src, &raw, 16);
    int over = 0;
#pragma unroll
    for (int g = 0; g < 2; g++) {
      int pos[2], n = 0;
#pragma unroll
      for (int i = 0; i < 4; i++)
        if (nz16(src[4 * g + i])) {
          if (n < 2) pos[n] = i; else over = 1;
          n++;
        }
      if (n == 0) { pos[0] = 0; pos[1] = 1; }
      else if (n == 1) {
        if (pos[0] == 3) { pos[1] = 3; pos[0] = 2; }
        else pos[1] = 3;
      }
      out[2 * g] = src[4 * g + pos[0]];
      out[2 * g + 1] = src[4 * g + pos[1]];
      bits |= (uint32_t)pos[0] << (4 * g);
      bits |= (uint32_t)pos[1] << (4 * g + 2);
    }
    if (over) atomicOr(overflow, 1);
  } else {
    bits = (0u) | (1u << 2) | (0u << 4) | (1u << 6);
  }
  *reinterpret_cast<uint2*>(sp + (b * 16 + r) * (BK / 2) + c * 4) = *reinterpret_cast<uint2*>(out);
  // chunk c holds compressed values 4c..4c+3 -> bits 8*(c%2) .. of word c/2
  uint8_t* ib = reinterpret_cast<uint8_t*>(idx + (b * 16 + r) * (BK / 16));
  ib[c] = (uint8_t)bits;
}

inline unsigned blocks_for(int64_t n, int t) {
  int64_t g = (n + t - 1) / t;
  return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace

namespace smat {
// Shared with the reference drop-in launcher.
smat_status compress_ref16(int64_t nnzb, const uint16_t* val, const int32_t* tags, uint16_t* sp,
                           uint8_t* meta, hipStream_t s, bool* overflowed) {
  DevBuf flag;
  SMAT_TRY(flag.alloc(4));
  SMAT_HIP_TRY(hipMemsetAsync(flag.p, 0, 4, s));
  if (nnzb > 0)
    hipLaunchKernelGGL(k_compress_ref16, blocks_for(nnzb * 32, 256), 256, 0, s, nnzb, val, tags, sp, meta,
                       flag.as<int32_t>());
  SMAT_LAUNCH_CHECK();
  int32_t h = 0;
  SMAT_HIP_TRY(hipMemcpyAsync(&h, flag.p, 4, hipMemcpyDeviceToHost, s));
  SMAT_HIP_TRY(hipStreamSynchronize(s));
  *overflowed = h != 0;
  return SMAT_OK;
}
}  // namespace smat

extern "C" {

smat_status smat_compress_24(const smat_bcsr_desc* A, int32_t layout, uint16_t* values24, void* metadata,
                             smat_stream stream) {
  SMAT_REQUIRE(A && values24 && metadata, "smat_compress_24: null argument");
  hipStream_t s = smat::to_hip(stream);
  bool over = false;
  if (layout == SMAT_24_REFERENCE_K16) {
    SMAT_REQUIRE(A->bk == 16, "smat_compress_24: the reference k=16 layout needs bk = 16");
    SMAT_TRY(smat::compress_ref16(A->nnzb, A->values, A->tags, values24, static_cast<uint8_t*>(metadata), s,
                                  &over));
  } else if (layout == SMAT_24_SMFMAC) {
    SMAT_REQUIRE(A->bk == 16 || A->bk == 32, "smat_compress_24: bk must be 16 or 32");
    smat::DevBuf flag;
    SMAT_TRY(flag.alloc(4));
    SMAT_HIP_TRY(hipMemsetAsync(flag.p, 0, 4, s));
    const int64_t n = A->nnzb * 16 * (A->bk / 8);
    if (A->nnzb > 0) {
      if (A->bk == 16)
        hipLaunchKernelGGL(k_compress_smfmac<16>, blocks_for(n, 256), 256, 0, s, A->nnzb, A->values, A->tags,
                           values24, static_cast<uint16_t*>(metadata), flag.as<int32_t>());
      else
        hipLaunchKernelGGL(k_compress_smfmac<32>, blocks_for(n, 256), 256, 0, s, A->nnzb, A->values, A->tags,
                           values24, static_cast<uint16_t*>(metadata), flag.as<int32_t>());
    }
    SMAT_LAUNCH_CHECK();
    int32_t h = 0;
    SMAT_HIP_TRY(hipMemcpyAsync(&h, flag.p, 4, hipMemcpyDeviceToHost, s));
    SMAT_HIP_TRY(hipStreamSynchronize(s));
    over = h != 0;
  } else {
    return smat::set_error(SMAT_ERR_INVALID, "smat_compress_24: unknown layout %d", layout);
  }
  if (over)
    return smat::set_error(SMAT_ERR_NOT_24,
                           "smat_compress_24: a block tagged 2:4 holds a 1x4 group with more than two nonzeros "
                           "(extra values dropped); re-tag with SMAT_TAG_EXACT24");
  return SMAT_OK;
}

}  // extern "C"
