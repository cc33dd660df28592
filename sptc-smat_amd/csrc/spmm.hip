// BCSR x dense SpMM on gfx950 matrix cores (FP16, BF16, FP8 e4m3 or split
// FP16 in, FP32 accumulate), its planner and work schedules, and batches.
//
// Reference semantics: mmaOBTKernelSparse (src/cuda_hgemm/src/mma/mmaOBT.cu:18-158)
// computes C = A*B over the blocks of one block row in BCSR order, and
// mmaOBTSKernelSparse[_large] (mmaOBTS.cu:18-205, mmaOBTS_large.cu:19-254)
// sends 2:4 blocks to the sparse tensor-core path.  The design here is
// MI355X-first:
//
//  * One 64-lane wave owns a task = (work item, N-tile).  A work item is a
//    block row, or a contiguous slice of a long block row.  Items come from a
//    plan built on the GPU with prefix sums over block-row lengths, so long
//    rows (R-MAT hubs) are split into slices; their partial 16 x N tiles go to
//    a scratch area and a second kernel reduces them in slice order
//    (deterministic, no atomics, no pre-zeroing of C).
//
//  * v_mfma_f32_16x16x32_{f16,bf16}: the 32-deep reduction of one MFMA is
//    filled by 32/bk blocks of width bk (bk = 1 .. 32).  The blocks need not
//    be adjacent in the matrix: each k-slot gathers the B row of its own
//    block column.  Operand layouts (lane l, element e):
//        A: A[l%16][8*(l/16)+e]      B: B[8*(l/16)+e][l%16]
//        D: D[4*(l/16)+e][l%16]
//
//  * v_smfmac_f32_16x16x64_{f16,bf16} for 2:4 blocks (bk = 16): one
//    instruction consumes four compressed 16x16 blocks.  Lane group g (lanes
//    16g..16g+15) carries block g of the quad: 8 kept values and one 16-bit
//    index word per row.  The matrix core pairs compressed value m of lane
//    group g with B lane group 2*(g%2) + m/4, element 8*(g/2) + 4*((m%4)/2)
//    + idx_m, so B lane group h needs the quad rows 8h..8h+7 (elements 0-7)
//    and 32+8h..32+8h+7 (elements 8-15) when the quad's 64 B rows are laid
//    out block by block.
//
//  * Row-major B is staged through LDS with 128-bit loads and read back as
//    operands with ds_read_b64_tr_b16 (see k_spmm_lds).
//
//  * Workgroups are 4 waves.  The 8 XCDs receive workgroups round-robin; the
//    block-row schedule remaps the workgroup index so that each XCD gets a
//    contiguous range of tasks (neighbouring block rows in one L2), the
//    tile-affinity schedule gives each XCD a subset of the N tiles (one
//    column slice of B per L2); see wave_task.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <mutex>
#include <vector>

#include "smat_internal.h"

namespace {

constexpr int kWaves = 4;  // waves per workgroup

struct Item {
  int32_t brow;
  int32_t pbeg;  // dense range (BCSR positions, or positions in the dense list)
  int32_t pend;
  int32_t slot;  // -1: write C; >= 0: partial tile index in scratch
  int32_t sbeg;  // 2:4 range (positions in the 2:4 list), mixed plans only
  int32_t send;
  int32_t pad0, pad1;
};

__device__ inline uint32_t xcd_remap(uint32_t bid, uint32_t nblocks) {
  // Hardware dispatches workgroup i to XCD i % 8.  Over the first
  // floor(n/8)*8 workgroups give XCD x the contiguous logical range
  // [x*per, (x+1)*per); the tail keeps its own index (a bijection).
  const uint32_t nx = 8;
  const uint32_t per = nblocks / nx;
  if (bid >= per * nx) return bid;
  return (bid % nx) * per + bid / nx;
}

// Work order of the one-shot kernels: wave task -> (item, N tile).
//   kOrderPlain: workgroup w runs tasks 4w..4w+3, task = item * ntiles + tile
//     (the waves of a workgroup take the N tiles of one item);
//   kOrderXcd: as kOrderPlain after xcd_remap (contiguous range per XCD);
//   kOrderTiles: XCD x (workgroups x, x+8, ...) runs only the N tiles
//     t = x mod 8.  With ntiles | 8, 8/ntiles XCDs share a tile and take its
//     item groups in turn; with 8 | ntiles, XCD x runs tiles x, x+8, ... one
//     after the other.  The 4 waves of a workgroup take 4 consecutive items of
//     one tile, so each L2 sees one column slice of B.
enum : int32_t { kOrderPlain = 0, kOrderXcd = 1, kOrderTiles = 2 };

__device__ inline bool wave_task(uint32_t bid, uint32_t nblocks, int wave, int64_t n_items, int32_t ntiles,
                                 int32_t order, int64_t& item, int& nt) {
  if (order == kOrderTiles) {
    const uint32_t x = bid & 7, r = bid >> 3;
    int64_t g;  // item group (kWaves consecutive items)
    if (ntiles <= 8) {
      const uint32_t per = 8u / (uint32_t)ntiles;
      nt = (int)(x % (uint32_t)ntiles);
      g = (int64_t)r * per + x / (uint32_t)ntiles;
    } else {
      const int64_t groups = (n_items + kWaves - 1) / kWaves;
      nt = (int)(x + 8 * (r / groups));
      g = r % groups;
      if (nt >= ntiles) return false;
    }
    item = g * kWaves + wave;
    return item < n_items;
  }
  const uint32_t wg = order == kOrderXcd ? xcd_remap(bid, nblocks) : bid;
  const int64_t task = (int64_t)wg * kWaves + wave;
  item = task / ntiles;
  nt = (int)(task % ntiles);
  return item < n_items;
}

// ---- element types: DT 0 = FP16, 1 = BF16; operands travel as raw words ---
template <int DT>
__device__ inline smat_f4 mma(uint4 a, uint4 b, smat_f4 c) {
  if constexpr (DT == 0)
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(smat_h8, a), __builtin_bit_cast(smat_h8, b),
                                                  c, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(smat_b8, a), __builtin_bit_cast(smat_b8, b),
                                                   c, 0, 0, 0);
}

struct U8 {
  uint4 lo, hi;
};

template <int DT>
__device__ inline smat_f4 smma(uint4 a, U8 b, int idx, smat_f4 c) {
  if constexpr (DT == 0)
    return __builtin_amdgcn_smfmac_f32_16x16x64_f16(__builtin_bit_cast(smat_h8, a), __builtin_bit_cast(smat_h16, b),
                                                    c, idx, 0, 0);
  else
    return __builtin_amdgcn_smfmac_f32_16x16x64_bf16(__builtin_bit_cast(smat_b8, a),
                                                     __builtin_bit_cast(smat_b16, b), c, idx, 0, 0);
}

// Two 4 x 16-bit transposed LDS reads -> 8 k-values of one B column.
__device__ inline uint4 tr_pair(const uint16_t* lo_addr, const uint16_t* hi_addr) {
  const smat_s4 lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((__attribute__((address_space(3))) smat_s4*)(lo_addr));
  const smat_s4 hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((__attribute__((address_space(3))) smat_s4*)(hi_addr));
  // Reinterpret the two 64-bit results as one operand.  (Extracting the 16-bit
  // elements one by one was lowered to a splat of element 0 by this compiler;
  // the whole-vector cast maps the loaded registers straight onto the operand.)
  const smat_s4 both[2] = {lo, hi};
  return __builtin_bit_cast(uint4, both);
}

template <int BLAYOUT>
__device__ inline uint32_t load_b(const uint16_t* __restrict__ B, int64_t ldb, int64_t k, int64_t n) {
  if (BLAYOUT == SMAT_B_ROW_MAJOR) return B[k * ldb + n];
  return B[n * ldb + k];
}

// The 32 k-slots of one MFMA step starting at block p are filled by the
// 32/BK blocks p, p+1, ...: k-slot s is column s % BK of block p + s / BK.
// A lane of group g (= lane / 16) holds k-slots 8g .. 8g+7 of row lane % 16;
// for BK >= 8 that is one 16-byte piece of one block row, for smaller BK it
// is 8/BK pieces of BK halves from consecutive blocks.  Blocks at or beyond
// `pend` contribute zeros.
// A operand of one MFMA step (k-slots 8*grp .. 8*grp+7 of row r16, from the
// 32/BK blocks starting at block p).  OFF32: byte offsets in 32 bits from the
// uniform base (the caller rebases `values` so that they fit).
template <int BK, bool OFF32 = false>
__device__ inline uint4 load_a(const uint16_t* __restrict__ values, int32_t p, int32_t pend, int r16,
                               int grp) {
  static_assert(BK == 1 || BK == 2 || BK == 4 || BK == 8 || BK == 16 || BK == 32, "block width");
  const uint8_t* __restrict__ base = reinterpret_cast<const uint8_t*>(values);
  // byte address of element e of block blk
  auto at = [&](int32_t blk, int e) -> const uint8_t* {
    if constexpr (OFF32) return base + (uint64_t)(((uint32_t)blk * (16u * BK) + (uint32_t)e) * 2u);
    else return base + ((int64_t)blk * 16 * BK + e) * 2;
  };
  if constexpr (BK >= 8) {
    const int32_t blk = p + (8 * grp) / BK;
    if (blk >= pend) return uint4{0, 0, 0, 0};
    return *reinterpret_cast<const uint4*>(at(blk, r16 * BK + (8 * grp) % BK));
  } else if constexpr (BK == 4) {
    uint2 w[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int32_t blk = p + 2 * grp + q;
      w[q] = blk < pend ? *reinterpret_cast<const uint2*>(at(blk, r16 * 4)) : uint2{0, 0};
    }
    return uint4{w[0].x, w[0].y, w[1].x, w[1].y};
  } else if constexpr (BK == 2) {
    uint32_t w[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int32_t blk = p + 4 * grp + q;
      w[q] = blk < pend ? *reinterpret_cast<const uint32_t*>(at(blk, r16 * 2)) : 0u;
    }
    return uint4{w[0], w[1], w[2], w[3]};
  } else {
    uint32_t w[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int32_t b0 = p + 8 * grp + 2 * q, b1 = b0 + 1;
      const uint32_t lo = b0 < pend ? *reinterpret_cast<const uint16_t*>(at(b0, r16)) : 0u;
      const uint32_t hi = b1 < pend ? *reinterpret_cast<const uint16_t*>(at(b1, r16)) : 0u;
      w[q] = lo | (hi << 16);
    }
    return uint4{w[0], w[1], w[2], w[3]};
  }
}

// With a row map (A's rows reordered) row r of A goes to C row rowmap[r].
// PRE (calls with a row map, 64-column FP16 / BF16 tiles): each lane read the
// C row of tile row lane & 15 when the task started (c_row_of_lane), so the
// load's latency hid behind the task's main loop, and store_tile gathers its
// four rows from lanes 0..15 with shuffles.  Otherwise rows are read from the
// map at the end (a dependent load before the stores).
__device__ inline int32_t c_row_of_lane(const Item& it, int lane, const int32_t* __restrict__ rowmap) {
  return rowmap && it.slot < 0 ? rowmap[(int64_t)it.brow * 16 + (lane & 15)] : 0;
}

// D layout: lane holds rows 4*grp + i, column 16*j + r16 of its tile.  crl:
// c_row_of_lane of this lane when PRE (then rowmap is set and the whole wave
// calls this).
template <int NSUB, bool OUT_F16, bool PRE = false>
__device__ inline void store_tile(const smat_f4 (&acc)[NSUB], const Item& it, int64_t n0, int r16, int grp,
                                  int64_t N, void* __restrict__ C, int64_t ldc, float* __restrict__ scratch,
                                  const int32_t* __restrict__ rowmap, int32_t crl = 0) {
  int64_t cr[4];
  if constexpr (PRE) {
#pragma unroll
    for (int i = 0; i < 4; i++) cr[i] = (int64_t)__shfl(crl, 4 * grp + i, 64);
  }
#pragma unroll
  for (int j = 0; j < NSUB; j++) {
    const int64_t n = n0 + 16 * j + r16;
    if (n >= N) continue;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int row = 4 * grp + i;
      if (it.slot >= 0) {
        scratch[((int64_t)it.slot * 16 + row) * N + n] = acc[j][i];
      } else {
        int64_t crow = (int64_t)it.brow * 16 + row;
        if constexpr (PRE) crow = cr[i];
        else if (rowmap) crow = rowmap[crow];
        if (OUT_F16)
          reinterpret_cast<_Float16*>(C)[crow * ldc + n] = (_Float16)acc[j][i];
        else
          reinterpret_cast<float*>(C)[crow * ldc + n] = acc[j][i];
      }
    }
  }
}

// One wave: C[16 x 16*NSUB] tile of one work item; B read straight into the
// MFMA operands (per-lane 16-bit gathers for row-major B; 16-byte loads for
// the reference layout when BK >= 8).
template <int BK, int NSUB, int BLAYOUT, int DT, bool OUT_F16>
__global__ __launch_bounds__(256) void k_spmm_dense(
    const Item* __restrict__ items, int64_t n_items, int32_t ntiles,
    const int32_t* __restrict__ colidx, const uint16_t* __restrict__ values,
    const uint16_t* __restrict__ B, int64_t ldb, int64_t N, void* __restrict__ C, int64_t ldc,
    float* __restrict__ scratch, const int32_t* __restrict__ rowmap, int32_t order) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  int64_t item_id;
  int nt;
  if (!wave_task(blockIdx.x, gridDim.x, wave, n_items, ntiles, order, item_id, nt)) return;
  const Item it = items[item_id];
  const int64_t n0 = (int64_t)nt * 16 * NSUB;

  smat_f4 acc[NSUB];
#pragma unroll
  for (int j = 0; j < NSUB; j++) acc[j] = smat_f4{0.f, 0.f, 0.f, 0.f};

  const int r16 = lane & 15;   // A row / B column within the tile
  const int grp = lane >> 4;   // 0..3: k-slot group of 8
  constexpr int STEP = 32 / BK;  // blocks per MFMA step

  for (int32_t p = it.pbeg; p < it.pend; p += STEP) {
    const uint4 a = load_a<BK>(values, p, it.pend, r16, grp);
    // B row (and validity) of each of this lane's 8 k-slots
    int32_t krow[8];
    bool kok[8];
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const int s = 8 * grp + e;
      const int32_t blk = p + s / BK;
      kok[e] = blk < it.pend;
      krow[e] = kok[e] ? colidx[blk] + s % BK : 0;
    }
#pragma unroll
    for (int j = 0; j < NSUB; j++) {
      const int64_t n = n0 + 16 * j + r16;
      uint4 b = uint4{0, 0, 0, 0};
      if (n < N) {
        if (BLAYOUT == SMAT_B_REFERENCE && BK >= 8 && (ldb % 8) == 0) {
          if (kok[0]) b = *reinterpret_cast<const uint4*>(B + n * ldb + krow[0]);
        } else {
          uint32_t w[4];
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const uint32_t lo = kok[2 * q] ? load_b<BLAYOUT>(B, ldb, krow[2 * q], n) : 0u;
            const uint32_t hi = kok[2 * q + 1] ? load_b<BLAYOUT>(B, ldb, krow[2 * q + 1], n) : 0u;
            w[q] = lo | (hi << 16);
          }
          b = uint4{w[0], w[1], w[2], w[3]};
        }
      }
      acc[j] = mma<DT>(a, b, acc[j]);
    }
  }
  store_tile<NSUB, OUT_F16>(acc, it, n0, r16, grp, N, C, ldc, scratch, rowmap);
}

// Element codecs of the LDS kernel.  DT 0 = FP16, 1 = BF16 (A values and B
// as given); DT 2 = FP8 e4m3 (one byte per element, A scaled per block row,
// B per column); DT 3 = split FP16 (every operand is a pair hi + lo of FP16
// arrays, lo stored after hi; C += Ahi*Bhi + Ahi*Blo + Alo*Bhi, with the
// same power-of-two scales as FP8 uses).
template <int DT>
struct Codec {
  static constexpr int ES = DT == 2 ? 1 : 2;  // bytes per element
  static constexpr int NARR = DT == 3 ? 2 : 1;  // arrays per operand (hi, lo)
  static constexpr bool SCALED = DT >= 2;
};

__device__ inline smat_f4 mma_fp8(uint2 a, uint2 b, smat_f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8((long)__builtin_bit_cast(uint64_t, a),
                                                    (long)__builtin_bit_cast(uint64_t, b), c, 0, 0, 0);
}

// One ds_read_b64_tr_b8: a 16-lane group reading an 8 x 16-byte sub-tile
// (lane L addressing row L/2, bytes 8*(L%2)..+7) receives column L of it,
// i.e. the 8 k-values of one B column (probed on gfx950).
__device__ inline uint2 tr8(const uint8_t* addr) {
  typedef int v2i __attribute__((ext_vector_type(2)));
  const v2i v = __builtin_amdgcn_ds_read_tr8_b64_v2i32((__attribute__((address_space(3))) v2i*)(addr));
  return __builtin_bit_cast(uint2, v);
}

// FP8 counterpart of load_a: 8 bytes (k-slots 8*grp .. 8*grp+7 of row r16)
// gathered from the 32/BK blocks of the step starting at block p.
template <int BK, bool OFF32 = false>
__device__ inline uint2 load_a8(const uint8_t* __restrict__ values, int32_t p, int32_t pend, int r16, int grp) {
  if constexpr (BK >= 8 || (BK == 1 && OFF32)) {
    // byte address of element e of block blk (OFF32: 32-bit offset from the
    // uniform base, which the caller rebases so that it fits)
    auto at = [&](int32_t blk, int e) -> const uint8_t* {
      if constexpr (OFF32) return values + (uint64_t)((uint32_t)blk * (16u * BK) + (uint32_t)e);
      else return values + (int64_t)blk * 16 * BK + e;
    };
    if constexpr (BK >= 8) {
      const int32_t blk = p + (8 * grp) / BK;
      if (blk >= pend) return uint2{0, 0};
      return *reinterpret_cast<const uint2*>(at(blk, r16 * BK + (8 * grp) % BK));
    } else {
      uint32_t w[8];
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const int32_t blk = p + 8 * grp + q;
        w[q] = blk < pend ? *at(blk, r16) : 0u;
      }
      return uint2{w[0] | (w[1] << 8) | (w[2] << 16) | (w[3] << 24),
                   w[4] | (w[5] << 8) | (w[6] << 16) | (w[7] << 24)};
    }
  } else if constexpr (BK == 4) {
    uint32_t w[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int32_t blk = p + 2 * grp + q;
      w[q] = blk < pend ? *reinterpret_cast<const uint32_t*>(values + (int64_t)blk * 64 + r16 * 4) : 0u;
    }
    return uint2{w[0], w[1]};
  } else if constexpr (BK == 2) {
    uint32_t w[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int32_t blk = p + 4 * grp + q;
      w[q] = blk < pend ? *reinterpret_cast<const uint16_t*>(values + (int64_t)blk * 32 + r16 * 2) : 0u;
    }
    return uint2{w[0] | (w[1] << 16), w[2] | (w[3] << 16)};
  } else {
    uint32_t w[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int32_t blk = p + 8 * grp + q;
      w[q] = blk < pend ? values[(int64_t)blk * 16 + r16] : 0u;
    }
    return uint2{w[0] | (w[1] << 8) | (w[2] << 16) | (w[3] << 24), w[4] | (w[5] << 8) | (w[6] << 16) | (w[7] << 24)};
  }
}

// A operand of one step for every codec (hi and, for split, lo).
template <int BK, int DT>
__device__ inline void load_a_any(const void* __restrict__ values, int64_t lo_off, int32_t pb, int32_t p,
                                  int32_t pend, int r16, int grp, uint4 (&a)[Codec<DT>::NARR]) {
  if constexpr (DT == 2 && (BK == 1 || BK >= 8)) {
    const uint8_t* v = static_cast<const uint8_t*>(values) + (int64_t)pb * 16 * BK;  // item's first block
    const uint2 w = load_a8<BK, true>(v, p - pb, pend - pb, r16, grp);
    a[0] = uint4{w.x, w.y, 0, 0};
  } else if constexpr (DT == 2) {
    const uint2 w = load_a8<BK>(static_cast<const uint8_t*>(values), p, pend, r16, grp);
    a[0] = uint4{w.x, w.y, 0, 0};
  } else if constexpr (BK == 1 || BK >= 8) {
    // 32-bit offsets from the item's first block (an item spans at most
    // 65536 blocks, 64 MiB at bk = 32): measured faster for bk = 1 and
    // bk >= 8, slower for 2 and 4
    const uint16_t* v = static_cast<const uint16_t*>(values) + (int64_t)pb * 16 * BK;
    a[0] = load_a<BK, true>(v, p - pb, pend - pb, r16, grp);
    if constexpr (DT == 3) a[Codec<DT>::NARR - 1] = load_a<BK, true>(v + lo_off, p - pb, pend - pb, r16, grp);
  } else {
    const uint16_t* v = static_cast<const uint16_t*>(values);
    a[0] = load_a<BK>(v, p, pend, r16, grp);
    if constexpr (DT == 3) a[Codec<DT>::NARR - 1] = load_a<BK>(v + lo_off, p, pend, r16, grp);
  }
}

// Row-major B through LDS.  Per MFMA step the 32 B rows it reduces over x the
// wave's 16*NSUB columns are fetched with 128-bit loads (each row segment is
// contiguous in HBM), stored to a per-wave LDS tile, and read back as MFMA B
// operands with transposing LDS reads:
//   * 16-bit elements, ds_read_b64_tr_b16: a 16-lane group reading a 4 x 16
//     sub-tile (lane L addressing row L/4, columns 4*(L%4)..+3) receives
//     column L of it, i.e. B[k0..k0+3][n0+L] — 4 of the 8 k-values the
//     operand layout wants; two such reads give a lane its 8 values;
//   * FP8, ds_read_b64_tr_b8: one read of an 8 x 16 sub-tile gives all 8.
// Waves never share LDS, so no barrier is needed: a wave's LDS accesses
// complete in issue order.  Two pipelines (both two register stages):
//   * lds_tile_narrow (16x1 blocks): a step's A values are one coalesced
//     128-bit load per lane, transposed through LDS like B; loads are
//     predicated by buffer resources, never by branches, and the order of
//     loads is fixed, so the compiler's wait counts keep two steps in flight;
//   * lds_tile_wide (wider blocks): A operands loaded straight into registers,
//     loads guarded by branches.  (Measured at R-MAT scale 18, N = 256: the
//     narrow pipeline takes 16x1 blocks from 0.353 to 0.247 ms and power-law
//     16x1 from 0.504 to 0.393 ms, while the same pipeline for 16x2, 16x4,
//     16x8 and 16x16 blocks was 4-28 % slower than this one, with any
//     pipeline depth, load form or occupancy tried.)
//
// Scaled codecs multiply the FP32 tile by a_scale[block row] * b_scale[col]
// before it is stored (split rows scale their partial tiles, so the slice
// reduction sums scaled values).
// Operands of one SpMM as the kernels see them (element offsets; for split
// operands the lo array starts a_lo / b_lo elements after the hi array).
struct Operands {
  const int32_t* colidx;
  const void* a;
  int64_t a_lo;
  const float* a_scale;
  const void* b;
  int64_t ldb;
  int64_t b_lo;
  const float* b_scale;
  int64_t N;
  void* c;
  int64_t ldc;
  float* scratch;
  const int32_t* rowmap;  // NULL, or the C row of each A row
  uint32_t b_bytes;       // bytes of one B array (the lo array of a split B is as long)
};

// LDS bytes one wave of the LDS kernel needs: the B tile of one step (32
// rows of 16*NSUB elements) and, for 16x1 blocks, behind it the A tile of
// one step (32 blocks of 16 elements); rows padded by 16 bytes.
template <int NSUB, int DT>
constexpr int lds_b_bytes() {
  return Codec<DT>::NARR * 32 * (16 * NSUB * Codec<DT>::ES + 16);
}
template <int DT>
constexpr int lds_a_bytes() {
  return Codec<DT>::NARR * 32 * (16 * Codec<DT>::ES + 16);
}
template <int BK, int NSUB, int DT>
constexpr int lds_wave_bytes() {
  return lds_b_bytes<NSUB, DT>() + (BK == 1 ? lds_a_bytes<DT>() : 0);
}

// Buffer loads of the LDS kernel.  Every load goes through a raw buffer
// resource (gfx9 layout) over exactly the bytes it may read; an offset at or
// past the end (kOob) returns zeros without a memory access.  That is how the
// kernel predicates its loads: no branch around a load and no select on
// loaded data, so nothing between a load's issue and its use makes the
// compiler drain the other stages' loads (a conditional load followed by a
// shift of its value did exactly that: a wait for every outstanding load after
// each step's issue).
typedef unsigned smat_u4v __attribute__((ext_vector_type(4)));
typedef unsigned smat_u2v __attribute__((ext_vector_type(2)));
constexpr uint32_t kOob = 0xFFFFFFF0u;          // past every resource (they are capped below it)
constexpr double kMaxResourceBytes = 4294967040.0;  // 2^32 - 256

__device__ inline __amdgpu_buffer_rsrc_t rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, 0x00020000);
}
__device__ inline uint4 bload16(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  const smat_u4v v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
  return uint4{v.x, v.y, v.z, v.w};
}
__device__ inline uint2 bload8(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  const smat_u2v v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
  return uint2{v.x, v.y};
}
__device__ inline uint32_t bload4(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
}

// A values of one step of 16x1 blocks as loaded: lane l holds elements
// 8*(l%2) .. 8*(l%2)+7 of block p + l/2 (16 bytes, or 8 for FP8), blocks past
// the item's end reading as zeros (their offsets fall outside the item's
// resource ra, whose offsets are relative to block pbeg; q0 = p - pbeg).
template <int ES>
__device__ inline void a_issue_bk1(__amdgpu_buffer_rsrc_t ra, int32_t q0, int lane, uint32_t (&w)[2 * ES]) {
  const uint32_t off = ((uint32_t)(q0 + (lane >> 1)) * 16u + (uint32_t)(lane & 1) * 8u) * ES;
  if constexpr (ES == 2) {
    const uint4 v = bload16(ra, off);
    w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
  } else {
    const uint2 v = bload8(ra, off);
    w[0] = v.x; w[1] = v.y;
  }
}

// One task (item, N tile) of 16x1 blocks: the narrow pipeline.
template <int NSUB, int DT, bool OUT_F16, bool PRE, int DEPTH = 2>
__device__ __forceinline__ void lds_tile_narrow(const Item& it, int64_t n0, const Operands& op,
                                                uint8_t* __restrict__ lds, int lane) {
  constexpr int BK = 1;
  using CD = Codec<DT>;
  constexpr int ES = CD::ES, NARR = CD::NARR;
  constexpr int NT = 16 * NSUB;           // columns per wave
  constexpr int RB = NT * ES;             // bytes per B row segment
  constexpr int CPR = RB / 16;            // 16-byte chunks per B row segment
  constexpr int RPL = 64 / CPR;           // B rows per load instruction
  static_assert(RPL <= 32, "a load instruction must not span more than one step");
  constexpr int NLD = 32 / RPL;           // load instructions per MFMA step and array
  constexpr int PITCH = RB + 16;          // LDS row pitch of the B tile in bytes (16-byte pad)
  constexpr int PA = 16 * ES + 16;        // LDS row pitch of the A tile
  constexpr int STEP = 32 / BK;           // blocks per MFMA step
  constexpr int AW = 2 * ES;              // 32-bit words of A per lane and step
  uint8_t* __restrict__ lds_a = lds + lds_b_bytes<NSUB, DT>();
  const uint8_t* __restrict__ B = static_cast<const uint8_t*>(op.b);
  const int64_t ldb = op.ldb, N = op.N;

  const int r16 = lane & 15, grp = lane >> 4;
  const int32_t crl = PRE ? c_row_of_lane(it, lane, op.rowmap) : 0;
  // load coordinates: instruction i covers rows i*RPL .. i*RPL+RPL-1
  const int lrow = lane / CPR, lch = lane % CPR;
  const int64_t ncol = n0 + lch * (16 / ES);
  const bool col_ok = ncol < N;  // N is a multiple of the chunk: a chunk is all in or all out

  smat_f4 acc[NSUB];
#pragma unroll
  for (int j = 0; j < NSUB; j++) acc[j] = smat_f4{0.f, 0.f, 0.f, 0.f};

  const int32_t pb = it.pbeg, pe = it.pend;
  // resources: the item's column starts and A blocks (offsets relative to
  // block pb), and each B array
  const __amdgpu_buffer_rsrc_t rc = rsrc(op.colidx + pb, (uint32_t)(pe - pb) * 4u);
  __amdgpu_buffer_rsrc_t ra[NARR], rb[NARR];
#pragma unroll
  for (int h = 0; h < NARR; h++) {
    ra[h] = rsrc(static_cast<const uint8_t*>(op.a) + ((int64_t)pb * 16 * BK + h * op.a_lo) * ES,
                 (uint32_t)(pe - pb) * (16u * BK * ES));
    rb[h] = rsrc(B + h * op.b_lo * ES, op.b_bytes);
  }
  // column starts of the blocks of step p: lane l holds colidx[p + l % STEP]
  // (zero past the end of the item)
  auto cload = [&](int32_t p) -> int32_t {
    return (int32_t)bload4(rc, (uint32_t)(p - pb + lane % STEP) * 4u);
  };
  // A words and B rows of step p (zeros past the end of the item), and the
  // column starts of step pc into cid once step p's are taken from it.
  // Issued for every step the loop reaches, so the loop body has no branch
  // around a load; the fixed order (A, column starts, B) is the same in the
  // prologue and the loop, so the compiler's wait counts stay exact.
  auto issue = [&](int32_t p, int32_t& cid, int32_t pc, uint4 (&bb)[NARR][NLD], uint32_t (&aw)[NARR][AW]) {
#pragma unroll
    for (int h = 0; h < NARR; h++) a_issue_bk1<ES>(ra[h], p - pb, lane, aw[h]);
    int32_t cb[NLD];
#pragma unroll
    for (int i = 0; i < NLD; i++) cb[i] = __shfl(cid, (i * RPL + lrow) / BK);
    cid = cload(pc);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < NLD; i++) {
      const int row = i * RPL + lrow;  // 0..31 = k-slot
      const int boff = row / BK;
      const int32_t c = cb[i];
      const bool ok = p + boff < pe && col_ok;
      // 32-bit byte offset from the array base (the launcher checks that B,
      // hi and lo arrays together, spans under 4 GiB)
      const uint32_t off = ok ? ((uint32_t)(c + row % BK) * (uint32_t)ldb + (uint32_t)ncol) * ES : kOob;
#pragma unroll
      for (int h = 0; h < NARR; h++) bb[h][i] = bload16(rb[h], off);
    }
  };
  // landed B rows and A values to their LDS tiles
  auto consume = [&](const uint4 (&bb)[NARR][NLD], const uint32_t (&aw)[NARR][AW]) {
#pragma unroll
    for (int h = 0; h < NARR; h++)
#pragma unroll
      for (int i = 0; i < NLD; i++)
        *reinterpret_cast<uint4*>(lds + (h * 32 + i * RPL + lrow) * PITCH + lch * 16) = bb[h][i];
#pragma unroll
    for (int h = 0; h < NARR; h++) {
      uint8_t* dst = lds_a + (h * 32 + (lane >> 1)) * PA + (lane & 1) * 8 * ES;
      if constexpr (ES == 2) *reinterpret_cast<uint4*>(dst) = uint4{aw[h][0], aw[h][1], aw[h][2], aw[h][3]};
      else *reinterpret_cast<uint2*>(dst) = uint2{aw[h][0], aw[h][1]};
    }
  };
  auto compute = [&]() {
    uint4 a[NARR];  // A operands from the A tile, transposed like B
#pragma unroll
    for (int h = 0; h < NARR; h++) {
      if constexpr (DT == 2) {
        const uint2 v = tr8(lds_a + (8 * grp + (r16 >> 1)) * PA + (r16 & 1) * 8);
        a[h] = uint4{v.x, v.y, 0, 0};
      } else {
        const uint8_t* ab = lds_a + (h * 32 + 8 * grp + (r16 >> 2)) * PA + (r16 & 3) * 8;
        a[h] = tr_pair(reinterpret_cast<const uint16_t*>(ab), reinterpret_cast<const uint16_t*>(ab + 4 * PA));
      }
    }
#pragma unroll
    for (int j = 0; j < NSUB; j++) {
      if constexpr (DT == 2) {
        const uint8_t* base = lds + (8 * grp + (r16 >> 1)) * PITCH + 16 * j + (r16 & 1) * 8;
        acc[j] = mma_fp8(uint2{a[0].x, a[0].y}, tr8(base), acc[j]);
      } else {
        const uint8_t* base = lds + (8 * grp + (r16 >> 2)) * PITCH + (16 * j + (r16 & 3) * 4) * 2;
        const uint4 bh = tr_pair(reinterpret_cast<const uint16_t*>(base),
                                 reinterpret_cast<const uint16_t*>(base + 4 * PITCH));
        if constexpr (DT == 3) {
          const uint8_t* lb = base + 32 * PITCH;
          const uint4 bl = tr_pair(reinterpret_cast<const uint16_t*>(lb),
                                   reinterpret_cast<const uint16_t*>(lb + 4 * PITCH));
          acc[j] = mma<0>(a[1], bh, acc[j]);  // small terms first
          acc[j] = mma<0>(a[0], bl, acc[j]);
          acc[j] = mma<0>(a[0], bh, acc[j]);
        } else {
          acc[j] = mma<DT>(a[0], bh, acc[j]);
        }
      }
    }
  };

  // DEPTH register stages, stage d holding steps d, d + DEPTH, ...; the round
  // loops are unrolled DEPTH times so every stage index is static.
  uint4 bS[DEPTH][NARR][NLD];
  uint32_t aS[DEPTH][NARR][AW];
  int32_t cS[DEPTH];
  // one step of stage d: its landed loads to LDS, the loads of the step
  // DEPTH further (when `more`), the matrix-core work.  The barrier keeps the
  // next stage's LDS stores (and their wait for its loads) behind this
  // stage's issue: hoisted above it they drain every load in flight.
  auto step = [&](int d, int32_t p_next, bool more) {
    consume(bS[d], aS[d]);
    if (more) issue(p_next, cS[d], p_next + DEPTH * STEP, bS[d], aS[d]);
    compute();
    __builtin_amdgcn_sched_barrier(0);
  };
#pragma unroll
  for (int d = 0; d < DEPTH; d++) cS[d] = cload(pb + d * STEP);
  if (pe - pb > DEPTH * STEP) {
    // prologue, then rounds in which every stage issues its next step: no
    // branch in the loop body and the same load order on entry as around
    // the loop, so the compiler's wait counts let DEPTH steps fly
#pragma unroll
    for (int d = 0; d < DEPTH; d++) {
      issue(pb + d * STEP, cS[d], pb + (d + DEPTH) * STEP, bS[d], aS[d]);
      __builtin_amdgcn_sched_barrier(0);
    }
    int32_t p = pb;
    for (; p + (2 * DEPTH - 1) * STEP < pe; p += DEPTH * STEP) {
#pragma unroll
      for (int d = 0; d < DEPTH; d++) step(d, p + (d + DEPTH) * STEP, true);
    }
    // drain: steps p .. p + (DEPTH-1)*STEP are in flight (all inside the
    // item), at most DEPTH - 1 more follow
#pragma unroll
    for (int d = 0; d < DEPTH; d++) step(d, p + (d + DEPTH) * STEP, p + (d + DEPTH) * STEP < pe);
#pragma unroll
    for (int d = 0; d < DEPTH; d++)
      if (p + (d + DEPTH) * STEP < pe) step(d, 0, false);
  } else {
    // short items (at most DEPTH steps): every load up front
#pragma unroll
    for (int d = 0; d < DEPTH; d++)
      if (pb + d * STEP < pe) issue(pb + d * STEP, cS[d], pe, bS[d], aS[d]);
#pragma unroll
    for (int d = 0; d < DEPTH; d++)
      if (pb + d * STEP < pe) step(d, 0, false);
  }
  if constexpr (CD::SCALED) {
    const float sa = op.a_scale[it.brow];
#pragma unroll
    for (int j = 0; j < NSUB; j++) {
      const int64_t n = n0 + 16 * j + r16;
      const float sc = sa * (n < N ? op.b_scale[n] : 0.f);
#pragma unroll
      for (int i = 0; i < 4; i++) acc[j][i] *= sc;
    }
  }
  store_tile<NSUB, OUT_F16, PRE>(acc, it, n0, r16, grp, N, op.c, op.ldc, op.scratch, op.rowmap, crl);
}

// One task of blocks 16 x BK, BK >= 2: the wide pipeline (two register
// stages X/Y, loop unrolled by two so the stage index is static; while step
// p is transposed and multiplied, the B rows and A values of steps p+STEP
// and p+2*STEP are in flight; a step's block column starts are fetched one
// phase before its B rows with one coalesced load, lane l holding
// colidx[p + l % STEP], and handed to the B-row lanes by a lane shuffle).
template <int BK, int NSUB, int DT, bool OUT_F16, bool PRE>
__device__ __forceinline__ void lds_tile_wide(const Item& it, int64_t n0, const Operands& op,
                                              uint8_t* __restrict__ lds, int lane) {
  using CD = Codec<DT>;
  constexpr int ES = CD::ES, NARR = CD::NARR;
  constexpr int NT = 16 * NSUB;           // columns per wave
  constexpr int RB = NT * ES;             // bytes per B row segment
  constexpr int CPR = RB / 16;            // 16-byte chunks per B row segment
  constexpr int RPL = 64 / CPR;           // B rows per load instruction
  static_assert(RPL <= 32, "a load instruction must not span more than one step");
  constexpr int NLD = 32 / RPL;           // load instructions per MFMA step and array
  constexpr int PITCH = RB + 16;          // LDS row pitch in bytes (16-byte pad)
  constexpr int STEP = 32 / BK;           // blocks per MFMA step
  const int32_t* __restrict__ colidx = op.colidx;
  const uint8_t* __restrict__ B = static_cast<const uint8_t*>(op.b);
  const int64_t ldb = op.ldb, N = op.N;

  const int r16 = lane & 15, grp = lane >> 4;
  const int32_t crl = PRE ? c_row_of_lane(it, lane, op.rowmap) : 0;
  // load coordinates: instruction i covers rows i*RPL .. i*RPL+RPL-1
  const int lrow = lane / CPR, lch = lane % CPR;
  const int64_t ncol = n0 + lch * (16 / ES);
  const bool col_ok = ncol < N;  // N is a multiple of the chunk: a chunk is all in or all out

  smat_f4 acc[NSUB];
#pragma unroll
  for (int j = 0; j < NSUB; j++) acc[j] = smat_f4{0.f, 0.f, 0.f, 0.f};

  const int32_t pe = it.pend;
  // column starts of the blocks of step p: lane l holds colidx[p + l % STEP]
  auto cload = [&](int32_t p) -> int32_t {
    const int32_t b = p + (lane % STEP);
    return b < pe ? colidx[b] : 0;
  };
  // B rows and A values of step p (zeros past the end of the item)
  auto issue = [&](int32_t p, int32_t cid, uint4 (&bb)[NARR][NLD], uint4 (&aa)[NARR]) {
    if (p >= pe) return;
#pragma unroll
    for (int i = 0; i < NLD; i++) {
      const int row = i * RPL + lrow;  // 0..31 = k-slot
      const int boff = row / BK;
      const int32_t c = __shfl(cid, boff);
      const bool ok = p + boff < pe && col_ok;
      // 32-bit byte offset from the uniform base (the launcher checks that B,
      // hi and lo arrays together, spans under 4 GiB): one VGPR per address
      // and the SGPR-base load form
      const uint32_t off = ((uint32_t)(c + row % BK) * (uint32_t)ldb + (uint32_t)ncol) * ES;
#pragma unroll
      for (int h = 0; h < NARR; h++)
        bb[h][i] = ok ? *reinterpret_cast<const uint4*>(B + (uint64_t)off + h * op.b_lo * ES) : uint4{0, 0, 0, 0};
    }
    load_a_any<BK, DT>(op.a, op.a_lo, it.pbeg, p, pe, r16, grp, aa);
  };
  auto consume = [&](const uint4 (&bb)[NARR][NLD]) {
#pragma unroll
    for (int h = 0; h < NARR; h++)
#pragma unroll
      for (int i = 0; i < NLD; i++)
        *reinterpret_cast<uint4*>(lds + (h * 32 + i * RPL + lrow) * PITCH + lch * 16) = bb[h][i];
  };
  auto compute = [&](const uint4 (&a)[NARR]) {
#pragma unroll
    for (int j = 0; j < NSUB; j++) {
      if constexpr (DT == 2) {
        const uint8_t* base = lds + (8 * grp + (r16 >> 1)) * PITCH + 16 * j + (r16 & 1) * 8;
        acc[j] = mma_fp8(uint2{a[0].x, a[0].y}, tr8(base), acc[j]);
      } else {
        const uint8_t* base = lds + (8 * grp + (r16 >> 2)) * PITCH + (16 * j + (r16 & 3) * 4) * 2;
        const uint4 bh = tr_pair(reinterpret_cast<const uint16_t*>(base),
                                 reinterpret_cast<const uint16_t*>(base + 4 * PITCH));
        if constexpr (DT == 3) {
          const uint8_t* lb = base + 32 * PITCH;
          const uint4 bl = tr_pair(reinterpret_cast<const uint16_t*>(lb),
                                   reinterpret_cast<const uint16_t*>(lb + 4 * PITCH));
          acc[j] = mma<0>(a[1], bh, acc[j]);  // small terms first
          acc[j] = mma<0>(a[0], bl, acc[j]);
          acc[j] = mma<0>(a[0], bh, acc[j]);
        } else {
          acc[j] = mma<DT>(a[0], bh, acc[j]);
        }
      }
    }
  };

  uint4 bX[NARR][NLD], bY[NARR][NLD], aX[NARR], aY[NARR];
  const int32_t pb = it.pbeg;
  int32_t cX = cload(pb);
  int32_t cY = cload(pb + STEP);
  issue(pb, cX, bX, aX);
  cX = cload(pb + 2 * STEP);
  issue(pb + STEP, cY, bY, aY);
  cY = cload(pb + 3 * STEP);
  for (int32_t p = pb; p < pe; p += 2 * STEP) {
    uint4 a[NARR];
    consume(bX);
#pragma unroll
    for (int h = 0; h < NARR; h++) a[h] = aX[h];
    issue(p + 2 * STEP, cX, bX, aX);
    cX = cload(p + 4 * STEP);
    compute(a);
    if (p + STEP >= pe) break;
    consume(bY);
#pragma unroll
    for (int h = 0; h < NARR; h++) a[h] = aY[h];
    issue(p + 3 * STEP, cY, bY, aY);
    cY = cload(p + 5 * STEP);
    compute(a);
  }
  if constexpr (CD::SCALED) {
    const float sa = op.a_scale[it.brow];
#pragma unroll
    for (int j = 0; j < NSUB; j++) {
      const int64_t n = n0 + 16 * j + r16;
      const float sc = sa * (n < N ? op.b_scale[n] : 0.f);
#pragma unroll
      for (int i = 0; i < 4; i++) acc[j][i] *= sc;
    }
  }
  store_tile<NSUB, OUT_F16, PRE>(acc, it, n0, r16, grp, N, op.c, op.ldc, op.scratch, op.rowmap, crl);
}

template <int BK, int NSUB, int DT, bool OUT_F16, bool PRE = false, int DEPTH = 2>
__device__ __forceinline__ void lds_tile(const Item& it, int64_t n0, const Operands& op, uint8_t* __restrict__ lds,
                                         int lane) {
  if constexpr (BK == 1) lds_tile_narrow<NSUB, DT, OUT_F16, PRE, DEPTH>(it, n0, op, lds, lane);
  else lds_tile_wide<BK, NSUB, DT, OUT_F16, PRE>(it, n0, op, lds, lane);
}

template <int BK, int NSUB, int DT, bool OUT_F16, bool PRE, int DEPTH = 2>
__device__ __forceinline__ void lds_kernel_body(const Item* __restrict__ items, int64_t n_items, int32_t ntiles,
                                                const Operands& op, int32_t order) {
  constexpr int WB = lds_wave_bytes<BK, NSUB, DT>();
  __shared__ __attribute__((aligned(16))) uint8_t lds_all[kWaves * WB];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  int64_t item_id;
  int nt;
  if (!wave_task(blockIdx.x, gridDim.x, wave, n_items, ntiles, order, item_id, nt)) return;
  const Item it = items[item_id];
  lds_tile<BK, NSUB, DT, OUT_F16, PRE, DEPTH>(it, (int64_t)nt * 16 * NSUB, op, lds_all + wave * WB, lane);
}

template <int BK, int NSUB, int DT, bool OUT_F16>
__global__ __launch_bounds__(256) void k_spmm_lds(const Item* __restrict__ items, int64_t n_items, int32_t ntiles,
                                                  Operands op, int32_t order) {
  lds_kernel_body<BK, NSUB, DT, OUT_F16, false>(items, n_items, ntiles, op, order);
}

// Calls with a row map on 64-column FP16 / BF16 tiles (NSUB = 4, DT <= 1):
// the row map read when a task starts (PRE), the kernel held to 4 waves per
// SIMD (128 VGPR + AGPR), which the row-map VGPR would otherwise cost.
template <int BK, int DT, bool OUT_F16>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_spmm_lds64(
    const Item* __restrict__ items, int64_t n_items, int32_t ntiles, Operands op, int32_t order) {
  lds_kernel_body<BK, 4, DT, OUT_F16, true>(items, n_items, ntiles, op, order);
}

// Mixed dense / 2:4 kernel for 16x16 blocks and row-major B.  Dense blocks
// (entries dense[pbeg..pend)) go through MFMA in pairs; 2:4 blocks (entries
// sparse[sbeg..send)) through SMFMAC in quads, from the compressed arrays of
// smat_compress_24 (SMAT_24_SMFMAC layout, indexed by BCSR block).  Both
// accumulate into the same FP32 tile.  One register stage of prefetch.  A
// list entry holds the block's BCSR position and its column start, so a B row
// address is one load away from the list (not two through colidx).
// 2:4 operands of a mixed plan.
struct Mixed24 {
  const int2* dense;   // {BCSR block, column start}
  const int2* sparse;
  const uint16_t* values24;
  const uint16_t* index24;
};

template <int NSUB>
constexpr int mixed_wave_bytes() {
  return 64 * (16 * NSUB + 8) * 2;
}

template <int NSUB, int DT, bool OUT_F16>
__device__ __forceinline__ void mixed_tile(const Item& it, int64_t n0, const Operands& op, const Mixed24& mx,
                                           uint8_t* __restrict__ lds_bytes, int lane) {
  constexpr int NT = 16 * NSUB;
  constexpr int CPR = NT / 8;
  constexpr int RPL = 64 / CPR;
  constexpr int NLD_D = 32 / RPL;  // loads per dense pair (32 rows)
  constexpr int NLD_S = 64 / RPL;  // loads per 2:4 quad (64 rows)
  constexpr int PITCH = NT + 8;
  uint16_t* lds = reinterpret_cast<uint16_t*>(lds_bytes);
  const uint16_t* __restrict__ values = static_cast<const uint16_t*>(op.a);
  const uint16_t* __restrict__ B = static_cast<const uint16_t*>(op.b);
  const int64_t ldb = op.ldb, N = op.N;
  const int2* __restrict__ dense = mx.dense;
  const int2* __restrict__ sparse = mx.sparse;
  const uint16_t* __restrict__ values24 = mx.values24;
  const uint16_t* __restrict__ index24 = mx.index24;

  const int r16 = lane & 15, grp = lane >> 4;
  const int lrow = lane / CPR, lch = lane % CPR;
  const int64_t ncol = n0 + lch * 8;
  const bool col_ok = ncol < N;

  smat_f4 acc[NSUB];
#pragma unroll
  for (int j = 0; j < NSUB; j++) acc[j] = smat_f4{0.f, 0.f, 0.f, 0.f};

  // ---- dense pairs
  {
    const int32_t pe = it.pend;
    uint4 bb[NLD_D], aa = uint4{0, 0, 0, 0};
    auto issue = [&](int32_t p) {
      if (p >= pe) return;
#pragma unroll
      for (int i = 0; i < NLD_D; i++) {
        const int row = i * RPL + lrow;
        const int32_t q = p + row / 16;
        if (q < pe && col_ok)
          bb[i] = *reinterpret_cast<const uint4*>(B + (int64_t)(dense[q].y + row % 16) * ldb + ncol);
        else
          bb[i] = uint4{0, 0, 0, 0};
      }
      const int32_t q = p + (lane >> 5);
      aa = q < pe ? *reinterpret_cast<const uint4*>(values + (int64_t)dense[q].x * 256 + r16 * 16 + (grp & 1) * 8)
                  : uint4{0, 0, 0, 0};
    };
    issue(it.pbeg);
    for (int32_t p = it.pbeg; p < pe; p += 2) {
#pragma unroll
      for (int i = 0; i < NLD_D; i++)
        *reinterpret_cast<uint4*>(lds + (i * RPL + lrow) * PITCH + lch * 8) = bb[i];
      const uint4 a = aa;
      issue(p + 2);
#pragma unroll
      for (int j = 0; j < NSUB; j++) {
        const uint16_t* base = lds + (8 * grp + (r16 >> 2)) * PITCH + 16 * j + (r16 & 3) * 4;
        acc[j] = mma<DT>(a, tr_pair(base, base + 4 * PITCH), acc[j]);
      }
    }
  }
  // ---- 2:4 quads
  {
    const int32_t se = it.send;
    uint4 bb[NLD_S], aa = uint4{0, 0, 0, 0};
    int idx = 0;
    auto issue = [&](int32_t p) {
      if (p >= se) return;
#pragma unroll
      for (int i = 0; i < NLD_S; i++) {
        const int row = i * RPL + lrow;  // 0..63: block row/16 of the quad
        const int32_t q = p + row / 16;
        if (q < se && col_ok)
          bb[i] = *reinterpret_cast<const uint4*>(B + (int64_t)(sparse[q].y + row % 16) * ldb + ncol);
        else
          bb[i] = uint4{0, 0, 0, 0};
      }
      const int32_t q = p + grp;  // lane group g carries block g of the quad
      if (q < se) {
        const int64_t id = sparse[q].x;
        aa = *reinterpret_cast<const uint4*>(values24 + (id * 16 + r16) * 8);
        idx = index24[id * 16 + r16];
      } else {
        aa = uint4{0, 0, 0, 0};
        idx = 0x4444;  // positions (0,1) in every group; values are zero
      }
    };
    issue(it.sbeg);
    for (int32_t p = it.sbeg; p < se; p += 4) {
#pragma unroll
      for (int i = 0; i < NLD_S; i++)
        *reinterpret_cast<uint4*>(lds + (i * RPL + lrow) * PITCH + lch * 8) = bb[i];
      const uint4 a = aa;
      const int ix = idx;
      issue(p + 4);
#pragma unroll
      for (int j = 0; j < NSUB; j++) {
        const uint16_t* b0 = lds + (8 * grp + (r16 >> 2)) * PITCH + 16 * j + (r16 & 3) * 4;
        const uint16_t* b1 = b0 + 32 * PITCH;
        U8 b;
        b.lo = tr_pair(b0, b0 + 4 * PITCH);
        b.hi = tr_pair(b1, b1 + 4 * PITCH);
        acc[j] = smma<DT>(a, b, ix, acc[j]);
      }
    }
  }
  store_tile<NSUB, OUT_F16>(acc, it, n0, r16, grp, N, op.c, op.ldc, op.scratch, op.rowmap);
}

template <int NSUB, int DT, bool OUT_F16>
__global__ __launch_bounds__(256) void k_spmm_mixed(const Item* __restrict__ items, int64_t n_items, int32_t ntiles,
                                                    Operands op, Mixed24 mx, int32_t order) {
  constexpr int WB = mixed_wave_bytes<NSUB>();
  __shared__ __attribute__((aligned(16))) uint8_t lds_all[kWaves * WB];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int64_t item_id;
  int nt;
  if (!wave_task(blockIdx.x, gridDim.x, wave, n_items, ntiles, order, item_id, nt)) return;
  const Item it = items[item_id];
  mixed_tile<NSUB, DT, OUT_F16>(it, (int64_t)nt * 16 * NSUB, op, mx, lds_all + wave * WB, lane);
}

struct Reduce {
  int32_t brow;
  int32_t slot_beg;
  int32_t slot_cnt;
  int32_t pad;
};

// Sums the partial tiles of each split block row in slice order.  With
// N % 4 == 0 (and 16-byte aligned C rows for FP32 output) a thread handles
// four consecutive columns with 16-byte loads; the summation order per
// element is the same either way.
template <bool OUT_F16, bool VEC4>
__global__ void k_reduce(const Reduce* __restrict__ red, int64_t n_red, int32_t chunks,
                         const float* __restrict__ scratch, int64_t N, void* __restrict__ C, int64_t ldc,
                         const int32_t* __restrict__ rowmap) {
  const int64_t r = blockIdx.x / chunks;
  if (r >= n_red) return;
  const Reduce rd = red[r];
  constexpr int V = VEC4 ? 4 : 1;
  const int64_t per_row = N / V, total = 16 * per_row;
  for (int64_t i = (int64_t)(blockIdx.x % chunks) * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)chunks * blockDim.x) {
    const int64_t row = i / per_row, n = (i % per_row) * V;
    // C row, looked up where it is needed: after the partial-tile loads
    auto c_row = [&]() { return rowmap ? (int64_t)rowmap[(int64_t)rd.brow * 16 + row] : (int64_t)rd.brow * 16 + row; };
    if constexpr (VEC4) {
      auto part = [&](int q) {
        return *reinterpret_cast<const smat_f4*>(scratch + ((int64_t)(rd.slot_beg + q) * 16 + row) * N + n);
      };
      smat_f4 s = smat_f4{0.f, 0.f, 0.f, 0.f};
      int q = 0;
      // four partial tiles in flight per thread, summed in slice order
      for (; q + 4 <= rd.slot_cnt; q += 4) {
        const smat_f4 t0 = part(q), t1 = part(q + 1), t2 = part(q + 2), t3 = part(q + 3);
        s += t0;
        s += t1;
        s += t2;
        s += t3;
      }
      for (; q < rd.slot_cnt; q++) s += part(q);
      const int64_t crow = c_row();
      if (OUT_F16) {
        _Float16* c = reinterpret_cast<_Float16*>(C) + crow * ldc + n;
#pragma unroll
        for (int e = 0; e < 4; e++) c[e] = (_Float16)s[e];
      } else {
        *reinterpret_cast<smat_f4*>(reinterpret_cast<float*>(C) + crow * ldc + n) = s;
      }
    } else {
      auto part = [&](int q) { return scratch[((int64_t)(rd.slot_beg + q) * 16 + row) * N + n]; };
      float s = 0.f;
      int q = 0;
      for (; q + 4 <= rd.slot_cnt; q += 4) {
        const float t0 = part(q), t1 = part(q + 1), t2 = part(q + 2), t3 = part(q + 3);
        s += t0;
        s += t1;
        s += t2;
        s += t3;
      }
      for (; q < rd.slot_cnt; q++) s += part(q);
      const int64_t crow = c_row();
      if (OUT_F16)
        reinterpret_cast<_Float16*>(C)[crow * ldc + n] = (_Float16)s;
      else
        reinterpret_cast<float*>(C)[crow * ldc + n] = s;
    }
  }
}

// ---- batched persistent launch -----------------------------------------------
// Many SpMMs (each with its own plan: block width, N-tile, work items) in one
// persistent launch per N-tile width present.  Every wave of a grid sized to
// fill the GPU once takes the next tasks (matrix, work item, N-tile) from a
// device counter, finds their matrix by binary search over the matrices'
// first task indices, and runs the LDS tile body of that matrix's block
// width.  Uneven matrices thus share one dynamic queue instead of one launch
// (and one tail) each.  The kernel is instantiated per tile width because one
// instance holding every (block width, tile width) body would take the
// register budget of the largest and halve the occupancy of the rest.
struct BatchMat {
  Operands op;
  Mixed24 mx;  // mixed (2:4) plans only
  const Item* items;
  int64_t task_beg;  // global index of this matrix's first task
  int32_t ntiles;
  int32_t kind;      // block-width code 0..5 for 1, 2, 4, 8, 16, 32; 6: mixed dense / 2:4 (16 x 16)
};

struct BatchRed {
  int32_t mat;
  int32_t brow;
  int32_t slot_beg;
  int32_t slot_cnt;
};

__global__ void k_batch_reset(unsigned long long* __restrict__ counters) {
  if (threadIdx.x < 32) counters[threadIdx.x] = 0ull;  // 4 tile widths x 8 XCD ranges
}

// Wave-uniform copies (scalar registers) of values every lane holds equally.
// readfirstlane returns a signed int: each half goes through uint32_t before
// widening, or a low word with bit 31 set would sign-extend into the high word.
__device__ inline int64_t uni(int64_t v) {
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  return (int64_t)(((uint64_t)hi << 32) | (uint64_t)lo);
}
template <typename T>
__device__ inline T* uni(T* p) {
  return reinterpret_cast<T*>(uni((int64_t)reinterpret_cast<uintptr_t>(p)));
}
__device__ inline Operands uni(const Operands& o) {
  Operands u;
  u.colidx = uni(o.colidx);
  u.a = uni(o.a);
  u.a_lo = uni(o.a_lo);
  u.a_scale = uni(o.a_scale);
  u.b = uni(o.b);
  u.ldb = uni(o.ldb);
  u.b_lo = uni(o.b_lo);
  u.b_scale = uni(o.b_scale);
  u.N = uni(o.N);
  u.c = uni(o.c);
  u.ldc = uni(o.ldc);
  u.scratch = uni(o.scratch);
  u.rowmap = uni(o.rowmap);
  u.b_bytes = (uint32_t)__builtin_amdgcn_readfirstlane((int)o.b_bytes);
  return u;
}

template <int DT, int NS>
__device__ __forceinline__ void batch_tile(int bkc, const Item& it, int nt, const Operands& op, const Mixed24& mx,
                                           uint8_t* lds, int lane) {
  const int64_t n0 = (int64_t)nt * 16 * NS;
  switch (bkc) {
    case 6:
      if constexpr (NS <= 4) mixed_tile<NS, DT, false>(it, n0, op, mx, lds, lane);  // 2:4 tiles are <= 64 wide
      break;
    case 0: lds_tile<1, NS, DT, false>(it, n0, op, lds, lane); break;
    case 1: lds_tile<2, NS, DT, false>(it, n0, op, lds, lane); break;
    case 2: lds_tile<4, NS, DT, false>(it, n0, op, lds, lane); break;
    case 3: lds_tile<8, NS, DT, false>(it, n0, op, lds, lane); break;
    case 4: lds_tile<16, NS, DT, false>(it, n0, op, lds, lane); break;
    default: lds_tile<32, NS, DT, false>(it, n0, op, lds, lane); break;
  }
}

// Persistent queue kernel.  The group's tasks are cut into 8 contiguous
// ranges, one per XCD (workgroup i runs on XCD i % 8), each with its own
// counter: neighbouring block rows then share an L2, and the atomics spread
// over 8 addresses.  A wave draws `chunk` consecutive tasks per atomic from
// its XCD's range, finds the first task's matrix by binary search and walks
// forward through the chunk; when its range is drained it takes chunks from
// the other ranges in turn.  Exit: counters only grow, so after at most one
// failed draw per range every wave leaves the loop.
template <int DT, int NS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NS == 8 ? 2 : NS == 4 ? 3 : 4)))
void k_batch(const BatchMat* __restrict__ mats, int32_t nmats, int64_t total, int64_t chunk,
             unsigned long long* __restrict__ counters) {
  // (the block width is only known per task: room for the A tile of 16x1 blocks)
  constexpr int WB = NS > 4 || lds_wave_bytes<1, NS, DT>() > mixed_wave_bytes<NS>() ? lds_wave_bytes<1, NS, DT>()
                                                                                     : mixed_wave_bytes<NS>();
  constexpr int NX = 8;
  __shared__ __attribute__((aligned(16))) uint8_t lds_all[kWaves * WB];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  uint8_t* lds = lds_all + wave * WB;
  const int xcd = (int)(blockIdx.x % NX);
  for (int q = 0; q < NX;) {
    const int r = (xcd + q) % NX;
    const int64_t rb = total * r / NX, re = total * (r + 1) / NX;
    unsigned long long t = 0;
    if (lane == 0) t = atomicAdd(&counters[r], (unsigned long long)chunk);
    const int64_t t0 = rb + uni((int64_t)t);  // lane 0's draw, wave-uniform
    if (t0 >= re) {
      q++;  // this range is drained
      continue;
    }
    const int64_t t1 = t0 + chunk < re ? t0 + chunk : re;
    int32_t lo = 0, hi = nmats - 1;
    while (lo < hi) {
      const int32_t mid = (lo + hi + 1) >> 1;
      if (mats[mid].task_beg <= t0) lo = mid;
      else hi = mid - 1;
    }
    // The chunk's tasks are consecutive: the matrix fields are loaded when the
    // walk enters a matrix and the item when it enters an item (the N tiles of
    // one item are consecutive tasks), not per task.
    int32_t cur = -1;
    int64_t tbeg = 0, next_beg = 0, cur_item = -1;
    uint32_t ntl = 1;
    int kind = 0;
    const Item* items = nullptr;
    Operands op{};
    Mixed24 mx{nullptr, nullptr, nullptr, nullptr};
    Item it{};
    for (int64_t task = t0; task < t1; task++) {
      if (cur < 0 || task >= next_beg) {
        while (lo + 1 < nmats && mats[lo + 1].task_beg <= task) lo++;
        const BatchMat& m = mats[lo];
        cur = lo;
        tbeg = uni(m.task_beg);
        next_beg = lo + 1 < nmats ? uni(mats[lo + 1].task_beg) : INT64_MAX;
        ntl = (uint32_t)__builtin_amdgcn_readfirstlane(m.ntiles);
        kind = __builtin_amdgcn_readfirstlane(m.kind);
        items = uni(m.items);
        op = uni(m.op);
        mx = kind == 6 ? Mixed24{uni(m.mx.dense), uni(m.mx.sparse), uni(m.mx.values24), uni(m.mx.index24)}
                       : Mixed24{nullptr, nullptr, nullptr, nullptr};
        cur_item = -1;
      }
      const uint32_t local = (uint32_t)(task - tbeg);  // < 2^31 per matrix
      const int64_t iid = local / ntl;
      if (iid != cur_item) {
        it = items[iid];
        cur_item = iid;
      }
      batch_tile<DT, NS>(kind, it, (int)(local % ntl), op, mx, lds, lane);
    }
  }
}

// Split-row reduction over every matrix of the batch (one workgroup per row).
__global__ void k_batch_reduce(const BatchRed* __restrict__ red, const BatchMat* __restrict__ mats) {
  const BatchRed rd = red[blockIdx.x];
  const Operands op = mats[rd.mat].op;
  const int64_t N = op.N, total = 16 * N;
  for (int64_t i = threadIdx.x; i < total; i += blockDim.x) {
    const int64_t row = i / N, n = i % N;
    float acc = 0.f;
    for (int q = 0; q < rd.slot_cnt; q++) acc += op.scratch[((int64_t)(rd.slot_beg + q) * 16 + row) * N + n];
    int64_t crow = (int64_t)rd.brow * 16 + row;
    if (op.rowmap) crow = op.rowmap[crow];
    static_cast<float*>(op.c)[crow * op.ldc + n] = acc;
  }
}

// ---- planning -------------------------------------------------------------
// A block row holds up to two lists: list 1 (dense, rp1) worked in steps of
// step1 blocks and list 2 (2:4, rp2, may be absent) in steps of step2.  A row
// whose total step count exceeds S is cut into k slices, each taking the
// same share of both lists.  Empty rows still get one item (it writes zeros).
__device__ inline int32_t ceil_div(int32_t a, int32_t b) { return (a + b - 1) / b; }

__global__ void k_plan_count(int64_t nbrows, const int32_t* __restrict__ rp1, int32_t step1,
                             const int32_t* __restrict__ rp2, int32_t step2, int32_t S,
                             int32_t* __restrict__ n_items, int32_t* __restrict__ n_slots,
                             int32_t* __restrict__ split_flag) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r > nbrows) return;
  if (r == nbrows) {
    n_items[r] = 0;
    n_slots[r] = 0;
    split_flag[r] = 0;
    return;
  }
  const int32_t n1 = rp1[r + 1] - rp1[r];
  const int32_t n2 = rp2 ? rp2[r + 1] - rp2[r] : 0;
  const int32_t steps = ceil_div(n1, step1) + ceil_div(n2, step2);
  const int32_t k = steps <= S ? 1 : ceil_div(steps, S);
  n_items[r] = k;
  n_slots[r] = k > 1 ? k : 0;
  split_flag[r] = k > 1 ? 1 : 0;
}

__global__ void k_plan_emit(int64_t nbrows, const int32_t* __restrict__ rp1, int32_t step1,
                            const int32_t* __restrict__ rp2, int32_t step2, const int32_t* __restrict__ item_off,
                            const int32_t* __restrict__ slot_off, const int32_t* __restrict__ red_off,
                            Item* __restrict__ items, Reduce* __restrict__ red) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nbrows) return;
  const int32_t b1 = rp1[r], e1 = rp1[r + 1];
  const int32_t b2 = rp2 ? rp2[r] : 0, e2 = rp2 ? rp2[r + 1] : 0;
  const int32_t k = item_off[r + 1] - item_off[r];
  const int32_t base = item_off[r];
  if (k == 1) {
    items[base] = Item{(int32_t)r, b1, e1, -1, b2, e2, 0, 0};
    return;
  }
  const int32_t per1 = ceil_div(ceil_div(e1 - b1, step1), k) * step1;
  const int32_t per2 = ceil_div(ceil_div(e2 - b2, step2), k) * step2;
  const int32_t sbase = slot_off[r];
  for (int32_t q = 0; q < k; q++) {
    const int32_t d0 = min(b1 + q * per1, e1), d1 = min(d0 + per1, e1);
    const int32_t s0 = min(b2 + q * per2, e2), s1 = min(s0 + per2, e2);
    items[base + q] = Item{(int32_t)r, d0, d1, sbase + q, s0, s1, 0, 0};
  }
  red[red_off[r]] = Reduce{(int32_t)r, sbase, k, 0};
}

// Longest and total block-row length (blocks) of a BCSR.
__global__ void k_row_len_stats(int64_t nbrows, const int32_t* __restrict__ rowptr,
                                unsigned long long* __restrict__ out) {
  __shared__ unsigned long long mx, tot;
  if (threadIdx.x == 0) mx = tot = 0;
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < nbrows) {
    const unsigned long long len = (unsigned long long)(rowptr[r + 1] - rowptr[r]);
    atomicMax(&mx, len);
    atomicAdd(&tot, len);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicMax(&out[0], mx);
    atomicAdd(&out[1], tot);
  }
}

__global__ void k_zero2(unsigned long long* p) {
  if (threadIdx.x < 2) p[threadIdx.x] = 0;
}

// Longest-first schedules: sort key (MFMA steps of the item) and identity index.
__global__ void k_item_keys(int64_t n, const Item* __restrict__ items, int32_t step1, int32_t step2,
                            int32_t* __restrict__ key, int32_t* __restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Item it = items[i];
  key[i] = ceil_div(it.pend - it.pbeg, step1) + ceil_div(it.send - it.sbeg, step2);
  idx[i] = (int32_t)i;
}

__global__ void k_item_gather(int64_t n, const Item* __restrict__ items, const int32_t* __restrict__ idx,
                              Item* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = items[idx[i]];
}

// Mixed plans: per-block flags and list positions of tag-2 and tag-1 blocks.
__global__ void k_list_flags(int64_t nnzb, const int32_t* __restrict__ tags, int32_t* __restrict__ fd,
                             int32_t* __restrict__ fs) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nnzb) return;
  const int t = b < nnzb ? tags[b] : 0;
  fd[b] = t == 2;
  fs[b] = t == 1;
}

__global__ void k_list_emit(int64_t nnzb, int64_t nbrows, const int32_t* __restrict__ tags,
                            const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
                            const int32_t* __restrict__ pd, const int32_t* __restrict__ ps,
                            int2* __restrict__ dense, int2* __restrict__ sparse, int32_t* __restrict__ drp,
                            int32_t* __restrict__ srp) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b <= nbrows) {
    drp[b] = pd[rowptr[b]];
    srp[b] = ps[rowptr[b]];
  }
  if (b >= nnzb) return;
  if (tags[b] == 2) dense[pd[b]] = int2{(int32_t)b, colidx[b]};
  else if (tags[b] == 1) sparse[ps[b]] = int2{(int32_t)b, colidx[b]};
}

inline unsigned blocks_for(int64_t n, int t) {
  int64_t g = (n + t - 1) / t;
  return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace

struct smat_plan_t {
  smat_bcsr_desc A{};
  int64_t N = 0;
  smat_spmm_opts opts{};
  int32_t nsub = 4;
  bool mixed = false;
  int32_t order = kOrderXcd;  // work order of the one-shot kernels (kOrder*)
  int64_t n_items = 0, n_red = 0, n_slots = 0;
  int64_t n_dense = 0, n_sparse = 0;
  smat::DevBuf items, red, scratch, dense, sparse;  // mixed plans: list entries {block, column start}
};

namespace {

inline int32_t ntiles_of(const smat_plan_t* P, int nsub) { return (int32_t)((P->N + 16 * nsub - 1) / (16 * nsub)); }

// Work order for a launch with ntiles N tiles (tile affinity needs ntiles | 8
// or 8 | ntiles; otherwise the plain order, which the sorted items keep
// longest first).
inline int32_t order_of(const smat_plan_t* P, int32_t ntiles) {
  if (P->order == kOrderTiles && !(8 % ntiles == 0 || ntiles % 8 == 0)) return kOrderPlain;
  return P->order;
}

// Workgroups of a one-shot launch in the given order.
inline unsigned grid_of(const smat_plan_t* P, int32_t ntiles, int32_t order) {
  if (order != kOrderTiles) return blocks_for(P->n_items * ntiles, kWaves);
  const int64_t groups = (P->n_items + kWaves - 1) / kWaves;
  const int64_t rounds = ntiles <= 8 ? (groups + 8 / ntiles - 1) / (8 / ntiles) : (int64_t)(ntiles / 8) * groups;
  return (unsigned)(rounds < 1 ? 8 : 8 * rounds);
}

template <int BK, int NSUB, int BLAYOUT, int DT, bool OUT_F16>
void launch_dense(const smat_plan_t* P, const Operands& op, hipStream_t s) {
  const int32_t ntiles = ntiles_of(P, NSUB);
  const int32_t order = order_of(P, ntiles);
  hipLaunchKernelGGL((k_spmm_dense<BK, NSUB, BLAYOUT, DT, OUT_F16>), grid_of(P, ntiles, order), 64 * kWaves, 0, s,
                     P->items.as<Item>(), P->n_items, ntiles, op.colidx, static_cast<const uint16_t*>(op.a),
                     static_cast<const uint16_t*>(op.b), op.ldb, op.N, op.c, op.ldc, op.scratch, op.rowmap, order);
}

template <int BK, int NSUB, int DT, bool OUT_F16>
void launch_lds(const smat_plan_t* P, const Operands& op, hipStream_t s) {
  const int32_t ntiles = ntiles_of(P, NSUB);
  const int32_t order = order_of(P, ntiles);
  if constexpr (NSUB == 4 && DT <= 1) {
    if (op.rowmap) {
      hipLaunchKernelGGL((k_spmm_lds64<BK, DT, OUT_F16>), grid_of(P, ntiles, order), 64 * kWaves, 0, s,
                         P->items.as<Item>(), P->n_items, ntiles, op, order);
      return;
    }
  }
  hipLaunchKernelGGL((k_spmm_lds<BK, NSUB, DT, OUT_F16>), grid_of(P, ntiles, order), 64 * kWaves, 0, s,
                     P->items.as<Item>(), P->n_items, ntiles, op, order);
}

template <int NSUB, int DT, bool OUT_F16>
void launch_mixed(const smat_plan_t* P, const uint16_t* v24, const uint16_t* i24, const Operands& op,
                  hipStream_t s) {
  const int32_t ntiles = ntiles_of(P, NSUB);
  const Mixed24 mx{P->dense.as<int2>(), P->sparse.as<int2>(), v24, i24};
  const int32_t order = order_of(P, ntiles);
  hipLaunchKernelGGL((k_spmm_mixed<NSUB, DT, OUT_F16>), grid_of(P, ntiles, order), 64 * kWaves, 0, s,
                     P->items.as<Item>(), P->n_items, ntiles, op, mx, order);
}

// LDS staging of row-major B: 16-byte chunks (N and ldb multiples of the
// chunk, B 16-byte aligned) and 32-bit byte offsets (every B array the call
// reads under 4 GiB).
// Bytes of one row-major B array of a call (the range of its buffer resource
// in the LDS kernel; 0 where the LDS kernel cannot run).
inline uint32_t b_array_bytes(const smat_plan_t* P, int64_t ldb) {
  const int es = P->opts.in_dtype == SMAT_FP8_E4M3 ? 1 : 2;
  const double bytes = (double)P->A.K * (double)ldb * es;
  return bytes <= kMaxResourceBytes ? (uint32_t)(P->A.K * ldb * es) : 0u;
}

inline bool lds_aligned(const smat_plan_t* P, const Operands& op, int elems_per_chunk) {
  const int es = P->opts.in_dtype == SMAT_FP8_E4M3 ? 1 : 2;
  const int narr = P->opts.in_dtype == SMAT_SPLIT_F16 ? 2 : 1;
  const double bytes = (double)P->A.K * (double)op.ldb * es * narr;
  return P->N % elems_per_chunk == 0 && op.ldb % elems_per_chunk == 0 &&
         (reinterpret_cast<uintptr_t>(op.b) & 15) == 0 && bytes <= kMaxResourceBytes;
}

template <int BK, int BLAYOUT, int DT, bool OUT_F16>
void dispatch_nsub(const smat_plan_t* P, const Operands& op, hipStream_t s) {
  if (BLAYOUT == SMAT_B_ROW_MAJOR && P->opts.b_staging != SMAT_BSTAGE_DIRECT && lds_aligned(P, op, 8)) {
    switch (P->nsub) {
      case 1: launch_lds<BK, 1, DT, OUT_F16>(P, op, s); return;
      case 2: launch_lds<BK, 2, DT, OUT_F16>(P, op, s); return;
      case 4: launch_lds<BK, 4, DT, OUT_F16>(P, op, s); return;
      default: launch_lds<BK, 8, DT, OUT_F16>(P, op, s); return;
    }
  }
  // (B not stageable at this call: 128-column plans run 64-column tiles; the
  // items do not depend on the tile width)
  switch (P->nsub) {
    case 1: launch_dense<BK, 1, BLAYOUT, DT, OUT_F16>(P, op, s); break;
    case 2: launch_dense<BK, 2, BLAYOUT, DT, OUT_F16>(P, op, s); break;
    default: launch_dense<BK, 4, BLAYOUT, DT, OUT_F16>(P, op, s); break;
  }
}

template <int BK, int DT>
void dispatch_layout(const smat_plan_t* P, const Operands& op, hipStream_t s) {
  const bool f16 = P->opts.out_dtype == SMAT_F16;
  if (P->opts.b_layout == SMAT_B_REFERENCE) {
    if (f16) dispatch_nsub<BK, SMAT_B_REFERENCE, DT, true>(P, op, s);
    else dispatch_nsub<BK, SMAT_B_REFERENCE, DT, false>(P, op, s);
  } else {
    if (f16) dispatch_nsub<BK, SMAT_B_ROW_MAJOR, DT, true>(P, op, s);
    else dispatch_nsub<BK, SMAT_B_ROW_MAJOR, DT, false>(P, op, s);
  }
}

// FP8 and split codecs: row-major B through LDS only, FP32 C.  FP8 tiles are
// at least 32 columns wide (a 16-byte load then spans at most 32 B rows) and
// 128 by default (each B row segment a whole 128-byte line).
template <int BK, int DT>
void dispatch_scaled(const smat_plan_t* P, const Operands& op, hipStream_t s) {
  if (DT == 2) {
    if (P->nsub == 8) launch_lds<BK, 8, 2, false>(P, op, s);
    else if (P->nsub == 4) launch_lds<BK, 4, 2, false>(P, op, s);
    else launch_lds<BK, 2, 2, false>(P, op, s);
  } else {
    switch (P->nsub) {
      case 1: launch_lds<BK, 1, 3, false>(P, op, s); break;
      case 2: launch_lds<BK, 2, 3, false>(P, op, s); break;
      default: launch_lds<BK, 4, 3, false>(P, op, s); break;
    }
  }
}

template <int BK>
void dispatch_dt(const smat_plan_t* P, const Operands& op, hipStream_t s) {
  switch (P->opts.in_dtype) {
    case SMAT_BF16: dispatch_layout<BK, 1>(P, op, s); break;
    case SMAT_FP8_E4M3: dispatch_scaled<BK, 2>(P, op, s); break;
    case SMAT_SPLIT_F16: dispatch_scaled<BK, 3>(P, op, s); break;
    default: dispatch_layout<BK, 0>(P, op, s); break;
  }
}

void dispatch_bk(const smat_plan_t* P, const Operands& op, hipStream_t s) {
  switch (P->A.bk) {
    case 1: dispatch_dt<1>(P, op, s); break;
    case 2: dispatch_dt<2>(P, op, s); break;
    case 4: dispatch_dt<4>(P, op, s); break;
    case 8: dispatch_dt<8>(P, op, s); break;
    case 16: dispatch_dt<16>(P, op, s); break;
    default: dispatch_dt<32>(P, op, s); break;
  }
}

template <int DT>
void dispatch_mixed(const smat_plan_t* P, const uint16_t* v24, const uint16_t* i24, const Operands& op,
                    hipStream_t s) {
  const bool f16 = P->opts.out_dtype == SMAT_F16;
  switch (P->nsub) {
    case 1: f16 ? launch_mixed<1, DT, true>(P, v24, i24, op, s) : launch_mixed<1, DT, false>(P, v24, i24, op, s); break;
    case 2: f16 ? launch_mixed<2, DT, true>(P, v24, i24, op, s) : launch_mixed<2, DT, false>(P, v24, i24, op, s); break;
    default: f16 ? launch_mixed<4, DT, true>(P, v24, i24, op, s) : launch_mixed<4, DT, false>(P, v24, i24, op, s); break;
  }
}

smat_status row_length_stats(const smat_bcsr_desc* A, int64_t* mx, int64_t* tot, hipStream_t s) {
  smat::DevBuf acc;
  SMAT_TRY(acc.alloc(16));
  hipLaunchKernelGGL(k_zero2, 1, 64, 0, s, acc.as<unsigned long long>());
  hipLaunchKernelGGL(k_row_len_stats, blocks_for(A->nbrows, 256), 256, 0, s, A->nbrows, A->rowptr,
                     acc.as<unsigned long long>());
  SMAT_LAUNCH_CHECK();
  unsigned long long h[2] = {0, 0};
  SMAT_HIP_TRY(hipMemcpyAsync(h, acc.p, 16, hipMemcpyDeviceToHost, s));
  SMAT_HIP_TRY(hipStreamSynchronize(s));
  *mx = (int64_t)h[0];
  *tot = (int64_t)h[1];
  return SMAT_OK;
}

// Exclusive scan helper over n int32 with a reusable temp buffer.
smat_status exscan(smat::DevBuf& tmp, const int32_t* in, int32_t* out, int64_t n, hipStream_t s) {
  size_t tb = 0;
  SMAT_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, in, out, (int)n, s));
  if (tb > tmp.bytes) SMAT_TRY(tmp.alloc(tb));
  tb = tmp.bytes;
  SMAT_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, in, out, (int)n, s));
  return SMAT_OK;
}

smat_status build_items(smat_plan_t* P, const int32_t* rp1, int32_t step1, const int32_t* rp2, int32_t step2,
                        int32_t S, hipStream_t s) {
  const int64_t nbrows = P->A.nbrows, nb1 = nbrows + 1;
  smat::DevBuf cnt_i, cnt_s, flag, off_i, off_s, off_f, tmp;
  SMAT_TRY(cnt_i.alloc(nb1 * 4));
  SMAT_TRY(cnt_s.alloc(nb1 * 4));
  SMAT_TRY(flag.alloc(nb1 * 4));
  SMAT_TRY(off_i.alloc(nb1 * 4));
  SMAT_TRY(off_s.alloc(nb1 * 4));
  SMAT_TRY(off_f.alloc(nb1 * 4));
  hipLaunchKernelGGL(k_plan_count, blocks_for(nb1, 256), 256, 0, s, nbrows, rp1, step1, rp2, step2, S,
                     cnt_i.as<int32_t>(), cnt_s.as<int32_t>(), flag.as<int32_t>());
  SMAT_LAUNCH_CHECK();
  SMAT_TRY(exscan(tmp, cnt_i.as<int32_t>(), off_i.as<int32_t>(), nb1, s));
  SMAT_TRY(exscan(tmp, cnt_s.as<int32_t>(), off_s.as<int32_t>(), nb1, s));
  SMAT_TRY(exscan(tmp, flag.as<int32_t>(), off_f.as<int32_t>(), nb1, s));
  int32_t tot[3] = {0, 0, 0};
  SMAT_HIP_TRY(hipMemcpyAsync(&tot[0], off_i.as<int32_t>() + nbrows, 4, hipMemcpyDeviceToHost, s));
  SMAT_HIP_TRY(hipMemcpyAsync(&tot[1], off_s.as<int32_t>() + nbrows, 4, hipMemcpyDeviceToHost, s));
  SMAT_HIP_TRY(hipMemcpyAsync(&tot[2], off_f.as<int32_t>() + nbrows, 4, hipMemcpyDeviceToHost, s));
  SMAT_HIP_TRY(hipStreamSynchronize(s));
  P->n_items = tot[0];
  P->n_slots = tot[1];
  P->n_red = tot[2];
  SMAT_TRY(P->items.alloc(P->n_items * sizeof(Item)));
  SMAT_TRY(P->red.alloc(P->n_red * sizeof(Reduce)));
  SMAT_TRY(P->scratch.alloc(P->n_slots * 16 * P->N * sizeof(float)));
  hipLaunchKernelGGL(k_plan_emit, blocks_for(nbrows, 256), 256, 0, s, nbrows, rp1, step1, rp2, step2,
                     off_i.as<int32_t>(), off_s.as<int32_t>(), off_f.as<int32_t>(), P->items.as<Item>(),
                     P->red.as<Reduce>());
  SMAT_LAUNCH_CHECK();
  if (P->order != kOrderXcd && P->n_items > 1) {
    // longest first: items sorted by their MFMA steps, descending (stable)
    const int64_t n = P->n_items;
    smat::DevBuf k0, k1, i0, i1, sorted, tmp2;
    SMAT_TRY(k0.alloc(n * 4));
    SMAT_TRY(k1.alloc(n * 4));
    SMAT_TRY(i0.alloc(n * 4));
    SMAT_TRY(i1.alloc(n * 4));
    SMAT_TRY(sorted.alloc(n * sizeof(Item)));
    hipLaunchKernelGGL(k_item_keys, blocks_for(n, 256), 256, 0, s, n, P->items.as<Item>(), step1, step2,
                       k0.as<int32_t>(), i0.as<int32_t>());
    size_t tb = 0;
    SMAT_HIP_TRY(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tb, k0.as<int32_t>(), k1.as<int32_t>(),
                                                              i0.as<int32_t>(), i1.as<int32_t>(), (int)n, 0, 32, s));
    SMAT_TRY(tmp2.alloc(tb));
    SMAT_HIP_TRY(hipcub::DeviceRadixSort::SortPairsDescending(tmp2.p, tb, k0.as<int32_t>(), k1.as<int32_t>(),
                                                              i0.as<int32_t>(), i1.as<int32_t>(), (int)n, 0, 32, s));
    hipLaunchKernelGGL(k_item_gather, blocks_for(n, 256), 256, 0, s, n, P->items.as<Item>(), i1.as<int32_t>(),
                       sorted.as<Item>());
    SMAT_LAUNCH_CHECK();
    SMAT_HIP_TRY(hipMemcpyAsync(P->items.p, sorted.p, n * sizeof(Item), hipMemcpyDeviceToDevice, s));
    SMAT_HIP_TRY(hipStreamSynchronize(s));
  }
  SMAT_HIP_TRY(hipStreamSynchronize(s));
  return SMAT_OK;
}

// Mixed plans: dense and 2:4 block-id lists grouped by block row, then items.
smat_status build_mixed(smat_plan_t* P, int32_t S, hipStream_t s) {
  const smat_bcsr_desc& A = P->A;
  const int64_t n = A.nnzb, nb1 = A.nbrows + 1;
  smat::DevBuf fd, fs, pd, ps, drp, srp, tmp;
  SMAT_TRY(fd.alloc((n + 1) * 4));
  SMAT_TRY(fs.alloc((n + 1) * 4));
  SMAT_TRY(pd.alloc((n + 1) * 4));
  SMAT_TRY(ps.alloc((n + 1) * 4));
  SMAT_TRY(drp.alloc(nb1 * 4));
  SMAT_TRY(srp.alloc(nb1 * 4));
  hipLaunchKernelGGL(k_list_flags, blocks_for(n + 1, 256), 256, 0, s, n, A.tags, fd.as<int32_t>(), fs.as<int32_t>());
  SMAT_LAUNCH_CHECK();
  SMAT_TRY(exscan(tmp, fd.as<int32_t>(), pd.as<int32_t>(), n + 1, s));
  SMAT_TRY(exscan(tmp, fs.as<int32_t>(), ps.as<int32_t>(), n + 1, s));
  int32_t cnt[2] = {0, 0};
  SMAT_HIP_TRY(hipMemcpyAsync(&cnt[0], pd.as<int32_t>() + n, 4, hipMemcpyDeviceToHost, s));
  SMAT_HIP_TRY(hipMemcpyAsync(&cnt[1], ps.as<int32_t>() + n, 4, hipMemcpyDeviceToHost, s));
  SMAT_HIP_TRY(hipStreamSynchronize(s));
  P->n_dense = cnt[0];
  P->n_sparse = cnt[1];
  SMAT_TRY(P->dense.alloc(P->n_dense * 8));
  SMAT_TRY(P->sparse.alloc(P->n_sparse * 8));
  const int64_t m = n > A.nbrows ? n : A.nbrows + 1;
  hipLaunchKernelGGL(k_list_emit, blocks_for(m, 256), 256, 0, s, n, A.nbrows, A.tags, A.rowptr, A.colidx,
                     pd.as<int32_t>(), ps.as<int32_t>(), P->dense.as<int2>(), P->sparse.as<int2>(),
                     drp.as<int32_t>(), srp.as<int32_t>());
  SMAT_LAUNCH_CHECK();
  return build_items(P, drp.as<int32_t>(), 2, srp.as<int32_t>(), 4, S, s);
}

// Validates the options and builds the work decomposition of an SpMM of A
// (N columns) into P.  Its device arrays come from DevBuf::alloc: owned
// allocations, or the static scratch under an ArenaScope.
smat_status plan_init(smat_plan_t* P, const smat_bcsr_desc* A, int64_t N, const smat_spmm_opts* opts,
                      hipStream_t s) {
  SMAT_REQUIRE(A, "smat_spmm_plan_create: null argument");
  SMAT_REQUIRE(A->bk == 1 || A->bk == 2 || A->bk == 4 || A->bk == 8 || A->bk == 16 || A->bk == 32,
               "smat_spmm_plan_create: bk must be 1, 2, 4, 8, 16 or 32");
  SMAT_REQUIRE(N > 0, "smat_spmm_plan_create: N must be positive");
  SMAT_REQUIRE(A->nbrows > 0 && A->nbrows < (1 << 30), "smat_spmm_plan_create: bad block-row count");
  smat_spmm_opts o{};
  if (opts) o = *opts;
  if (o.out_dtype == 0) o.out_dtype = SMAT_F32;
  if (o.in_dtype == 0) o.in_dtype = SMAT_F16;
  SMAT_REQUIRE(o.out_dtype == SMAT_F32 || o.out_dtype == SMAT_F16, "smat_spmm_plan_create: out_dtype");
  SMAT_REQUIRE(o.in_dtype == SMAT_F16 || o.in_dtype == SMAT_BF16 || o.in_dtype == SMAT_FP8_E4M3 ||
                   o.in_dtype == SMAT_SPLIT_F16,
               "smat_spmm_plan_create: unknown in_dtype %d", o.in_dtype);
  if (o.in_dtype == SMAT_FP8_E4M3 || o.in_dtype == SMAT_SPLIT_F16)
    SMAT_REQUIRE(o.kernel != SMAT_KERNEL_MIXED24 && o.b_layout == SMAT_B_ROW_MAJOR && o.out_dtype == SMAT_F32 &&
                     o.b_staging != SMAT_BSTAGE_DIRECT,
                 "smat_spmm_plan_create: FP8 and split inputs need the dense kernel, row-major B staged "
                 "through LDS and FP32 C");
  SMAT_REQUIRE(o.b_layout == SMAT_B_ROW_MAJOR || o.b_layout == SMAT_B_REFERENCE,
               "smat_spmm_plan_create: b_layout");
  SMAT_REQUIRE(o.kernel == SMAT_KERNEL_AUTO || o.kernel == SMAT_KERNEL_DENSE || o.kernel == SMAT_KERNEL_MIXED24,
               "smat_spmm_plan_create: unknown kernel %d", o.kernel);
  if (o.kernel == SMAT_KERNEL_MIXED24)
    SMAT_REQUIRE(A->bk == 16 && o.b_layout == SMAT_B_ROW_MAJOR,
                 "smat_spmm_plan_create: the 2:4 kernel needs bk = 16 and row-major B");
  SMAT_REQUIRE(o.n_tile == 0 || o.n_tile == 16 || o.n_tile == 32 || o.n_tile == 64 || o.n_tile == 128,
               "smat_spmm_plan_create: n_tile must be 0, 16, 32, 64 or 128");
  if (o.n_tile == 128)
    SMAT_REQUIRE(o.kernel != SMAT_KERNEL_MIXED24 &&
                     (o.in_dtype == SMAT_F16 || o.in_dtype == SMAT_BF16 || o.in_dtype == SMAT_FP8_E4M3) &&
                     o.b_layout == SMAT_B_ROW_MAJOR && o.b_staging != SMAT_BSTAGE_DIRECT,
                 "smat_spmm_plan_create: n_tile 128 is for the dense kernel with FP16 / BF16 / FP8 row-major B "
                 "staged through LDS");
  SMAT_REQUIRE(o.b_staging >= SMAT_BSTAGE_AUTO && o.b_staging <= SMAT_BSTAGE_LDS,
               "smat_spmm_plan_create: unknown b_staging %d", o.b_staging);
  SMAT_REQUIRE(o.schedule >= SMAT_SCHED_AUTO && o.schedule <= SMAT_SCHED_TILE_AFFINITY,
               "smat_spmm_plan_create: unknown schedule %d", o.schedule);
  int nt = o.n_tile ? o.n_tile : (N >= 64 ? 64 : (N > 16 ? 32 : 16));
  if (o.in_dtype == SMAT_FP8_E4M3) {
    // FP8 tiles are at least 32 columns wide; by default 128 (a whole
    // 128-byte line of each B row) when N allows
    if (nt < 32) nt = 32;
    if (o.n_tile == 0 && N >= 128) nt = 128;
  }
  if (o.schedule == SMAT_SCHED_AUTO) {
    // skewed block rows (the longest over 8x the mean) go longest first,
    // with tile affinity when every N tile is shared by 2+ XCDs (with more
    // tiles each XCD would re-read all of A for its one slice of B); 16x1
    // blocks only with 4 tiles (with 2, longest first alone measured
    // faster: R-MAT s18 N = 64 in 32-column tiles 0.098 -> 0.068 ms,
    // power-law N = 128 in 64-column tiles 0.337 -> 0.332 ms)
    int64_t mx = 0, tot = 0;
    SMAT_TRY(row_length_stats(A, &mx, &tot, s));
    const int64_t ntl = (N + nt - 1) / nt;
    const bool affinity = ntl == 4 || (ntl == 2 && A->bk != 1);
    o.schedule = !(tot > 0 && mx * A->nbrows > 8 * tot) ? SMAT_SCHED_BLOCK_ROWS
                 : affinity                            ? SMAT_SCHED_TILE_AFFINITY
                                                       : SMAT_SCHED_LONGEST_FIRST;
  }

  P->A = *A;
  P->N = N;
  P->opts = o;
  P->mixed = o.kernel == SMAT_KERNEL_MIXED24;
  P->order = o.schedule == SMAT_SCHED_BLOCK_ROWS ? kOrderXcd
             : o.schedule == SMAT_SCHED_TILE_AFFINITY ? kOrderTiles : kOrderPlain;
  P->nsub = nt / 16;
  const int32_t step = 32 / A->bk;  // blocks per MFMA step
  int32_t S = o.max_blocks_per_item;  // in blocks
  if (S <= 0) {
    const int64_t ntl = (N + nt - 1) / nt;
    const int64_t s_min = 12 * (int64_t)step;  // finer slices only add partial tiles to reduce
    int64_t s_auto;
    if (P->order == kOrderXcd) {
      // block-row order: aim for >= 64k wave tasks so the tail stays short
      const int64_t target_items = (65536 + ntl - 1) / ntl;
      s_auto = (A->nnzb + target_items - 1) / target_items;
    } else {
      // longest first: a slice may be as long as half of one resident
      // wave's fair share of the work (total steps x tiles / resident
      // waves); longer rows start first and end within the kernel's span
      // without being cut
      int dev = 0, cus = 256;
      if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        (void)hipGetLastError();
        cus = 256;
      }
      // (half of it: the last rows to start still end with the rest)
      const int64_t waves = (int64_t)cus * 16;
      const int64_t steps = (A->nnzb + step - 1) / step;
      if (A->bk == 1) {
        // 16x1 blocks: every slice refills the two-step load pipeline and
        // writes a partial tile, so a slice runs a whole fair share plus 32
        // steps, and at least 64 steps.  Measured best slice lengths: R-MAT
        // scale 18 N = 256 about 128 steps (fair share 93), N = 64 about 64
        // (fair share 23), power-law N = 128 about 170-214 (fair share 170).
        s_auto = std::max<int64_t>((steps * ntl + waves - 1) / waves + 32, 64) * step;
      } else {
        s_auto = (steps * ntl + 2 * waves - 1) / (2 * waves) * step;
      }
    }
    S = (int32_t)(s_auto < s_min ? s_min : (s_auto > 65536 ? 65536 : s_auto));
  }
  // an item spans at most 65536 blocks (the LDS kernels address an item's A
  // values with 32-bit offsets from its first block)
  if (S > 65536) S = 65536;
  if (P->mixed) return build_mixed(P, (S + 1) / 2, s);  // S in MFMA steps (a dense pair or a 2:4 quad each)
  return build_items(P, A->rowptr, step, nullptr, 1, (S + step - 1) / step, s);
}

}  // namespace

extern "C" {

smat_status smat_spmm_plan_create(const smat_bcsr_desc* A, int64_t N, const smat_spmm_opts* opts,
                                  smat_stream stream, smat_plan* out) {
  SMAT_REQUIRE(A && out, "smat_spmm_plan_create: null argument");
  *out = nullptr;
  smat_plan_t* P = new smat_plan_t();
  const smat_status st = plan_init(P, A, N, opts, smat::to_hip(stream));
  if (st != SMAT_OK) {
    delete P;
    return st;
  }
  *out = P;
  return SMAT_OK;
}

}  // extern "C"

namespace {

// Checks the operands of one call against its plan and fills the kernels'
// view of them (no launch).
smat_status prepare_operands(smat_plan P, const smat_spmm_operands* ops, Operands* out) {
  SMAT_REQUIRE(P && ops && ops->b && ops->c, "smat_spmm_execute: null argument");
  const int64_t ldb = ops->ldb, ldc = ops->ldc;
  const int dt = P->opts.in_dtype;
  const bool scaled = dt == SMAT_FP8_E4M3 || dt == SMAT_SPLIT_F16;
  SMAT_REQUIRE(ldc >= P->N, "smat_spmm_execute: ldc < N");
  if (P->opts.b_layout == SMAT_B_ROW_MAJOR)
    SMAT_REQUIRE(ldb >= P->N, "smat_spmm_execute: ldb < N for row-major B");
  else
    SMAT_REQUIRE(ldb >= P->A.K, "smat_spmm_execute: ldb < K for reference-layout B");
  const void* a = ops->a_values ? ops->a_values : P->A.values;
  SMAT_REQUIRE(a, "smat_spmm_execute: no A values");
  Operands op{P->A.colidx, a, 0, ops->a_scale, ops->b, ldb, 0, ops->b_scale, P->N, ops->c, ldc,
              P->scratch.as<float>(), ops->c_rowmap, b_array_bytes(P, ldb)};
  const int chunk = dt == SMAT_FP8_E4M3 ? 16 : 8;  // elements per 16-byte B load
  const bool aligned = lds_aligned(P, op, chunk);
  if ((P->opts.b_staging == SMAT_BSTAGE_LDS && P->opts.b_layout == SMAT_B_ROW_MAJOR) || P->mixed || scaled)
    SMAT_REQUIRE(aligned,
                 "smat_spmm_execute: LDS staging needs N and ldb multiples of %d, a 16-byte aligned B and B "
                 "arrays under 4 GiB",
                 chunk);
  if (scaled) {
    SMAT_REQUIRE(ops->a_scale && ops->b_scale, "smat_spmm_execute: FP8 and split operands need a_scale and b_scale");
    SMAT_REQUIRE((reinterpret_cast<uintptr_t>(a) & 15) == 0, "smat_spmm_execute: A values must be 16-byte aligned");
    if (dt == SMAT_SPLIT_F16) {
      op.a_lo = P->A.nnzb * 16 * P->A.bk;  // [hi | lo], smat_pack_a
      op.b_lo = P->A.K * ldb;               // [hi | lo], smat_pack_b
    }
  }
  if (P->mixed)
    SMAT_REQUIRE(ops->values24 && ops->index24 && (reinterpret_cast<uintptr_t>(ops->values24) & 15) == 0,
                 "smat_spmm_execute: the 2:4 kernel needs 16-byte aligned values24 and index24");
  *out = op;
  return SMAT_OK;
}

// Launches one call's kernels (main kernel, then the split-row reduction).
smat_status launch_call(smat_plan P, const smat_spmm_operands* ops, const Operands& op, hipStream_t s) {
  const int dt = P->opts.in_dtype;
  const int64_t ldc = ops->ldc;
  if (P->mixed) {
    const uint16_t* i24 = static_cast<const uint16_t*>(ops->index24);
    if (dt == SMAT_BF16) dispatch_mixed<1>(P, ops->values24, i24, op, s);
    else dispatch_mixed<0>(P, ops->values24, i24, op, s);
  } else {
    dispatch_bk(P, op, s);
  }
  if (P->n_red > 0) {
    const bool f16 = P->opts.out_dtype == SMAT_F16;
    // 16-byte scratch loads need N % 4 == 0; FP32 C also needs 16-byte aligned rows
    const bool vec = P->N % 4 == 0 &&
                     (f16 || (ldc % 4 == 0 && (reinterpret_cast<uintptr_t>(ops->c) & 15) == 0));
    const int32_t chunks = (int32_t)((16 * P->N / (vec ? 4 : 1) + 255) / 256);
    const unsigned g = (unsigned)(chunks * P->n_red);
    const Reduce* rd = P->red.as<Reduce>();
    const float* sc = P->scratch.as<float>();
    if (vec) {
      if (f16) hipLaunchKernelGGL((k_reduce<true, true>), g, 256, 0, s, rd, P->n_red, chunks, sc, P->N, ops->c, ldc, ops->c_rowmap);
      else hipLaunchKernelGGL((k_reduce<false, true>), g, 256, 0, s, rd, P->n_red, chunks, sc, P->N, ops->c, ldc, ops->c_rowmap);
    } else {
      if (f16) hipLaunchKernelGGL((k_reduce<true, false>), g, 256, 0, s, rd, P->n_red, chunks, sc, P->N, ops->c, ldc, ops->c_rowmap);
      else hipLaunchKernelGGL((k_reduce<false, false>), g, 256, 0, s, rd, P->n_red, chunks, sc, P->N, ops->c, ldc, ops->c_rowmap);
    }
  }
  SMAT_LAUNCH_CHECK();
  return SMAT_OK;
}

}  // namespace

extern "C" {

smat_status smat_spmm_execute_ex(smat_plan P, const smat_spmm_operands* ops, smat_stream stream) {
  Operands op{};
  SMAT_TRY(prepare_operands(P, ops, &op));
  return launch_call(P, ops, op, smat::to_hip(stream));
}

smat_status smat_spmm_execute(smat_plan P, const uint16_t* values24, const void* index24, const uint16_t* B,
                              int64_t ldb, void* C, int64_t ldc, smat_stream stream) {
  SMAT_REQUIRE(P, "smat_spmm_execute: null plan");
  SMAT_REQUIRE(P->opts.in_dtype == SMAT_F16 || P->opts.in_dtype == SMAT_BF16,
               "smat_spmm_execute: FP8 and split plans take their scales through smat_spmm_execute_ex");
  smat_spmm_operands ops{};
  ops.b = B;
  ops.ldb = ldb;
  ops.c = C;
  ops.ldc = ldc;
  ops.values24 = values24;
  ops.index24 = index24;
  return smat_spmm_execute_ex(P, &ops, stream);
}

smat_status smat_spmm_plan_destroy(smat_plan P) {
  delete P;
  return SMAT_OK;
}

smat_status smat_spmm_plan_info(smat_plan P, int64_t* n_items, int64_t* n_split_rows, int64_t* scratch_bytes) {
  SMAT_REQUIRE(P, "smat_spmm_plan_info: null plan");
  if (n_items) *n_items = P->n_items;
  if (n_split_rows) *n_split_rows = P->n_red;
  if (scratch_bytes) *scratch_bytes = (int64_t)P->scratch.bytes;
  return SMAT_OK;
}

}  // extern "C"

// ---- SpMM without allocation ------------------------------------------------
// Static device scratch of smat_spmm_noalloc: the planner's counters and scan
// space, the work items, the split-row records and the partial tiles of
// split block rows.  Each GPU holds its own copy for the life of the
// process.  Host code never uses the array's name as an address (on the host
// it is only a stand-in): it uses the per-device address hipGetSymbolAddress
// returns (noalloc_base).
static __device__ __align__(256) char scratch_noalloc[SMAT_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_base[kMaxDevices] = {};

// Address of the current device's copy of scratch_noalloc (looked up once per device).
smat_status noalloc_base(void** out) {
  int dev = 0;
  SMAT_HIP_TRY(hipGetDevice(&dev));
  SMAT_REQUIRE(dev >= 0 && dev < kMaxDevices, "smat_spmm_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_base[dev]) {
    void* p = nullptr;
    SMAT_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch_noalloc)));
    SMAT_REQUIRE(p != nullptr, "smat_spmm_noalloc: no device address for the static scratch");
    g_noalloc_base[dev] = p;
  }
  *out = g_noalloc_base[dev];
  return SMAT_OK;
}

}  // namespace

namespace smat {
smat_status noalloc_scratch(void** base) { return noalloc_base(base); }
}  // namespace smat

extern "C" {

smat_status smat_spmm_noalloc(const smat_bcsr_desc* A, int64_t N, const smat_spmm_opts* opts,
                              const smat_spmm_operands* ops, smat_stream stream) {
  SMAT_REQUIRE(A && ops, "smat_spmm_noalloc: null argument");
  void* base = nullptr;
  SMAT_TRY(noalloc_base(&base));
  const hipStream_t s = smat::to_hip(stream);
  // every device array of the plan is carved from the scratch, from its start
  smat::Arena arena{static_cast<char*>(base), (size_t)SMAT_NOALLOC_SCRATCH_BYTES, 0};
  smat::ArenaScope scope(&arena);
  smat_plan_t P;
  SMAT_TRY(plan_init(&P, A, N, opts, s));
  Operands op{};
  SMAT_TRY(prepare_operands(&P, ops, &op));
  return launch_call(&P, ops, op, s);
}

}  // extern "C"

struct smat_batch_t {
  int32_t count = 0;
  int32_t dt = 0;
  int64_t total = 0, n_red = 0;
  int32_t g_beg[4] = {0, 0, 0, 0}, g_cnt[4] = {0, 0, 0, 0};  // matrices per N-tile width 16, 32, 64, 128
  int64_t g_tasks[4] = {0, 0, 0, 0};
  unsigned grid[4] = {0, 0, 0, 0};
  smat::DevBuf mats, red, counter;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // groups after the first run on side streams forked from the caller's
  // stream (parallel graph branches when captured), joined before the
  // split-row reduction
  hipStream_t side[3] = {nullptr, nullptr, nullptr};
  hipEvent_t fork = nullptr, join[3] = {nullptr, nullptr, nullptr};
  // SMAT_BATCH_BRANCHES: the entries, and per branch the entries it runs in
  // order (lane l: lane_order[lane_beg[l] .. lane_beg[l + 1]))
  int32_t mode = SMAT_BATCH_PERSISTENT;
  int32_t lanes = 0;
  std::vector<smat_plan> plans;
  std::vector<smat_spmm_operands> ops;
  std::vector<Operands> prepared;
  std::vector<int32_t> lane_order, lane_beg;
  ~smat_batch_t() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (int i = 0; i < 3; i++) {
      if (side[i]) (void)hipStreamDestroy(side[i]);
      if (join[i]) (void)hipEventDestroy(join[i]);
    }
    if (fork) (void)hipEventDestroy(fork);
  }
};

namespace {

template <int DT, int NS>
unsigned batch_grid() {
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_batch<DT, NS>, 64 * kWaves, 0) != hipSuccess) {
    (void)hipGetLastError();
    return 1024;
  }
  return (unsigned)(cus * (per < 1 ? 1 : per));
}

template <int DT>
void batch_group(const smat_batch_t* b, int g, hipStream_t s) {
  // about 16 draws per wave: few atomics, still fine-grained balance at the end
  const int64_t waves = (int64_t)b->grid[g] * kWaves;
  const int64_t chunk = std::max<int64_t>(1, b->g_tasks[g] / (waves * 16));
  const BatchMat* m = b->mats.as<BatchMat>() + b->g_beg[g];
  unsigned long long* c = b->counter.as<unsigned long long>() + 8 * g;
  if (g == 0) hipLaunchKernelGGL((k_batch<DT, 1>), b->grid[g], 64 * kWaves, 0, s, m, b->g_cnt[g], b->g_tasks[g], chunk, c);
  else if (g == 1) hipLaunchKernelGGL((k_batch<DT, 2>), b->grid[g], 64 * kWaves, 0, s, m, b->g_cnt[g], b->g_tasks[g], chunk, c);
  else if (g == 2) hipLaunchKernelGGL((k_batch<DT, 4>), b->grid[g], 64 * kWaves, 0, s, m, b->g_cnt[g], b->g_tasks[g], chunk, c);
  else hipLaunchKernelGGL((k_batch<DT, 8>), b->grid[g], 64 * kWaves, 0, s, m, b->g_cnt[g], b->g_tasks[g], chunk, c);
}

inline int nsub_group(int nsub) { return nsub == 1 ? 0 : nsub == 2 ? 1 : nsub == 4 ? 2 : 3; }

// Branches: lane 0 runs on s, lane l > 0 on side stream l - 1, forked from s
// before any work and joined back at the end.
smat_status batch_launch_branches(smat_batch_t* b, hipStream_t s) {
  if (b->lanes > 1) SMAT_HIP_TRY(hipEventRecord(b->fork, s));
  for (int l = 0; l < b->lanes; l++) {
    hipStream_t ls = l == 0 ? s : b->side[l - 1];
    if (l > 0) SMAT_HIP_TRY(hipStreamWaitEvent(ls, b->fork, 0));
    for (int32_t k = b->lane_beg[l]; k < b->lane_beg[l + 1]; k++) {
      const int32_t i = b->lane_order[k];
      SMAT_TRY(launch_call(b->plans[i], &b->ops[i], b->prepared[i], ls));
    }
    if (l > 0) SMAT_HIP_TRY(hipEventRecord(b->join[l - 1], ls));
  }
  for (int l = 1; l < b->lanes; l++) SMAT_HIP_TRY(hipStreamWaitEvent(s, b->join[l - 1], 0));
  return SMAT_OK;
}

smat_status batch_launch(smat_batch_t* b, hipStream_t s) {
  if (b->mode == SMAT_BATCH_BRANCHES) return batch_launch_branches(b, s);
  hipLaunchKernelGGL(k_batch_reset, 1, 64, 0, s, b->counter.as<unsigned long long>());
  // the first non-empty group runs on s, the others on side streams that
  // start right after the counter reset (the fork is recorded before any
  // group is launched); s waits for them before the reduction
  int groups = 0, used = 0, nside = 0;
  for (int g = 0; g < 4; g++) groups += b->g_cnt[g] > 0;
  if (groups > 1) SMAT_HIP_TRY(hipEventRecord(b->fork, s));
  for (int g = 0; g < 4; g++) {
    if (b->g_cnt[g] == 0) continue;
    hipStream_t gs = s;
    if (used++ > 0) {
      gs = b->side[nside];
      SMAT_HIP_TRY(hipStreamWaitEvent(gs, b->fork, 0));
    }
    if (b->dt == SMAT_BF16) batch_group<1>(b, g, gs);
    else batch_group<0>(b, g, gs);
    if (gs != s) SMAT_HIP_TRY(hipEventRecord(b->join[nside++], gs));
  }
  for (int i = 0; i < nside; i++) SMAT_HIP_TRY(hipStreamWaitEvent(s, b->join[i], 0));
  if (b->n_red > 0)
    hipLaunchKernelGGL(k_batch_reduce, (unsigned)b->n_red, 256, 0, s, b->red.as<BatchRed>(), b->mats.as<BatchMat>());
  SMAT_LAUNCH_CHECK();
  return SMAT_OK;
}

int bk_code(int bk) { return bk == 1 ? 0 : bk == 2 ? 1 : bk == 4 ? 2 : bk == 8 ? 3 : bk == 16 ? 4 : 5; }

smat_status batch_create_branches(int32_t count, const smat_plan* plans, const smat_spmm_operands* ops,
                                  smat_batch* out) {
  smat_batch_t* B = new smat_batch_t();
  B->mode = SMAT_BATCH_BRANCHES;
  B->count = count;
  B->plans.assign(plans, plans + count);
  B->ops.assign(ops, ops + count);
  B->prepared.resize((size_t)count);
  std::vector<double> work((size_t)count);
  for (int32_t i = 0; i < count; i++) {
    const smat_status st = prepare_operands(plans[i], &ops[i], &B->prepared[(size_t)i]);
    if (st != SMAT_OK) {
      delete B;
      return st;
    }
    const smat_plan_t* P = plans[i];
    B->total += P->n_items * (int64_t)ntiles_of(P, P->nsub);
    B->n_red += P->n_red;
    // stored-block MFMA work plus the C tile written, in elements
    work[(size_t)i] = (double)P->A.nnzb * P->A.bk * 16.0 * (double)P->N + (double)P->A.nbrows * 16.0 * (double)P->N;
  }
  // longest first onto the least loaded branch
  B->lanes = count < 4 ? count : 4;
  std::vector<int32_t> idx((size_t)count);
  for (int32_t i = 0; i < count; i++) idx[(size_t)i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) { return work[(size_t)a] > work[(size_t)b]; });
  std::vector<std::vector<int32_t>> lane((size_t)B->lanes);
  std::vector<double> load((size_t)B->lanes, 0.0);
  for (int32_t i : idx) {
    const size_t l = (size_t)(std::min_element(load.begin(), load.end()) - load.begin());
    lane[l].push_back(i);
    load[l] += work[(size_t)i];
  }
  B->lane_beg.push_back(0);
  for (auto& v : lane) {
    B->lane_order.insert(B->lane_order.end(), v.begin(), v.end());
    B->lane_beg.push_back((int32_t)B->lane_order.size());
  }
  hipError_t ce = hipEventCreateWithFlags(&B->fork, hipEventDisableTiming);
  for (int i = 0; i + 1 < B->lanes && ce == hipSuccess; i++) {
    ce = hipStreamCreateWithFlags(&B->side[i], hipStreamNonBlocking);
    if (ce == hipSuccess) ce = hipEventCreateWithFlags(&B->join[i], hipEventDisableTiming);
  }
  if (ce != hipSuccess) {
    delete B;
    return smat::set_error(SMAT_ERR_HIP, "smat_batch_create: stream setup failed: %s", hipGetErrorString(ce));
  }
  *out = B;
  return SMAT_OK;
}

}  // namespace

extern "C" {

smat_status smat_batch_create(int32_t count, const smat_plan* plans, const smat_spmm_operands* ops,
                              smat_stream stream, smat_batch* out) {
  return smat_batch_create_ex(count, plans, ops, SMAT_BATCH_AUTO, stream, out);
}

smat_status smat_batch_create_ex(int32_t count, const smat_plan* plans, const smat_spmm_operands* ops,
                                 int32_t mode, smat_stream stream, smat_batch* out) {
  SMAT_REQUIRE(out && count > 0 && plans && ops, "smat_batch_create: bad argument");
  SMAT_REQUIRE(mode == SMAT_BATCH_AUTO || mode == SMAT_BATCH_PERSISTENT || mode == SMAT_BATCH_BRANCHES,
               "smat_batch_create: unknown mode %d", mode);
  *out = nullptr;
  for (int32_t i = 0; i < count; i++) {
    SMAT_REQUIRE(plans[i], "smat_batch_create: entry %d: null plan", i);
    for (int32_t j = 0; j < i; j++)
      SMAT_REQUIRE(plans[j] != plans[i] || plans[i]->n_red == 0,
                   "smat_batch_create: entries %d and %d share a plan with split rows (one scratch)", j, i);
  }
  if (mode != SMAT_BATCH_PERSISTENT) return batch_create_branches(count, plans, ops, out);
  hipStream_t s = smat::to_hip(stream);
  std::vector<BatchMat> mats((size_t)count);
  std::vector<BatchRed> reds;
  const int dt = plans[0] ? plans[0]->opts.in_dtype : 0;
  // matrices ordered by N-tile width (16, 32, 64, 128), input order inside a group
  std::vector<int32_t> order;
  int32_t g_cnt[4] = {0, 0, 0, 0};
  for (int g = 0; g < 4; g++)
    for (int32_t i = 0; i < count; i++)
      if (plans[i] && nsub_group(plans[i]->nsub) == g) {
        order.push_back(i);
        g_cnt[g]++;
      }
  SMAT_REQUIRE((int32_t)order.size() == count, "smat_batch_create: null plan");
  int64_t g_tasks[4] = {0, 0, 0, 0};
  for (int32_t pos = 0; pos < count; pos++) {
    const int32_t i = order[(size_t)pos];
    const smat_plan_t* P = plans[i];
    const smat_spmm_operands& o = ops[i];
    const int g = nsub_group(P->nsub);
    SMAT_REQUIRE(P && o.b && o.c, "smat_batch_create: entry %d: null plan or operand", i);
    SMAT_REQUIRE(P->opts.in_dtype == dt && (dt == SMAT_F16 || dt == SMAT_BF16) &&
                     P->opts.b_layout == SMAT_B_ROW_MAJOR && P->opts.out_dtype == SMAT_F32 &&
                     P->opts.b_staging != SMAT_BSTAGE_DIRECT,
                 "smat_batch_create: entry %d: batches take plans with one 16-bit input type, row-major B "
                 "staged through LDS and FP32 C",
                 i);
    if (P->mixed)
      SMAT_REQUIRE(o.values24 && o.index24 && (reinterpret_cast<uintptr_t>(o.values24) & 15) == 0,
                   "smat_batch_create: entry %d: a 2:4 plan needs 16-byte aligned values24 and index24", i);
    SMAT_REQUIRE(o.ldb >= P->N && o.ldc >= P->N, "smat_batch_create: entry %d: ldb / ldc < N", i);
    const void* a = o.a_values ? o.a_values : P->A.values;
    BatchMat& m = mats[(size_t)pos];
    m.op = Operands{P->A.colidx, a, 0, nullptr, o.b, o.ldb, 0, nullptr, P->N, o.c, o.ldc, P->scratch.as<float>(),
                    o.c_rowmap, b_array_bytes(P, o.ldb)};
    SMAT_REQUIRE(lds_aligned(P, m.op, 8), "smat_batch_create: entry %d: N and ldb must be multiples of 8, B "
                                          "16-byte aligned and under 4 GiB", i);
    m.items = P->items.as<Item>();
    m.ntiles = ntiles_of(P, P->nsub);
    m.kind = P->mixed ? 6 : bk_code(P->A.bk);
    m.mx = P->mixed ? Mixed24{P->dense.as<int2>(), P->sparse.as<int2>(), o.values24,
                              static_cast<const uint16_t*>(o.index24)}
                    : Mixed24{nullptr, nullptr, nullptr, nullptr};
    m.task_beg = g_tasks[g];  // within its group
    g_tasks[g] += P->n_items * (int64_t)m.ntiles;
    if (P->n_red > 0) {
      std::vector<Reduce> r((size_t)P->n_red);
      SMAT_HIP_TRY(hipMemcpyAsync(r.data(), P->red.p, r.size() * sizeof(Reduce), hipMemcpyDeviceToHost, s));
      SMAT_HIP_TRY(hipStreamSynchronize(s));
      for (const Reduce& x : r) reds.push_back(BatchRed{pos, x.brow, x.slot_beg, x.slot_cnt});
    }
  }
  smat_batch_t* B = new smat_batch_t();
  B->count = count;
  B->dt = dt;
  B->total = g_tasks[0] + g_tasks[1] + g_tasks[2] + g_tasks[3];
  for (int g = 0, beg = 0; g < 4; g++) {
    B->g_beg[g] = beg;
    B->g_cnt[g] = g_cnt[g];
    B->g_tasks[g] = g_tasks[g];
    beg += g_cnt[g];
  }
  B->n_red = (int64_t)reds.size();
  smat_status st = B->mats.alloc(mats.size() * sizeof(BatchMat));
  if (st == SMAT_OK) st = B->red.alloc((reds.size() + 1) * sizeof(BatchRed));
  if (st == SMAT_OK) st = B->counter.alloc(32 * sizeof(unsigned long long));
  if (st != SMAT_OK) {
    delete B;
    return st;
  }
  hipError_t e = hipMemcpyAsync(B->mats.p, mats.data(), mats.size() * sizeof(BatchMat), hipMemcpyHostToDevice, s);
  if (e == hipSuccess && !reds.empty())
    e = hipMemcpyAsync(B->red.p, reds.data(), reds.size() * sizeof(BatchRed), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    delete B;
    return smat::set_error(SMAT_ERR_HIP, "smat_batch_create: copy failed: %s", hipGetErrorString(e));
  }
  B->grid[0] = dt == SMAT_BF16 ? batch_grid<1, 1>() : batch_grid<0, 1>();
  B->grid[1] = dt == SMAT_BF16 ? batch_grid<1, 2>() : batch_grid<0, 2>();
  B->grid[2] = dt == SMAT_BF16 ? batch_grid<1, 4>() : batch_grid<0, 4>();
  B->grid[3] = dt == SMAT_BF16 ? batch_grid<1, 8>() : batch_grid<0, 8>();
  int used = 0;
  for (int g = 0; g < 4; g++) used += B->g_cnt[g] > 0;
  hipError_t ce = hipEventCreateWithFlags(&B->fork, hipEventDisableTiming);
  for (int i = 0; i + 1 < used && ce == hipSuccess; i++) {
    ce = hipStreamCreateWithFlags(&B->side[i], hipStreamNonBlocking);
    if (ce == hipSuccess) ce = hipEventCreateWithFlags(&B->join[i], hipEventDisableTiming);
  }
  if (ce != hipSuccess) {
    delete B;
    return smat::set_error(SMAT_ERR_HIP, "smat_batch_create: stream setup failed: %s", hipGetErrorString(ce));
  }
  *out = B;
  return SMAT_OK;
}

smat_status smat_batch_capture(smat_batch B, smat_stream stream) {
  SMAT_REQUIRE(B, "smat_batch_capture: null batch");
  SMAT_REQUIRE(stream, "smat_batch_capture: capture needs a non-default stream");
  hipStream_t s = smat::to_hip(stream);
  if (B->exec) {
    (void)hipGraphExecDestroy(B->exec);
    B->exec = nullptr;
  }
  if (B->graph) {
    (void)hipGraphDestroy(B->graph);
    B->graph = nullptr;
  }
  SMAT_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const smat_status st = batch_launch(B, s);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(s, &g);
  if (st != SMAT_OK) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  if (e != hipSuccess) return smat::set_error(SMAT_ERR_HIP, "smat_batch_capture: %s", hipGetErrorString(e));
  B->graph = g;
  SMAT_HIP_TRY(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
  return SMAT_OK;
}

smat_status smat_batch_execute(smat_batch B, smat_stream stream) {
  SMAT_REQUIRE(B, "smat_batch_execute: null batch");
  hipStream_t s = smat::to_hip(stream);
  if (B->exec) {
    SMAT_HIP_TRY(hipGraphLaunch(B->exec, s));
    return SMAT_OK;
  }
  return batch_launch(B, s);
}

smat_status smat_batch_info(smat_batch B, int64_t* tasks, int64_t* split_rows, int32_t* grid, int32_t* captured) {
  SMAT_REQUIRE(B, "smat_batch_info: null batch");
  if (tasks) *tasks = B->total;
  if (split_rows) *split_rows = B->n_red;
  if (grid && B->mode == SMAT_BATCH_BRANCHES) {
    *grid = B->lanes;
  } else if (grid) {  // workgroups the largest group's launch keeps resident
    int32_t gmax = 0;
    for (int g = 0; g < 4; g++)
      if (B->g_cnt[g] > 0 && (int32_t)B->grid[g] > gmax) gmax = (int32_t)B->grid[g];
    *grid = gmax;
  }
  if (captured) *captured = B->exec != nullptr;
  return SMAT_OK;
}

smat_status smat_batch_destroy(smat_batch B) {
  delete B;
  return SMAT_OK;
}

}  // extern "C"
