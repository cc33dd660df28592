// BCSR x dense SpMM on gfx950 MFMA (FP16 in, FP32 accumulate).
//
// Reference semantics: mmaOBTKernelSparse (src/cuda_hgemm/src/mma/mmaOBT.cu:18-158)
// computes C = A*B over the blocks of one block row in BCSR order.  The
// design here is MI355X-first:
//
//  * One 64-lane wave owns a task = (work item, N-tile).  A work item is a
//    block row, or a contiguous slice of a long block row.  Items come from a
//    plan built on the GPU with a prefix sum over block-row lengths, so long
//    rows (R-MAT hubs) are split into slices of at most `S` blocks; their
//    partial 16 x N tiles go to a scratch area and a second kernel reduces
//    them in slice order (deterministic, no atomics, no pre-zeroing of C).
//
//  * v_mfma_f32_16x16x32_f16: the 32-deep reduction of one MFMA is filled by
//    two 16x16 blocks (bk = 16) or one 16x32 block (bk = 32).  The two blocks
//    need not be adjacent in the matrix: lanes 0-31 carry block p, lanes
//    32-63 block p+1, and each lane gathers the B rows of its own block.
//    Operand layouts (lane l, element e):
//        A: A[l%16][8*(l/16)+e]      B: B[8*(l/16)+e][l%16]
//        D: D[4*(l/16)+e][l%16]
//    A lane's 8 A values are 16 contiguous bytes of one block row, so one
//    128-bit load per lane fetches a whole block pair (1 KiB, coalesced).
//
//  * N is covered by NSUB 16-column MFMA tiles per wave (N-tile = 16*NSUB).
//    The wave keeps the A pair in registers across its NSUB tiles.
//
//  * Workgroups are 4 waves; the workgroup index is remapped so that the 8
//    XCDs (which receive workgroups round-robin) each get a contiguous range
//    of tasks, keeping neighbouring block rows in one L2.
#include <hipcub/hipcub.hpp>

#include "smat_internal.h"

namespace {

constexpr int kWaves = 4;  // waves per workgroup

struct Item {
  int32_t brow;
  int32_t pbeg;
  int32_t pend;
  int32_t slot;  // -1: write C; >= 0: partial tile index in scratch
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

__device__ inline _Float16 u16_as_h(uint16_t v) { return __builtin_bit_cast(_Float16, v); }

template <int BLAYOUT>
__device__ inline uint16_t load_b(const uint16_t* __restrict__ B, int64_t ldb, int64_t k, int64_t n) {
  if (BLAYOUT == SMAT_B_ROW_MAJOR) return B[k * ldb + n];
  return B[n * ldb + k];
}

// The 32 k-slots of one MFMA step starting at block p are filled by the
// 32/BK blocks p, p+1, ...: k-slot s is column s % BK of block p + s / BK.
// A lane of group g (= lane / 16) holds k-slots 8g .. 8g+7 of row lane % 16;
// for BK >= 8 that is one 16-byte piece of one block row, for smaller BK it
// is 8/BK pieces of BK halves from consecutive blocks.  Blocks at or beyond
// `pend` contribute zeros.
template <int BK>
__device__ inline uint4 load_a(const uint16_t* __restrict__ values, int32_t p, int32_t pend, int r16,
                               int grp) {
  static_assert(BK == 1 || BK == 2 || BK == 4 || BK == 8 || BK == 16 || BK == 32, "block width");
  if constexpr (BK >= 8) {
    const int32_t blk = p + (8 * grp) / BK;
    if (blk >= pend) return uint4{0, 0, 0, 0};
    return *reinterpret_cast<const uint4*>(values + (int64_t)blk * 16 * BK + r16 * BK + (8 * grp) % BK);
  } else if constexpr (BK == 4) {
    uint2 w[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int32_t blk = p + 2 * grp + q;
      w[q] = blk < pend ? *reinterpret_cast<const uint2*>(values + (int64_t)blk * 64 + r16 * 4) : uint2{0, 0};
    }
    return uint4{w[0].x, w[0].y, w[1].x, w[1].y};
  } else if constexpr (BK == 2) {
    uint32_t w[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int32_t blk = p + 4 * grp + q;
      w[q] = blk < pend ? *reinterpret_cast<const uint32_t*>(values + (int64_t)blk * 32 + r16 * 2) : 0u;
    }
    return uint4{w[0], w[1], w[2], w[3]};
  } else {
    uint32_t w[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int32_t b0 = p + 8 * grp + 2 * q, b1 = b0 + 1;
      const uint32_t lo = b0 < pend ? values[(int64_t)b0 * 16 + r16] : 0u;
      const uint32_t hi = b1 < pend ? values[(int64_t)b1 * 16 + r16] : 0u;
      w[q] = lo | (hi << 16);
    }
    return uint4{w[0], w[1], w[2], w[3]};
  }
}

// One wave: C[16 x 16*NSUB] tile of one work item; B read straight into the
// MFMA operands (per-lane 16-bit gathers for row-major B; 16-byte loads for
// the reference layout when BK >= 8).
template <int BK, int NSUB, int BLAYOUT, bool OUT_F16>
__global__ __launch_bounds__(256) void k_spmm_dense(
    const Item* __restrict__ items, int64_t n_items, int32_t ntiles,
    const int32_t* __restrict__ colidx, const uint16_t* __restrict__ values,
    const uint16_t* __restrict__ B, int64_t ldb, int64_t N, void* __restrict__ C, int64_t ldc,
    float* __restrict__ scratch) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  const uint32_t wg = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t task = (int64_t)wg * kWaves + wave;
  const int64_t item_id = task / ntiles;
  if (item_id >= n_items) return;
  const int nt = (int)(task % ntiles);
  const Item it = items[item_id];
  const int64_t n0 = (int64_t)nt * 16 * NSUB;

  smat_f4 acc[NSUB];
#pragma unroll
  for (int j = 0; j < NSUB; j++) acc[j] = smat_f4{0.f, 0.f, 0.f, 0.f};

  const int r16 = lane & 15;   // A row / B column within the tile
  const int grp = lane >> 4;   // 0..3: k-slot group of 8
  constexpr int STEP = 32 / BK;  // blocks per MFMA step

  for (int32_t p = it.pbeg; p < it.pend; p += STEP) {
    const smat_h8 a = __builtin_bit_cast(smat_h8, load_a<BK>(values, p, it.pend, r16, grp));
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
      smat_h8 b = smat_h8{0, 0, 0, 0, 0, 0, 0, 0};
      if (n < N) {
        if (BLAYOUT == SMAT_B_REFERENCE && BK >= 8 && (ldb % 8) == 0) {
          if (kok[0]) b = __builtin_bit_cast(smat_h8, *reinterpret_cast<const uint4*>(B + n * ldb + krow[0]));
        } else {
#pragma unroll
          for (int e = 0; e < 8; e++)
            if (kok[e]) b[e] = u16_as_h(load_b<BLAYOUT>(B, ldb, krow[e], n));
        }
      }
      acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, acc[j], 0, 0, 0);
    }
  }

  // D layout: lane holds rows 4*grp + i, column 16*j + r16.
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
        const int64_t crow = (int64_t)it.brow * 16 + row;
        if (OUT_F16)
          reinterpret_cast<_Float16*>(C)[crow * ldc + n] = (_Float16)acc[j][i];
        else
          reinterpret_cast<float*>(C)[crow * ldc + n] = acc[j][i];
      }
    }
  }
}

// Row-major B through LDS.  Per MFMA step the 32 B rows it reduces over
// (two 16x16 blocks or one 16x32 block) x the wave's 16*NSUB columns are
// fetched with 128-bit loads (each row segment is contiguous in HBM), stored
// to a per-wave LDS tile, and read back as MFMA B operands with
// ds_read_b64_tr_b16: a 16-lane group reading a 4 x 16 sub-tile (lane L
// addressing row L/4, columns 4*(L%4)..+3) receives column L of it, i.e.
// B[k0..k0+3][n0+L] — exactly 4 of the 8 k-values the operand layout wants.
// Two such reads give a lane its 8 values.
//
// Pipeline (per wave, two register stages X/Y, loop unrolled by two so the
// stage index is static): while step p is transposed and multiplied, the B
// rows and A values of steps p+STEP and p+2*STEP are in flight.  A step's
// block column starts are fetched one phase before its B rows, with one
// coalesced load (lane l holds colidx[p + l % STEP]) and handed to the B-row
// lanes by a lane shuffle; because each phase issues the column-index load
// before the B loads, the in-order vmcnt wait for an index never drains the
// other stage's B loads.  Waves never share LDS, so no barrier is needed: a
// wave's LDS accesses complete in issue order.
template <int BK, int NSUB, bool OUT_F16>
__global__ __launch_bounds__(256) void k_spmm_lds(
    const Item* __restrict__ items, int64_t n_items, int32_t ntiles,
    const int32_t* __restrict__ colidx, const uint16_t* __restrict__ values,
    const uint16_t* __restrict__ B, int64_t ldb, int64_t N, void* __restrict__ C, int64_t ldc,
    float* __restrict__ scratch) {
  constexpr int NT = 16 * NSUB;           // columns per wave
  constexpr int CPR = NT / 8;             // 16-byte chunks per B row segment
  constexpr int RPL = 64 / CPR;           // B rows per load instruction
  constexpr int NLD = 32 / RPL;           // load instructions per MFMA step
  constexpr int PITCH = NT + 8;           // LDS row pitch in halves (16-byte pad)
  constexpr int STEP = 32 / BK;           // blocks per MFMA step
  __shared__ __attribute__((aligned(16))) uint16_t lds_all[kWaves * 32 * PITCH];

  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  const uint32_t wg = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t task = (int64_t)wg * kWaves + wave;
  const int64_t item_id = task / ntiles;
  if (item_id >= n_items) return;
  const int nt = (int)(task % ntiles);
  const Item it = items[item_id];
  const int64_t n0 = (int64_t)nt * NT;
  uint16_t* lds = lds_all + wave * 32 * PITCH;

  const int r16 = lane & 15, grp = lane >> 4;
  // load coordinates: instruction i covers rows i*RPL .. i*RPL+RPL-1
  const int lrow = lane / CPR, lch = lane % CPR;
  const int64_t ncol = n0 + lch * 8;
  const bool col_ok = ncol < N;  // N % 8 == 0: a chunk is all in or all out

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
  auto issue = [&](int32_t p, int32_t cid, uint4 (&bb)[NLD], uint4& aa) {
    if (p >= pe) return;
#pragma unroll
    for (int i = 0; i < NLD; i++) {
      const int row = i * RPL + lrow;  // 0..31 = k-slot
      const int boff = row / BK;
      const int32_t c = __shfl(cid, boff);
      if (p + boff < pe && col_ok)
        bb[i] = *reinterpret_cast<const uint4*>(B + (int64_t)(c + row % BK) * ldb + ncol);
      else
        bb[i] = uint4{0, 0, 0, 0};
    }
    aa = load_a<BK>(values, p, pe, r16, grp);
  };
  auto consume = [&](const uint4 (&bb)[NLD], const uint4& aa) -> smat_h8 {
#pragma unroll
    for (int i = 0; i < NLD; i++)
      *reinterpret_cast<uint4*>(lds + (i * RPL + lrow) * PITCH + lch * 8) = bb[i];
    return __builtin_bit_cast(smat_h8, aa);
  };
  auto compute = [&](const smat_h8& a) {
#pragma unroll
    for (int j = 0; j < NSUB; j++) {
      const uint16_t* base = lds + (8 * grp + (r16 >> 2)) * PITCH + 16 * j + (r16 & 3) * 4;
      const smat_s4 lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16(
          (__attribute__((address_space(3))) smat_s4*)(base));
      const smat_s4 hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16(
          (__attribute__((address_space(3))) smat_s4*)(base + 4 * PITCH));
      // Reinterpret the two 64-bit results as one operand.  (Extracting the
      // 16-bit elements one by one was lowered to a splat of element 0 by
      // this compiler; the whole-vector cast maps the loaded registers
      // straight onto the MFMA operand.)
      const smat_s4 both[2] = {lo, hi};
      const smat_h8 b = __builtin_bit_cast(smat_h8, both);
      acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, acc[j], 0, 0, 0);
    }
  };

  uint4 bX[NLD], bY[NLD], aX, aY;
  const int32_t pb = it.pbeg;
  int32_t cX = cload(pb);
  int32_t cY = cload(pb + STEP);
  issue(pb, cX, bX, aX);
  cX = cload(pb + 2 * STEP);
  issue(pb + STEP, cY, bY, aY);
  cY = cload(pb + 3 * STEP);
  for (int32_t p = pb; p < pe; p += 2 * STEP) {
    smat_h8 a = consume(bX, aX);
    issue(p + 2 * STEP, cX, bX, aX);
    cX = cload(p + 4 * STEP);
    compute(a);
    if (p + STEP >= pe) break;
    a = consume(bY, aY);
    issue(p + 3 * STEP, cY, bY, aY);
    cY = cload(p + 5 * STEP);
    compute(a);
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
        const int64_t crow = (int64_t)it.brow * 16 + row;
        if (OUT_F16)
          reinterpret_cast<_Float16*>(C)[crow * ldc + n] = (_Float16)acc[j][i];
        else
          reinterpret_cast<float*>(C)[crow * ldc + n] = acc[j][i];
      }
    }
  }
}

struct Reduce {
  int32_t brow;
  int32_t slot_beg;
  int32_t slot_cnt;
  int32_t pad;
};

// Sums the partial tiles of each split block row in slice order.
template <bool OUT_F16>
__global__ void k_reduce(const Reduce* __restrict__ red, int64_t n_red, int32_t chunks,
                         const float* __restrict__ scratch, int64_t N, void* __restrict__ C, int64_t ldc) {
  const int64_t r = blockIdx.x / chunks;
  if (r >= n_red) return;
  const Reduce rd = red[r];
  const int64_t total = 16 * N;
  for (int64_t i = (int64_t)(blockIdx.x % chunks) * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)chunks * blockDim.x) {
    const int64_t row = i / N, n = i % N;
    float s = 0.f;
    for (int q = 0; q < rd.slot_cnt; q++) s += scratch[((int64_t)(rd.slot_beg + q) * 16 + row) * N + n];
    const int64_t crow = (int64_t)rd.brow * 16 + row;
    if (OUT_F16)
      reinterpret_cast<_Float16*>(C)[crow * ldc + n] = (_Float16)s;
    else
      reinterpret_cast<float*>(C)[crow * ldc + n] = s;
  }
}

// ---- planning -------------------------------------------------------------
// Per block row: number of items (>= 1: empty rows still write zeros) and of
// scratch slots (items if split, else 0).
__global__ void k_plan_count(int64_t nbrows, const int32_t* __restrict__ rowptr, int32_t S,
                             int32_t* __restrict__ n_items, int32_t* __restrict__ n_slots) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r > nbrows) return;
  if (r == nbrows) {
    n_items[r] = 0;
    n_slots[r] = 0;
    return;
  }
  int32_t nb = rowptr[r + 1] - rowptr[r];
  int32_t k = nb <= S ? 1 : (nb + S - 1) / S;
  n_items[r] = k;
  n_slots[r] = k > 1 ? k : 0;
}

__global__ void k_plan_emit(int64_t nbrows, const int32_t* __restrict__ rowptr, int32_t S,
                            const int32_t* __restrict__ item_off, const int32_t* __restrict__ slot_off,
                            const int32_t* __restrict__ red_flag_off, Item* __restrict__ items,
                            Reduce* __restrict__ red) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nbrows) return;
  int32_t pb = rowptr[r], pe = rowptr[r + 1];
  int32_t k = item_off[r + 1] - item_off[r];
  int32_t base = item_off[r];
  if (k == 1) {
    items[base] = Item{(int32_t)r, pb, pe, -1};
    return;
  }
  int32_t sbase = slot_off[r];
  for (int32_t q = 0; q < k; q++) {
    int32_t b0 = pb + q * S;
    int32_t b1 = b0 + S < pe ? b0 + S : pe;
    items[base + q] = Item{(int32_t)r, b0, b1, sbase + q};
  }
  red[red_flag_off[r]] = Reduce{(int32_t)r, sbase, k, 0};
}

__global__ void k_split_flag(int64_t nbrows, const int32_t* __restrict__ n_slots, int32_t* __restrict__ f) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r > nbrows) return;
  f[r] = (r < nbrows && n_slots[r] > 0) ? 1 : 0;
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
  int64_t n_items = 0, n_red = 0, n_slots = 0;
  smat::DevBuf items, red, scratch;
};

namespace {

template <int BK, int NSUB, int BLAYOUT, bool OUT_F16>
void launch_dense(const smat_plan_t* P, const uint16_t* B, int64_t ldb, void* C, int64_t ldc,
                  hipStream_t s) {
  const int32_t ntiles = (int32_t)((P->N + 16 * NSUB - 1) / (16 * NSUB));
  const int64_t tasks = P->n_items * ntiles;
  const unsigned grid = blocks_for(tasks, kWaves);
  hipLaunchKernelGGL((k_spmm_dense<BK, NSUB, BLAYOUT, OUT_F16>), grid, 64 * kWaves, 0, s,
                     P->items.as<Item>(), P->n_items, ntiles, P->A.colidx, P->A.values, B, ldb, P->N,
                     C, ldc, P->scratch.as<float>());
}

template <int BK, int NSUB, bool OUT_F16>
void launch_lds(const smat_plan_t* P, const uint16_t* B, int64_t ldb, void* C, int64_t ldc, hipStream_t s) {
  const int32_t ntiles = (int32_t)((P->N + 16 * NSUB - 1) / (16 * NSUB));
  const int64_t tasks = P->n_items * ntiles;
  const unsigned grid = blocks_for(tasks, kWaves);
  hipLaunchKernelGGL((k_spmm_lds<BK, NSUB, OUT_F16>), grid, 64 * kWaves, 0, s, P->items.as<Item>(),
                     P->n_items, ntiles, P->A.colidx, P->A.values, B, ldb, P->N, C, ldc,
                     P->scratch.as<float>());
}

template <int BK, int BLAYOUT, bool OUT_F16>
void dispatch_nsub(const smat_plan_t* P, const uint16_t* B, int64_t ldb, void* C, int64_t ldc,
                   hipStream_t s) {
  const bool lds_ok = (P->N % 8 == 0) && (ldb % 8 == 0) && ((reinterpret_cast<uintptr_t>(B) & 15) == 0);
  if (BLAYOUT == SMAT_B_ROW_MAJOR && P->opts.b_staging != SMAT_BSTAGE_DIRECT && lds_ok) {
    switch (P->nsub) {
      case 1: launch_lds<BK, 1, OUT_F16>(P, B, ldb, C, ldc, s); return;
      case 2: launch_lds<BK, 2, OUT_F16>(P, B, ldb, C, ldc, s); return;
      default: launch_lds<BK, 4, OUT_F16>(P, B, ldb, C, ldc, s); return;
    }
  }
  switch (P->nsub) {
    case 1: launch_dense<BK, 1, BLAYOUT, OUT_F16>(P, B, ldb, C, ldc, s); break;
    case 2: launch_dense<BK, 2, BLAYOUT, OUT_F16>(P, B, ldb, C, ldc, s); break;
    default: launch_dense<BK, 4, BLAYOUT, OUT_F16>(P, B, ldb, C, ldc, s); break;
  }
}

template <int BK>
void dispatch_layout(const smat_plan_t* P, const uint16_t* B, int64_t ldb, void* C, int64_t ldc,
                     hipStream_t s) {
  const bool f16 = P->opts.out_dtype == SMAT_F16;
  if (P->opts.b_layout == SMAT_B_REFERENCE) {
    if (f16) dispatch_nsub<BK, SMAT_B_REFERENCE, true>(P, B, ldb, C, ldc, s);
    else dispatch_nsub<BK, SMAT_B_REFERENCE, false>(P, B, ldb, C, ldc, s);
  } else {
    if (f16) dispatch_nsub<BK, SMAT_B_ROW_MAJOR, true>(P, B, ldb, C, ldc, s);
    else dispatch_nsub<BK, SMAT_B_ROW_MAJOR, false>(P, B, ldb, C, ldc, s);
  }
}

}  // namespace

extern "C" {

smat_status smat_spmm_plan_create(const smat_bcsr_desc* A, int64_t N, const smat_spmm_opts* opts,
                                  smat_stream stream, smat_plan* out) {
  SMAT_REQUIRE(A && out, "smat_spmm_plan_create: null argument");
  *out = nullptr;
  SMAT_REQUIRE(A->bk == 1 || A->bk == 2 || A->bk == 4 || A->bk == 8 || A->bk == 16 || A->bk == 32,
               "smat_spmm_plan_create: bk must be 1, 2, 4, 8, 16 or 32");
  SMAT_REQUIRE(N > 0, "smat_spmm_plan_create: N must be positive");
  SMAT_REQUIRE(A->nbrows > 0 && A->nbrows < (1 << 30), "smat_spmm_plan_create: bad block-row count");
  smat_spmm_opts o{};
  if (opts) o = *opts;
  if (o.out_dtype == 0) o.out_dtype = SMAT_F32;
  SMAT_REQUIRE(o.out_dtype == SMAT_F32 || o.out_dtype == SMAT_F16, "smat_spmm_plan_create: out_dtype");
  SMAT_REQUIRE(o.b_layout == SMAT_B_ROW_MAJOR || o.b_layout == SMAT_B_REFERENCE,
               "smat_spmm_plan_create: b_layout");
  SMAT_REQUIRE(o.kernel == SMAT_KERNEL_AUTO || o.kernel == SMAT_KERNEL_DENSE,
               "smat_spmm_plan_create: kernel %d not available in this build", o.kernel);
  SMAT_REQUIRE(o.n_tile == 0 || o.n_tile == 16 || o.n_tile == 32 || o.n_tile == 64,
               "smat_spmm_plan_create: n_tile must be 0, 16, 32 or 64");
  SMAT_REQUIRE(o.b_staging >= SMAT_BSTAGE_AUTO && o.b_staging <= SMAT_BSTAGE_LDS,
               "smat_spmm_plan_create: unknown b_staging %d", o.b_staging);
  hipStream_t s = smat::to_hip(stream);

  smat_plan_t* P = new smat_plan_t();
  P->A = *A;
  P->N = N;
  P->opts = o;
  int nt = o.n_tile ? o.n_tile : (N >= 64 ? 64 : (N > 16 ? 32 : 16));
  P->nsub = nt / 16;
  int32_t S = o.max_blocks_per_item;
  if (S <= 0) {
    // aim for >= 64k wave tasks, never below 64 blocks per slice
    int64_t ntl = (N + nt - 1) / nt;
    int64_t target_items = (65536 + ntl - 1) / ntl;
    int64_t s_auto = (A->nnzb + target_items - 1) / (target_items > 0 ? target_items : 1);
    S = (int32_t)(s_auto < 64 ? 64 : (s_auto > 4096 ? 4096 : s_auto));
  }
  {
    const int32_t step = 32 / A->bk;  // blocks per MFMA step: keep steps inside one slice
    S = (S + step - 1) / step * step;
  }
  auto fail = [&](smat_status st) {
    delete P;
    return st;
  };
  const int64_t nb1 = A->nbrows + 1;
  smat::DevBuf cnt_i, cnt_s, off_i, off_s, flag, off_f, tmp;
  smat_status st;
  if ((st = cnt_i.alloc(nb1 * 4)) || (st = cnt_s.alloc(nb1 * 4)) || (st = off_i.alloc(nb1 * 4)) ||
      (st = off_s.alloc(nb1 * 4)) || (st = flag.alloc(nb1 * 4)) || (st = off_f.alloc(nb1 * 4)))
    return fail(st);
  hipLaunchKernelGGL(k_plan_count, blocks_for(nb1, 256), 256, 0, s, A->nbrows, A->rowptr, S,
                     cnt_i.as<int32_t>(), cnt_s.as<int32_t>());
  hipLaunchKernelGGL(k_split_flag, blocks_for(nb1, 256), 256, 0, s, A->nbrows, cnt_s.as<int32_t>(),
                     flag.as<int32_t>());
  size_t tb = 0;
  if (hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt_i.as<int32_t>(), off_i.as<int32_t>(), (int)nb1, s) !=
      hipSuccess)
    return fail(smat::set_error(SMAT_ERR_HIP, "plan scan sizing failed"));
  if ((st = tmp.alloc(tb))) return fail(st);
  if (hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, cnt_i.as<int32_t>(), off_i.as<int32_t>(), (int)nb1, s) !=
          hipSuccess ||
      hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, cnt_s.as<int32_t>(), off_s.as<int32_t>(), (int)nb1, s) !=
          hipSuccess ||
      hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, flag.as<int32_t>(), off_f.as<int32_t>(), (int)nb1, s) !=
          hipSuccess)
    return fail(smat::set_error(SMAT_ERR_HIP, "plan scan failed"));
  int32_t tot[3] = {0, 0, 0};
  if (hipMemcpyAsync(&tot[0], off_i.as<int32_t>() + A->nbrows, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(&tot[1], off_s.as<int32_t>() + A->nbrows, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(&tot[2], off_f.as<int32_t>() + A->nbrows, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return fail(smat::set_error(SMAT_ERR_HIP, "plan readback failed"));
  P->n_items = tot[0];
  P->n_slots = tot[1];
  P->n_red = tot[2];
  if ((st = P->items.alloc(P->n_items * sizeof(Item))) || (st = P->red.alloc(P->n_red * sizeof(Reduce))) ||
      (st = P->scratch.alloc(P->n_slots * 16 * N * sizeof(float))))
    return fail(st);
  hipLaunchKernelGGL(k_plan_emit, blocks_for(A->nbrows, 256), 256, 0, s, A->nbrows, A->rowptr, S,
                     off_i.as<int32_t>(), off_s.as<int32_t>(), off_f.as<int32_t>(), P->items.as<Item>(),
                     P->red.as<Reduce>());
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return fail(smat::set_error(SMAT_ERR_HIP, "plan emit failed"));
  *out = P;
  return SMAT_OK;
}

smat_status smat_spmm_execute(smat_plan P, const uint16_t* values24, const void* index24, const uint16_t* B,
                              int64_t ldb, void* C, int64_t ldc, smat_stream stream) {
  (void)values24;
  (void)index24;
  SMAT_REQUIRE(P && B && C, "smat_spmm_execute: null argument");
  SMAT_REQUIRE(ldc >= P->N, "smat_spmm_execute: ldc < N");
  if (P->opts.b_layout == SMAT_B_ROW_MAJOR)
    SMAT_REQUIRE(ldb >= P->N, "smat_spmm_execute: ldb < N for row-major B");
  else
    SMAT_REQUIRE(ldb >= P->A.K, "smat_spmm_execute: ldb < K for reference-layout B");
  if (P->opts.b_staging == SMAT_BSTAGE_LDS && P->opts.b_layout == SMAT_B_ROW_MAJOR)
    SMAT_REQUIRE(P->N % 8 == 0 && ldb % 8 == 0 && (reinterpret_cast<uintptr_t>(B) & 15) == 0,
                 "smat_spmm_execute: LDS staging needs N %% 8 == 0, ldb %% 8 == 0 and a 16-byte aligned B");
  hipStream_t s = smat::to_hip(stream);
  switch (P->A.bk) {
    case 1: dispatch_layout<1>(P, B, ldb, C, ldc, s); break;
    case 2: dispatch_layout<2>(P, B, ldb, C, ldc, s); break;
    case 4: dispatch_layout<4>(P, B, ldb, C, ldc, s); break;
    case 8: dispatch_layout<8>(P, B, ldb, C, ldc, s); break;
    case 16: dispatch_layout<16>(P, B, ldb, C, ldc, s); break;
    default: dispatch_layout<32>(P, B, ldb, C, ldc, s); break;
  }
  if (P->n_red > 0) {
    const int32_t chunks = (int32_t)((16 * P->N + 255) / 256);
    const unsigned g = (unsigned)(chunks * P->n_red);
    if (P->opts.out_dtype == SMAT_F16)
      hipLaunchKernelGGL(k_reduce<true>, g, 256, 0, s, P->red.as<Reduce>(), P->n_red, chunks,
                         P->scratch.as<float>(), P->N, C, ldc);
    else
      hipLaunchKernelGGL(k_reduce<false>, g, 256, 0, s, P->red.as<Reduce>(), P->n_red, chunks,
                         P->scratch.as<float>(), P->N, C, ldc);
  }
  SMAT_LAUNCH_CHECK();
  return SMAT_OK;
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
