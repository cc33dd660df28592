// Merged tiles on the GPU: 4 x 4 sub-blocks of 16 x bk (64 x 64 tiles for
// bk = 16, 64 x 128 for bk = 32) built from a BCSR, their 2:4 compression and
// the SpMM that reads them.
//
// Reference: SparseMatrix::bcsrBlocking (src/cuda_hgemm/src/common/matrix.h:
// 794-1137, BLOCK = 4), read by mmaOBTKernelSparse_tiled (mma/mmaOBT_tiled.cu:
// 23-206) and, after preprocessing_mmaOBTSKernelSparse_tiled_large
// (mma/mmaOBTS_tiled_large.cu:21-150), by mmaOBTSKernelSparse_tiled_large
// (mmaOBTS_tiled_large.cu:152-414).  Kept from the reference: sub-block
// sb = 4*sr + sc of tile t is stored 16 x bk row-major at (16*t + sb)*16*bk
// (the offset 4bk*py + 16*px of matrix.h:848-852 with py = 16*sr, px = bk*sc),
// colidx holds the tile's first column, info has one entry per sub-block
// (0 empty, 1 2:4, 2 dense; -0 counts as zero, as the half comparison of
// matrix.h:1061 does).  Changed, each to make the producer agree with its
// consumers: tiles are sorted by (tile row, tile column), which is the order
// the reference's relative index map numbers them in (matrix.h:919-927),
// rather than by first appearance (matrix.h:934-949); rowptr has an entry for
// every tile row, as the kernels index it (mmaOBT_tiled.cu:67), where
// bcsrBlocking pushes one only for tile rows that hold tiles; info is computed
// from each sub-block's own values, where matrix.h:1042-1050 reads the tile
// as one 64 x 4bk row-major matrix.
//
// Construction: every block gets the key tile_row * ntcols + tile_col; a
// radix sort groups blocks by tile, a head-flag scan numbers the tiles, and
// the blocks' 16 x bk values are copied to their sub-block slots — O(nnzb log
// nnzb), where bcsrBlocking compares each block with every earlier tile.
//
// SpMM (k_spmm_tiles): one workgroup of 4 waves per (tile row, 64 columns of
// C); wave w owns sub-row w (rows 16w .. 16w+15 of the tile row).  Per tile,
// the 4bk x 64 panel of B rows under the tile is staged once in LDS by the
// whole workgroup and read by all four waves (the reuse that merging buys:
// each B row is fetched once per tile, not once per 16-row block row); B rows
// under sub-block columns that are empty in all four sub-rows are not
// fetched.  The panel is double-buffered: while a tile is multiplied, the
// next tile's B rows and A operands are in flight, and one barrier per tile
// separates the buffers.  Each wave multiplies its four sub-blocks: on MFMA
// 16x16x32 per 32-deep step (two 16-wide sub-blocks, or one 32-wide), or,
// when a 64-deep window of the sub-row holds no dense sub-block and needs
// both of its steps, on one SMFMAC 16x16x64; empty steps are skipped.
// Which path a wave takes depends on the tile's sub-block types only, so
// every branch is wave-uniform.  Per tile row, each wave reads the tile
// columns, tile indices and packed sub-block types of up to 64 tiles with one
// load per lane and takes them from there with readlane.
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "mfma_ops.h"
#include "smat_internal.h"

struct smat_tiles_t {
  smat_tiles_desc d{};
  smat::DevBuf rowptr, colidx, values, info;
};

namespace {

using namespace smat_mfma;

constexpr int kThreads = 256;

inline unsigned grid_for(int64_t n) {
  int64_t g = (n + kThreads - 1) / kThreads;
  if (g < 1) g = 1;
  return (unsigned)g;
}

__device__ inline int64_t find_row(const int32_t* __restrict__ rowptr, int64_t n, int64_t e) {
  // largest r in [0, n) with rowptr[r] <= e
  int64_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (rowptr[mid] <= e) lo = mid;
    else hi = mid;
  }
  return lo;
}

// ---- construction ---------------------------------------------------------

// Block b: key (block row / 4) * ntc + column / tw, and its sub-block slot.
__global__ void k_tile_keys(int64_t nnzb, int64_t nbrows, int64_t ntc, int bk, const int32_t* __restrict__ rowptr,
                            const int32_t* __restrict__ colidx, uint64_t* __restrict__ keys,
                            int32_t* __restrict__ idx, uint8_t* __restrict__ slot) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nnzb) return;
  const int64_t br = find_row(rowptr, nbrows, b);
  const int32_t col = colidx[b];
  const int tw = 4 * bk;
  keys[b] = (uint64_t)(br >> 2) * (uint64_t)ntc + (uint64_t)(col / tw);
  idx[b] = (int32_t)b;
  slot[b] = (uint8_t)(4 * (br & 3) + (col % tw) / bk);
}

__global__ void k_heads(int64_t n, const uint64_t* __restrict__ sk, int32_t* __restrict__ head) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  head[i] = (i == 0 || sk[i] != sk[i - 1]) ? 1 : 0;
}

// At each head: the tile's first column and tile row.
__global__ void k_tile_meta(int64_t n, int64_t ntc, int tw, const uint64_t* __restrict__ sk,
                            const int32_t* __restrict__ head, const int32_t* __restrict__ tid,
                            int32_t* __restrict__ t_colidx, int32_t* __restrict__ t_row) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !head[i]) return;
  const int64_t t = tid[i] - 1;
  const uint64_t k = sk[i];
  t_colidx[t] = (int32_t)((k % (uint64_t)ntc) * (uint64_t)tw);
  t_row[t] = (int32_t)(k / (uint64_t)ntc);
}

// rowptr[r] = number of tiles whose tile row is < r.
__global__ void k_tile_rowptr(int64_t ntiles, int64_t ntrows, const int32_t* __restrict__ t_row,
                              int32_t* __restrict__ rowptr) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t > ntiles) return;
  const int64_t prev = t == 0 ? -1 : t_row[t - 1];
  const int64_t cur = t == ntiles ? ntrows : t_row[t];
  for (int64_t r = prev + 1; r <= cur; r++) rowptr[r] = (int32_t)t;
}

__global__ void k_zero_u16(int64_t n, uint16_t* __restrict__ p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// One thread per (sorted block i, 16-byte chunk): block values -> sub-block slot.
__global__ void k_tile_scatter(int64_t nnzb, int bk, const uint16_t* __restrict__ val,
                               const int32_t* __restrict__ sidx, const int32_t* __restrict__ tid,
                               const uint8_t* __restrict__ slot, uint16_t* __restrict__ t_val) {
  const int ch = 2 * bk;  // 16-byte chunks per 16 x bk block
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = g / ch;
  const int c = (int)(g % ch);
  if (i >= nnzb) return;
  const int64_t b = sidx[i], t = tid[i] - 1;
  const int64_t bsz = 16 * (int64_t)bk;
  *reinterpret_cast<uint4*>(t_val + (t * 16 + slot[b]) * bsz + c * 8) =
      *reinterpret_cast<const uint4*>(val + b * bsz + c * 8);
}

// One thread per sub-block: 0 all zero, 1 every 1x4 group <= 2 nonzeros, else 2.
__global__ void k_tile_info(int64_t nsub, int bk, const uint16_t* __restrict__ t_val, int32_t* __restrict__ info) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsub) return;
  const uint4* v = reinterpret_cast<const uint4*>(t_val + s * 16 * bk);
  bool any = false, dense = false;
  for (int q = 0; q < 2 * bk; q++) {  // 8 halves = two 1x4 groups per load
    const uint4 w = v[q];
    const uint32_t u[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int g = 0; g < 2; g++) {
      int c = 0;
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const uint32_t x = u[2 * g + h];
        c += ((x & 0x7fffu) != 0) + ((x & 0x7fff0000u) != 0);
      }
      any |= c > 0;
      dense |= c > 2;
    }
  }
  info[s] = !any ? 0 : dense ? 2 : 1;
}

__global__ void k_fill_i32(int64_t n, int32_t* __restrict__ p, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_tile_grid(int64_t ntiles, int64_t ntrows, int64_t ntc, int tw, const int32_t* __restrict__ rowptr,
                            const int32_t* __restrict__ colidx, int32_t* __restrict__ info,
                            int32_t* __restrict__ rel) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  const int64_t g = find_row(rowptr, ntrows, t) * ntc + colidx[t] / tw;
  if (info) info[g] = 1;
  if (rel) rel[g] = (int32_t)t;
}

// ---- SpMM -----------------------------------------------------------------

constexpr int kNT = 64;                // C columns per workgroup
constexpr int kPitch = kNT * 2 + 16;   // LDS bytes per staged B row (16-byte pad)

template <int BK>
constexpr int tiles_lds_bytes() {
  return 2 * 4 * BK * kPitch;  // two panels of 4*BK rows
}

struct TileArgs {
  const int32_t* rowptr;
  const int32_t* colidx;
  const int32_t* relmap;  // REL: tile index = relmap[tile_row * ntc + colidx / tw]
  int64_t ntc;
  const uint16_t* values;
  const int32_t* info;  // NULL: every sub-block dense
  const uint16_t* values24;
  const uint16_t* index24;
  const uint16_t* b;
  int64_t ldb;
  void* c;
  int64_t ldc;
  int64_t M, N, K;
  int32_t nslices;  // ceil(N / 64)
};

// Sub-block types of tile t, 2 bits each at bit 2*sb (values other than 0
// and 1 count as dense).
__device__ inline uint32_t pack_info(const int32_t* __restrict__ info, int64_t t) {
  if (!info) return 0xAAAAAAAAu;
  const int4* p = reinterpret_cast<const int4*>(info + t * 16);
  uint32_t w = 0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int4 v = p[q];
    const int e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int i = 0; i < 4; i++) w |= (uint32_t)(e[i] == 0 ? 0 : e[i] == 1 ? 1 : 2) << (2 * (4 * q + i));
  }
  return w;
}

// A 64-deep window of one sub-row (sub-block types ws, 2 bits each) runs on
// SMFMAC when it holds no dense sub-block and both of its 32-deep MFMA steps
// would be needed; otherwise its non-empty steps run on MFMA (one SMFMAC
// reads as many B operands as two MFMA steps, so it only pays when it
// replaces two).
template <int BK>
__device__ inline bool window_on_smfmac(uint32_t ws) {
  if constexpr (BK == 32) return (ws & 0x3u) && (ws & 0xCu) && !(ws & 0xAu);
  else return (ws & 0x0Fu) && (ws & 0xF0u) && !(ws & 0xAAu);
}

// Bit sc set when sub-block column sc is non-empty in some sub-row.
__device__ inline uint32_t column_mask(uint32_t w) {
  const uint32_t nz = (w | (w >> 1)) & 0x55555555u;  // bit 2*sb: sub-block sb non-empty
  const uint32_t r = nz | (nz >> 8) | (nz >> 16) | (nz >> 24);  // bits 0, 2, 4, 6: columns 0..3
  return (r & 1) | ((r >> 1) & 2) | ((r >> 2) & 4) | ((r >> 3) & 8);
}

template <int BK, int DT, bool OUT_F16, bool MIXED, bool REL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_spmm_tiles(TileArgs a) {
  static_assert(BK == 16 || BK == 32, "sub-block width");
  constexpr int ROWS = 4 * BK;      // panel rows (tile width)
  constexpr int NL = ROWS / 32;     // B rows each thread stages per tile
  constexpr int BUF = ROWS * kPitch;
  __shared__ __attribute__((aligned(16))) uint8_t lds[2 * BUF];

  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int r16 = lane & 15, grp = lane >> 4;
  const uint32_t task = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t tr = task / (uint32_t)a.nslices;
  const int64_t n0 = (int64_t)(task % (uint32_t)a.nslices) * kNT;
  const int32_t pb = a.rowptr[tr], pe = a.rowptr[tr + 1];
  const int64_t K = a.K, N = a.N, ldb = a.ldb;
  const uint16_t* __restrict__ values = a.values;

  // B staging: thread -> (panel row lr + 32 i, 16-byte chunk lch)
  const int lr = tid >> 3, lch = tid & 7;
  const int64_t ncol = n0 + lch * 8;
  const bool col_ok = ncol < N;  // a chunk is read whole: ldb >= N rounded up to 8

  // per-lane slice of the tile row: column, tile index, packed types of tile cb + lane
  int32_t cbase = pb, l_col = 0, l_t = 0;
  uint32_t l_w = 0;
  auto load_chunk = [&](int32_t cb) {
    cbase = cb;
    const int32_t pos = cb + lane;
    l_col = 0;
    l_t = 0;
    l_w = 0;
    if (pos < pe) {
      l_col = a.colidx[pos];
      l_t = REL ? a.relmap[tr * a.ntc + l_col / ROWS] : pos;
      l_w = pack_info(a.info, l_t);
    }
  };
  auto fetch = [&](int32_t p, int32_t& col, int64_t& t, uint32_t& w) {
    if (p >= cbase + 64) load_chunk(p);
    const int q = p - cbase;
    col = __builtin_amdgcn_readlane(l_col, q);
    t = __builtin_amdgcn_readlane(l_t, q);
    w = (uint32_t)__builtin_amdgcn_readlane((int)l_w, q);
  };

  uint4 bq[NL];
  auto issue_b = [&](int32_t col, uint32_t w) {
    const uint32_t cm = column_mask(w);
#pragma unroll
    for (int i = 0; i < NL; i++) {
      const int row = i * 32 + lr;
      const int64_t k = (int64_t)col + row;
      const bool ok = ((cm >> (row / BK)) & 1) && k < K && col_ok;
      bq[i] = ok ? *reinterpret_cast<const uint4*>(a.b + k * ldb + ncol) : uint4{0, 0, 0, 0};
    }
  };
  auto stage_b = [&](uint8_t* buf) {
#pragma unroll
    for (int i = 0; i < NL; i++) *reinterpret_cast<uint4*>(buf + (i * 32 + lr) * kPitch + lch * 16) = bq[i];
  };

  // A operands of this wave's sub-row of tile t.  Dense steps: a[s] holds
  // k-slots 8*grp .. of 32-deep step s.  2:4 windows (MIXED): a[2*v] and
  // ix[v] hold lane group grp's compressed values and index word of window v.
  const uint4 zero4 = uint4{0, 0, 0, 0};
  auto sub_ptr = [&](int64_t t, int sc) { return values + ((t * 16 + 4 * wave + sc) * 16 + r16) * BK; };
  auto issue_a = [&](int64_t t, uint32_t w, uint4 (&av)[4], int (&ix)[2]) {
    const uint32_t sw = (w >> (8 * wave)) & 0xffu;
    auto dense_steps = [&](int s0, int s1) {
#pragma unroll
      for (int s = s0; s < s1; s++) {
        if constexpr (BK == 32) {
          av[s] = ((sw >> (2 * s)) & 3) ? *reinterpret_cast<const uint4*>(sub_ptr(t, s) + 8 * grp) : zero4;
        } else {
          const int sc = 2 * s + (grp >> 1);
          av[s] = ((sw >> (2 * sc)) & 3) ? *reinterpret_cast<const uint4*>(sub_ptr(t, sc) + 8 * (grp & 1)) : zero4;
        }
      }
    };
    if constexpr (!MIXED) {
      dense_steps(0, BK == 32 ? 4 : 2);
    } else {
      constexpr int NWIN = BK == 32 ? 2 : 1;  // 64-deep windows per tile
      constexpr int SPW = 64 / BK;            // sub-blocks per window
#pragma unroll
      for (int v = 0; v < NWIN; v++) {
        if (!window_on_smfmac<BK>((sw >> (2 * SPW * v)) & ((1u << (2 * SPW)) - 1))) {
          dense_steps(2 * v, 2 * v + 2);
        } else {
          // lane group grp meets window rows 16*grp .. 16*grp+15
          const int sc = BK == 32 ? 2 * v + (grp >> 1) : grp;
          const int hf = BK == 32 ? (grp & 1) : 0;  // 16-column half of a 32-wide sub-block
          const int64_t sb = (t * 16 + 4 * wave + sc) * 16 + r16;
          if ((sw >> (2 * sc)) & 3) {
            av[2 * v] = *reinterpret_cast<const uint4*>(a.values24 + sb * (BK / 2) + 8 * hf);
            ix[v] = a.index24[sb * (BK / 16) + hf];
          } else {
            av[2 * v] = zero4;
            ix[v] = 0x4444;  // positions (0, 1) of every group; the values are zero
          }
        }
      }
    }
  };

  smat_f4 acc[4];
#pragma unroll
  for (int j = 0; j < 4; j++) acc[j] = smat_f4{0.f, 0.f, 0.f, 0.f};

  auto dense_step = [&](const uint8_t* buf, int s, uint4 av) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const uint8_t* base = buf + (32 * s + 8 * grp + (r16 >> 2)) * kPitch + (16 * j + (r16 & 3) * 4) * 2;
      acc[j] = mma<DT>(av, tr_pair(base, base + 4 * kPitch), acc[j]);
    }
  };
  auto sparse_window = [&](const uint8_t* buf, int v, uint4 av, int ix) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const uint8_t* b0 = buf + (64 * v + 8 * grp + (r16 >> 2)) * kPitch + (16 * j + (r16 & 3) * 4) * 2;
      const uint8_t* b1 = b0 + 32 * kPitch;
      U8 bb;
      bb.lo = tr_pair(b0, b0 + 4 * kPitch);
      bb.hi = tr_pair(b1, b1 + 4 * kPitch);
      acc[j] = smma<DT>(av, bb, ix, acc[j]);
    }
  };
  auto compute = [&](const uint8_t* buf, uint32_t w, const uint4 (&av)[4], const int (&ix)[2]) {
    const uint32_t sw = (w >> (8 * wave)) & 0xffu;
    auto steps = [&](int s0, int s1) {
#pragma unroll
      for (int s = s0; s < s1; s++) {
        const uint32_t need = BK == 32 ? (sw >> (2 * s)) & 3 : (sw >> (4 * s)) & 15;
        if (need) dense_step(buf, s, av[s]);
      }
    };
    if constexpr (!MIXED) {
      steps(0, BK == 32 ? 4 : 2);
    } else {
      constexpr int NWIN = BK == 32 ? 2 : 1;
      constexpr int SPW = 64 / BK;
#pragma unroll
      for (int v = 0; v < NWIN; v++) {
        const uint32_t ws = (sw >> (2 * SPW * v)) & ((1u << (2 * SPW)) - 1);
        if (!window_on_smfmac<BK>(ws)) steps(2 * v, 2 * v + 2);
        else sparse_window(buf, v, av[2 * v], ix[v]);
      }
    }
  };

  uint4 aC[4] = {zero4, zero4, zero4, zero4}, aN[4] = {zero4, zero4, zero4, zero4};
  int ixC[2] = {0x4444, 0x4444}, ixN[2] = {0x4444, 0x4444};
  uint32_t wC = 0, wN = 0;
  if (pb < pe) {
    load_chunk(pb);
    int32_t col;
    int64_t t;
    fetch(pb, col, t, wC);
    issue_b(col, wC);
    issue_a(t, wC, aC, ixC);
    stage_b(lds);
  }
  __syncthreads();
  for (int32_t p = pb; p < pe; p++) {
    uint8_t* cur = lds + ((p - pb) & 1) * BUF;
    uint8_t* nxt = lds + (((p - pb) & 1) ^ 1) * BUF;
    const bool more = p + 1 < pe;
    if (more) {
      int32_t col;
      int64_t t;
      fetch(p + 1, col, t, wN);
      issue_b(col, wN);
      issue_a(t, wN, aN, ixN);
    }
    compute(cur, wC, aC, ixC);
    if (more) stage_b(nxt);
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 4; s++) aC[s] = aN[s];
    ixC[0] = ixN[0];
    ixC[1] = ixN[1];
    wC = wN;
  }

  // D layout: lane holds rows 4*grp + i of the sub-row, column 16*j + r16
  const int64_t row0 = tr * 64 + 16 * wave + 4 * grp;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int64_t n = n0 + 16 * j + r16;
    if (n >= N) continue;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int64_t row = row0 + i;
      if (row >= a.M) continue;
      if constexpr (OUT_F16)
        static_cast<_Float16*>(a.c)[row * a.ldc + n] = (_Float16)acc[j][i];
      else
        static_cast<float*>(a.c)[row * a.ldc + n] = acc[j][i];
    }
  }
}

template <int BK, int DT, bool OUT_F16, bool MIXED, bool REL>
void launch_tiles(const TileArgs& a, int64_t ntrows, hipStream_t s) {
  const int64_t grid = ntrows * (int64_t)a.nslices;
  hipLaunchKernelGGL((k_spmm_tiles<BK, DT, OUT_F16, MIXED, REL>), (unsigned)grid, 256, 0, s, a);
}

template <int BK, bool REL>
smat_status dispatch_tiles(const TileArgs& a, int64_t ntrows, int dt, bool out_f16, bool mixed, hipStream_t s) {
  if (ntrows <= 0 || a.nslices <= 0) return SMAT_OK;
#define SMAT_TILES_CASE(D, O, X)                                            \
  if (dt == (D) && out_f16 == (O) && mixed == (X)) {                        \
    launch_tiles<BK, D, O, X, REL>(a, ntrows, s);                           \
    SMAT_LAUNCH_CHECK();                                                    \
    return SMAT_OK;                                                         \
  }
  SMAT_TILES_CASE(0, true, false)
  SMAT_TILES_CASE(0, true, true)
  SMAT_TILES_CASE(0, false, false)
  SMAT_TILES_CASE(0, false, true)
  if constexpr (!REL) {
    SMAT_TILES_CASE(1, true, false)
    SMAT_TILES_CASE(1, true, true)
    SMAT_TILES_CASE(1, false, false)
    SMAT_TILES_CASE(1, false, true)
  }
#undef SMAT_TILES_CASE
  return smat::set_error(SMAT_ERR_UNSUPPORTED, "tile SpMM: no kernel for this element type");
}

}  // namespace

namespace smat {

// Tile SpMM on the reference drop-in's arrays: FP16 in and out, tiles found
// through the relative index map of the tile grid.  Array consistency is the
// caller's (see ref_launchers.hip).
smat_status tiles_spmm_rel(int bk, bool mixed, int64_t M, int64_t K, int64_t N, const int32_t* rowptr,
                           const int32_t* colidx, const int32_t* relmap, const uint16_t* values,
                           const int32_t* info, const uint16_t* values24, const uint16_t* index24,
                           const uint16_t* B, int64_t ldb, uint16_t* C, int64_t ldc, hipStream_t s) {
  TileArgs a{};
  a.rowptr = rowptr;
  a.colidx = colidx;
  a.relmap = relmap;
  a.ntc = (K + 4 * bk - 1) / (4 * bk);
  a.values = values;
  a.info = info;
  a.values24 = values24;
  a.index24 = index24;
  a.b = B;
  a.ldb = ldb;
  a.c = C;
  a.ldc = ldc;
  a.M = M;
  a.N = N;
  a.K = K;
  a.nslices = (int32_t)((N + kNT - 1) / kNT);
  const int64_t ntrows = (M + 63) / 64;
  if (bk == 16) return dispatch_tiles<16, true>(a, ntrows, 0, true, mixed, s);
  return dispatch_tiles<32, true>(a, ntrows, 0, true, mixed, s);
}

}  // namespace smat

extern "C" {

smat_status smat_tiles_from_bcsr(const smat_bcsr_desc* A, smat_stream stream, smat_tiles* out) {
  SMAT_REQUIRE(A && out, "smat_tiles_from_bcsr: null argument");
  *out = nullptr;
  SMAT_REQUIRE(A->bm == 16 && (A->bk == 16 || A->bk == 32),
               "smat_tiles_from_bcsr: tiles are built from 16 x 16 or 16 x 32 blocks (got 16 x %d)", A->bk);
  SMAT_REQUIRE(A->M > 0 && A->M % 16 == 0 && A->K > 0 && A->K % A->bk == 0,
               "smat_tiles_from_bcsr: bad shape %lld x %lld", (long long)A->M, (long long)A->K);
  SMAT_REQUIRE(A->nnzb >= 0 && A->nnzb < (int64_t)INT32_MAX && A->rowptr && (A->nnzb == 0 || (A->colidx && A->values)),
               "smat_tiles_from_bcsr: bad BCSR arrays");
  hipStream_t s = smat::to_hip(stream);
  const int bk = A->bk, tw = 4 * bk;
  const int64_t nnzb = A->nnzb, nbrows = A->M / 16;
  const int64_t ntrows = (A->M + 63) / 64, ntc = (A->K + tw - 1) / tw;

  smat_tiles_t* T = new smat_tiles_t();
  auto fail = [&](smat_status st) {
    delete T;
    return st;
  };
  smat::DevBuf keys, keys_sorted, idx, idx_sorted, slot, head, tid, trow, tmp;
  int64_t ntiles = 0;
  smat_status st;
  if (nnzb > 0) {
    if ((st = keys.alloc(nnzb * 8)) || (st = keys_sorted.alloc(nnzb * 8)) || (st = idx.alloc(nnzb * 4)) ||
        (st = idx_sorted.alloc(nnzb * 4)) || (st = slot.alloc(nnzb)) || (st = head.alloc(nnzb * 4)) ||
        (st = tid.alloc(nnzb * 4)))
      return fail(st);
    hipLaunchKernelGGL(k_tile_keys, grid_for(nnzb), kThreads, 0, s, nnzb, nbrows, ntc, bk, A->rowptr, A->colidx,
                       keys.as<uint64_t>(), idx.as<int32_t>(), slot.as<uint8_t>());
    if (hipGetLastError() != hipSuccess) return fail(smat::set_error(SMAT_ERR_HIP, "k_tile_keys launch failed"));
    const uint64_t max_key = (uint64_t)ntrows * (uint64_t)ntc;
    int end_bit = 1;
    while (end_bit < 64 && (max_key >> end_bit) != 0) end_bit++;
    size_t tbytes = 0, sbytes = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, tbytes, keys.as<uint64_t>(), keys_sorted.as<uint64_t>(),
                                           idx.as<int32_t>(), idx_sorted.as<int32_t>(), (int)nnzb, 0, end_bit,
                                           s) != hipSuccess ||
        hipcub::DeviceScan::InclusiveSum(nullptr, sbytes, head.as<int32_t>(), tid.as<int32_t>(), (int)nnzb, s) !=
            hipSuccess)
      return fail(smat::set_error(SMAT_ERR_HIP, "tile sort sizing failed"));
    if ((st = tmp.alloc(std::max(tbytes, sbytes)))) return fail(st);
    if (hipcub::DeviceRadixSort::SortPairs(tmp.p, tbytes, keys.as<uint64_t>(), keys_sorted.as<uint64_t>(),
                                           idx.as<int32_t>(), idx_sorted.as<int32_t>(), (int)nnzb, 0, end_bit,
                                           s) != hipSuccess)
      return fail(smat::set_error(SMAT_ERR_HIP, "tile sort failed"));
    hipLaunchKernelGGL(k_heads, grid_for(nnzb), kThreads, 0, s, nnzb, keys_sorted.as<uint64_t>(),
                       head.as<int32_t>());
    sbytes = tmp.bytes;
    if (hipcub::DeviceScan::InclusiveSum(tmp.p, sbytes, head.as<int32_t>(), tid.as<int32_t>(), (int)nnzb, s) !=
        hipSuccess)
      return fail(smat::set_error(SMAT_ERR_HIP, "tile scan failed"));
    int32_t last = 0;
    if (hipMemcpyAsync(&last, tid.as<int32_t>() + (nnzb - 1), 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return fail(smat::set_error(SMAT_ERR_HIP, "tile count readback failed"));
    ntiles = last;
  }
  const int64_t tsz = 16 * 16 * (int64_t)bk;  // halves per tile
  if ((st = T->rowptr.alloc((ntrows + 1) * 4)) || (st = T->colidx.alloc(ntiles * 4)) ||
      (st = T->values.alloc(ntiles * tsz * 2)) || (st = T->info.alloc(ntiles * 16 * 4)) ||
      (st = trow.alloc(ntiles * 4 + 4)))
    return fail(st);
  if (ntiles > 0) {
    hipLaunchKernelGGL(k_tile_meta, grid_for(nnzb), kThreads, 0, s, nnzb, ntc, tw, keys_sorted.as<uint64_t>(),
                       head.as<int32_t>(), tid.as<int32_t>(), T->colidx.as<int32_t>(), trow.as<int32_t>());
    hipLaunchKernelGGL(k_zero_u16, grid_for(ntiles * tsz), kThreads, 0, s, ntiles * tsz, T->values.as<uint16_t>());
    hipLaunchKernelGGL(k_tile_scatter, grid_for(nnzb * 2 * bk), kThreads, 0, s, nnzb, bk, A->values,
                       idx_sorted.as<int32_t>(), tid.as<int32_t>(), slot.as<uint8_t>(), T->values.as<uint16_t>());
    hipLaunchKernelGGL(k_tile_info, grid_for(ntiles * 16), kThreads, 0, s, ntiles * 16, bk, T->values.as<uint16_t>(),
                       T->info.as<int32_t>());
  }
  hipLaunchKernelGGL(k_tile_rowptr, grid_for(ntiles + 1), kThreads, 0, s, ntiles, ntrows, trow.as<int32_t>(),
                     T->rowptr.as<int32_t>());
  if (hipGetLastError() != hipSuccess) return fail(smat::set_error(SMAT_ERR_HIP, "tile build launch failed"));
  // temporaries are freed on return
  if (hipStreamSynchronize(s) != hipSuccess) return fail(smat::set_error(SMAT_ERR_HIP, "tile build failed"));

  T->d.M = A->M;
  T->d.K = A->K;
  T->d.bk = bk;
  T->d.tw = tw;
  T->d.ntrows = ntrows;
  T->d.ntcols = ntc;
  T->d.ntiles = ntiles;
  T->d.rowptr = T->rowptr.as<int32_t>();
  T->d.colidx = T->colidx.as<int32_t>();
  T->d.values = T->values.as<uint16_t>();
  T->d.info = T->info.as<int32_t>();
  *out = T;
  return SMAT_OK;
}

smat_status smat_tiles_describe(smat_tiles T, smat_tiles_desc* out) {
  SMAT_REQUIRE(T && out, "smat_tiles_describe: null argument");
  *out = T->d;
  return SMAT_OK;
}

smat_status smat_tiles_destroy(smat_tiles T) {
  delete T;
  return SMAT_OK;
}

smat_status smat_tiles_grid_maps(const smat_tiles_desc* T, int32_t* tile_info_grid, int32_t* relative_tile_index,
                                 smat_stream stream) {
  SMAT_REQUIRE(T && (tile_info_grid || relative_tile_index), "smat_tiles_grid_maps: null argument");
  hipStream_t s = smat::to_hip(stream);
  const int64_t ng = T->ntrows * T->ntcols;
  if (tile_info_grid) hipLaunchKernelGGL(k_fill_i32, grid_for(ng), kThreads, 0, s, ng, tile_info_grid, 0);
  if (relative_tile_index) hipLaunchKernelGGL(k_fill_i32, grid_for(ng), kThreads, 0, s, ng, relative_tile_index, -1);
  if (T->ntiles > 0)
    hipLaunchKernelGGL(k_tile_grid, grid_for(T->ntiles), kThreads, 0, s, T->ntiles, T->ntrows, T->ntcols, T->tw,
                       T->rowptr, T->colidx, tile_info_grid, relative_tile_index);
  SMAT_LAUNCH_CHECK();
  return SMAT_OK;
}

smat_status smat_tiles_compress_24(const smat_tiles_desc* T, uint16_t* values24, uint16_t* index24,
                                   smat_stream stream) {
  SMAT_REQUIRE(T && values24 && index24, "smat_tiles_compress_24: null argument");
  // the sub-blocks are 16 x bk blocks in a row, tagged by info
  smat_bcsr_desc d{};
  d.M = 16;
  d.K = T->bk;
  d.bm = 16;
  d.bk = T->bk;
  d.nnzb = T->ntiles * 16;
  d.values = T->values;
  d.tags = T->info;
  return smat_compress_24(&d, SMAT_24_SMFMAC, values24, index24, stream);
}

smat_status smat_tiles_spmm(const smat_tiles_desc* T, const smat_tiles_operands* ops, smat_stream stream) {
  SMAT_REQUIRE(T && ops, "smat_tiles_spmm: null argument");
  SMAT_REQUIRE(T->bk == 16 || T->bk == 32, "smat_tiles_spmm: bk must be 16 or 32");
  SMAT_REQUIRE(ops->dtype == SMAT_F16 || ops->dtype == SMAT_BF16, "smat_tiles_spmm: dtype must be F16 or BF16");
  SMAT_REQUIRE(ops->out_dtype == SMAT_F16 || ops->out_dtype == SMAT_F32,
               "smat_tiles_spmm: out_dtype must be F16 or F32");
  SMAT_REQUIRE(ops->N > 0 && ops->b && ops->c, "smat_tiles_spmm: empty B or C");
  SMAT_REQUIRE(ops->ldb >= ops->N && ops->ldb % 8 == 0 && (reinterpret_cast<uintptr_t>(ops->b) & 15) == 0,
               "smat_tiles_spmm: B needs ldb >= N, ldb a multiple of 8 and 16-byte alignment");
  SMAT_REQUIRE(ops->ldc >= ops->N, "smat_tiles_spmm: ldc < N");
  SMAT_REQUIRE((ops->values24 == nullptr) == (ops->index24 == nullptr),
               "smat_tiles_spmm: values24 and index24 go together");
  TileArgs a{};
  a.rowptr = T->rowptr;
  a.colidx = T->colidx;
  a.relmap = nullptr;
  a.ntc = T->ntcols;
  a.values = ops->values ? static_cast<const uint16_t*>(ops->values) : T->values;
  a.info = T->info;
  a.values24 = ops->values24;
  a.index24 = ops->index24;
  a.b = static_cast<const uint16_t*>(ops->b);
  a.ldb = ops->ldb;
  a.c = ops->c;
  a.ldc = ops->ldc;
  a.M = T->M;
  a.N = ops->N;
  a.K = T->K;
  const int64_t ns = (ops->N + kNT - 1) / kNT;
  SMAT_REQUIRE(ns < (int64_t)INT32_MAX && T->ntrows * ns < (int64_t)UINT32_MAX, "smat_tiles_spmm: grid too large");
  a.nslices = (int32_t)ns;
  const int dt = ops->dtype == SMAT_BF16 ? 1 : 0;
  const bool out_f16 = ops->out_dtype == SMAT_F16, mixed = ops->values24 != nullptr;
  hipStream_t s = smat::to_hip(stream);
  if (T->bk == 16) return dispatch_tiles<16, false>(a, T->ntrows, dt, out_f16, mixed, s);
  return dispatch_tiles<32, false>(a, T->ntrows, dt, out_f16, mixed, s);
}

}  // extern "C"
