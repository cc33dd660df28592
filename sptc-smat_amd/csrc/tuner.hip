// Learned tuner: matrix statistics and features (GPU), and the MLP cost model
// (forward, backward, Adam and a fused inference kernel, FP32).
//
// The reference picks its configuration by hand or by an exhaustive search
// over the Jaccard threshold (reorder_matrix.cpp:17-40, collect_taus.py:37-44);
// the learned model has no reference counterpart.  The exhaustive sweep stays
// the fallback and the yardstick of the model (host side, smat/tuner.py).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "smat_internal.h"

namespace {

inline unsigned grid_of(int64_t n, int t) {
  const int64_t g = (n + t - 1) / t;
  return (unsigned)(g < 1 ? 1 : (g > (1 << 30) ? (1 << 30) : g));
}

// ---- statistics ---------------------------------------------------------------
// Row-length histogram bucket: 0 -> 0, 1 -> 1, [2^(b-1), 2^b) -> b, capped at 11.
__device__ inline int len_bucket(int64_t len) {
  if (len <= 0) return 0;
  int b = 64 - __clzll((unsigned long long)len);  // floor(log2 len) + 1
  return b > 11 ? 11 : b;
}

// Rows longer than this get a wave each for their per-entry statistics.
constexpr int32_t kLongRow = 256;

// Per row: histogram, longest row; per entry (rows up to kLongRow entries):
// |row - col| sum, entries near the diagonal, and the sort key
// (block row << colbits | column).  Rows with
// ascending columns also report their violations of the exact 2:4 rule (an
// aligned 1x4 group with more than two nonzero values, +-0 counting as zero):
// one key (block row << (colbits - 4) | column / 16) per violating group,
// written to vkeys at positions reserved with one atomic per workgroup; a row
// whose columns do not ascend sets *unsorted instead.  Block-local bins, then
// one atomic per bin and workgroup.
template <typename KT>
__global__ __launch_bounds__(256) void k_row_stats(int64_t M, const int32_t* __restrict__ rp,
                                                   const int32_t* __restrict__ ci, const uint16_t* __restrict__ val,
                                                   int colbits, unsigned long long* __restrict__ acc,
                                                   KT* __restrict__ keys, KT* __restrict__ vkeys,
                                                   unsigned long long* __restrict__ vcount,
                                                   int32_t* __restrict__ unsorted) {
  // acc: [0..11] histogram, [12] max row, [13] sum |r-c|, [14] near diagonal
  using Scan = hipcub::BlockScan<uint32_t, 256>;
  __shared__ typename Scan::TempStorage scan_tmp;
  __shared__ unsigned long long hist[12], sums[2];
  __shared__ unsigned long long mx, vbase;
  if (threadIdx.x < 12) hist[threadIdx.x] = 0;
  if (threadIdx.x < 2) sums[threadIdx.x] = 0;
  if (threadIdx.x == 0) mx = 0;
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t nv = 0;  // violating groups of this row
  int32_t b = 0, e = 0;
  if (r < M) {
    b = rp[r];
    e = rp[r + 1];
    const int64_t len = e - b;
    atomicAdd(&hist[len_bucket(len)], 1ull);
    atomicMax(&mx, (unsigned long long)len);
    if (len > kLongRow) e = b;  // entries of long rows: k_row_stats_long
    unsigned long long s = 0, nd = 0;
    int32_t prev = -1, g_prev = -1, cnt = 0;
    bool sorted = true;
    for (int32_t q = b; q < e; q++) {
      const int32_t c = ci[q];
      const int64_t d = r > c ? r - (int64_t)c : (int64_t)c - r;
      s += (unsigned long long)d;
      nd += d <= 64;
      keys[q] = ((KT)(r / 16) << colbits) | (KT)(uint32_t)c;
      sorted = sorted && c > prev;
      prev = c;
      if ((c >> 2) != g_prev) {
        g_prev = c >> 2;
        cnt = 0;
      }
      if ((val[q] & 0x7fff) != 0 && ++cnt == 3) nv++;
    }
    if (!sorted) {
      atomicOr(unsorted, 1);
      nv = 0;
    }
    atomicAdd(&sums[0], s);
    atomicAdd(&sums[1], nd);
  }
  uint32_t off = 0, total = 0;
  Scan(scan_tmp).ExclusiveSum(nv, off, total);
  if (threadIdx.x == 0) vbase = total ? atomicAdd(vcount, (unsigned long long)total) : 0ull;
  __syncthreads();
  if (nv) {  // second walk over the row: write its violation keys
    KT* out = vkeys + vbase + off;
    int32_t g_prev = -1, cnt = 0;
    for (int32_t q = b; q < e; q++) {
      const int32_t c = ci[q];
      if ((c >> 2) != g_prev) {
        g_prev = c >> 2;
        cnt = 0;
      }
      if ((val[q] & 0x7fff) != 0 && ++cnt == 3) *out++ = ((KT)(r / 16) << (colbits - 4)) | (KT)(uint32_t)(c >> 4);
    }
  }
  if (threadIdx.x < 12 && hist[threadIdx.x]) atomicAdd(&acc[threadIdx.x], hist[threadIdx.x]);
  if (threadIdx.x == 0) {
    atomicMax(&acc[12], mx);
    atomicAdd(&acc[13], sums[0]);
    atomicAdd(&acc[14], sums[1]);
  }
}

// Per-entry statistics of rows longer than kLongRow, one wave per row (a
// workgroup of 4 waves takes 4 rows; short rows exit): the same sums, sort
// keys and strict-ascending check as k_row_stats.  In a row whose columns
// ascend, an entry is the third nonzero value of its aligned 1x4 group
// exactly when two of the (at most three) entries before it are nonzero
// values of the same group, so each lane finds its violations by looking
// back three entries; slots in vkeys are reserved per 64-entry chunk.  Rows
// that do not ascend only set *unsorted (their keys then go unused), and
// no write passes vcap.
template <typename KT>
__global__ __launch_bounds__(256) void k_row_stats_long(int64_t M, const int32_t* __restrict__ rp,
                                                        const int32_t* __restrict__ ci,
                                                        const uint16_t* __restrict__ val, int colbits,
                                                        unsigned long long* __restrict__ acc,
                                                        KT* __restrict__ keys, KT* __restrict__ vkeys, int64_t vcap,
                                                        unsigned long long* __restrict__ vcount,
                                                        int32_t* __restrict__ unsorted) {
  const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= M) return;
  const int32_t b = rp[r], e = rp[r + 1];
  if (e - b <= kLongRow) return;
  unsigned long long s = 0, nd = 0;
  bool sorted = true;
  for (int32_t q0 = b; q0 < e; q0 += 64) {
    const int32_t q = q0 + lane;
    bool viol = false;
    int32_t c = 0;
    if (q < e) {
      c = ci[q];
      const int64_t d = r > c ? r - (int64_t)c : (int64_t)c - r;
      s += (unsigned long long)d;
      nd += d <= 64;
      keys[q] = ((KT)(r / 16) << colbits) | (KT)(uint32_t)c;
      if (q > b) sorted = sorted && c > ci[q - 1];
      if ((val[q] & 0x7fff) != 0) {
        int before = 0;
        for (int k = 1; k <= 3 && q - k >= b; k++)
          before += (ci[q - k] >> 2) == (c >> 2) && (val[q - k] & 0x7fff) != 0;
        viol = before == 2;
      }
    }
    const unsigned long long mask = __ballot(viol);
    if (mask) {
      unsigned long long base = 0;
      if (lane == 0) base = atomicAdd(vcount, (unsigned long long)__popcll(mask));
      base = __shfl(base, 0, 64);
      const unsigned long long below = lane ? mask & (~0ull >> (64 - lane)) : 0ull;
      const int64_t pos = (int64_t)(base + (unsigned long long)__popcll(below));
      if (viol && pos < vcap) vkeys[pos] = ((KT)(r / 16) << (colbits - 4)) | (KT)(uint32_t)(c >> 4);
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    s += __shfl_down(s, o, 64);
    nd += __shfl_down(nd, o, 64);
  }
  if (__any(!sorted) && lane == 0) atomicOr(unsorted, 1);
  if (lane == 0) {
    atomicAdd(&acc[13], s);
    atomicAdd(&acc[14], nd);
  }
}

// On keys sorted by (block row, column): a key starts a new 16 x bk block iff
// its block row or its column / bk differs from the previous key's.
template <typename KT>
__global__ __launch_bounds__(256) void k_block_heads(int64_t n, const KT* __restrict__ keys, int colbits,
                                                     unsigned long long* __restrict__ nnzb) {
  __shared__ unsigned long long cnt[6];
  if (threadIdx.x < 6) cnt[threadIdx.x] = 0;
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const KT k = keys[i];
    const KT p = i > 0 ? keys[i - 1] : k;
    const bool new_row = i == 0 || (k >> colbits) != (p >> colbits);
    const KT mask = ((KT)1 << colbits) - 1;
    const KT c = k & mask, pc = p & mask;
#pragma unroll
    for (int w = 0; w < 6; w++)
      if (new_row || (c >> w) != (pc >> w)) atomicAdd(&cnt[w], 1ull);
  }
  __syncthreads();
  if (threadIdx.x < 6 && cnt[threadIdx.x]) atomicAdd(&nnzb[threadIdx.x], cnt[threadIdx.x]);
}

// Distinct keys of a sorted array (run heads).
template <typename KT>
__global__ void k_count_heads(int64_t n, const KT* __restrict__ keys, unsigned long long* __restrict__ out) {
  __shared__ unsigned long long c;
  if (threadIdx.x == 0) c = 0;
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && (i == 0 || keys[i] != keys[i - 1])) atomicAdd(&c, 1ull);
  __syncthreads();
  if (threadIdx.x == 0 && c) atomicAdd(out, c);
}

__global__ void k_count_tag1(int64_t n, const int32_t* __restrict__ tags, unsigned long long* __restrict__ out) {
  __shared__ unsigned long long c;
  if (threadIdx.x == 0) c = 0;
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && tags[i] == 1) atomicAdd(&c, 1ull);
  __syncthreads();
  if (threadIdx.x == 0 && c) atomicAdd(out, c);
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// ---- MLP kernels --------------------------------------------------------------
// Y[b][o] = act(sum_i X[b][i] W[o][i] + bias[o]); act = ReLU or identity.
__global__ void k_linear_fwd(const float* __restrict__ X, int64_t B, int32_t n_in, const float* __restrict__ W,
                             const float* __restrict__ bias, int32_t n_out, int32_t relu, float* __restrict__ Y) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * n_out) return;
  const int64_t b = t / n_out;
  const int32_t o = (int32_t)(t % n_out);
  float s = bias[o];
  for (int32_t i = 0; i < n_in; i++) s = fmaf(X[b * n_in + i], W[(int64_t)o * n_in + i], s);
  Y[t] = relu && s < 0.f ? 0.f : s;
}

// dZ = dY, masked by the ReLU (Y > 0) when the layer has one.
__device__ inline float dz(const float* __restrict__ Y, const float* __restrict__ dY, int64_t idx, int32_t relu) {
  return relu && !(Y[idx] > 0.f) ? 0.f : dY[idx];
}

// dX[b][i] = sum_o dZ[b][o] W[o][i]
__global__ void k_linear_bwd_x(const float* __restrict__ Y, const float* __restrict__ dY, int64_t B, int32_t n_in,
                               const float* __restrict__ W, int32_t n_out, int32_t relu, float* __restrict__ dX) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * n_in) return;
  const int64_t b = t / n_in;
  const int32_t i = (int32_t)(t % n_in);
  float s = 0.f;
  for (int32_t o = 0; o < n_out; o++) s = fmaf(dz(Y, dY, b * n_out + o, relu), W[(int64_t)o * n_in + i], s);
  dX[t] = s;
}

// dW[o][i] = sum_b dZ[b][o] X[b][i]; db[o] = sum_b dZ[b][o] (thread per weight,
// rows of the batch in order: deterministic).
__global__ void k_linear_bwd_w(const float* __restrict__ X, const float* __restrict__ Y, const float* __restrict__ dY,
                               int64_t B, int32_t n_in, int32_t n_out, int32_t relu, float* __restrict__ dW,
                               float* __restrict__ db) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nw = (int64_t)n_out * (n_in + 1);
  if (t >= nw) return;
  const int32_t o = (int32_t)(t / (n_in + 1)), i = (int32_t)(t % (n_in + 1));  // i == n_in: the bias
  float s = 0.f;
  for (int64_t b = 0; b < B; b++) {
    const float g = dz(Y, dY, b * n_out + o, relu);
    s = i < n_in ? fmaf(g, X[b * n_in + i], s) : s + g;
  }
  if (i < n_in) dW[(int64_t)o * n_in + i] = s;
  else db[o] = s;
}

// Mean squared error of pred against y over B rows: dpred = 2 (pred - y) / B,
// loss[0] = mean (pred - y)^2 (one workgroup, fixed summation order).
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ pred, const float* __restrict__ y, int64_t B,
                                             float* __restrict__ dpred, float* __restrict__ loss) {
  __shared__ float part[256];
  float s = 0.f;
  for (int64_t b = threadIdx.x; b < B; b += 256) {
    const float d = pred[b] - y[b];
    dpred[b] = 2.f * d / (float)B;
    s = fmaf(d, d, s);
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[0] = part[0] / (float)B;
}

// Adam (Kingma & Ba 2015, bias-corrected), step t >= 1.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Inference: one workgroup scores every candidate configuration of one matrix
// with the 3-layer network (inputs standardised in the kernel) and writes the
// scores and the index of the lowest available one.  Layer widths <= 64.
constexpr int kMaxIn = 64, kMaxHidden = 128, kMaxCfg = 256;

// Inference: one workgroup per candidate configuration (blockIdx.x), the
// three layers with one thread per output and a fixed summation order;
// k_mlp_argmin then takes the first minimum over the available candidates.
__global__ __launch_bounds__(64) void k_mlp_scores(
    const float* __restrict__ feats, int32_t n_feat, const float* __restrict__ cfg_enc, int32_t n_enc,
    const float* __restrict__ mu, const float* __restrict__ sigma, const float* __restrict__ W1,
    const float* __restrict__ b1, const float* __restrict__ W2, const float* __restrict__ b2,
    const float* __restrict__ W3, const float* __restrict__ b3, int32_t hidden, float* __restrict__ scores) {
  __shared__ float x[kMaxIn], h1[kMaxHidden], h2[kMaxHidden];
  const int c = blockIdx.x;
  const int n_in = n_feat + n_enc;
  for (int i = threadIdx.x; i < n_in; i += blockDim.x) {
    const float raw = i < n_feat ? feats[i] : cfg_enc[c * n_enc + (i - n_feat)];
    x[i] = (raw - mu[i]) / sigma[i];
  }
  __syncthreads();
  for (int o = threadIdx.x; o < hidden; o += blockDim.x) {
    float s = b1[o];
    for (int i = 0; i < n_in; i++) s = fmaf(x[i], W1[o * n_in + i], s);
    h1[o] = s > 0.f ? s : 0.f;
  }
  __syncthreads();
  for (int o = threadIdx.x; o < hidden; o += blockDim.x) {
    float s = b2[o];
    for (int i = 0; i < hidden; i++) s = fmaf(h1[i], W2[o * hidden + i], s);
    h2[o] = s > 0.f ? s : 0.f;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = b3[0];
    for (int i = 0; i < hidden; i++) s = fmaf(h2[i], W3[i], s);
    scores[c] = s;
  }
}

__global__ void k_mlp_argmin(const float* __restrict__ scores, const uint8_t* __restrict__ avail, int32_t n_cfg,
                             int32_t* __restrict__ best) {
  if (threadIdx.x != 0) return;
  int bi = -1;
  for (int c = 0; c < n_cfg; c++)
    if (avail[c] && (bi < 0 || scores[c] < scores[bi])) bi = c;
  best[0] = bi;
}

inline int bits_for(int64_t x) {  // bits needed for values 0..x (x >= 0)
  int b = 1;
  while (b < 63 && (x >> b) != 0) b++;
  return b;
}

// Statistics pass with KT-wide sort keys: row statistics, the stored-block
// counts of the six block widths (one radix sort over the key bits in use),
// and the number of distinct 16 x 16 blocks that violate the exact 2:4 rule
// (*violated; -1 when some row's columns do not ascend).
template <typename KT>
smat_status stats_pass(int64_t M, int64_t nnz, const int32_t* rowptr, const int32_t* colidx, const uint16_t* values,
                       int colbits, int keybits, unsigned long long* a, int64_t* violated, hipStream_t s) {
  smat::DevBuf keys, keys2, vkeys, vkeys2, tmp, tmp2, misc;
  SMAT_TRY(keys.alloc((nnz + 1) * sizeof(KT)));
  SMAT_TRY(keys2.alloc((nnz + 1) * sizeof(KT)));
  SMAT_TRY(vkeys.alloc((nnz / 3 + 1) * sizeof(KT)));  // at most one violation per three entries
  SMAT_TRY(misc.alloc(32));                            // [0] violation count (u64), [1] unsorted flag
  SMAT_HIP_TRY(hipMemsetAsync(misc.p, 0, 32, s));
  unsigned long long* vcount = misc.as<unsigned long long>();
  int32_t* unsorted = reinterpret_cast<int32_t*>(misc.as<unsigned long long>() + 1);
  hipLaunchKernelGGL(k_row_stats<KT>, grid_of(M, 256), 256, 0, s, M, rowptr, colidx, values, colbits, a,
                     keys.as<KT>(), vkeys.as<KT>(), vcount, unsorted);
  hipLaunchKernelGGL(k_row_stats_long<KT>, grid_of(M, 4), 256, 0, s, M, rowptr, colidx, values, colbits, a,
                     keys.as<KT>(), vkeys.as<KT>(), nnz / 3 + 1, vcount, unsorted);
  SMAT_LAUNCH_CHECK();
  if (nnz > 0) {
    size_t tb = 0;
    SMAT_HIP_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, keys.as<KT>(), keys2.as<KT>(), (int)nnz, 0, keybits,
                                                   s));
    SMAT_TRY(tmp.alloc(tb));
    SMAT_HIP_TRY(hipcub::DeviceRadixSort::SortKeys(tmp.p, tb, keys.as<KT>(), keys2.as<KT>(), (int)nnz, 0, keybits, s));
    hipLaunchKernelGGL(k_block_heads<KT>, grid_of(nnz, 256), 256, 0, s, nnz, keys2.as<KT>(), colbits, a + 15);
    SMAT_LAUNCH_CHECK();
  }
  unsigned long long h[2] = {0, 0};
  SMAT_HIP_TRY(hipMemcpyAsync(h, misc.p, 16, hipMemcpyDeviceToHost, s));
  SMAT_HIP_TRY(hipStreamSynchronize(s));
  if ((h[1] & 0xffffffffull) != 0) {
    *violated = -1;
    return SMAT_OK;
  }
  const int64_t nv = (int64_t)h[0];
  *violated = 0;
  if (nv == 0) return SMAT_OK;
  SMAT_TRY(vkeys2.alloc(nv * sizeof(KT)));
  size_t tb = 0;
  SMAT_HIP_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, vkeys.as<KT>(), vkeys2.as<KT>(), (int)nv, 0,
                                                 keybits - 4, s));
  SMAT_TRY(tmp2.alloc(tb));
  SMAT_HIP_TRY(hipcub::DeviceRadixSort::SortKeys(tmp2.p, tb, vkeys.as<KT>(), vkeys2.as<KT>(), (int)nv, 0,
                                                 keybits - 4, s));
  SMAT_HIP_TRY(hipMemsetAsync(misc.p, 0, 8, s));
  hipLaunchKernelGGL(k_count_heads<KT>, grid_of(nv, 256), 256, 0, s, nv, vkeys2.as<KT>(), vcount);
  SMAT_LAUNCH_CHECK();
  SMAT_HIP_TRY(hipMemcpyAsync(h, misc.p, 8, hipMemcpyDeviceToHost, s));
  SMAT_HIP_TRY(hipStreamSynchronize(s));
  *violated = (int64_t)h[0];
  return SMAT_OK;
}

}  // namespace

extern "C" {

smat_status smat_matrix_stats(int64_t M, int64_t K, int64_t nnz, const int32_t* rowptr, const int32_t* colidx,
                              const uint16_t* values, smat_matrix_stats_t* out, smat_stream stream) {
  SMAT_REQUIRE(out && rowptr && M > 0 && K > 0 && nnz >= 0, "smat_matrix_stats: bad argument");
  SMAT_REQUIRE(M % 16 == 0 && K % 32 == 0, "smat_matrix_stats: M must be a multiple of 16 and K of 32");
  SMAT_REQUIRE(nnz == 0 || (colidx && values), "smat_matrix_stats: null CSR arrays");
  SMAT_REQUIRE(K <= (int64_t)INT32_MAX + 1, "smat_matrix_stats: K too large");
  hipStream_t s = smat::to_hip(stream);
  smat::DevBuf acc;
  SMAT_TRY(acc.alloc(32 * sizeof(unsigned long long)));  // 15 row stats, 6 block counts
  unsigned long long* a = acc.as<unsigned long long>();
  hipLaunchKernelGGL(k_zero_u64, 1, 32, 0, s, a, 32);
  // sort keys: block row above the column; 32-bit keys when both fit
  const int colbits = std::max(5, bits_for(K - 1));
  const int keybits = bits_for(M / 16 - 1) + colbits;
  int64_t violated = 0;
  if (keybits <= 32) SMAT_TRY(stats_pass<uint32_t>(M, nnz, rowptr, colidx, values, colbits, keybits, a, &violated, s));
  else SMAT_TRY(stats_pass<uint64_t>(M, nnz, rowptr, colidx, values, colbits, keybits, a, &violated, s));
  unsigned long long h[32];
  SMAT_HIP_TRY(hipMemcpyAsync(h, a, sizeof(h), hipMemcpyDeviceToHost, s));
  SMAT_HIP_TRY(hipStreamSynchronize(s));
  const int64_t nnzb16 = (int64_t)h[15 + 4];
  int64_t blocks24 = nnzb16 - violated;
  if (violated < 0) {
    // rows with unsorted columns: tag the 16 x 16 BCSR under the exact rule
    smat_bcsr A16 = nullptr;
    SMAT_TRY(smat_bcsr_from_csr(M, K, nnz, rowptr, colidx, values, 16, SMAT_TAG_EXACT24, stream, &A16));
    smat_bcsr_desc d{};
    smat_status st = smat_bcsr_describe(A16, &d);
    unsigned long long c = 0;
    if (st == SMAT_OK && d.nnzb > 0) {
      hipLaunchKernelGGL(k_zero_u64, 1, 32, 0, s, a, 1);
      hipLaunchKernelGGL(k_count_tag1, grid_of(d.nnzb, 256), 256, 0, s, d.nnzb, d.tags, a);
      hipError_t e = hipGetLastError();
      if (e == hipSuccess) e = hipMemcpyAsync(&c, a, 8, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      if (e != hipSuccess) st = smat::set_error(SMAT_ERR_HIP, "smat_matrix_stats: %s", hipGetErrorString(e));
    }
    smat_bcsr_destroy(A16);
    if (st != SMAT_OK) return st;
    blocks24 = (int64_t)c;
  }
  out->rows = M;
  out->cols = K;
  out->nnz = nnz;
  for (int b = 0; b < 12; b++) out->row_hist[b] = (int64_t)h[b];
  out->max_row = (int64_t)h[12];
  out->sum_abs_offset = (int64_t)h[13];
  out->near_diag = (int64_t)h[14];
  for (int w = 0; w < 6; w++) out->nnzb[w] = (int64_t)h[15 + w];
  out->nnzb16 = nnzb16;
  out->blocks24 = blocks24;
  return SMAT_OK;
}

smat_status smat_tuner_features(const smat_matrix_stats_t* st, int64_t N, float* out) {
  SMAT_REQUIRE(st && out && N > 0 && st->rows > 0, "smat_tuner_features: bad argument");
  const double rows = (double)st->rows, nnz = (double)st->nnz;
  const double mean = nnz / rows;
  int k = 0;
  out[k++] = (float)std::log2(rows);
  out[k++] = (float)std::log2(nnz + 1.0);
  out[k++] = (float)std::log2(mean + 1.0);
  out[k++] = (float)std::log2((double)st->max_row + 1.0) - (float)std::log2(mean + 1.0);
  for (int b = 0; b < 12; b++) out[k++] = (float)((double)st->row_hist[b] / rows);
  out[k++] = (float)(nnz > 0 ? (double)st->sum_abs_offset / nnz / rows : 0.0);
  out[k++] = (float)(nnz > 0 ? (double)st->near_diag / nnz : 0.0);
  for (int w = 0; w < 6; w++) {  // log2 fill of 16 x bk blocks, bk = 1 << w
    const double stored = (double)st->nnzb[w] * 16.0 * (double)(1 << w);
    out[k++] = (float)(stored > 0 ? std::log2(nnz / stored) : 0.0);
  }
  out[k++] = (float)(st->nnzb16 > 0 ? (double)st->blocks24 / (double)st->nnzb16 : 0.0);
  out[k++] = (float)std::log2((double)N);
  return k == SMAT_TUNER_NF ? SMAT_OK : smat::set_error(SMAT_ERR_INVALID, "smat_tuner_features: internal count");
}

smat_status smat_linear_forward(const float* X, int64_t B, int32_t n_in, const float* W, const float* bias,
                                int32_t n_out, int32_t relu, float* Y, smat_stream stream) {
  SMAT_REQUIRE(X && W && bias && Y && B > 0 && n_in > 0 && n_out > 0, "smat_linear_forward: bad argument");
  hipLaunchKernelGGL(k_linear_fwd, grid_of(B * n_out, 256), 256, 0, smat::to_hip(stream), X, B, n_in, W, bias, n_out,
                     relu, Y);
  SMAT_LAUNCH_CHECK();
  return SMAT_OK;
}

smat_status smat_linear_backward(const float* X, const float* Y, const float* dY, int64_t B, int32_t n_in,
                                 const float* W, int32_t n_out, int32_t relu, float* dX, float* dW, float* db,
                                 smat_stream stream) {
  SMAT_REQUIRE(X && Y && dY && W && dW && db && B > 0 && n_in > 0 && n_out > 0, "smat_linear_backward: bad argument");
  hipStream_t s = smat::to_hip(stream);
  hipLaunchKernelGGL(k_linear_bwd_w, grid_of((int64_t)n_out * (n_in + 1), 64), 64, 0, s, X, Y, dY, B, n_in, n_out,
                     relu, dW, db);
  if (dX) hipLaunchKernelGGL(k_linear_bwd_x, grid_of(B * n_in, 256), 256, 0, s, Y, dY, B, n_in, W, n_out, relu, dX);
  SMAT_LAUNCH_CHECK();
  return SMAT_OK;
}

smat_status smat_mse_grad(const float* pred, const float* y, int64_t B, float* dpred, float* loss,
                          smat_stream stream) {
  SMAT_REQUIRE(pred && y && dpred && loss && B > 0, "smat_mse_grad: bad argument");
  hipLaunchKernelGGL(k_mse, 1, 256, 0, smat::to_hip(stream), pred, y, B, dpred, loss);
  SMAT_LAUNCH_CHECK();
  return SMAT_OK;
}

smat_status smat_adam_step(float* p, const float* g, float* m, float* v, int64_t n, float lr, float beta1,
                           float beta2, float eps, int32_t t, smat_stream stream) {
  SMAT_REQUIRE(p && g && m && v && n > 0 && t >= 1, "smat_adam_step: bad argument");
  const float c1 = 1.f - std::pow(beta1, (float)t), c2 = 1.f - std::pow(beta2, (float)t);
  hipLaunchKernelGGL(k_adam, grid_of(n, 256), 256, 0, smat::to_hip(stream), p, g, m, v, n, lr, beta1, beta2, eps, c1,
                     c2);
  SMAT_LAUNCH_CHECK();
  return SMAT_OK;
}

smat_status smat_tuner_predict(const float* feats, int32_t n_feat, const float* cfg_enc, int32_t n_enc,
                               const uint8_t* avail, int32_t n_cfg, const float* mu, const float* sigma,
                               const float* W1, const float* b1, const float* W2, const float* b2, const float* W3,
                               const float* b3, int32_t hidden, float* scores, int32_t* best, smat_stream stream) {
  SMAT_REQUIRE(feats && cfg_enc && avail && mu && sigma && W1 && b1 && W2 && b2 && W3 && b3 && scores && best,
               "smat_tuner_predict: null argument");
  SMAT_REQUIRE(n_feat + n_enc <= kMaxIn && hidden <= kMaxHidden && hidden > 0 && n_cfg > 0 && n_cfg <= kMaxCfg,
               "smat_tuner_predict: sizes exceed the kernel's limits (inputs %d, hidden %d, configs %d)",
               kMaxIn, kMaxHidden, kMaxCfg);
  hipStream_t s = smat::to_hip(stream);
  hipLaunchKernelGGL(k_mlp_scores, (unsigned)n_cfg, 64, 0, s, feats, n_feat, cfg_enc, n_enc, mu, sigma, W1, b1, W2, b2,
                     W3, b3, hidden, scores);
  hipLaunchKernelGGL(k_mlp_argmin, 1, 64, 0, s, scores, avail, n_cfg, best);
  SMAT_LAUNCH_CHECK();
  return SMAT_OK;
}

}  // extern "C"
