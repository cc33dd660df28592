// Jaccard row reordering (SPARTA "Keeper", the grouping of the SMaT reorder
// pipeline: src/reorder.py:131 runs Matrix_Blocking with -a 5 -b 8 -B 16 -F 1).
// Bit-exact with IterativeBlockingKeeper
// (preprocess/src/general/blocking.cpp:431-547), get_fixed_size_grouping and
// get_permutation (preprocess/src/general/utilities.cpp:8-54) and
// CSR::permute_rows (csr.cpp:65-74).
//
// Work split.  Seeds are sequential (each seed's outcome decides the next
// seed and which rows are still free).  Inside a seed the candidates j > i
// are independent until the first merge, so the GPU evaluates one round of
// the seed at a time:
//   * k_scan: for every candidate of [j0, rows) the prune test (seed-start
//     distances; a candidate's distance is only written when it is visited),
//     the "already grouped" test and, for eligible rows, the Jaccard
//     distance to the current group pattern (block-level, count_zeros = 1,
//     double division cast to float as blocking.cpp:974 does);
//   * a block-wide scan compacts the compared candidates, in row order, to
//     (row, distance) pairs and finds the first one with distance <= tau;
//   * k_commit applies the visit side effects of [j0, merge] (pruned rows get
//     distance -1, compared rows their distance) once the host knows where
//     the round stopped.
// The host replays the order-dependent part on the compacted list: the
// top-up candidate set with the reference's exact truncation
// (`advance(best.end(), k)` followed by `erase(it, end())`, whose effect
// depends on the shape of libstdc++'s red-black tree — the same tree
// algorithm of the same library reproduces it), merge_rows on the pattern,
// the top-up and the group numbering.  That part is pointer chasing through a
// set of thousands of elements per seed, latency-bound on a GPU lane.  Two
// exact shortcuts keep it cheap: the truncation is computed from the shape
// at the set's largest node in O(1) (truncate_from_end), and a candidate
// above the current maximum whose insertion the truncation would undo
// without any rebalancing is skipped (about 80 % of them on R-MAT graphs).
#include <algorithm>
#include <iterator>
#include <numeric>
#include <set>
#include <utility>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "smat_internal.h"

namespace {

constexpr int kScanThreads = 256;
constexpr int32_t kLongRow = 64;  // rows with more entries are compared by a whole wave
constexpr uint8_t ST_SKIP = 0, ST_PRUNED = 1, ST_COMPARED = 2;

// JaccardDistanceGroup(pattern, gA, row, 1, block) over sorted column lists.
__device__ inline float jaccard(const int32_t* __restrict__ A, int32_t nA, int64_t gA,
                                const int32_t* __restrict__ Bv, int32_t nB, int64_t block) {
  if (nA == 0 && nB == 0) return 0.f;
  if (nA == 0 || nB == 0) return 1.f;
  const int64_t addA = 1, addB = gA;  // count_zeros: |A\B| weighted by g_B (=1), |B\A| by g_A
  int32_t i = 0, j = 0;
  int64_t count = 0, bA = 0, bB = 0;
  while (i < nA && j < nB) {
    const int64_t pa = A[i] / block, pb = Bv[j] / block;
    if (pa < pb) {
      count += addA;
      bA++;
      while (i < nA && A[i] / block == pa) i++;
    } else if (pa > pb) {
      count += addB;
      bB++;
      while (j < nB && Bv[j] / block == pb) j++;
    } else {
      bA++;
      bB++;
      while (i < nA && A[i] / block == pa) i++;
      while (j < nB && Bv[j] / block == pb) j++;
    }
  }
  while (i < nA) {
    const int64_t pa = A[i] / block;
    count += addA;
    bA++;
    while (i < nA && A[i] / block == pa) i++;
  }
  while (j < nB) {
    const int64_t pb = Bv[j] / block;
    count += addB;
    bB++;
    while (j < nB && Bv[j] / block == pb) j++;
  }
  return (float)((2.0 * (double)count) / (double)(bA * gA + bB * 1 + count));
}

// Round, kernel 1 (kScanThreads candidates per block): status, distance,
// block-local exclusive position of each compared candidate, per-block
// count, and the first merging row (atomicMin).
__global__ __launch_bounds__(kScanThreads) void k_round_scan(
    int64_t rows, int64_t seed, int64_t j0, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
    const int32_t* __restrict__ pattern, int32_t plen, int64_t gsize, float tau, int64_t block,
    const float* __restrict__ d, const uint8_t* __restrict__ grouped, uint8_t* __restrict__ status,
    float* __restrict__ dist, int32_t* __restrict__ local_pos, int32_t* __restrict__ block_count,
    unsigned long long* __restrict__ first_merge) {
  __shared__ int32_t wsum[kScanThreads / 64];
  const int t = threadIdx.x;
  const int64_t j = j0 + (int64_t)blockIdx.x * kScanThreads + t;
  int flag = 0;
  if (j < rows) {
    const float di = d[seed], dj = d[j];
    uint8_t st;
    if (di != -1.f && dj != -1.f && fabsf(di - dj) > tau) {
      st = ST_PRUNED;
    } else if (grouped[j]) {
      st = ST_SKIP;
    } else {
      const int32_t b = rowptr[j], e = rowptr[j + 1];
      st = ST_COMPARED;
      flag = 1;
      if (e - b <= kLongRow) {  // longer rows get a whole wave in k_round_long
        const float x = jaccard(pattern, plen, gsize, colidx + b, e - b, block);
        dist[j] = x;
        if (x <= tau) atomicMin(first_merge, (unsigned long long)j);
      }
    }
    status[j] = st;
  }
  // block-local exclusive prefix of the compared flags (wave ballots + LDS)
  const int lane = t & 63, w = t >> 6;
  const unsigned long long bal = __ballot(flag);
  const int in_wave = __popcll(bal & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
  if (lane == 0) wsum[w] = __popcll(bal);
  __syncthreads();
  int before = 0, total = 0;
#pragma unroll
  for (int q = 0; q < kScanThreads / 64; q++) {
    before += (q < w) ? wsum[q] : 0;
    total += wsum[q];
  }
  if (flag) local_pos[j] = before + in_wave;
  if (t == 0) block_count[blockIdx.x] = total;
}

// Round, long candidates: one wave per row of `long_rows` (ascending row
// indices, each longer than kLongRow) that k_round_scan marked compared.  The
// distance needs only three counts — distinct column blocks of the pattern
// (n_blk_pat, from the host), of the row, and of the row's blocks present in
// the pattern — and the same final expression as jaccard(), so it is
// bit-identical to the per-thread path.
__global__ __launch_bounds__(256) void k_round_long(const int64_t* __restrict__ long_rows, int64_t n_long,
                                                    const int32_t* __restrict__ rowptr,
                                                    const int32_t* __restrict__ colidx,
                                                    const int32_t* __restrict__ pattern, int32_t plen,
                                                    int64_t n_blk_pat, int64_t gsize, float tau, int64_t block,
                                                    const uint8_t* __restrict__ status, float* __restrict__ dist,
                                                    unsigned long long* __restrict__ first_merge) {
  const int64_t w = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (w >= n_long) return;
  const int64_t j = long_rows[w];
  if (status[j] != ST_COMPARED) return;  // wave-uniform
  const int32_t b = rowptr[j], e = rowptr[j + 1];
  int64_t nb = 0, inter = 0;
  for (int32_t q = b + lane; q < e; q += 64) {
    const int64_t blk = colidx[q] / block;
    if (q > b && colidx[q - 1] / block == blk) continue;  // count each block at its first entry
    nb++;
    // does the pattern hold a column of this block?  (first pattern column >= blk * block)
    int32_t lo = 0, hi = plen;
    const int64_t c0 = blk * block;
    while (lo < hi) {
      const int32_t mid = (lo + hi) >> 1;
      if (pattern[mid] < c0) lo = mid + 1;
      else hi = mid;
    }
    if (lo < plen && pattern[lo] / block == blk) inter++;
  }
  for (int off = 32; off > 0; off >>= 1) {
    nb += __shfl_xor(nb, off);
    inter += __shfl_xor(inter, off);
  }
  if (lane != 0) return;
  float x;
  if (plen == 0) x = 1.f;  // the row is not empty (longer than kLongRow)
  else {
    const int64_t count = (n_blk_pat - inter) * 1 + (nb - inter) * gsize;
    x = (float)((2.0 * (double)count) / (double)(n_blk_pat * gsize + nb * 1 + count));
  }
  dist[j] = x;
  if (x <= tau) atomicMin(first_merge, (unsigned long long)j);
}

// Round, kernel 2 (one block): block offsets, total and first merge ->
// header in mapped host memory; resets first_merge for the next round.
__global__ __launch_bounds__(1024) void k_round_offsets(int32_t nb, const int32_t* __restrict__ block_count,
                                                        int32_t* __restrict__ block_off,
                                                        unsigned long long* __restrict__ first_merge,
                                                        int64_t* __restrict__ hdr) {
  __shared__ int32_t part[1024];
  const int t = threadIdx.x;
  int32_t carry = 0;
  for (int32_t base = 0; base < nb; base += 1024) {
    const int32_t v = (base + t < nb) ? block_count[base + t] : 0;
    part[t] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // Hillis-Steele inclusive scan
      const int32_t x = (t >= off) ? part[t - off] : 0;
      __syncthreads();
      part[t] += x;
      __syncthreads();
    }
    if (base + t < nb) block_off[base + t] = carry + part[t] - v;
    carry += part[1023];
    __syncthreads();
  }
  if (t == 0) {
    hdr[0] = carry;
    hdr[1] = (int64_t)*first_merge;
    *first_merge = (unsigned long long)INT64_MAX;
  }
}

// Round, kernel 3: compared candidates in row order -> mapped host memory.
__global__ __launch_bounds__(kScanThreads) void k_round_write(int64_t rows, int64_t j0,
                                                              const uint8_t* __restrict__ status,
                                                              const float* __restrict__ dist,
                                                              const int32_t* __restrict__ local_pos,
                                                              const int32_t* __restrict__ block_off,
                                                              int32_t* __restrict__ out_j, float* __restrict__ out_d) {
  const int64_t j = j0 + (int64_t)blockIdx.x * kScanThreads + threadIdx.x;
  if (j >= rows || status[j] != ST_COMPARED) return;
  const int32_t q = block_off[blockIdx.x] + local_pos[j];
  out_j[q] = (int32_t)j;
  out_d[q] = dist[j];
}

__global__ void k_reset_first(unsigned long long* first_merge) { *first_merge = (unsigned long long)INT64_MAX; }

// Visit side effects for [j0, jend): pruned -> -1, compared -> its distance.
__global__ void k_commit(int64_t j0, int64_t jend, const uint8_t* __restrict__ status,
                         const float* __restrict__ dist, float* __restrict__ d) {
  const int64_t j = j0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= jend) return;
  const uint8_t st = status[j];
  if (st == ST_PRUNED) d[j] = -1.f;
  else if (st == ST_COMPARED) d[j] = dist[j];
}

__global__ void k_mark(const int64_t* __restrict__ rows_list, int32_t n, uint8_t* __restrict__ grouped) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) grouped[rows_list[t]] = 1;
}

__global__ void k_init(int64_t rows, float* __restrict__ d, uint8_t* __restrict__ grouped) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= rows) return;
  d[j] = (j == 0) ? -1.f : 0.f;  // the reference's VLA initialiser {-1} sets element 0 only
  grouped[j] = 0;
}

// Row permutation of a CSR: new row r is old row perm[r].
__global__ void k_perm_lengths(int64_t rows, const int64_t* __restrict__ perm, const int32_t* __restrict__ rowptr,
                               int32_t* __restrict__ len) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r > rows) return;
  len[r] = r < rows ? rowptr[perm[r] + 1] - rowptr[perm[r]] : 0;
}

__global__ void k_perm_gather(int64_t rows, const int64_t* __restrict__ perm, const int32_t* __restrict__ rowptr,
                              const int32_t* __restrict__ colidx, const uint16_t* __restrict__ val,
                              const int32_t* __restrict__ new_rowptr, int32_t* __restrict__ out_col,
                              uint16_t* __restrict__ out_val) {
  // one 64-lane group per row
  const int64_t r = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (r >= rows) return;
  const int32_t src = rowptr[perm[r]], n = rowptr[perm[r] + 1] - src, dst = new_rowptr[r];
  for (int32_t k = lane; k < n; k += 64) {
    out_col[dst + k] = colidx[src + k];
    if (val) out_val[dst + k] = val[src + k];
  }
}

// Column relabel: new column of an entry = inv[old column].
__global__ void k_relabel_cols(int64_t nnz, const int32_t* __restrict__ colidx, const int64_t* __restrict__ inv,
                               int32_t* __restrict__ out) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nnz) out[q] = (int32_t)inv[colidx[q]];
}

// Dense row gather: out row r = in row perm[r] (elem_bytes-wide elements,
// 16-byte chunks where the row is 16-byte aligned, else 2-byte units).
__global__ void k_gather_rows(int64_t K, int64_t row_bytes, const uint8_t* __restrict__ in, int64_t ld_in_bytes,
                              const int64_t* __restrict__ perm, uint8_t* __restrict__ out, int64_t ld_out_bytes) {
  const int64_t r = blockIdx.y;
  if (r >= K) return;
  const uint8_t* src = in + perm[r] * ld_in_bytes;
  uint8_t* dst = out + r * ld_out_bytes;
  const bool wide = ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst) | row_bytes) & 15) == 0;
  if (wide) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < row_bytes / 16;
         i += (int64_t)gridDim.x * blockDim.x)
      reinterpret_cast<uint4*>(dst)[i] = reinterpret_cast<const uint4*>(src)[i];
  } else {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < row_bytes / 2;
         i += (int64_t)gridDim.x * blockDim.x)
      reinterpret_cast<uint16_t*>(dst)[i] = reinterpret_cast<const uint16_t*>(src)[i];
  }
}

inline unsigned nblk(int64_t n, int t) {
  const int64_t g = (n + t - 1) / t;
  return (unsigned)(g < 1 ? 1 : g);
}

// merge_rows (utilities.cpp:145-173) on host column lists: not a plain
// union — when lower_bound reaches the end of A the loop stops, dropping A's
// remaining elements, and B's tail is appended.
std::vector<int32_t> merge_rows(const std::vector<int32_t>& A, const int32_t* B, int64_t nB) {
  std::vector<int32_t> out;
  out.reserve(A.size() + (size_t)nB);
  auto i = A.begin();
  int64_t j = 0;
  while (j < nB) {
    const int32_t bv = B[j];
    auto ni = std::lower_bound(i, A.end(), bv);
    if (ni == A.end()) break;
    out.insert(out.end(), i, ni);
    out.push_back(bv);
    if (*ni == bv) ++ni;
    i = ni;
    j++;
  }
  out.insert(out.end(), B + j, B + nB);
  return out;
}

// Fixed-size node pool for the top-up set.  The set keeps libstdc++'s
// red-black tree algorithm (so its shape, which the truncation depends on,
// is unchanged); only node allocation skips malloc.
struct NodePool {
  std::vector<void*> free_;
  std::vector<std::vector<char>> chunks_;
  size_t size_ = 0;
  void* get(size_t sz) {
    if (size_ == 0) size_ = sz;
    if (sz != size_) return ::operator new(sz);
    if (free_.empty()) {
      chunks_.emplace_back(size_ * 4096);
      char* b = chunks_.back().data();
      for (size_t k = 0; k < 4096; k++) free_.push_back(b + k * size_);
    }
    void* p = free_.back();
    free_.pop_back();
    return p;
  }
  void put(void* p, size_t sz) {
    if (sz != size_) ::operator delete(p);
    else free_.push_back(p);
  }
};

template <class T>
struct PoolAlloc {
  using value_type = T;
  NodePool* pool;
  explicit PoolAlloc(NodePool* p) : pool(p) {}
  template <class U>
  PoolAlloc(const PoolAlloc<U>& o) : pool(o.pool) {}
  T* allocate(size_t n) { return static_cast<T*>(pool->get(n * sizeof(T))); }
  void deallocate(T* p, size_t n) { pool->put(p, n * sizeof(T)); }
  template <class U>
  bool operator==(const PoolAlloc<U>& o) const { return pool == o.pool; }
  template <class U>
  bool operator!=(const PoolAlloc<U>& o) const { return pool != o.pool; }
};

// The Keeper's top-up set is a libstdc++ std::set, and its truncation
// (blocking.cpp:505-507: it = end(); advance(it, keep); erase(it, end()))
// walks forward from end(), which depends on the red-black tree's shape.
// These helpers read that shape through the iterator's node (libstdc++'s
// _Rb_tree_node_base: _M_color, _M_left), the library the reference is built
// with and whose tree the replay reproduces.
template <class It>
inline bool node_is_red(It it) {
  return it._M_node->_M_color == std::_S_red;
}

// Same effect as the reference's truncation, in O(1) instead of keep steps
// (half of them climbs from the largest node to the root).  ++end() is the
// leftmost node of the rightmost node R's subtree: R itself, or R's left
// child L, which is then a red leaf (R has no right child).  The walk from
// end() therefore cycles end, R (period 2) or end, L, R (period 3), so
// advance(end(), keep) stops at R for odd keep or end() for even keep without
// L, and at L / R / end() for keep % 3 == 1 / 2 / 0 with L.  erase(it, end())
// removes the elements from there on, smaller first (when `it` is begin()
// the range erase clears the set instead, with the same, empty, result).
template <class Set>
inline void truncate_from_end(Set& best, size_t keep) {
  auto R = std::prev(best.end());
  if (R._M_node->_M_left != nullptr) {
    const size_t m = keep % 3;
    if (m == 1) {
      auto L = std::prev(R);
      best.erase(L);
      best.erase(R);
    } else if (m == 2) {
      best.erase(R);
    }
  } else if (keep & 1) {
    best.erase(R);
  }
}

// ---- VBR blocking statistics (BlockingEngine::CollectBlockingInfo) --------
// A block row is the set of rows sharing one grouping value g; its height
// h[g] comes from a histogram over the values (which lie in [0, 2 rows) for
// GetGrouping and get_fixed_size_grouping output).  Its nonzero blocks are
// the distinct keys g * nbc + column block among its entries: one radix sort
// of the per-entry keys, then the heads of equal runs.  Every figure is an
// integer sum or maximum, so the result does not depend on execution order.
__global__ void k_group_heights(int64_t rows, const int64_t* __restrict__ grouping, int32_t* __restrict__ h,
                                int32_t* __restrict__ bad) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= rows) return;
  const int64_t g = grouping[r];
  if (g < 0 || g >= 2 * rows) {
    atomicOr(bad, 1);
    return;
  }
  atomicAdd(&h[g], 1);
}

__global__ void k_block_keys(int64_t rows, int64_t nnz, int64_t cols, const int32_t* __restrict__ rowptr,
                             const int32_t* __restrict__ colidx, const int64_t* __restrict__ grouping, int64_t nbc,
                             int64_t col_block, uint64_t* __restrict__ keys, int32_t* __restrict__ bad) {
  const int64_t nz = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (nz >= nnz) return;
  // row of entry nz: the last r with rowptr[r] <= nz (rowptr[0] = 0 <= nz < rowptr[rows] = nnz)
  int64_t lo = 0, hi = rows;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (rowptr[mid] <= nz) lo = mid;
    else hi = mid;
  }
  const int64_t c = colidx[nz];
  const int64_t g = grouping[lo];
  if (c < 0 || c >= cols) atomicOr(bad, 2);
  const bool ok = c >= 0 && c < cols && g >= 0 && g < 2 * rows;
  keys[nz] = ok ? (uint64_t)g * (uint64_t)nbc + (uint64_t)(c / col_block) : ~0ull;  // invalid keys sort last
}

__device__ inline int64_t wave_sum(int64_t v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// sums: [0] VBR_nzcount, [1] VBR_nzblocks_count, [2] total block height
__global__ void k_vbr_heads(int64_t n, const uint64_t* __restrict__ keys, int64_t nbc, int64_t col_block,
                            int64_t cols, int64_t rows, const int32_t* __restrict__ h, int32_t* __restrict__ cntb,
                            unsigned long long* __restrict__ sums) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t nzc = 0, nzb = 0, th = 0;
  if (i < n && keys[i] != ~0ull && (i == 0 || keys[i] != keys[i - 1])) {
    const int64_t g = (int64_t)(keys[i] / (uint64_t)nbc);
    const int64_t cb = (int64_t)(keys[i] % (uint64_t)nbc);
    const int64_t hh = h[g];
    nzb = 1;
    th = hh;
    nzc = col_block * hh;
    if (cb == nbc - 1 && cols % col_block != 0) nzc -= hh * (col_block - cols % col_block);  // short last block
    atomicAdd(&cntb[g], 1);
  }
  nzc = wave_sum(nzc);
  nzb = wave_sum(nzb);
  th = wave_sum(th);
  if ((threadIdx.x & 63) == 0 && (nzb | nzc | th) != 0) {
    atomicAdd(&sums[0], (unsigned long long)nzc);
    atomicAdd(&sums[1], (unsigned long long)nzb);
    atomicAdd(&sums[2], (unsigned long long)th);
  }
}

// Locality order key of each row: (63 - floor(log2 length)) above the first
// stored column above the second one (`cols` where the row has fewer); the
// second column is left out when three fields would not fit in 64 bits.
__global__ void k_locality_keys(int64_t rows, int64_t cols, const int32_t* __restrict__ rowptr,
                                const int32_t* __restrict__ colidx, int cbits, int with_second,
                                uint64_t* __restrict__ keys, int64_t* __restrict__ ids) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= rows) return;
  const int32_t b = rowptr[r], e = rowptr[r + 1];
  const int64_t len = e - b;
  const uint64_t lb = len > 1 ? (uint64_t)(63 - __clzll((unsigned long long)len)) : 0ull;
  const uint64_t c0 = len > 0 ? (uint64_t)(uint32_t)colidx[b] : (uint64_t)cols;
  const uint64_t c1 = len > 1 ? (uint64_t)(uint32_t)colidx[b + 1] : (uint64_t)cols;
  keys[r] = with_second ? ((63ull - lb) << (2 * cbits)) | (c0 << cbits) | c1 : ((63ull - lb) << cbits) | c0;
  ids[r] = r;
}

__global__ void k_max_i32(int64_t n, const int32_t* __restrict__ v, unsigned int* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int32_t m = i < n ? v[i] : 0;
  for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_down(m, o, 64));
  if ((threadIdx.x & 63) == 0 && m > 0) atomicMax(out, (unsigned int)m);
}

}  // namespace

extern "C" {

smat_status smat_reorder_grouping(int64_t rows, const int32_t* rowptr, const int32_t* colidx,
                                  const int32_t* rowptr_host, const int32_t* colidx_host, float tau,
                                  int64_t col_block, int64_t max_rows, int64_t* grouping_host,
                                  int64_t* counters_host, smat_stream stream) {
  SMAT_REQUIRE(rows > 0 && rowptr && colidx && rowptr_host && colidx_host && grouping_host,
               "smat_reorder_grouping: null argument or empty matrix");
  SMAT_REQUIRE(rows < INT32_MAX, "smat_reorder_grouping: too many rows");
  SMAT_REQUIRE(col_block > 0 && max_rows > 0, "smat_reorder_grouping: col_block and max_rows must be positive");
  SMAT_REQUIRE(tau >= 0.f, "smat_reorder_grouping: tau must be >= 0");
  hipStream_t s = smat::to_hip(stream);
  const int64_t nb_max = (rows + kScanThreads - 1) / kScanThreads;
  smat::DevBuf d, grouped, status, dist, lpos, bcount, boff, first, pat;
  SMAT_TRY(d.alloc(rows * 4));
  SMAT_TRY(grouped.alloc(rows));
  SMAT_TRY(status.alloc(rows));
  SMAT_TRY(dist.alloc(rows * 4));
  SMAT_TRY(lpos.alloc(rows * 4));
  SMAT_TRY(bcount.alloc(nb_max * 4));
  SMAT_TRY(boff.alloc(nb_max * 4));
  SMAT_TRY(first.alloc(8));
  // Pinned, device-mapped host buffers: the round kernels write the compared
  // list and its header straight into host memory (one sync per round);
  // marks hold the rows to flag as grouped (each row is written once).
  // Two result areas: while the host replays round r from one, the GPU
  // already runs round r + 1 into the other.
  smat::HostBuf h_j[2], h_d[2], h_hdr, h_marks, h_pat;
  for (int b = 0; b < 2; b++) {
    SMAT_TRY(h_j[b].alloc(rows * 4));
    SMAT_TRY(h_d[b].alloc(rows * 4));
  }
  SMAT_TRY(h_hdr.alloc(32));
  SMAT_TRY(h_marks.alloc(rows * 8));
  // rows longer than kLongRow, ascending: their distances take a wave each
  std::vector<int64_t> long_rows;
  for (int64_t r = 0; r < rows; r++)
    if (rowptr_host[r + 1] - rowptr_host[r] > kLongRow) long_rows.push_back(r);
  smat::DevBuf d_long;
  SMAT_TRY(d_long.alloc((long_rows.size() + 1) * 8));
  if (!long_rows.empty())
    SMAT_HIP_TRY(hipMemcpyAsync(d_long.p, long_rows.data(), long_rows.size() * 8, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_init, nblk(rows, 256), 256, 0, s, rows, d.as<float>(), grouped.as<uint8_t>());
  hipLaunchKernelGGL(k_reset_first, 1, 1, 0, s, first.as<unsigned long long>());
  SMAT_LAUNCH_CHECK();

  volatile int64_t* hdr = h_hdr.as<int64_t>();
  int64_t* marks = h_marks.as<int64_t>();
  int64_t marks_used = 0;
  auto mark = [&](const int64_t* list, int32_t n) {
    std::copy(list, list + n, marks + marks_used);
    hipLaunchKernelGGL(k_mark, nblk(n, 64), 64, 0, s, h_marks.dev<int64_t>() + marks_used, n,
                       grouped.as<uint8_t>());
    marks_used += n;
  };

  std::vector<int64_t> grouping((size_t)rows, -1);
  int64_t comparisons = 0, merges = 0;
  size_t pat_cap = 0;
  NodePool pool;
  using BestSet = std::set<std::pair<float, int64_t>, std::less<std::pair<float, int64_t>>,
                           PoolAlloc<std::pair<float, int64_t>>>;

  for (int64_t i = 0; i < rows; i++) {
    if (grouping[i] != -1) continue;
    const int64_t group = i + rows;
    BestSet best{std::less<std::pair<float, int64_t>>(), PoolAlloc<std::pair<float, int64_t>>(&pool)};
    std::vector<int64_t> merged{i};
    std::vector<int32_t> pattern(colidx_host + rowptr_host[i], colidx_host + rowptr_host[i + 1]);
    int64_t gsize = 1;
    grouping[i] = group;
    mark(&i, 1);
    bool pattern_dirty = true;
    int64_t n_blk_pat = 0;  // distinct column blocks of the pattern
    // One round: every candidate of [j0, rows) against the current pattern,
    // results into area b.
    auto launch_round = [&](int b, int64_t j0) -> smat_status {
      if (pattern_dirty) {
        // the pattern goes through a pinned staging area so the copy is asynchronous
        if (pattern.size() + 1 > pat_cap) {
          pat_cap = 2 * (pattern.size() + 1);
          SMAT_TRY(pat.alloc(pat_cap * 4));
          SMAT_TRY(h_pat.alloc(pat_cap * 4));
        }
        std::copy(pattern.begin(), pattern.end(), h_pat.as<int32_t>());
        if (!pattern.empty())
          SMAT_HIP_TRY(hipMemcpyAsync(pat.p, h_pat.p, pattern.size() * 4, hipMemcpyHostToDevice, s));
        n_blk_pat = 0;
        for (size_t q = 0; q < pattern.size(); q++)
          if (q == 0 || pattern[q] / col_block != pattern[q - 1] / col_block) n_blk_pat++;
        pattern_dirty = false;
      }
      const int64_t n = rows - j0;
      const int32_t nb = (int32_t)((n + kScanThreads - 1) / kScanThreads);
      hipLaunchKernelGGL(k_round_scan, nb, kScanThreads, 0, s, rows, i, j0, rowptr, colidx, pat.as<int32_t>(),
                         (int32_t)pattern.size(), gsize, tau, col_block, d.as<float>(), grouped.as<uint8_t>(),
                         status.as<uint8_t>(), dist.as<float>(), lpos.as<int32_t>(), bcount.as<int32_t>(),
                         first.as<unsigned long long>());
      const int64_t lb = std::lower_bound(long_rows.begin(), long_rows.end(), j0) - long_rows.begin();
      const int64_t n_long = (int64_t)long_rows.size() - lb;
      if (n_long > 0)
        hipLaunchKernelGGL(k_round_long, nblk(n_long, 4), 256, 0, s, d_long.as<int64_t>() + lb, n_long, rowptr,
                           colidx, pat.as<int32_t>(), (int32_t)pattern.size(), n_blk_pat, gsize, tau, col_block,
                           status.as<uint8_t>(), dist.as<float>(), first.as<unsigned long long>());
      hipLaunchKernelGGL(k_round_offsets, 1, 1024, 0, s, nb, bcount.as<int32_t>(), boff.as<int32_t>(),
                         first.as<unsigned long long>(), h_hdr.dev<int64_t>() + 2 * b);
      hipLaunchKernelGGL(k_round_write, nb, kScanThreads, 0, s, rows, j0, status.as<uint8_t>(), dist.as<float>(),
                         lpos.as<int32_t>(), boff.as<int32_t>(), h_j[b].dev<int32_t>(), h_d[b].dev<float>());
      SMAT_LAUNCH_CHECK();
      return SMAT_OK;
    };
    int64_t j0 = i + 1;
    int b = 0;
    bool launched = false;
    if (j0 < rows && gsize != max_rows) {
      SMAT_TRY(launch_round(b, j0));
      launched = true;
    }
    while (launched) {
      SMAT_HIP_TRY(hipStreamSynchronize(s));
      const int64_t ncmp = hdr[2 * b];
      const int64_t jm = hdr[2 * b + 1];
      // the truncation bound of this round's candidates: before its merge
      const size_t keep = (size_t)(max_rows - (int64_t)merged.size());
      const int64_t jend = jm == INT64_MAX ? rows : jm + 1;
      hipLaunchKernelGGL(k_commit, nblk(jend - j0, 256), 256, 0, s, j0, jend, status.as<uint8_t>(),
                         dist.as<float>(), d.as<float>());
      SMAT_LAUNCH_CHECK();
      launched = false;
      if (jm != INT64_MAX) {  // merge row jm, then start the next round at once
        comparisons++;
        merges++;
        grouping[jm] = group;
        merged.push_back(jm);
        pattern = merge_rows(pattern, colidx_host + rowptr_host[jm], rowptr_host[jm + 1] - rowptr_host[jm]);
        pattern_dirty = true;
        gsize++;
        mark(&jm, 1);
        if (jm + 1 < rows && gsize != max_rows) {
          SMAT_TRY(launch_round(1 - b, jm + 1));
          launched = true;
        }
      }
      // replay this round's compared candidates strictly before the merge (or
      // all of them) while the GPU works on the next round
      const int32_t* oj = h_j[b].as<int32_t>();
      const float* od = h_d[b].as<float>();
      for (int64_t q = 0; q < ncmp && oj[q] < jm; q++) {
        comparisons++;
        const std::pair<float, int64_t> key(od[q], (int64_t)oj[q]);
        if (!best.empty() && *best.rbegin() < key) {
          // A key above the current maximum becomes the red right child of
          // the rightmost node R.  When R is black no rebalancing follows;
          // with keep odd and the set over keep the truncation below then
          // erases exactly that new leaf (it has no left child), again
          // without rebalancing: the tree ends as it started, so the pair is
          // skipped.  Otherwise the key is inserted with the end() hint,
          // which only skips the descent (the node lands in the same place).
          if ((keep & 1) && best.size() + 1 > keep && !node_is_red(std::prev(best.end()))) continue;
          best.insert(best.end(), key);
        } else {
          best.insert(key);
        }
        if (best.size() > keep) truncate_from_end(best, keep);  // blocking.cpp:505-507
      }
      j0 = jend;
      b = 1 - b;
    }
    // top-up from the candidate set (blocking.cpp:513-521)
    std::vector<int64_t> topped;
    if (gsize < max_rows) {
      for (auto it = best.begin(); it != best.end() && gsize != max_rows; ++it) {
        grouping[it->second] = group;
        merged.push_back(it->second);
        topped.push_back(it->second);
        gsize++;
      }
    }
    if (gsize == max_rows)
      for (int64_t idx : merged) grouping[idx] -= rows;  // complete groups first
    if (!topped.empty()) mark(topped.data(), (int32_t)topped.size());
  }
  SMAT_HIP_TRY(hipStreamSynchronize(s));
  std::copy(grouping.begin(), grouping.end(), grouping_host);
  if (counters_host) {
    counters_host[0] = comparisons;
    counters_host[1] = merges;
  }
  return SMAT_OK;
}

// get_permutation (utilities.cpp:8-20): iota stably-unspecified sorted by key
// with std::sort and an int comparator, exactly as the reference does.
smat_status smat_grouping_to_permutation(int64_t rows, const int64_t* grouping_host, int64_t* perm_host) {
  SMAT_REQUIRE(rows >= 0 && grouping_host && perm_host, "smat_grouping_to_permutation: null argument");
  std::vector<int64_t> v((size_t)rows);
  std::iota(v.begin(), v.end(), 0);
  auto comp = [&](int a, int b) { return grouping_host[a] < grouping_host[b]; };
  std::sort(v.begin(), v.end(), comp);
  std::copy(v.begin(), v.end(), perm_host);
  return SMAT_OK;
}

// get_fixed_size_grouping (utilities.cpp:45-54).
smat_status smat_fixed_size_grouping(int64_t rows, const int64_t* grouping_host, int64_t row_block,
                                     int64_t* out_host) {
  SMAT_REQUIRE(rows >= 0 && grouping_host && out_host && row_block > 0, "smat_fixed_size_grouping: bad argument");
  std::vector<int64_t> perm((size_t)rows);
  smat_grouping_to_permutation(rows, grouping_host, perm.data());
  for (int64_t i = 0; i < rows; i++) out_host[perm[i]] = i / row_block;
  return SMAT_OK;
}

// CSR::reorder / permute_rows (csr.cpp:65-74, 97-105) on the GPU: new row r is
// old row perm[r]; the output arrays are caller-allocated (nnz entries).
smat_status smat_permute_csr_rows(int64_t rows, const int32_t* rowptr, const int32_t* colidx,
                                  const uint16_t* values, const int64_t* perm, int32_t* out_rowptr,
                                  int32_t* out_colidx, uint16_t* out_values, smat_stream stream) {
  SMAT_REQUIRE(rows > 0 && rowptr && colidx && perm && out_rowptr && out_colidx,
               "smat_permute_csr_rows: null argument");
  hipStream_t s = smat::to_hip(stream);
  smat::DevBuf len, tmp;
  SMAT_TRY(len.alloc((rows + 1) * 4));
  hipLaunchKernelGGL(k_perm_lengths, nblk(rows + 1, 256), 256, 0, s, rows, perm, rowptr, len.as<int32_t>());
  size_t tb = 0;
  SMAT_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, len.as<int32_t>(), out_rowptr, (int)(rows + 1), s));
  SMAT_TRY(tmp.alloc(tb));
  SMAT_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, len.as<int32_t>(), out_rowptr, (int)(rows + 1), s));
  hipLaunchKernelGGL(k_perm_gather, nblk(rows, 4), 256, 0, s, rows, perm, rowptr, colidx, values, out_rowptr,
                     out_colidx, out_values);
  SMAT_LAUNCH_CHECK();
  SMAT_HIP_TRY(hipStreamSynchronize(s));
  return SMAT_OK;
}

// Locality row order (no reference counterpart): rows sorted by descending
// log2 length, then by their first and second stored columns, ties in row
// order (a stable radix sort).  perm[r] = old row of new row r.
smat_status smat_locality_order(int64_t rows, int64_t cols, const int32_t* rowptr, const int32_t* colidx,
                                int64_t* perm, smat_stream stream) {
  SMAT_REQUIRE(rows > 0 && cols > 0 && rowptr && colidx && perm, "smat_locality_order: bad argument");
  SMAT_REQUIRE(rows <= INT32_MAX && cols <= INT32_MAX, "smat_locality_order: too many rows or columns");
  hipStream_t s = smat::to_hip(stream);
  int cbits = 1;
  while (cbits < 32 && (cols >> cbits) != 0) cbits++;  // bits for 0..cols
  const int with_second = 6 + 2 * cbits <= 64;
  const int end_bit = 6 + (with_second ? 2 : 1) * cbits;
  smat::DevBuf keys, keys2, ids, tmp;
  SMAT_TRY(keys.alloc(rows * 8));
  SMAT_TRY(keys2.alloc(rows * 8));
  SMAT_TRY(ids.alloc(rows * 8));
  hipLaunchKernelGGL(k_locality_keys, nblk(rows, 256), 256, 0, s, rows, cols, rowptr, colidx, cbits, with_second,
                     keys.as<uint64_t>(), ids.as<int64_t>());
  SMAT_LAUNCH_CHECK();
  size_t tb = 0;
  SMAT_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys.as<uint64_t>(), keys2.as<uint64_t>(),
                                                  ids.as<int64_t>(), perm, (int)rows, 0, end_bit, s));
  SMAT_TRY(tmp.alloc(tb));
  SMAT_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, keys.as<uint64_t>(), keys2.as<uint64_t>(),
                                                  ids.as<int64_t>(), perm, (int)rows, 0, end_bit, s));
  SMAT_HIP_TRY(hipStreamSynchronize(s));
  return SMAT_OK;
}

// Column pass of the reference's reorder driver (src/reorder.py:143-183):
// transpose the pattern (transpose_el_matrix sorts the swapped edges, so each
// transposed row lists its original rows in ascending order), run the same
// Keeper grouping on it (Matrix_Blocking -a 5 -F 1), and turn the grouping
// into the order of the columns.
smat_status smat_reorder_columns(int64_t rows, int64_t cols, const int32_t* rowptr_host, const int32_t* colidx_host,
                                 float tau, int64_t col_block, int64_t max_rows, int64_t* col_perm_host,
                                 int64_t* counters_host, smat_stream stream) {
  SMAT_REQUIRE(rows > 0 && cols > 0 && rowptr_host && colidx_host && col_perm_host,
               "smat_reorder_columns: bad argument");
  SMAT_REQUIRE(cols < INT32_MAX, "smat_reorder_columns: too many columns");
  const int64_t nnz = rowptr_host[rows];
  std::vector<int32_t> trp((size_t)cols + 1, 0), tci((size_t)(nnz > 0 ? nnz : 1));
  for (int64_t q = 0; q < nnz; q++) {
    SMAT_REQUIRE(colidx_host[q] >= 0 && colidx_host[q] < cols, "smat_reorder_columns: column out of range");
    trp[(size_t)colidx_host[q] + 1]++;
  }
  for (int64_t c = 0; c < cols; c++) trp[(size_t)c + 1] += trp[(size_t)c];
  std::vector<int32_t> fill(trp.begin(), trp.end() - 1);
  for (int64_t r = 0; r < rows; r++)  // rows in ascending order: each transposed row comes out sorted
    for (int32_t q = rowptr_host[r]; q < rowptr_host[r + 1]; q++) tci[(size_t)fill[(size_t)colidx_host[q]]++] = (int32_t)r;
  hipStream_t s = smat::to_hip(stream);
  smat::DevBuf d_rp, d_ci;
  SMAT_TRY(d_rp.alloc(trp.size() * 4));
  SMAT_TRY(d_ci.alloc(tci.size() * 4));
  SMAT_HIP_TRY(hipMemcpyAsync(d_rp.p, trp.data(), trp.size() * 4, hipMemcpyHostToDevice, s));
  SMAT_HIP_TRY(hipMemcpyAsync(d_ci.p, tci.data(), tci.size() * 4, hipMemcpyHostToDevice, s));
  std::vector<int64_t> g((size_t)cols), fixed((size_t)cols);
  SMAT_TRY(smat_reorder_grouping(cols, d_rp.as<int32_t>(), d_ci.as<int32_t>(), trp.data(), tci.data(), tau, col_block,
                                 max_rows, g.data(), counters_host, stream));
  SMAT_TRY(smat_fixed_size_grouping(cols, g.data(), max_rows, fixed.data()));
  return smat_grouping_to_permutation(cols, fixed.data(), col_perm_host);
}

smat_status smat_permute_csr_cols(int64_t nnz, const int32_t* colidx, const int64_t* inv_perm, int32_t* out_colidx,
                                  smat_stream stream) {
  SMAT_REQUIRE(nnz >= 0 && (nnz == 0 || (colidx && inv_perm && out_colidx)), "smat_permute_csr_cols: bad argument");
  if (nnz == 0) return SMAT_OK;
  hipLaunchKernelGGL(k_relabel_cols, nblk(nnz, 256), 256, 0, smat::to_hip(stream), nnz, colidx, inv_perm, out_colidx);
  SMAT_LAUNCH_CHECK();
  return SMAT_OK;
}

smat_status smat_permute_dense_rows(int64_t K, int64_t N, int32_t elem_bytes, const void* in, int64_t ld_in,
                                    const int64_t* perm, void* out, int64_t ld_out, smat_stream stream) {
  SMAT_REQUIRE(K > 0 && N > 0 && in && perm && out && ld_in >= N && ld_out >= N &&
                   (elem_bytes == 2 || elem_bytes == 4),
               "smat_permute_dense_rows: bad argument");
  SMAT_REQUIRE(K < 65536 * 32768LL, "smat_permute_dense_rows: too many rows");
  const int64_t row_bytes = N * elem_bytes;
  const unsigned gx = (unsigned)std::min<int64_t>(8, (row_bytes / 16 + 255) / 256 + 1);
  // one row per grid.y index, in slabs of 65535 rows (the grid.y limit)
  hipStream_t s = smat::to_hip(stream);
  const int64_t step = 65535;
  for (int64_t r0 = 0; r0 < K; r0 += step) {
    const int64_t nr = std::min(step, K - r0);
    hipLaunchKernelGGL(k_gather_rows, dim3(gx, (unsigned)nr), 256, 0, s, nr, row_bytes,
                       static_cast<const uint8_t*>(in), ld_in * elem_bytes, perm + r0,
                       static_cast<uint8_t*>(out) + r0 * ld_out * elem_bytes, ld_out * elem_bytes);
  }
  SMAT_LAUNCH_CHECK();
  return SMAT_OK;
}

// BlockingEngine::CollectBlockingInfo (blocking.cpp:562-614) with the row
// partition of get_partition (utilities.cpp:23-43): out[0] = VBR_nzcount,
// out[1] = VBR_nzblocks_count, out[2] = VBR_longest_row; *avg_height =
// VBR_average_height, (float)total height / (float)blocks as the reference
// divides (NaN with no blocks).  Column blocks are counted with the exact
// ceil(cols / col_block); the reference's float ceil agrees below 2^24 columns.
smat_status smat_blocking_info(int64_t rows, int64_t cols, int64_t nnz, const int32_t* rowptr, const int32_t* colidx,
                               const int64_t* grouping, int64_t col_block, int64_t* out_host, float* avg_height_host,
                               smat_stream stream) {
  SMAT_REQUIRE(rows > 0 && cols > 0 && nnz >= 0 && col_block > 0 && rowptr && grouping && out_host &&
                   avg_height_host && (nnz == 0 || colidx),
               "smat_blocking_info: bad argument");
  SMAT_REQUIRE(2 * rows < (int64_t)INT32_MAX && nnz < (int64_t)INT32_MAX, "smat_blocking_info: too large");
  const int64_t nbc = (cols + col_block - 1) / col_block;
  SMAT_REQUIRE(nbc <= (int64_t)(~0ull >> 2) / (2 * rows), "smat_blocking_info: rows x column blocks too large");
  hipStream_t s = smat::to_hip(stream);
  const int64_t ng = 2 * rows;
  smat::DevBuf h, cntb, flags, sums, keys, keys2, tmp;
  SMAT_TRY(h.alloc(ng * 4));
  SMAT_TRY(cntb.alloc(ng * 4));
  SMAT_TRY(flags.alloc(64));
  SMAT_TRY(sums.alloc(64));
  SMAT_HIP_TRY(hipMemsetAsync(h.p, 0, ng * 4, s));
  SMAT_HIP_TRY(hipMemsetAsync(cntb.p, 0, ng * 4, s));
  SMAT_HIP_TRY(hipMemsetAsync(flags.p, 0, 64, s));
  SMAT_HIP_TRY(hipMemsetAsync(sums.p, 0, 64, s));
  hipLaunchKernelGGL(k_group_heights, nblk(rows, 256), 256, 0, s, rows, grouping, h.as<int32_t>(),
                     flags.as<int32_t>());
  SMAT_LAUNCH_CHECK();
  if (nnz > 0) {
    SMAT_TRY(keys.alloc(nnz * 8));
    SMAT_TRY(keys2.alloc(nnz * 8));
    hipLaunchKernelGGL(k_block_keys, nblk(nnz, 256), 256, 0, s, rows, nnz, cols, rowptr, colidx, grouping, nbc,
                       col_block, keys.as<uint64_t>(), flags.as<int32_t>());
    SMAT_LAUNCH_CHECK();
    size_t tb = 0;
    SMAT_HIP_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, keys.as<uint64_t>(), keys2.as<uint64_t>(), (int)nnz,
                                                   0, 64, s));
    SMAT_TRY(tmp.alloc(tb));
    SMAT_HIP_TRY(hipcub::DeviceRadixSort::SortKeys(tmp.p, tb, keys.as<uint64_t>(), keys2.as<uint64_t>(), (int)nnz, 0,
                                                   64, s));
    hipLaunchKernelGGL(k_vbr_heads, nblk(nnz, 256), 256, 0, s, nnz, keys2.as<uint64_t>(), nbc, col_block, cols,
                       rows, h.as<int32_t>(), cntb.as<int32_t>(), sums.as<unsigned long long>());
    hipLaunchKernelGGL(k_max_i32, nblk(ng, 256), 256, 0, s, ng, cntb.as<int32_t>(),
                       reinterpret_cast<unsigned int*>(sums.as<unsigned long long>() + 3));
    SMAT_LAUNCH_CHECK();
  }
  unsigned long long hs[4] = {0, 0, 0, 0};
  int32_t hf = 0;
  SMAT_HIP_TRY(hipMemcpyAsync(hs, sums.p, sizeof hs, hipMemcpyDeviceToHost, s));
  SMAT_HIP_TRY(hipMemcpyAsync(&hf, flags.p, 4, hipMemcpyDeviceToHost, s));
  SMAT_HIP_TRY(hipStreamSynchronize(s));
  SMAT_REQUIRE(!(hf & 1), "smat_blocking_info: grouping value outside [0, 2 * rows)");
  SMAT_REQUIRE(!(hf & 2), "smat_blocking_info: column index outside [0, cols)");
  out_host[0] = (int64_t)hs[0];
  out_host[1] = (int64_t)hs[1];
  out_host[2] = (int64_t)(hs[3] & 0xffffffffull);
  *avg_height_host = (float)(int64_t)hs[2] / (float)(int64_t)hs[1];
  return SMAT_OK;
}

}  // extern "C"
