// CSR -> BCSR on the GPU, block tagging, dense-grid maps and 2:4/dense list
// split.  Integer work only; bit-identical to the reference's host code
// (SparseMatrix::csrToBcsr, src/cuda_hgemm/src/common/matrix.h:666-792, and
// filterBcsrBlocks, matrix.h:1140-1288) without its dense block grid:
//
//   1. every CSR entry gets the 64-bit key (row/16)*ncb + col/bk, where ncb
//      is the number of block columns, and a flag for the reference's dense
//      rule (three consecutive CSR entries c-2, c-1, c with c%4 in {2,3});
//   2. a radix sort of (key, entry) groups entries by block in row-major
//      block order — the order csrToBcsr assigns relative block indices;
//   3. a head-flag scan numbers the blocks; row pointers come from the block
//      rows of the heads, column starts from the keys;
//   4. values are scattered to block*16*bk + (row%16)*bk + col%bk.
#include <hipcub/hipcub.hpp>

#include "smat_internal.h"

struct smat_bcsr_t {
  smat_bcsr_desc d{};
  smat::DevBuf rowptr, colidx, values, tags;
};

namespace {

constexpr int kThreads = 256;

inline unsigned grid_for(int64_t n) {
  int64_t g = (n + kThreads - 1) / kThreads;
  if (g < 1) g = 1;
  return (unsigned)g;
}

__device__ inline int64_t find_row(const int32_t* __restrict__ rowptr, int64_t M, int64_t e) {
  // largest r with rowptr[r] <= e, r in [0, M)
  int64_t lo = 0, hi = M;
  while (hi - lo > 1) {
    int64_t mid = (lo + hi) >> 1;
    if (rowptr[mid] <= e) lo = mid; else hi = mid;
  }
  return lo;
}

__global__ void k_entry_keys(int64_t M, int64_t nnz, int64_t ncb, int bk,
                             const int32_t* __restrict__ rowptr,
                             const int32_t* __restrict__ colidx,
                             uint64_t* __restrict__ keys, int32_t* __restrict__ idx,
                             uint8_t* __restrict__ dense_flag) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  int64_t row = find_row(rowptr, M, e);
  int32_t col = colidx[e];
  keys[e] = (uint64_t)(row >> 4) * (uint64_t)ncb + (uint64_t)(col / bk);
  idx[e] = (int32_t)e;
  int r4 = col & 3;
  uint8_t f = 0;
  if ((r4 == 2 || r4 == 3) && e >= (int64_t)rowptr[row] + 2 && colidx[e - 1] == col - 1 &&
      colidx[e - 2] == col - 2)
    f = 1;
  dense_flag[e] = f;
}

__global__ void k_heads(int64_t n, const uint64_t* __restrict__ sk, int32_t* __restrict__ head) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  head[i] = (i == 0 || sk[i] != sk[i - 1]) ? 1 : 0;
}

// For each head: block column start and block row of block b = blkid-1.
__global__ void k_block_meta(int64_t n, int64_t ncb, int bk, const uint64_t* __restrict__ sk,
                             const int32_t* __restrict__ head, const int32_t* __restrict__ blkid,
                             int32_t* __restrict__ colidx_out, int32_t* __restrict__ brow_out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !head[i]) return;
  int64_t b = blkid[i] - 1;
  uint64_t k = sk[i];
  colidx_out[b] = (int32_t)((k % (uint64_t)ncb) * (uint64_t)bk);
  brow_out[b] = (int32_t)(k / (uint64_t)ncb);
}

// rowptr[r] = number of blocks whose block row is < r.
__global__ void k_rowptr(int64_t nnzb, int64_t nbrows, const int32_t* __restrict__ brow,
                         int32_t* __restrict__ rowptr) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nnzb) return;
  int64_t prev = (b == 0) ? -1 : brow[b - 1];
  int64_t cur = (b == nnzb) ? nbrows : brow[b];
  for (int64_t r = prev + 1; r <= cur; r++) rowptr[r] = (int32_t)b;
}

__global__ void k_fill_u16(int64_t n, uint16_t* __restrict__ p, uint16_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_i32(int64_t n, int32_t* __restrict__ p, int32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_scatter(int64_t M, int64_t n, int bk, const int32_t* __restrict__ rowptr,
                          const int32_t* __restrict__ colidx, const uint16_t* __restrict__ val,
                          const int32_t* __restrict__ sidx, const int32_t* __restrict__ blkid,
                          const uint8_t* __restrict__ dense_flag, uint16_t* __restrict__ out_val,
                          int32_t* __restrict__ tags) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t e = sidx[i];
  int64_t b = blkid[i] - 1;
  int64_t row = find_row(rowptr, M, e);
  int32_t col = colidx[e];
  out_val[b * 16 * bk + (row & 15) * bk + (col % bk)] = val[e];
  if (dense_flag[e]) tags[b] = 2;
}

__global__ void k_tag_exact24(int64_t nnzb, int bk, const uint16_t* __restrict__ val,
                              int32_t* __restrict__ tags) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nnzb) return;
  const uint16_t* v = val + b * 16 * bk;
  int t = 1;
  for (int g = 0; g < 16 * bk / 4; g++) {
    int c = ((v[4 * g] & 0x7fff) != 0) + ((v[4 * g + 1] & 0x7fff) != 0) +
            ((v[4 * g + 2] & 0x7fff) != 0) + ((v[4 * g + 3] & 0x7fff) != 0);
    if (c > 2) t = 2;
  }
  tags[b] = t;
}

__global__ void k_grid_maps(int64_t nnzb, int64_t nbrows, int64_t ncb, int bk,
                            const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
                            const int32_t* __restrict__ tags, int32_t* __restrict__ info,
                            int32_t* __restrict__ rel) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nnzb) return;
  int64_t br = find_row(rowptr, nbrows, b);
  int64_t g = br * ncb + colidx[b] / bk;
  info[g] = tags[b];
  rel[g] = (int32_t)b;
}

__global__ void k_flags(int64_t nnzb, const int32_t* __restrict__ tags, int32_t* __restrict__ fs,
                        int32_t* __restrict__ fd) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nnzb) return;
  int t = (b < nnzb) ? tags[b] : 0;
  fs[b] = (t == 1);
  fd[b] = (t == 2);
}

__global__ void k_filter_rowptr(int64_t nbrows, const int32_t* __restrict__ rowptr,
                                const int32_t* __restrict__ ps, const int32_t* __restrict__ pd,
                                int32_t* __restrict__ s_rowptr, int32_t* __restrict__ d_rowptr) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r > nbrows) return;
  int32_t p = rowptr[r];
  if (s_rowptr) s_rowptr[r] = ps[p];
  if (d_rowptr) d_rowptr[r] = pd[p];
}

__global__ void k_filter_copy(int64_t nnzb, int bk, const int32_t* __restrict__ tags,
                              const int32_t* __restrict__ colidx, const uint16_t* __restrict__ val,
                              const int32_t* __restrict__ ps, const int32_t* __restrict__ pd,
                              int32_t* __restrict__ s_colidx, uint16_t* __restrict__ s_val,
                              int32_t* __restrict__ d_colidx, uint16_t* __restrict__ d_val) {
  // one 64-thread group per block; each thread copies 16*bk/64 halves
  int64_t b = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  int lane = threadIdx.x & 63;
  if (b >= nnzb) return;
  int t = tags[b];
  int32_t* oc = (t == 1) ? s_colidx : d_colidx;
  uint16_t* ov = (t == 1) ? s_val : d_val;
  int64_t pos = (t == 1) ? ps[b] : pd[b];
  if (t != 1 && t != 2) return;
  if (oc == nullptr || ov == nullptr) return;
  int bsz = 16 * bk;
  if (lane == 0) oc[pos] = colidx[b];
  for (int i = lane; i < bsz; i += 64) ov[pos * bsz + i] = val[b * bsz + i];
}

}  // namespace

extern "C" {

smat_status smat_bcsr_from_csr(int64_t M, int64_t K, int64_t nnz, const int32_t* csr_rowptr,
                               const int32_t* csr_colidx, const uint16_t* csr_values, int32_t bk,
                               int32_t tag_rule, smat_stream stream, smat_bcsr* out) {
  SMAT_REQUIRE(out != nullptr, "smat_bcsr_from_csr: out is NULL");
  *out = nullptr;
  SMAT_REQUIRE(bk == 1 || bk == 2 || bk == 4 || bk == 8 || bk == 16 || bk == 32,
               "smat_bcsr_from_csr: bk must be 1, 2, 4, 8, 16 or 32 (got %d)", bk);
  SMAT_REQUIRE(M > 0 && M % 16 == 0, "smat_bcsr_from_csr: M=%lld must be a positive multiple of 16",
               (long long)M);
  SMAT_REQUIRE(K > 0 && K % bk == 0, "smat_bcsr_from_csr: K=%lld must be a positive multiple of bk",
               (long long)K);
  SMAT_REQUIRE(nnz >= 0 && nnz < (int64_t)INT32_MAX, "smat_bcsr_from_csr: nnz out of range");
  SMAT_REQUIRE(csr_rowptr && (nnz == 0 || (csr_colidx && csr_values)),
               "smat_bcsr_from_csr: null CSR array");
  SMAT_REQUIRE(tag_rule == SMAT_TAG_REFERENCE || tag_rule == SMAT_TAG_EXACT24,
               "smat_bcsr_from_csr: unknown tag rule %d", tag_rule);
  hipStream_t s = smat::to_hip(stream);
  const int64_t nbrows = M / 16, ncb = K / bk;

  smat_bcsr_t* A = new smat_bcsr_t();
  auto fail = [&](smat_status st) {
    delete A;
    return st;
  };

  smat::DevBuf keys, keys_sorted, idx, idx_sorted, flag, head, blkid, brow, tmp;
  int64_t nnzb = 0;
  if (nnz > 0) {
    smat_status st;
    if ((st = keys.alloc(nnz * 8)) || (st = keys_sorted.alloc(nnz * 8)) || (st = idx.alloc(nnz * 4)) ||
        (st = idx_sorted.alloc(nnz * 4)) || (st = flag.alloc(nnz)) || (st = head.alloc(nnz * 4)) ||
        (st = blkid.alloc(nnz * 4)))
      return fail(st);
    hipLaunchKernelGGL(k_entry_keys, grid_for(nnz), kThreads, 0, s, M, nnz, ncb, (int)bk,
                       csr_rowptr, csr_colidx, keys.as<uint64_t>(), idx.as<int32_t>(),
                       flag.as<uint8_t>());
    if (hipGetLastError() != hipSuccess) return fail(smat::set_error(SMAT_ERR_HIP, "k_entry_keys launch failed"));
    uint64_t max_key = (uint64_t)nbrows * (uint64_t)ncb;
    int end_bit = 1;
    while (end_bit < 64 && (max_key >> end_bit) != 0) end_bit++;
    size_t tbytes = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, tbytes, keys.as<uint64_t>(), keys_sorted.as<uint64_t>(),
                                           idx.as<int32_t>(), idx_sorted.as<int32_t>(), (int)nnz, 0,
                                           end_bit, s) != hipSuccess)
      return fail(smat::set_error(SMAT_ERR_HIP, "radix sort sizing failed"));
    size_t sbytes = 0;
    if (hipcub::DeviceScan::InclusiveSum(nullptr, sbytes, head.as<int32_t>(), blkid.as<int32_t>(), (int)nnz,
                                         s) != hipSuccess)
      return fail(smat::set_error(SMAT_ERR_HIP, "scan sizing failed"));
    if ((st = tmp.alloc(tbytes > sbytes ? tbytes : sbytes))) return fail(st);
    if (hipcub::DeviceRadixSort::SortPairs(tmp.p, tbytes, keys.as<uint64_t>(), keys_sorted.as<uint64_t>(),
                                           idx.as<int32_t>(), idx_sorted.as<int32_t>(), (int)nnz, 0,
                                           end_bit, s) != hipSuccess)
      return fail(smat::set_error(SMAT_ERR_HIP, "radix sort failed"));
    hipLaunchKernelGGL(k_heads, grid_for(nnz), kThreads, 0, s, nnz, keys_sorted.as<uint64_t>(),
                       head.as<int32_t>());
    sbytes = tmp.bytes;
    if (hipcub::DeviceScan::InclusiveSum(tmp.p, sbytes, head.as<int32_t>(), blkid.as<int32_t>(), (int)nnz,
                                         s) != hipSuccess)
      return fail(smat::set_error(SMAT_ERR_HIP, "scan failed"));
    int32_t last = 0;
    if (hipMemcpyAsync(&last, blkid.as<int32_t>() + (nnz - 1), 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return fail(smat::set_error(SMAT_ERR_HIP, "block count readback failed"));
    nnzb = last;
  }
  smat_status st;
  if ((st = A->rowptr.alloc((nbrows + 1) * 4)) || (st = A->colidx.alloc(nnzb * 4)) ||
      (st = A->values.alloc(nnzb * 16 * bk * 2)) || (st = A->tags.alloc(nnzb * 4)) ||
      (st = brow.alloc(nnzb * 4 + 4)))
    return fail(st);
  hipLaunchKernelGGL(k_fill_u16, grid_for(nnzb * 16 * bk), kThreads, 0, s, nnzb * 16 * bk,
                     A->values.as<uint16_t>(), (uint16_t)0);
  // blocks narrower than one 1x4 group cannot be 2:4: tag them dense
  hipLaunchKernelGGL(k_fill_i32, grid_for(nnzb), kThreads, 0, s, nnzb, A->tags.as<int32_t>(), bk >= 4 ? 1 : 2);
  if (nnz > 0) {
    hipLaunchKernelGGL(k_block_meta, grid_for(nnz), kThreads, 0, s, nnz, ncb, (int)bk,
                       keys_sorted.as<uint64_t>(), head.as<int32_t>(), blkid.as<int32_t>(),
                       A->colidx.as<int32_t>(), brow.as<int32_t>());
    hipLaunchKernelGGL(k_scatter, grid_for(nnz), kThreads, 0, s, M, nnz, (int)bk, csr_rowptr,
                       csr_colidx, csr_values, idx_sorted.as<int32_t>(), blkid.as<int32_t>(),
                       flag.as<uint8_t>(), A->values.as<uint16_t>(), A->tags.as<int32_t>());
  }
  hipLaunchKernelGGL(k_rowptr, grid_for(nnzb + 1), kThreads, 0, s, nnzb, nbrows, brow.as<int32_t>(),
                     A->rowptr.as<int32_t>());
  if (tag_rule == SMAT_TAG_EXACT24 && nnzb > 0 && bk >= 4)
    hipLaunchKernelGGL(k_tag_exact24, grid_for(nnzb), kThreads, 0, s, nnzb, (int)bk,
                       A->values.as<uint16_t>(), A->tags.as<int32_t>());
  if (hipGetLastError() != hipSuccess) return fail(smat::set_error(SMAT_ERR_HIP, "bcsr build launch failed"));
  // temporaries are freed on return; make sure the stream is done with them
  if (hipStreamSynchronize(s) != hipSuccess)
    return fail(smat::set_error(SMAT_ERR_HIP, "bcsr build failed on the device"));

  A->d.M = M;
  A->d.K = K;
  A->d.bm = 16;
  A->d.bk = bk;
  A->d.nbrows = nbrows;
  A->d.nnzb = nnzb;
  A->d.rowptr = A->rowptr.as<int32_t>();
  A->d.colidx = A->colidx.as<int32_t>();
  A->d.values = A->values.as<uint16_t>();
  A->d.tags = A->tags.as<int32_t>();
  *out = A;
  return SMAT_OK;
}

smat_status smat_bcsr_describe(smat_bcsr A, smat_bcsr_desc* out) {
  SMAT_REQUIRE(A && out, "smat_bcsr_describe: null argument");
  *out = A->d;
  return SMAT_OK;
}

smat_status smat_bcsr_destroy(smat_bcsr A) {
  delete A;
  return SMAT_OK;
}

smat_status smat_bcsr_retag_exact24(smat_bcsr A, smat_stream stream) {
  SMAT_REQUIRE(A, "smat_bcsr_retag_exact24: null BCSR");
  if (A->d.nnzb == 0 || A->d.bk < 4) return SMAT_OK;
  hipLaunchKernelGGL(k_tag_exact24, grid_for(A->d.nnzb), kThreads, 0, smat::to_hip(stream),
                     A->d.nnzb, A->d.bk, A->d.values, A->d.tags);
  SMAT_LAUNCH_CHECK();
  return SMAT_OK;
}

smat_status smat_bcsr_grid_maps(const smat_bcsr_desc* A, int32_t* block_info,
                                int32_t* relative_block_index, smat_stream stream) {
  SMAT_REQUIRE(A && block_info && relative_block_index, "smat_bcsr_grid_maps: null argument");
  hipStream_t s = smat::to_hip(stream);
  int64_t ncb = A->K / A->bk;
  int64_t ng = A->nbrows * ncb;
  hipLaunchKernelGGL(k_fill_i32, grid_for(ng), kThreads, 0, s, ng, block_info, 0);
  hipLaunchKernelGGL(k_fill_i32, grid_for(ng), kThreads, 0, s, ng, relative_block_index, -1);
  if (A->nnzb > 0)
    hipLaunchKernelGGL(k_grid_maps, grid_for(A->nnzb), kThreads, 0, s, A->nnzb, A->nbrows, ncb, A->bk,
                       A->rowptr, A->colidx, A->tags, block_info, relative_block_index);
  SMAT_LAUNCH_CHECK();
  return SMAT_OK;
}

smat_status smat_bcsr_filter(const smat_bcsr_desc* A, int64_t* s_nnzb_host, int64_t* d_nnzb_host,
                             int32_t* s_rowptr, int32_t* s_colidx, uint16_t* s_values,
                             int32_t* d_rowptr, int32_t* d_colidx, uint16_t* d_values,
                             smat_stream stream) {
  SMAT_REQUIRE(A, "smat_bcsr_filter: null BCSR");
  hipStream_t s = smat::to_hip(stream);
  const int64_t n = A->nnzb;
  smat::DevBuf fs, fd, ps, pd, tmp;
  SMAT_TRY(fs.alloc((n + 1) * 4));
  SMAT_TRY(fd.alloc((n + 1) * 4));
  SMAT_TRY(ps.alloc((n + 1) * 4));
  SMAT_TRY(pd.alloc((n + 1) * 4));
  hipLaunchKernelGGL(k_flags, grid_for(n + 1), kThreads, 0, s, n, A->tags, fs.as<int32_t>(),
                     fd.as<int32_t>());
  size_t tb = 0;
  SMAT_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, fs.as<int32_t>(), ps.as<int32_t>(),
                                                (int)(n + 1), s));
  SMAT_TRY(tmp.alloc(tb));
  SMAT_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, fs.as<int32_t>(), ps.as<int32_t>(),
                                                (int)(n + 1), s));
  SMAT_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, fd.as<int32_t>(), pd.as<int32_t>(),
                                                (int)(n + 1), s));
  int32_t cnt[2] = {0, 0};
  SMAT_HIP_TRY(hipMemcpyAsync(&cnt[0], ps.as<int32_t>() + n, 4, hipMemcpyDeviceToHost, s));
  SMAT_HIP_TRY(hipMemcpyAsync(&cnt[1], pd.as<int32_t>() + n, 4, hipMemcpyDeviceToHost, s));
  SMAT_HIP_TRY(hipStreamSynchronize(s));
  if (s_nnzb_host) *s_nnzb_host = cnt[0];
  if (d_nnzb_host) *d_nnzb_host = cnt[1];
  if (s_rowptr || d_rowptr)
    hipLaunchKernelGGL(k_filter_rowptr, grid_for(A->nbrows + 1), kThreads, 0, s, A->nbrows, A->rowptr,
                       ps.as<int32_t>(), pd.as<int32_t>(), s_rowptr, d_rowptr);
  if (n > 0 && ((s_colidx && s_values) || (d_colidx && d_values)))
    hipLaunchKernelGGL(k_filter_copy, (unsigned)((n + 3) / 4), 256, 0, s, n, A->bk, A->tags, A->colidx,
                       A->values, ps.as<int32_t>(), pd.as<int32_t>(), s_colidx, s_values, d_colidx,
                       d_values);
  SMAT_LAUNCH_CHECK();
  SMAT_HIP_TRY(hipStreamSynchronize(s));
  return SMAT_OK;
}

}  // extern "C"
