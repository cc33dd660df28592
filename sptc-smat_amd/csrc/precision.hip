// Precision modes: FP32 values into the BCSR block layout, operand packing for
// FP16 / BF16 / FP8 e4m3 (per-block-row and per-column scales) / split FP16,
// and the FP64 error-norm reduction (the only FP64 arithmetic on the GPU).
//
// The reference computes in FP16 only (values converted from the MTX doubles
// at mmio_highlevel.h:608, FP16 tensor-core accumulation, ptx.h:20-41); these
// modes are the precision ladder of this library and have no reference
// counterpart beyond the FP16 one.
#include "smat_internal.h"

namespace {

inline unsigned grid_for(int64_t n, int t) {
  const int64_t g = (n + t - 1) / t;
  return (unsigned)(g < 1 ? 1 : (g > (1 << 30) ? (1 << 30) : g));
}

__global__ void k_fill_f32(float* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.f;
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

// One wave per CSR row: each entry finds its block by binary search over the
// block row's (ascending) block column starts.
__global__ void k_scatter_values(int64_t M, int32_t bk, const int32_t* __restrict__ brp,
                                 const int32_t* __restrict__ bcol, const int32_t* __restrict__ rp,
                                 const int32_t* __restrict__ ci, const float* __restrict__ v,
                                 float* __restrict__ out) {
  const int64_t r = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (r >= M) return;
  const int64_t br = r / 16;
  const int32_t lo0 = brp[br], hi0 = brp[br + 1];
  for (int32_t q = rp[r] + lane; q < rp[r + 1]; q += 64) {
    const int32_t c = ci[q];
    const int32_t start = c - c % bk;
    int32_t lo = lo0, hi = hi0;  // first position with bcol >= start
    while (lo < hi) {
      const int32_t mid = (lo + hi) >> 1;
      if (bcol[mid] < start) lo = mid + 1;
      else hi = mid;
    }
    if (lo < hi0 && bcol[lo] == start) out[((int64_t)lo * 16 + (r % 16)) * bk + c % bk] = v[q];
  }
}

// ---- scales ---------------------------------------------------------------
// FP8: s = amax / 448 (1 for amax == 0); split: s = 2^(e - 15) for
// amax in [2^(e-1), 2^e), so amax / s lies in [2^14, 2^15).
__device__ inline float scale_of(float amax, int dt) {
  if (!(amax > 0.f)) return 1.f;
  if (dt == SMAT_FP8_E4M3) return amax / 448.f;
  int e;
  (void)frexpf(amax, &e);
  return ldexpf(1.f, e - 15);
}

// e4m3 of four scaled values, round to nearest even, clamped to +-448.
__device__ inline uint32_t fp8x4(float a, float b, float c, float d) {
  a = fminf(fmaxf(a, -448.f), 448.f);
  b = fminf(fmaxf(b, -448.f), 448.f);
  c = fminf(fmaxf(c, -448.f), 448.f);
  d = fminf(fmaxf(d, -448.f), 448.f);
  int w = __builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false);
  w = __builtin_amdgcn_cvt_pk_fp8_f32(c, d, w, true);
  return (uint32_t)w;
}

__device__ inline uint16_t to_bits(float x, int dt) {
  if (dt == SMAT_BF16) return __builtin_bit_cast(uint16_t, (__bf16)x);
  return __builtin_bit_cast(uint16_t, (_Float16)x);
}

// Elementwise F16 / BF16 conversion of n values.
__global__ void k_convert16(const float* __restrict__ x, int64_t n, int dt, uint16_t* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = to_bits(x[i], dt);
}

// One workgroup per block row: amax over its values, scale, then the packed
// operand (FP8 bytes, or split hi / lo halves).
__global__ __launch_bounds__(256) void k_pack_a_rows(const int32_t* __restrict__ brp, int32_t bk,
                                                     const float* __restrict__ x, int dt, int64_t lo_off,
                                                     void* __restrict__ out, float* __restrict__ scale) {
  __shared__ float red[256];
  const int64_t r = blockIdx.x;
  const int64_t beg = (int64_t)brp[r] * 16 * bk, end = (int64_t)brp[r + 1] * 16 * bk;
  float m = 0.f;
  for (int64_t i = beg + threadIdx.x; i < end; i += 256) m = fmaxf(m, fabsf(x[i]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  const float s = scale_of(red[0], dt);
  if (threadIdx.x == 0) scale[r] = s;
  if (dt == SMAT_FP8_E4M3) {
    uint32_t* o4 = static_cast<uint32_t*>(out);  // 16 * bk is a multiple of 4
    for (int64_t i = beg / 4 + threadIdx.x; i < end / 4; i += 256)
      o4[i] = fp8x4(x[4 * i] / s, x[4 * i + 1] / s, x[4 * i + 2] / s, x[4 * i + 3] / s);
  } else {
    uint16_t* o = static_cast<uint16_t*>(out);
    for (int64_t i = beg + threadIdx.x; i < end; i += 256) {
      const float y = x[i] / s;  // exact: s is a power of two
      const _Float16 hi = (_Float16)y;
      o[i] = __builtin_bit_cast(uint16_t, hi);
      o[lo_off + i] = __builtin_bit_cast(uint16_t, (_Float16)(y - (float)hi));
    }
  }
}

// Column amax of B: thread = column, gridDim.y slabs of rows; float bits of
// non-negative values order like unsigned integers.
__global__ void k_col_amax(const float* __restrict__ B, int64_t K, int64_t N, int64_t ldb, int64_t rows_per,
                           uint32_t* __restrict__ amax) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const int64_t k0 = (int64_t)blockIdx.y * rows_per, k1 = min(K, k0 + rows_per);
  float m = 0.f;
  for (int64_t k = k0; k < k1; k++) m = fmaxf(m, fabsf(B[k * ldb + n]));
  atomicMax(&amax[n], __float_as_uint(m));
}

__global__ void k_pack_b(const float* __restrict__ B, int64_t K, int64_t N, int64_t ldb, int dt,
                         const uint32_t* __restrict__ amax, void* __restrict__ out) {
  if (dt == SMAT_FP8_E4M3) {  // four columns per thread (N % 4 == 0)
    const int64_t nq = N / 4, total = K * nq;
    uint32_t* o4 = static_cast<uint32_t*>(out);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
      const int64_t k = i / nq, n = 4 * (i % nq);
      const float* row = B + k * ldb + n;
      o4[i] = fp8x4(row[0] / scale_of(__uint_as_float(amax[n]), dt),
                    row[1] / scale_of(__uint_as_float(amax[n + 1]), dt),
                    row[2] / scale_of(__uint_as_float(amax[n + 2]), dt),
                    row[3] / scale_of(__uint_as_float(amax[n + 3]), dt));
    }
  } else {
    const int64_t total = K * N;
    uint16_t* o = static_cast<uint16_t*>(out);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
      const int64_t k = i / N, n = i % N;
      const float y = B[k * ldb + n] / scale_of(__uint_as_float(amax[n]), dt);
      const _Float16 hi = (_Float16)y;
      o[i] = __builtin_bit_cast(uint16_t, hi);
      o[total + i] = __builtin_bit_cast(uint16_t, (_Float16)(y - (float)hi));
    }
  }
}

__global__ void k_scale_from_amax(int64_t N, int dt, const uint32_t* __restrict__ amax, float* __restrict__ scale) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < N) scale[n] = scale_of(__uint_as_float(amax[n]), dt);
}

// ---- FP64 error sums (fixed-order two-level reduction, deterministic) ------
constexpr int kErrBlocks = 1024;

__global__ __launch_bounds__(256) void k_err_partial(const float* __restrict__ C, int64_t ldc,
                                                     const double* __restrict__ R, int64_t ldr, int64_t rows,
                                                     int64_t N, double* __restrict__ work) {
  __shared__ double se[256], sr[256];
  double e2 = 0.0, r2 = 0.0;
  const int64_t total = rows * N;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (int64_t)kErrBlocks * 256) {
    const int64_t r = i / N, n = i % N;
    const double ref = R[r * ldr + n];
    const double d = (double)C[r * ldc + n] - ref;
    e2 += d * d;
    r2 += ref * ref;
  }
  se[threadIdx.x] = e2;
  sr[threadIdx.x] = r2;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      se[threadIdx.x] += se[threadIdx.x + o];
      sr[threadIdx.x] += sr[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    work[2 * blockIdx.x] = se[0];
    work[2 * blockIdx.x + 1] = sr[0];
  }
}

__global__ __launch_bounds__(kErrBlocks) void k_err_final(const double* __restrict__ work, double* __restrict__ out) {
  __shared__ double se[kErrBlocks], sr[kErrBlocks];
  se[threadIdx.x] = work[2 * threadIdx.x];
  sr[threadIdx.x] = work[2 * threadIdx.x + 1];
  __syncthreads();
  for (int o = kErrBlocks / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      se[threadIdx.x] += se[threadIdx.x + o];
      sr[threadIdx.x] += sr[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = se[0];
    out[1] = sr[0];
  }
}

}  // namespace

extern "C" {

smat_status smat_bcsr_scatter_values(const smat_bcsr_desc* A, const int32_t* csr_rowptr, const int32_t* csr_colidx,
                                     const float* csr_values, float* out_values, smat_stream stream) {
  SMAT_REQUIRE(A && csr_rowptr && out_values, "smat_bcsr_scatter_values: null argument");
  hipStream_t s = smat::to_hip(stream);
  const int64_t n = A->nnzb * 16 * A->bk;
  hipLaunchKernelGGL(k_fill_f32, grid_for(n, 256 * 8), 256, 0, s, out_values, n);
  if (A->nnzb > 0) {
    SMAT_REQUIRE(csr_colidx && csr_values, "smat_bcsr_scatter_values: null CSR arrays");
    hipLaunchKernelGGL(k_scatter_values, grid_for(A->M, 4), 256, 0, s, A->M, A->bk, A->rowptr, A->colidx, csr_rowptr,
                       csr_colidx, csr_values, out_values);
  }
  SMAT_LAUNCH_CHECK();
  return SMAT_OK;
}

smat_status smat_pack_a(const smat_bcsr_desc* A, const float* values_f32, int32_t dtype, void* a_out, float* a_scale,
                        smat_stream stream) {
  SMAT_REQUIRE(A && values_f32 && a_out, "smat_pack_a: null argument");
  SMAT_REQUIRE(dtype == SMAT_F16 || dtype == SMAT_BF16 || dtype == SMAT_FP8_E4M3 || dtype == SMAT_SPLIT_F16,
               "smat_pack_a: unknown dtype %d", dtype);
  hipStream_t s = smat::to_hip(stream);
  const int64_t n = A->nnzb * 16 * A->bk;
  if (dtype == SMAT_F16 || dtype == SMAT_BF16) {
    if (n > 0)
      hipLaunchKernelGGL(k_convert16, grid_for(n, 256 * 8), 256, 0, s, values_f32, n, dtype,
                         static_cast<uint16_t*>(a_out));
  } else {
    SMAT_REQUIRE(a_scale, "smat_pack_a: FP8 and split need a_scale[nbrows]");
    SMAT_REQUIRE(A->nbrows < (1 << 30), "smat_pack_a: too many block rows");
    hipLaunchKernelGGL(k_pack_a_rows, (unsigned)A->nbrows, 256, 0, s, A->rowptr, A->bk, values_f32, dtype, n, a_out,
                       a_scale);
  }
  SMAT_LAUNCH_CHECK();
  return SMAT_OK;
}

smat_status smat_pack_b(const float* B, int64_t K, int64_t N, int64_t ldb, int32_t dtype, void* b_out, float* b_scale,
                        smat_stream stream) {
  SMAT_REQUIRE(B && b_out && K > 0 && N > 0 && ldb >= N, "smat_pack_b: bad argument");
  SMAT_REQUIRE(dtype == SMAT_F16 || dtype == SMAT_BF16 || dtype == SMAT_FP8_E4M3 || dtype == SMAT_SPLIT_F16,
               "smat_pack_b: unknown dtype %d", dtype);
  hipStream_t s = smat::to_hip(stream);
  if (dtype == SMAT_F16 || dtype == SMAT_BF16) {
    SMAT_REQUIRE(ldb == N, "smat_pack_b: F16/BF16 packing needs ldb == N");
    hipLaunchKernelGGL(k_convert16, grid_for(K * N, 256 * 8), 256, 0, s, B, K * N, dtype, static_cast<uint16_t*>(b_out));
    SMAT_LAUNCH_CHECK();
    return SMAT_OK;
  }
  SMAT_REQUIRE(b_scale, "smat_pack_b: FP8 and split need b_scale[N]");
  if (dtype == SMAT_FP8_E4M3) SMAT_REQUIRE(N % 4 == 0, "smat_pack_b: FP8 needs N %% 4 == 0");
  // amax bits live in b_scale until the last kernel turns them into scales
  uint32_t* amax = reinterpret_cast<uint32_t*>(b_scale);
  hipLaunchKernelGGL(k_fill_u32, grid_for(N, 256), 256, 0, s, amax, N);
  const int64_t rows_per = 512;
  const int64_t slabs = (K + rows_per - 1) / rows_per;
  SMAT_REQUIRE(slabs < 65536, "smat_pack_b: K too large");
  hipLaunchKernelGGL(k_col_amax, dim3(grid_for(N, 256), (unsigned)slabs), 256, 0, s, B, K, N, ldb, rows_per, amax);
  hipLaunchKernelGGL(k_pack_b, grid_for(K * N, 256 * 8), 256, 0, s, B, K, N, ldb, dtype, amax, b_out);
  hipLaunchKernelGGL(k_scale_from_amax, grid_for(N, 256), 256, 0, s, N, dtype, amax, b_scale);
  SMAT_LAUNCH_CHECK();
  return SMAT_OK;
}

smat_status smat_error_sums(const float* C, int64_t ldc, const double* R, int64_t ldr, int64_t rows, int64_t N,
                            double* out, double* work, smat_stream stream) {
  SMAT_REQUIRE(C && R && out && work && rows >= 0 && N > 0 && ldc >= N && ldr >= N, "smat_error_sums: bad argument");
  hipStream_t s = smat::to_hip(stream);
  hipLaunchKernelGGL(k_err_partial, kErrBlocks, 256, 0, s, C, ldc, R, ldr, rows, N, work);
  hipLaunchKernelGGL(k_err_final, 1, kErrBlocks, 0, s, work, out);
  SMAT_LAUNCH_CHECK();
  return SMAT_OK;
}

}  // extern "C"
