// gfx950 matrix-core operand helpers shared by the tile kernels.
//
//  v_mfma_f32_16x16x32_{f16,bf16}: lane l carries A[l%16][8*(l/16) .. +7] and
//    B[8*(l/16) .. +7][l%16]; the result D[4*(l/16) + e][l%16] (e = 0..3).
//  v_smfmac_f32_16x16x64_{f16,bf16}: lane group g (lanes 16g .. 16g+15)
//    carries 8 kept values of 16 k-positions and one 16-bit index word (2 bits
//    per value); with the B operand read as below, group g's values meet the
//    B rows 16g .. 16g+15 of the 64-deep window.
//  ds_read_b64_tr_b16: a 16-lane group reading a 4 x 16 sub-tile of 16-bit
//    elements (lane L addressing row L/4, columns 4*(L%4) .. +3) receives
//    column L, i.e. 4 k-values of one B column; two reads give the 8 k-values
//    of an MFMA B operand.
#pragma once

#include "smat_internal.h"

namespace smat_mfma {

// Hardware dispatches workgroup i to XCD i % 8.  Over the first
// floor(n/8)*8 workgroups this gives XCD x the contiguous logical range
// [x*per, (x+1)*per); the tail keeps its own index (a bijection).
__device__ inline uint32_t xcd_remap(uint32_t bid, uint32_t nblocks) {
  const uint32_t nx = 8;
  const uint32_t per = nblocks / nx;
  if (bid >= per * nx) return bid;
  return (bid % nx) * per + bid / nx;
}

// DT 0 = FP16, 1 = BF16; operands travel as raw 16-bit words.
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

// Two transposed LDS reads (rows lo_addr.. and hi_addr.., 4 each) -> the 8
// k-values of one B column as one operand.
__device__ inline uint4 tr_pair(const uint8_t* lo_addr, const uint8_t* hi_addr) {
  const smat_s4 lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((__attribute__((address_space(3))) smat_s4*)(lo_addr));
  const smat_s4 hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((__attribute__((address_space(3))) smat_s4*)(hi_addr));
  // the whole-vector cast maps the two loaded register pairs onto the operand
  const smat_s4 both[2] = {lo, hi};
  return __builtin_bit_cast(uint4, both);
}

}  // namespace smat_mfma
