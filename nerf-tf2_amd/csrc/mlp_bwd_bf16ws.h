// Helpers of the weight-stationary BF16 dX chain (mlp_bwd_bf16ws8.h): dz row
// stores from the accumulator layout, the forward's ReLU bit masks applied to
// packed bf16, and buffer loads and stores whose whole offset is in the VGPR
// offset.  dz rows are written as 16-byte chunks of 8 consecutive features, swapped
// out of the accumulator layout across lane halves (swap_halves).
#pragma once
#include "mlp_fwd_bf16ws.h"

namespace wsb {

using namespace ws;

// Stores take their whole offset in the VGPR offset (scalar offset 0): with a
// scalar offset register the compiler puts no wait state between a 16-byte
// store and a VALU write of its data registers, and on gfx950 such stores have
// written another value for some lanes.
__device__ __forceinline__ void st4(bf16x4 v, __amdgpu_buffer_rsrc_t r, uint32_t voff, int64_t soff) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff + (uint32_t)soff, 0, kNT);
}

// v with the ReLU mask applied: the forward's bit mask w holds, for the pairs of
// accumulator elements kk = 4 h + k (dword k of the half h), bits shift + kk and
// 16 + shift + kk.
__device__ __forceinline__ bf16x8 mask_bits(bf16x8 v, uint32_t w, int shift) {
  u32x4 d = __builtin_bit_cast(u32x4, v);
#pragma unroll
  for (int k = 0; k < 4; ++k) d[k] &= ((w >> (shift + k)) & 0x00010001u) * 0xFFFFu;
  return __builtin_bit_cast(bf16x8, d);
}

__device__ __forceinline__ uint32_t ld_u32(__amdgpu_buffer_rsrc_t r, uint32_t voff, int64_t soff) {
  return __builtin_amdgcn_raw_buffer_load_b32(r, voff, (int)soff, 0);
}

// Stores a 16-feature group given in the accumulator layout as row chunks.
__device__ __forceinline__ void st_group(bf16x8 v, __amdgpu_buffer_rsrc_t r, uint32_t voff, int64_t soff) {
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, swap_halves(v)), r, voff + (uint32_t)soff, 0, kNT);
}

}  // namespace wsb
