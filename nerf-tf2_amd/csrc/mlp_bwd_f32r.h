// Register-resident FP32 dX chain of the reference topology (W = 256 with the
// rgb branch): the backward twin of mlp_fwd_f32r.h.
//
// A wave owns 16 samples and walks the layers top-down on
// v_mfma_f32_16x16x4f32.  dh_{l-1} = W_l[:W] dz_l: the output blocks are the 16
// blocks of 16 input features of layer l, the k-steps run over its outputs, and
// dz_l is the B operand in the accumulator layout of the step before (lane (j,
// q) holds features 16 nb + 4 q + r of sample j), with the transposed weights
// packed in that k order.  A wave needs under 200 registers, so two waves
// share each SIMD.  The ReLU masks are the forward's branch bits (two words per
// lane and trunk layer, one for the rgb branch, all loaded up front), so no
// activation row is read; every dz_l is written row-major for the weight GEMMs
// as 16-byte buffer stores over the wave's 16 rows (rows past M fall outside
// the resources).
//   rgb = sigmoid(z):  dz_rgb = d_rgb * rgb (1 - rgb)
//   sigma = relu(z):   dz_sigma = d_sigma * [sigma > 0]
//   dz = dh * [h > 0] for the ReLU layers; the bottleneck is linear.
#pragma once
#include "mlp_fwd_f32r.h"

namespace f32r {

__device__ __forceinline__ void store_f4(const floatx4& v, __amdgpu_buffer_rsrc_t r, uint32_t off) {
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4_cvt, v), r, off, 0, 0);
}

// dz[nb] = dh[nb] * [h > 0] over blocks 0 .. NBO-1, with the forward's branch bits
// (bit 4 (nb % 8) + r of w[nb / 8]); then the dz row at dz_off.
template <int NBO>
__device__ __forceinline__ void mask_store(floatx4 (&dz)[16], const uint32_t (&w)[2], __amdgpu_buffer_rsrc_t rdz,
                                           uint32_t row_z, int64_t dz_off) {
#pragma unroll
  for (int nb = 0; nb < NBO; ++nb) {
#pragma unroll
    for (int r = 0; r < 4; ++r) dz[nb][r] = ((w[nb >> 3] >> (4 * (nb & 7) + r)) & 1u) ? dz[nb][r] : 0.f;
    store_f4(dz[nb], rdz, row_z + (uint32_t)((dz_off + 16 * nb) * 4));
  }
}

}  // namespace f32r

template <int W, int D, int SKIP, int LX, int LD>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void mlp_bwd_f32r_kernel(
    MlpPlan P, const char* __restrict__ pk, int64_t M, const float* __restrict__ d_rgb,
    const float* __restrict__ d_sigma, const float* __restrict__ rgb, const float* __restrict__ sigma,
    const float* __restrict__ stash, float* __restrict__ dz) {
  using namespace f32r;
  static_assert(W == 256 && LD > 0 && SKIP >= 0 && SKIP + 1 < D, "register-resident kernel is for the 8x256 NeRF MLP");
  constexpr int NH = W / 16;   // k-step groups over a W-wide dz
  constexpr int NR = W / 32;   // blocks of 16 in the W/2-wide rgb branch
  const int lane = threadIdx.x & 63, j = lane & 15, q = lane >> 4;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t m0 = ((int64_t)blockIdx.x * kWaves + wv) * 16;
  if (m0 >= M) return;  // no workgroup barrier in this kernel
  const int64_t m = m0 + j;
  const bool valid = m < M;
  const int64_t mm = valid ? m : M - 1;
  const __amdgpu_buffer_rsrc_t rs = ws::packed_rsrc(pk, P.packed_bytes);
  const uint32_t lane_off = 16u * (uint32_t)lane, q_off = 16u * (uint32_t)q;
  const int64_t w_rows = M - m0 < 16 ? M - m0 : 16;
  const __amdgpu_buffer_rsrc_t rst = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<float*>(stash) + m0 * P.stash_per_sample, 0, (int)(w_rows * P.stash_per_sample * 4), 0x00020000);
  const __amdgpu_buffer_rsrc_t rdz =
      __builtin_amdgcn_make_buffer_rsrc(dz + m0 * P.dz_per_sample, 0, (int)(w_rows * P.dz_per_sample * 4), 0x00020000);
  const uint32_t row_s = (uint32_t)(j * P.stash_per_sample * 4);
  const uint32_t row_z0 = (uint32_t)(j * P.dz_per_sample * 4), row_z = row_z0 + q_off;
  // the forward's ReLU bits of every trunk layer and of the rgb branch
  uint32_t mw[D][2], mr[2] = {0u, 0u};
#pragma unroll
  for (int l = 0; l < D; ++l) {
    const ws::u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(rst, row_s + (uint32_t)((P.st_mask + 8 * l + 2 * q) * 4), 0, 0);
    mw[l][0] = v[0];
    mw[l][1] = v[1];
  }
  mr[0] = __builtin_amdgcn_raw_buffer_load_b32(rst, row_s + (uint32_t)((P.st_mask + 8 * D + q) * 4), 0, 0);

  const float sg = sigma[mm];
  const float ds = valid ? d_sigma[mm] : 0.f;
  const float dzs = sg > 0.f ? ds : 0.f;
  float dzr[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float y = rgb[3 * mm + c];
    const float dy = valid ? d_rgb[3 * mm + c] : 0.f;
    dzr[c] = dy * y * (1.f - y);  // sigmoid'
  }
  if (q == 0) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, dzr[c]), rdz,
                                            row_z0 + (uint32_t)((P.dz_rgb + c) * 4), 0, 0);
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, dzs), rdz,
                                          row_z0 + (uint32_t)(P.dz_sigma * 4), 0, 0);
  }
  floatx4 act[16], acc[16];
  // rgb head: dh_rbr[f] = sum_c W_rgb[f][c] dz_rgb[c] (weights [feature][4]), masked
#pragma unroll
  for (int nb = 0; nb < NR; ++nb)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      // feature 16 nb + 4 q + r: byte 16 (16 nb + r) + 64 q
      const floatx4 w = ws::ldf4(rs, 4u * q_off, P.p_hr_nat + 16 * (16 * nb + r));
      act[nb][r] = dzr[0] * w[0] + dzr[1] * w[1] + dzr[2] * w[2];
    }
  mask_store<NR>(act, mr, rdz, row_z, P.dz_rbr);
  // dh_bott = W_{D+1}[:W] dz_rbr; the bottleneck is linear, so dz_bott = dh_bott
  layer<16, 0, NR>(acc, act, nullptr, rs, lane_off, P.p_bwd4[D + 1]);
#pragma unroll
  for (int nb = 0; nb < 16; ++nb) {
    act[nb] = acc[nb];
    store_f4(act[nb], rdz, row_z + (uint32_t)((P.dz_bott + 16 * nb) * 4));
  }
  // dh_{D-1} = W_D dz_bott + W_sigma dz_sigma
  layer<16, 0, NH>(acc, act, nullptr, rs, lane_off, P.p_bwd4[D]);
#pragma unroll
  for (int nb = 0; nb < 16; ++nb) {
    const floatx4 w = ws::ldf4(rs, q_off, P.p_hs_nat + 64 * nb);
#pragma unroll
    for (int r = 0; r < 4; ++r) act[nb][r] = acc[nb][r] + w[r] * dzs;
  }
  mask_store<16>(act, mw[D - 1], rdz, row_z, P.dz_h[D - 1]);
  // trunk, top-down: dh_{l-1} = W_l[:W] dz_l, dz_{l-1} = dh_{l-1} [h_{l-1} > 0]
  ws::static_for<1, D>([&](auto Lc) {
    constexpr int l = D - decltype(Lc)::value;  // D-1 .. 1
    layer<16, 0, NH>(acc, act, nullptr, rs, lane_off, P.p_bwd4[l]);
#pragma unroll
    for (int nb = 0; nb < 16; ++nb) act[nb] = acc[nb];
    mask_store<16>(act, mw[l - 1], rdz, row_z, P.dz_h[l - 1]);
  });
}

template <int W, int D, int SKIP, int LX, int LD>
void launch_bwd_f32r(const MlpPlan& P, const void* packed, int64_t M, const float* d_rgb, const float* d_sigma,
                     const float* rgb, const float* sigma, const void* stash, void* dz, hipStream_t s) {
  const int64_t nwg = (M + 16 * f32r::kWaves - 1) / (16 * f32r::kWaves);
  mlp_bwd_f32r_kernel<W, D, SKIP, LX, LD><<<(unsigned)nwg, 64 * f32r::kWaves, 0, s>>>(
      P, reinterpret_cast<const char*>(packed), M, d_rgb, d_sigma, rgb, sigma,
      reinterpret_cast<const float*>(stash), reinterpret_cast<float*>(dz));
}
