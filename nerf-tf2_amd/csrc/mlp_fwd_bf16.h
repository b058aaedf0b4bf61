// BF16 fused forward: same structure as the fp32 kernel, with
// v_mfma_f32_32x32x16_bf16 (fp32 accumulate), bf16 weights and bf16
// activations in the LDS slab (16 KB per wave for W = 256).  Bias, activation
// functions and the sigma / rgb heads stay fp32.
#pragma once
#include "mlp_kernels.h"

template <int W, int D, int SKIP, int LX, int LD>
__global__ __launch_bounds__(256) void mlp_fwd_bf16_kernel(
    MlpPlan P, const char* __restrict__ pk, int64_t M, const float* __restrict__ xyz,
    const float* __restrict__ dirs, const float* __restrict__ ro, const float* __restrict__ rd,
    const float* __restrict__ tv, int spr, float* __restrict__ rgb_out,
    float* __restrict__ sig_out, __bf16* __restrict__ stash) {
  using C = NetC<W, D, SKIP, LX, LD>;
  constexpr int TH = W / 16;  // bf16 k-steps over a hidden input
  __shared__ bf16x8 s_act[4 * TH * 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, j = lane & 31, g = lane >> 5;
  bf16x8* act = s_act + wv * TH * 64;
  const int64_t nblk = (M + 31) / 32;
  const int64_t blk = (int64_t)blockIdx.x * 4 + wv;
  if (blk >= nblk) return;
  const int64_t m = blk * 32 + j;
  const bool valid = m < M;
  const int64_t mm = valid ? m : M - 1;
  // this sample's row of the stash, blocked by 32 samples (the wave's block)
  const BlkRow<__bf16> srow{stash ? stash + blk * 32 * P.stash_per_sample : nullptr,
                            (int)(M - blk * 32 < 32 ? M - blk * 32 : 32), j};
  const bool st = stash && valid;
  float x[3], d[3];
  load_sample(mm, xyz, dirs, ro, rd, tv, spr, x, d);
  bf16x8 exb[C::KX / 8], edb[C::KDA / 8];
  {
    float ex[C::KX];
    encode_slots<LX, C::KX>(x, g, ex);
    slots_to_bf16(ex, exb);
    if (st) store_slots_bf16<C::KX>(ex, srow, P.st_encx + g * C::KX);
  }
  if constexpr (C::RGB) {
    float ed[C::KDA];
    encode_slots<LD, C::KDA>(d, g, ed);
    slots_to_bf16(ed, edb);
    if (st) store_slots_bf16<C::KDA>(ed, srow, P.st_encd + g * C::KDA);
  }

  floatx16 acc[C::NB];
  zero_acc(acc);
  mfma_regs_bf16<C::NB, C::KX / 8>(acc, reinterpret_cast<const bf16x8*>(pk + P.p_fwd[0]), exb, lane);
  bias_act<C::NB, 1>(acc, reinterpret_cast<const float*>(pk + P.p_bias[0]), g);
  wave_sync();
  store_act_bf16(acc, act, lane);
  if (st) store_rows_bf16(acc, srow, P.st_h[0], g);
#pragma unroll
  for (int l = 1; l < D; ++l) {
    const bf16x8* wq = reinterpret_cast<const bf16x8*>(pk + P.p_fwd[l]);
    zero_acc(acc);
    wave_sync();
    mfma_lds_bf16<C::NB, TH>(acc, wq, act, lane);
    if (l - 1 == SKIP) mfma_regs_bf16<C::NB, C::KX / 8>(acc, wq + TH * C::NB * 64, exb, lane);
    bias_act<C::NB, 1>(acc, reinterpret_cast<const float*>(pk + P.p_bias[l]), g);
    wave_sync();
    store_act_bf16(acc, act, lane);
    if (st) store_rows_bf16(acc, srow, P.st_h[l], g);
  }
  // heads see the bf16-rounded activations, like the following layers
  round_acc_bf16(acc);
  const float* hs = reinterpret_cast<const float*>(pk + P.p_head_sigma);
  float zs = head_dot_f32<C::NB>(acc, hs + g * C::KH);
  zs = zs + shfl_xor(zs, 32);
  zs = zs + hs[2 * C::KH];
  const float sigma = C::RGB ? relu(zs) : softplus(zs);
  float rgb[3] = {0.f, 0.f, 0.f};
  if constexpr (C::RGB) {
    const int mb = D;
    zero_acc(acc);
    wave_sync();
    mfma_lds_bf16<C::NB, TH>(acc, reinterpret_cast<const bf16x8*>(pk + P.p_fwd[mb]), act, lane);
    bias_act<C::NB, 0>(acc, reinterpret_cast<const float*>(pk + P.p_bias[mb]), g);
    wave_sync();
    store_act_bf16(acc, act, lane);
    if (st) store_rows_bf16(acc, srow, P.st_bott, g);
    const int mr = D + 1;
    const bf16x8* wq = reinterpret_cast<const bf16x8*>(pk + P.p_fwd[mr]);
    floatx16 accr[C::NBR];
    zero_acc(accr);
    wave_sync();
    mfma_lds_bf16<C::NBR, TH>(accr, wq, act, lane);
    mfma_regs_bf16<C::NBR, C::KDA / 8>(accr, wq + TH * C::NBR * 64, edb, lane);
    bias_act<C::NBR, 1>(accr, reinterpret_cast<const float*>(pk + P.p_bias[mr]), g);
    round_acc_bf16(accr);
    if (st) store_rows_bf16(accr, srow, P.st_rbr, g);
    rgb_head_f32<C::NBR>(accr, reinterpret_cast<const float*>(pk + P.p_head_rgb), g, rgb);
  }
  if (valid && g == 0) {
    sig_out[m] = sigma;
    if constexpr (C::RGB) {
      rgb_out[3 * m + 0] = rgb[0];
      rgb_out[3 * m + 1] = rgb[1];
      rgb_out[3 * m + 2] = rgb[2];
    }
  }
}

template <int W, int D, int SKIP, int LX, int LD>
int launch_fwd_bf16(const MlpPlan& P, const void* packed, int64_t M, const float* xyz,
                    const float* dirs, const float* ro, const float* rd, const float* tv, int spr,
                    float* rgb, float* sigma, void* stash, hipStream_t s) {
  const int64_t nblk = (M + 31) / 32;
  mlp_fwd_bf16_kernel<W, D, SKIP, LX, LD><<<(unsigned)((nblk + 3) / 4), 256, 0, s>>>(
      P, reinterpret_cast<const char*>(packed), M, xyz, dirs, ro, rd, tv, spr, rgb, sigma,
      reinterpret_cast<__bf16*>(stash));
  return 0;
}
