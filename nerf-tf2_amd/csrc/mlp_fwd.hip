// Fused positional-encoding + MLP forward (model.py:289-394) on MFMA.
//
// The reference topology (8 x 256 with the rgb branch) runs the register-
// resident kernel of mlp_fwd_f32r.h in fp32 and the eight-wave weight-
// stationary kernel of mlp_fwd_bf16ws8.h in bf16.  The fp32 kernel below serves the other topologies
// (the proposal and finite-difference nets).  One wave owns 32 samples and
// carries them through every layer:
//   * the encoding is computed in registers directly in B-operand slots;
//   * each layer's output accumulates in registers (floatx16 per 32 features),
//     gets bias + activation, and is written to the wave's LDS slab as the next
//     layer's B operand (no transpose: see mlp_plan.h);
//   * weights (A operand) stream from L2 with one k-group of prefetch;
//   * the 1-wide sigma head and 3-wide rgb head are VALU dot products over the
//     lane's features plus one cross-half shuffle.
// No cross-wave synchronisation: each wave only touches its own LDS slab.
#include "common.h"
#include "mlp_plan.h"
#include "mlp_kernels.h"
#include <type_traits>

#include "mlp_fwd_bf16.h"
#include "mlp_fwd_bf16ws.h"
#include "mlp_fwd_f32r.h"
#include "mlp_fwd_bf16ws8.h"

namespace {

template <int W, int D, int SKIP, int LX, int LD>
__global__ __launch_bounds__(256) void mlp_fwd_f32_kernel(
    MlpPlan P, const char* __restrict__ pk, int64_t M, const float* __restrict__ xyz,
    const float* __restrict__ dirs, const float* __restrict__ ro, const float* __restrict__ rd,
    const float* __restrict__ tv, int spr, float* __restrict__ rgb_out,
    float* __restrict__ sig_out, float* __restrict__ stash) {
  using C = NetC<W, D, SKIP, LX, LD>;
  __shared__ floatx4 s_act[4 * C::QH * 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, j = lane & 31, g = lane >> 5;
  floatx4* act = s_act + wv * C::QH * 64;
  const int64_t nblk = (M + 31) / 32;
  {
    // one sample block per wave, no grid-stride loop: a persistent loop would
    // let the compiler hoist every (loop-invariant) weight load out of it
    const int64_t blk = (int64_t)blockIdx.x * 4 + wv;
    if (blk >= nblk) return;
    const int64_t m = blk * 32 + j;
    const bool valid = m < M;
    const int64_t mm = valid ? m : M - 1;
    float x[3], d[3];
    load_sample(mm, xyz, dirs, ro, rd, tv, spr, x, d);
    float ex[C::KX];
    encode_slots<LX, C::KX>(x, g, ex);
    float ed[C::KDA];
    if constexpr (C::RGB) encode_slots<LD, C::KDA>(d, g, ed);
    float* srow = (stash && valid) ? stash + mm * P.stash_per_sample : nullptr;
    if (srow) store_slots_f32<C::KX>(ex, srow + P.st_encx + g * C::KX);

    floatx16 acc[C::NB];
    // dense_0 on the xyz encoding
    zero_acc(acc);
    mfma_regs_f32<C::NB, C::KX>(acc, reinterpret_cast<const floatx4*>(pk + P.p_fwd[0]), ex, lane);
    bias_act<C::NB, 1>(acc, reinterpret_cast<const float*>(pk + P.p_bias[0]), g);
    wave_sync();
    store_act_f32(acc, act, lane);
    if (srow) store_rows_f32(acc, srow + P.st_h[0], g);
    // dense_1 .. dense_{D-1}; the encoding is re-concatenated after dense_SKIP
#pragma unroll
    for (int l = 1; l < D; ++l) {
      const floatx4* wq = reinterpret_cast<const floatx4*>(pk + P.p_fwd[l]);
      zero_acc(acc);
      wave_sync();
      mfma_lds_f32<C::NB, C::QH>(acc, wq, act, lane);
      if (l - 1 == SKIP) mfma_regs_f32<C::NB, C::KX>(acc, wq + C::QH * C::NB * 64, ex, lane);
      bias_act<C::NB, 1>(acc, reinterpret_cast<const float*>(pk + P.p_bias[l]), g);
      wave_sync();
      store_act_f32(acc, act, lane);
      if (srow) store_rows_f32(acc, srow + P.st_h[l], g);
    }
    // sigma head (model.py:374-375)
    const float* hs = reinterpret_cast<const float*>(pk + P.p_head_sigma);
    float zs = head_dot_f32<C::NB>(acc, hs + g * C::KH);
    zs = zs + shfl_xor(zs, 32);
    zs = zs + hs[2 * C::KH];
    const float sigma = C::RGB ? relu(zs) : softplus(zs);
    float rgb[3] = {0.f, 0.f, 0.f};
    if constexpr (C::RGB) {
      // bottleneck dense_D, linear (model.py:377-378)
      const int mb = D;
      zero_acc(acc);
      wave_sync();
      mfma_lds_f32<C::NB, C::QH>(acc, reinterpret_cast<const floatx4*>(pk + P.p_fwd[mb]), act,
                                 lane);
      bias_act<C::NB, 0>(acc, reinterpret_cast<const float*>(pk + P.p_bias[mb]), g);
      wave_sync();
      store_act_f32(acc, act, lane);
      if (srow) store_rows_f32(acc, srow + P.st_bott, g);
      // dense_{D+1} on concat(bottleneck, direction encoding) (model.py:380-384)
      const int mr = D + 1;
      const floatx4* wq = reinterpret_cast<const floatx4*>(pk + P.p_fwd[mr]);
      floatx16 accr[C::NBR];
      zero_acc(accr);
      wave_sync();
      mfma_lds_f32<C::NBR, C::QH>(accr, wq, act, lane);
      mfma_regs_f32<C::NBR, C::KDA>(accr, wq + C::QH * C::NBR * 64, ed, lane);
      bias_act<C::NBR, 1>(accr, reinterpret_cast<const float*>(pk + P.p_bias[mr]), g);
      if (srow) {
        store_rows_f32(accr, srow + P.st_rbr, g);
        store_slots_f32<C::KDA>(ed, srow + P.st_encd + g * C::KDA);
      }
      // rgb head, sigmoid (model.py:386-387)
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
}

template <int W, int D, int SKIP, int LX, int LD>
int launch_fwd_f32(const MlpPlan& P, const void* packed, int64_t M, const float* xyz,
                   const float* dirs, const float* ro, const float* rd, const float* tv, int spr,
                   float* rgb, float* sigma, void* stash, hipStream_t s) {
  if constexpr (W == 256 && LD > 0 && SKIP >= 0 && SKIP + 1 < D) {
    // the reference topology: the register-resident kernel (16 samples per wave)
    launch_fwd_f32r<W, D, SKIP, LX, LD>(P, packed, M, xyz, dirs, ro, rd, tv, spr, rgb, sigma, stash, s);
  } else {
    const int64_t nblk = (M + 31) / 32;
    const int64_t grid = (nblk + 3) / 4;
    mlp_fwd_f32_kernel<W, D, SKIP, LX, LD><<<(unsigned)grid, 256, 0, s>>>(
        P, reinterpret_cast<const char*>(packed), M, xyz, dirs, ro, rd, tv, spr, rgb, sigma,
        reinterpret_cast<float*>(stash));
  }
  return 0;
}

}  // namespace

extern "C" int32_t nerf_mlp_forward(const nerf_mlp_desc* desc, int32_t precision,
                                    const void* packed, int64_t M, const float* xyz,
                                    const float* dirs, const float* rays_o, const float* rays_d,
                                    const float* t_vals, int32_t samples_per_ray, float* rgb,
                                    float* sigma, void* stash, void* stream) {
  MlpPlan P;
  NERF_CHECK_ARG(nerf_plan_from_desc(desc, precision, &P) == 0,
                 "nerf_mlp_forward: unsupported MLP description or precision");
  NERF_CHECK_ARG(M >= 0, "nerf_mlp_forward: M < 0");
  if (M == 0) return NERF_OK;  // empty batch
  NERF_CHECK_ARG(packed && sigma, "nerf_mlp_forward: null pointer");
  NERF_CHECK_ARG(!P.rgb || rgb, "nerf_mlp_forward: rgb output required for this MLP");
  if (rays_o) {
    NERF_CHECK_ARG(rays_d && t_vals && samples_per_ray > 0 && M % samples_per_ray == 0,
                   "nerf_mlp_forward: ray input needs rays_d, t_vals, samples_per_ray | M");
  } else {
    NERF_CHECK_ARG(xyz && (dirs || !P.rgb), "nerf_mlp_forward: xyz/dirs input required");
  }
  if (M == 0) return NERF_OK;
  hipStream_t s = as_stream(stream);
  int rc = nerf_dispatch_fwd(P, packed, M, xyz, dirs, rays_o, rays_d, t_vals, samples_per_ray,
                             rgb, sigma, stash, s);
  NERF_CHECK_ARG(rc == 0, "nerf_mlp_forward: no kernel instantiated for D=%d W=%d skip=%d "
                          "L_xyz=%d L_dir=%d precision=%d", P.D, P.W, P.skip, P.LX, P.LD,
                 precision);
  NERF_LAUNCH_CHECK("nerf_mlp_forward");
  return NERF_OK;
}

// Instantiated topologies: the reference NeRF MLP, the small finite-difference
// check MLP and the proposal density MLP.
int nerf_dispatch_fwd(const MlpPlan& P, const void* packed, int64_t M, const float* xyz,
                      const float* dirs, const float* ro, const float* rd, const float* tv,
                      int spr, float* rgb, float* sigma, void* stash, hipStream_t s) {
#define NERF_FWD_INST(W_, D_, S_, LX_, LD_)                                                   \
  if (P.W == W_ && P.D == D_ && P.skip == S_ && P.LX == LX_ && P.LD == LD_) {                \
    /* the stash layout (MlpPlan::sblk) must be the one the kernel writes */                 \
    if (P.prec == NERF_PREC_F32)                                                              \
      return P.sblk != 1 ? -1 : launch_fwd_f32<W_, D_, S_, LX_, LD_>(P, packed, M, xyz, dirs, \
                                                  ro, rd, tv, spr, rgb, sigma, stash, s);     \
    if constexpr (W_ == 256 && LX_ == 10 && LD_ == 4) {                                       \
      if (P.sblk != 32) return -1;                                                            \
      launch_fwd_bf16ws8<W_, D_, S_, LX_, LD_>(P, packed, M, xyz, dirs, ro, rd, tv, spr, rgb,   \
                                               sigma, stash, s);                              \
      return 0;                                                                               \
    }                                                                                         \
    if (P.sblk != 32) return -1;                                                              \
    return launch_fwd_bf16<W_, D_, S_, LX_, LD_>(P, packed, M, xyz, dirs, ro, rd, tv, spr,    \
                                                 rgb, sigma, stash, s);                       \
  }
  NERF_MLP_INSTANCES(NERF_FWD_INST)
#undef NERF_FWD_INST
  return -1;
}
