// Learned-proposal support: the histogram upper-bound ("interlevel") loss of
// mip-NeRF 360 (Barron et al. 2022) between the proposal's weights on its
// own bins and the weights the fine model produces, and the per-ray missed mass
// that decides the fallback to the plain coarse pass.
//
// For fine interval i = [t_i, t_{i+1}] (the last one ends at `far`), the
// proposal's outer measure is bound_i = sum of w_p over the proposal bins that
// overlap it, widened by `dilate` bins on each side (a blurred histogram; 0 is
// the plain outer measure), summed directly over the bin range (a difference
// of two float32 cdf values would cancel for the many tiny fine weights).  missed_i = max(0, w_i - bound_i); the ray's missed mass is
// sum_i missed_i and its loss sum_i missed_i^2 / (w_i + 1e-7), with w_i held
// constant.  d loss / d w_p[k] = sum over intervals i whose bin range holds k of
// -2 missed_i / (w_i + 1e-7), summed by one lane per bin in a fixed order.
#include "common.h"

namespace {

constexpr int kW = 4;  // waves (rays) per block
constexpr int kMaxSp = 64, kMaxSf = 256;

__global__ __launch_bounds__(256) void interlevel_kernel(
    int64_t n_rays, int Sp, int Sf, const float* __restrict__ edges, const float* __restrict__ wp,
    const float* __restrict__ tf, const float* __restrict__ wf, const float* __restrict__ far,
    float* __restrict__ missed, float* __restrict__ loss, float* __restrict__ g_wp, float gscale,
    int dilate) {
  __shared__ float s_e[kW][kMaxSp + 1], s_w[kW][kMaxSp], s_coef[kW][kMaxSf];
  __shared__ int s_lo[kW][kMaxSf], s_hi[kW][kMaxSf];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t ray = (int64_t)blockIdx.x * kW + w;
  if (ray >= n_rays) return;
  // proposal edges and weights
  if (lane <= Sp) s_e[w][lane] = edges[ray * (Sp + 1) + lane];
  if (lane < Sp) s_w[w][lane] = wp[ray * Sp + lane];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const float fr = far[ray];
  // this lane's fine intervals i = lane + 64 it, loaded together up front
  constexpr int kIt = kMaxSf / 64;
  float lo_v[kIt], hi_v[kIt], w_v[kIt];
#pragma unroll
  for (int it = 0; it < kIt; ++it) {
    const int i = lane + 64 * it;
    lo_v[it] = i < Sf ? tf[ray * Sf + i] : 0.f;
    hi_v[it] = i + 1 < Sf ? tf[ray * Sf + i + 1] : fr;
    w_v[it] = i < Sf ? wf[ray * Sf + i] : 0.f;
  }
  float ms = 0.f, ls = 0.f;
#pragma unroll
  for (int it = 0; it < kIt; ++it) {
    const int i = lane + 64 * it;
    if (i >= Sf) continue;
    const float lo_t = lo_v[it];
    const float hi_t = hi_v[it];
    // k_lo: last edge <= lo_t ; k_hi: first edge >= hi_t
    int a = 0, b = Sp + 1;
    while (a < b) {
      const int mid = (a + b) >> 1;
      if (s_e[w][mid] <= lo_t) a = mid + 1; else b = mid;
    }
    int klo = a - 1;
    klo = klo < 0 ? 0 : (klo > Sp - 1 ? Sp - 1 : klo);
    a = 0; b = Sp + 1;
    while (a < b) {
      const int mid = (a + b) >> 1;
      if (s_e[w][mid] < hi_t) a = mid + 1; else b = mid;
    }
    int khi = a;
    khi = khi > Sp ? Sp : (khi < klo + 1 ? klo + 1 : khi);
    klo = klo - dilate < 0 ? 0 : klo - dilate;
    khi = khi + dilate > Sp ? Sp : khi + dilate;
    float bound = 0.f;
    for (int k = klo; k < khi; ++k) bound += s_w[w][k];
    const float wi = w_v[it];
    const float mi = fmaxf(wi - bound, 0.f);
    ms += mi;
    ls += mi * mi / (wi + 1e-7f);
    s_coef[w][i] = -2.f * mi / (wi + 1e-7f);
    s_lo[w][i] = klo;
    s_hi[w][i] = khi;
  }
  ms = wave_sum(ms);
  ls = wave_sum(ls);
  if (lane == 0) {
    if (missed) missed[ray] = ms;
    if (loss) loss[ray] = ls;
  }
  if (!g_wp) return;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (lane < Sp) {
    float g = 0.f;
#pragma unroll 8
    for (int i = 0; i < Sf; ++i)
      if (s_lo[w][i] <= lane && lane < s_hi[w][i]) g += s_coef[w][i];
    g_wp[ray * Sp + lane] = g * gscale;
  }
}

}  // namespace

extern "C" int32_t nerf_proposal_interlevel(int64_t n_rays, int32_t n_prop, int32_t n_fine,
                                            const float* prop_edges, const float* prop_weights,
                                            const float* t_fine, const float* fine_weights,
                                            const float* far, float* missed, float* loss,
                                            float* g_prop_weights, float grad_scale, int32_t dilate,
                                            void* stream) {
  NERF_CHECK_ARG(n_rays >= 0 && n_prop >= 1 && n_prop <= kMaxSp && n_fine >= 1 && n_fine <= kMaxSf,
                 "nerf_proposal_interlevel: n_prop=%d (1..%d) n_fine=%d (1..%d)", n_prop, kMaxSp,
                 n_fine, kMaxSf);
  if (n_rays == 0) return NERF_OK;  // empty batch
  NERF_CHECK_ARG(prop_edges && prop_weights && t_fine && fine_weights && far,
                 "nerf_proposal_interlevel: null pointer");
  NERF_CHECK_ARG(dilate >= 0 && dilate < n_prop, "nerf_proposal_interlevel: dilate=%d", dilate);
  if (n_rays == 0) return NERF_OK;
  interlevel_kernel<<<(unsigned)((n_rays + kW - 1) / kW), 64 * kW, 0, as_stream(stream)>>>(
      n_rays, n_prop, n_fine, prop_edges, prop_weights, t_fine, fine_weights, far, missed, loss,
      g_prop_weights, grad_scale, dilate);
  NERF_LAUNCH_CHECK("nerf_proposal_interlevel");
  return NERF_OK;
}
