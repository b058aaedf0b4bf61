// Alpha compositing: post_process_model_output / compute_weights / sigma_to_alpha
// (nerf/utils/ray_utils.py:408-551) and its backward.
//
// One wave64 per ray; lane l owns the P = ceil(S/64) consecutive samples
// [l*P, l*P + P).  With delta_i = t_{i+1} - t_i (1e10 for the last sample),
// alpha_i = 1 - exp(-sigma_i delta_i) and the transmittance factor
// f_i = (1 - alpha_i) + 1e-10 (ray_utils.py:480), the transmittance
// T_i = prod_{j<i} f_j is computed as exp of an EXCLUSIVE LOG-SPACE SCAN
// sum_{j<i} log f_j, carried as a compensated (hi, lo) pair both inside a
// lane and across lanes, so its rounding error does not grow with S.
// Reductions over the ray (rgb, depth, acc) are wave reductions; no atomics.
#include "common.h"

namespace {

constexpr int kMaxP = 8;  // S <= 512

template <typename T>
__device__ __forceinline__ T tlog(T x);
template <>
__device__ __forceinline__ float tlog<float>(float x) { return logf(x); }
template <>
__device__ __forceinline__ double tlog<double>(double x) { return log(x); }
template <typename T>
__device__ __forceinline__ T texp(T x);
template <>
__device__ __forceinline__ float texp<float>(float x) { return expf(x); }
template <>
__device__ __forceinline__ double texp<double>(double x) { return exp(x); }

// Exclusive wave-wide scan of compensated pairs: returns the sum over lanes < lane.
template <typename T>
__device__ __forceinline__ void wave_exclusive_scan_dd(T hi, T lo, int lane, T& ehi, T& elo) {
  // inclusive Hillis-Steele
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T oh = shfl_up(hi, d), ol = shfl_up(lo, d);
    if (lane >= d) dd_add(oh, ol, hi, lo, hi, lo);
  }
  // shift by one lane
  ehi = shfl_up(hi, 1);
  elo = shfl_up(lo, 1);
  if (lane == 0) {
    ehi = T(0);
    elo = T(0);
  }
}

// Per-lane forward quantities of one ray.
template <typename T>
struct RayState {
  T t[kMaxP], sig[kMaxP], dlt[kMaxP], e[kMaxP], alpha[kMaxP], f[kMaxP], Tr[kMaxP];
};

// Computes delta, exp(-sigma delta), alpha, the factor f and the transmittance
// T for the P samples of this lane.  Samples beyond S get alpha = 0, f = 1.
template <typename T, int P>
__device__ __forceinline__ void ray_forward(const T* __restrict__ sig_r, const T* __restrict__ t_r,
                                            int S, int lane, RayState<T>& st) {
  const int i0 = lane * P;
  T lh = T(0), ll = T(0);  // compensated local sum of log f
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int i = i0 + p;
    if (i < S) {
      st.t[p] = t_r[i];
      st.sig[p] = sig_r[i];
      // diffs = t[1:] - t[:-1], concat 1e10                 (ray_utils.py:462-468)
      st.dlt[p] = (i + 1 < S) ? fsub(t_r[i + 1], st.t[p]) : T(1e10);
      // alpha = 1 - exp(-sigma * diffs)                      (ray_utils.py:423)
      st.e[p] = texp<T>(-fmul(st.sig[p], st.dlt[p]));
      st.alpha[p] = fsub(T(1), st.e[p]);
      // 1 - alpha + 1e-10                                    (ray_utils.py:480)
      st.f[p] = fadd(fsub(T(1), st.alpha[p]), T(1e-10));
    } else {
      st.t[p] = T(0);
      st.sig[p] = T(0);
      st.dlt[p] = T(0);
      st.e[p] = T(1);
      st.alpha[p] = T(0);
      st.f[p] = T(1);
    }
  }
  // local partial sums (lane-exclusive prefix inside the lane after the scan)
  T pre_h[kMaxP], pre_l[kMaxP];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    pre_h[p] = lh;
    pre_l[p] = ll;
    dd_add(lh, ll, tlog<T>(st.f[p]), T(0), lh, ll);
  }
  T eh, el;
  wave_exclusive_scan_dd(lh, ll, lane, eh, el);
#pragma unroll
  for (int p = 0; p < P; ++p) {
    T sh, sl;
    dd_add(eh, el, pre_h[p], pre_l[p], sh, sl);
    st.Tr[p] = texp<T>(fadd(sh, sl));
  }
}

template <typename T, int P>
__global__ __launch_bounds__(256) void composite_fwd_kernel(
    int64_t n_rays, int S, int white_bg, const T* __restrict__ sigma, const T* __restrict__ rgb,
    const T* __restrict__ t_vals, T* __restrict__ weights, T* __restrict__ pred_rgb,
    T* __restrict__ pred_depth, T* __restrict__ acc_map) {
  const int lane = threadIdx.x & 63;
  const int64_t ray = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (ray >= n_rays) return;  // wave-uniform
  const T* sig_r = sigma + ray * S;
  const T* t_r = t_vals + ray * S;
  const T* rgb_r = rgb ? rgb + ray * S * 3 : nullptr;
  RayState<T> st;
  ray_forward<T, P>(sig_r, t_r, S, lane, st);
  T cr = T(0), cg = T(0), cb = T(0), dep = T(0), acc = T(0);
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int i = lane * P + p;
    if (i < S) {
      const T w = fmul(st.alpha[p], st.Tr[p]);  // ray_utils.py:480
      weights[ray * S + i] = w;
      if (rgb) {
        cr = fadd(cr, fmul(w, rgb_r[3 * i + 0]));  // ray_utils.py:534
        cg = fadd(cg, fmul(w, rgb_r[3 * i + 1]));
        cb = fadd(cb, fmul(w, rgb_r[3 * i + 2]));
      }
      dep = fadd(dep, fmul(st.t[p], w));  // ray_utils.py:537
      acc = fadd(acc, w);                 // ray_utils.py:540
    }
  }
  cr = wave_sum(cr);
  cg = wave_sum(cg);
  cb = wave_sum(cb);
  dep = wave_sum(dep);
  acc = wave_sum(acc);
  if (lane == 0) {
    if (white_bg) {  // pred_rgb + (1 - acc)                   (ray_utils.py:542-544)
      const T bg = fsub(T(1), acc);
      cr = fadd(cr, bg);
      cg = fadd(cg, bg);
      cb = fadd(cb, bg);
    }
    pred_rgb[3 * ray + 0] = cr;
    pred_rgb[3 * ray + 1] = cg;
    pred_rgb[3 * ray + 2] = cb;
    pred_depth[ray] = dep;
    acc_map[ray] = acc;
  }
}

// Backward.  With G_i = dL/dw_i (all routes: weights, rgb, depth, acc, white
// background) and R_i = sum_{k>i} G_k w_k:
//   dL/dalpha_i = G_i T_i - R_i / f_i     (T_k, k>i, holds the factor f_i)
//   dL/dsigma_i = dL/dalpha_i * delta_i * exp(-sigma_i delta_i)
//   dL/drgb_i   = g_rgb * w_i
// The suffix sum R is a reverse exclusive wave scan.
template <typename T, int P>
__global__ __launch_bounds__(256) void composite_bwd_kernel(
    int64_t n_rays, int S, int white_bg, const T* __restrict__ sigma, const T* __restrict__ rgb,
    const T* __restrict__ t_vals, const T* __restrict__ g_rgb, const T* __restrict__ g_depth,
    const T* __restrict__ g_acc, const T* __restrict__ g_w, T* __restrict__ d_sigma,
    T* __restrict__ d_rgb) {
  const int lane = threadIdx.x & 63;
  const int64_t ray = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (ray >= n_rays) return;
  const T* sig_r = sigma + ray * S;
  const T* t_r = t_vals + ray * S;
  const T* rgb_r = rgb ? rgb + ray * S * 3 : nullptr;
  RayState<T> st;
  ray_forward<T, P>(sig_r, t_r, S, lane, st);
  const T gr = g_rgb[3 * ray + 0], gg = g_rgb[3 * ray + 1], gb = g_rgb[3 * ray + 2];
  const T gd = g_depth ? g_depth[ray] : T(0);
  const T ga = g_acc ? g_acc[ray] : T(0);
  const T gbg = white_bg ? fadd(fadd(gr, gg), gb) : T(0);
  T G[kMaxP], w[kMaxP];
  T lsum = T(0);
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int i = lane * P + p;
    w[p] = fmul(st.alpha[p], st.Tr[p]);
    G[p] = T(0);
    if (i < S) {
      T g = rgb_r ? fadd(fadd(fmul(gr, rgb_r[3 * i]), fmul(gg, rgb_r[3 * i + 1])), fmul(gb, rgb_r[3 * i + 2]))
                  : T(0);
      g = fsub(g, gbg);
      g = fadd(g, fmul(gd, st.t[p]));
      g = fadd(g, ga);
      if (g_w) g = fadd(g, g_w[ray * S + i]);
      G[p] = g;
    }
    lsum = fadd(lsum, fmul(G[p], w[p]));
  }
  // reverse exclusive scan over lanes: sum over lanes > lane
  T incl = lsum;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T o = shfl_down(incl, d);
    if (lane + d < 64) incl = fadd(incl, o);
  }
  T R = shfl_down(incl, 1);  // sum over lanes > lane
  if (lane == 63) R = T(0);
#pragma unroll
  for (int p = P - 1; p >= 0; --p) {
    const int i = lane * P + p;
    if (i < S) {
      const T da = fsub(fmul(G[p], st.Tr[p]), fdiv(R, st.f[p]));
      d_sigma[ray * S + i] = fmul(fmul(da, st.dlt[p]), st.e[p]);
      if (d_rgb) {
        d_rgb[(ray * S + i) * 3 + 0] = fmul(gr, w[p]);
        d_rgb[(ray * S + i) * 3 + 1] = fmul(gg, w[p]);
        d_rgb[(ray * S + i) * 3 + 2] = fmul(gb, w[p]);
      }
    }
    R = fadd(R, fmul(G[p], w[p]));
  }
}

template <typename T>
int32_t launch_fwd(int64_t n_rays, int S, int white_bg, const T* sigma, const T* rgb,
                   const T* t_vals, T* weights, T* pred_rgb, T* pred_depth, T* acc_map,
                   void* stream, const char* name) {
  NERF_CHECK_ARG(n_rays >= 0 && S >= 1 && S <= 64 * kMaxP, "%s: S=%d must be in [1, %d]", name,
                 S, 64 * kMaxP);
  if (n_rays == 0) return NERF_OK;  // empty batch: nothing to read or write
  NERF_CHECK_ARG(sigma && t_vals && weights && pred_rgb && pred_depth && acc_map,
                 "%s: null pointer", name);
  if (n_rays == 0) return NERF_OK;
  const int P = (S + 63) / 64;
  dim3 grid((unsigned)((n_rays + 3) / 4)), block(256);
  hipStream_t s = as_stream(stream);
#define NERF_FWD_CASE(PP)                                                                      \
  case PP:                                                                                     \
    composite_fwd_kernel<T, PP><<<grid, block, 0, s>>>(n_rays, S, white_bg, sigma, rgb, t_vals, \
                                                       weights, pred_rgb, pred_depth, acc_map); \
    break;
  switch (P) {
    NERF_FWD_CASE(1)
    NERF_FWD_CASE(2)
    NERF_FWD_CASE(3)
    NERF_FWD_CASE(4)
    NERF_FWD_CASE(5)
    NERF_FWD_CASE(6)
    NERF_FWD_CASE(7)
    NERF_FWD_CASE(8)
  }
#undef NERF_FWD_CASE
  NERF_LAUNCH_CHECK(name);
  return NERF_OK;
}

template <typename T>
int32_t launch_bwd(int64_t n_rays, int S, int white_bg, const T* sigma, const T* rgb,
                   const T* t_vals, const T* g_rgb, const T* g_depth, const T* g_acc,
                   const T* g_w, T* d_sigma, T* d_rgb, void* stream, const char* name) {
  NERF_CHECK_ARG(n_rays >= 0 && S >= 1 && S <= 64 * kMaxP, "%s: S=%d must be in [1, %d]", name,
                 S, 64 * kMaxP);
  if (n_rays == 0) return NERF_OK;
  NERF_CHECK_ARG(sigma && t_vals && g_rgb && d_sigma && (d_rgb || !rgb), "%s: null pointer", name);
  if (n_rays == 0) return NERF_OK;
  const int P = (S + 63) / 64;
  dim3 grid((unsigned)((n_rays + 3) / 4)), block(256);
  hipStream_t s = as_stream(stream);
#define NERF_BWD_CASE(PP)                                                                    \
  case PP:                                                                                   \
    composite_bwd_kernel<T, PP><<<grid, block, 0, s>>>(n_rays, S, white_bg, sigma, rgb,      \
                                                       t_vals, g_rgb, g_depth, g_acc, g_w,   \
                                                       d_sigma, d_rgb);                      \
    break;
  switch (P) {
    NERF_BWD_CASE(1)
    NERF_BWD_CASE(2)
    NERF_BWD_CASE(3)
    NERF_BWD_CASE(4)
    NERF_BWD_CASE(5)
    NERF_BWD_CASE(6)
    NERF_BWD_CASE(7)
    NERF_BWD_CASE(8)
  }
#undef NERF_BWD_CASE
  NERF_LAUNCH_CHECK(name);
  return NERF_OK;
}

}  // namespace

extern "C" int32_t nerf_composite_forward(int64_t n_rays, int32_t S, int32_t white_bg,
                                          const float* sigma, const float* sample_rgb,
                                          const float* t_vals, float* weights, float* pred_rgb,
                                          float* pred_depth, float* acc_map, void* stream) {
  return launch_fwd<float>(n_rays, S, white_bg, sigma, sample_rgb, t_vals, weights, pred_rgb,
                           pred_depth, acc_map, stream, "nerf_composite_forward");
}

extern "C" int32_t nerf_composite_forward_f64(int64_t n_rays, int32_t S, int32_t white_bg,
                                              const double* sigma, const double* sample_rgb,
                                              const double* t_vals, double* weights,
                                              double* pred_rgb, double* pred_depth,
                                              double* acc_map, void* stream) {
  return launch_fwd<double>(n_rays, S, white_bg, sigma, sample_rgb, t_vals, weights, pred_rgb,
                            pred_depth, acc_map, stream, "nerf_composite_forward_f64");
}

extern "C" int32_t nerf_composite_backward(int64_t n_rays, int32_t S, int32_t white_bg,
                                           const float* sigma, const float* sample_rgb,
                                           const float* t_vals, const float* g_rgb,
                                           const float* g_depth, const float* g_acc,
                                           const float* g_weights, float* d_sigma, float* d_rgb,
                                           void* stream) {
  return launch_bwd<float>(n_rays, S, white_bg, sigma, sample_rgb, t_vals, g_rgb, g_depth, g_acc,
                           g_weights, d_sigma, d_rgb, stream, "nerf_composite_backward");
}

extern "C" int32_t nerf_composite_backward_f64(int64_t n_rays, int32_t S, int32_t white_bg,
                                               const double* sigma, const double* sample_rgb,
                                               const double* t_vals, const double* g_rgb,
                                               const double* g_depth, const double* g_acc,
                                               const double* g_weights, double* d_sigma,
                                               double* d_rgb, void* stream) {
  return launch_bwd<double>(n_rays, S, white_bg, sigma, sample_rgb, t_vals, g_rgb, g_depth, g_acc,
                            g_weights, d_sigma, d_rgb, stream, "nerf_composite_backward_f64");
}
