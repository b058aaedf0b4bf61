// Loss gradient, Adam (TF ApplyAdam form) and the finite check of the loss scaler.
//   MeanSquaredError        nerf/core/model.py:32, 157-168
//   Adam + ExponentialDecay nerf/core/model.py:413-418
#include "common.h"

namespace {

// One block: deterministic reduction of sum((pred - true)^2), plus the gradient.
__global__ __launch_bounds__(1024) void mse_kernel(int64_t n, const float* __restrict__ pred,
                                                   const float* __restrict__ truth, float scale,
                                                   float* __restrict__ g, float* __restrict__ loss) {
  __shared__ float part[1024 / 64];
  const float inv = 1.0f / (float)(3 * n);
  float acc = 0.f;
  for (int64_t i = threadIdx.x; i < 3 * n; i += blockDim.x) {
    const float d = pred[i] - truth[i];
    acc += d * d;
    if (g) g[i] = scale * (2.0f * d * inv);
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.f;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += part[w];
    if (loss) *loss = s * inv;
  }
}

__global__ void adam_kernel(int64_t n, float* __restrict__ var, float* __restrict__ m,
                            float* __restrict__ v, const float* __restrict__ grad, float lr_t,
                            float omb1, float omb2, float eps, const float* __restrict__ inv_scale,
                            const int32_t* __restrict__ skip) {
  if (skip && *skip) return;
  const float s = inv_scale ? *inv_scale : 1.0f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float gr = inv_scale ? fmul(grad[i], s) : grad[i];
    float mi = m[i], vi = v[i];
    mi = fadd(mi, fmul(fsub(gr, mi), omb1));                 // m += (g - m) * (1 - b1)
    vi = fadd(vi, fmul(fsub(fmul(gr, gr), vi), omb2));       // v += (g^2 - v) * (1 - b2)
    m[i] = mi;
    v[i] = vi;
    var[i] = fsub(var[i], fdiv(fmul(mi, lr_t), fadd(fsqrt(vi), eps)));
  }
}

__global__ void finite_kernel(int64_t n, const float* __restrict__ x, int32_t* __restrict__ flag) {
  int bad = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    bad |= !isfinite(x[i]);
  if (__any(bad) && (threadIdx.x & 63) == 0) *flag = 1;
}

// Device-side loss-scaled step state (layout of nerf_step_state).
struct StepState {
  float scale, inv_scale;
  int32_t overflow, good, iterations, skipped;
  float lr_t, omb1, omb2;
  int32_t reserved[7];
};
static_assert(sizeof(StepState) == 64, "nerf_step_state is 64 bytes");

// float32(base ** e) with the power evaluated in float64 and rounded once.
__device__ float pow_f32(float base, float e) { return (float)pow((double)base, (double)e); }

__global__ void step_begin_kernel(StepState* st, float lr0, float decay_steps, float rate, float b1, float b2) {
  if (threadIdx.x != 0) return;
  st->overflow = 0;
  st->inv_scale = fdiv(1.0f, st->scale);
  const int it = st->iterations;
  const float lr = fmul(lr0, pow_f32(rate, fdiv((float)it, decay_steps)));
  const float t = (float)(it + 1);
  const float b1p = pow_f32(b1, t), b2p = pow_f32(b2, t);
  st->lr_t = fdiv(fmul(lr, fsqrt(fsub(1.0f, b2p))), fsub(1.0f, b1p));
  st->omb1 = fsub(1.0f, b1);
  st->omb2 = fsub(1.0f, b2);
}

__global__ void adam_state_kernel(int64_t n, float* __restrict__ var, float* __restrict__ m,
                                  float* __restrict__ v, const float* __restrict__ grad, float eps,
                                  const StepState* __restrict__ st) {
  if (st->overflow) return;
  const float s = st->inv_scale, lr_t = st->lr_t, omb1 = st->omb1, omb2 = st->omb2;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float gr = fmul(grad[i], s);
    float mi = m[i], vi = v[i];
    mi = fadd(mi, fmul(fsub(gr, mi), omb1));
    vi = fadd(vi, fmul(fsub(fmul(gr, gr), vi), omb2));
    m[i] = mi;
    v[i] = vi;
    var[i] = fsub(var[i], fdiv(fmul(mi, lr_t), fadd(fsqrt(vi), eps)));
  }
}

__global__ void step_end_kernel(StepState* st, int interval, float factor, float min_scale) {
  if (threadIdx.x != 0) return;
  if (st->overflow) {
    const float s = fdiv(st->scale, factor);
    st->scale = s > min_scale ? s : min_scale;
    st->good = 0;
    st->skipped += 1;
  } else {
    st->iterations += 1;
    st->good += 1;
    if (st->good >= interval) {
      st->scale = fmul(st->scale, factor);
      st->good = 0;
    }
  }
}

}  // namespace

extern "C" int32_t nerf_step_begin(nerf_step_state* state, float lr0, int32_t decay_steps, float decay_rate,
                                   float b1, float b2, void* stream) {
  NERF_CHECK_ARG(state && decay_steps > 0, "nerf_step_begin: bad arguments");
  step_begin_kernel<<<1, 64, 0, as_stream(stream)>>>(reinterpret_cast<StepState*>(state), lr0,
                                                    (float)decay_steps, decay_rate, b1, b2);
  NERF_LAUNCH_CHECK("nerf_step_begin");
  return NERF_OK;
}

extern "C" int32_t nerf_adam_tf_state(int64_t n, float* var, float* m, float* v, const float* grad, float eps,
                                      const nerf_step_state* state, void* stream) {
  NERF_CHECK_ARG(n >= 0 && var && m && v && grad && state, "nerf_adam_tf_state: bad arguments");
  if (n == 0) return NERF_OK;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  adam_state_kernel<<<(unsigned)blocks, 256, 0, as_stream(stream)>>>(
      n, var, m, v, grad, eps, reinterpret_cast<const StepState*>(state));
  NERF_LAUNCH_CHECK("nerf_adam_tf_state");
  return NERF_OK;
}

extern "C" int32_t nerf_step_end(nerf_step_state* state, int32_t interval, float factor, float min_scale,
                                 void* stream) {
  NERF_CHECK_ARG(state && interval > 0 && factor > 0.f, "nerf_step_end: bad arguments");
  step_end_kernel<<<1, 64, 0, as_stream(stream)>>>(reinterpret_cast<StepState*>(state), interval, factor,
                                                  min_scale);
  NERF_LAUNCH_CHECK("nerf_step_end");
  return NERF_OK;
}

extern "C" int32_t nerf_mse_grad(int64_t n_rays, const float* pred_rgb, const float* true_rgb,
                                 float scale, float* g_rgb, float* loss_out, void* stream) {
  NERF_CHECK_ARG(n_rays > 0 && pred_rgb && true_rgb, "nerf_mse_grad: bad arguments");
  mse_kernel<<<1, 1024, 0, as_stream(stream)>>>(n_rays, pred_rgb, true_rgb, scale, g_rgb, loss_out);
  NERF_LAUNCH_CHECK("nerf_mse_grad");
  return NERF_OK;
}

extern "C" int32_t nerf_adam_tf(int64_t n, float* var, float* m, float* v, const float* grad,
                                float lr_t, float one_minus_b1, float one_minus_b2, float eps,
                                const float* inv_scale, const int32_t* skip_flag, void* stream) {
  NERF_CHECK_ARG(n >= 0 && var && m && v && grad, "nerf_adam_tf: bad arguments");
  if (n == 0) return NERF_OK;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  adam_kernel<<<(unsigned)blocks, 256, 0, as_stream(stream)>>>(n, var, m, v, grad, lr_t, one_minus_b1,
                                                              one_minus_b2, eps, inv_scale, skip_flag);
  NERF_LAUNCH_CHECK("nerf_adam_tf");
  return NERF_OK;
}

extern "C" int32_t nerf_check_finite(int64_t n, const float* x, int32_t* flag, void* stream) {
  NERF_CHECK_ARG(n >= 0 && x && flag, "nerf_check_finite: bad arguments");
  if (n == 0) return NERF_OK;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  finite_kernel<<<(unsigned)blocks, 256, 0, as_stream(stream)>>>(n, x, flag);
  NERF_LAUNCH_CHECK("nerf_check_finite");
  return NERF_OK;
}
