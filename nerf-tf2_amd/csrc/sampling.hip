// Ray generation and stratified coarse sampling.
//   get_rays / get_rays_tf            nerf/utils/ray_utils.py:6-106
//   create_input_batch_coarse_model   nerf/utils/ray_utils.py:137-274
// Every float op is an explicitly rounded intrinsic, so the results do not
// depend on contraction: bin edges and t values are reproduced bit for bit.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "common.h"

static thread_local char g_err[512] = "";

void nerf_set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

extern "C" const char* nerf_last_error(void) { return g_err; }
extern "C" int32_t nerf_abi_version(void) { return 2; }

// ---------------------------------------------------------------------------
// get_rays: float64 pinhole rays, cast to float32.
// ---------------------------------------------------------------------------
struct Cam64 {
  double K[9];
  double c2w[16];
};
struct Cam32 {
  float K[9];
  float c2w[16];
};

__global__ void get_rays_f64_kernel(int H, int W, Cam64 cam, float* __restrict__ ro,
                                    float* __restrict__ rd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)H * W) return;
  const int v = (int)(i / W), u = (int)(i % W);
  // (u - cx) / fx, (v - cy) / fy, z = 1                      (ray_utils.py:37-39)
  const double x = fdiv(fsub((double)u, cam.K[2]), cam.K[0]);
  const double y = fdiv(fsub((double)v, cam.K[5]), cam.K[4]);
  double d[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {  // R @ dir                   (pose_utils.py:104)
    const double* R = cam.c2w + 4 * r;
    d[r] = fadd(fadd(fmul(R[0], x), fmul(R[1], y)), R[2]);
  }
  // / (|d| + 1e-8)                                           (pose_utils.py:71-73)
  const double mag = sqrt(fadd(fadd(fmul(d[0], d[0]), fmul(d[1], d[1])), fmul(d[2], d[2])));
  const double den = fadd(mag, 1e-8);
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    rd[3 * i + r] = (float)fdiv(d[r], den);
    ro[3 * i + r] = (float)cam.c2w[4 * r + 3];
  }
}

__global__ void get_rays_f32_kernel(int H, int W, Cam32 cam, float* __restrict__ ro,
                                    float* __restrict__ rd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)H * W) return;
  const int v = (int)(i / W), u = (int)(i % W);
  const float x = fdiv(fsub((float)u, cam.K[2]), cam.K[0]);  // ray_utils.py:87
  const float y = fdiv(fsub((float)v, cam.K[5]), cam.K[4]);  // ray_utils.py:88
  float d[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {  // directions @ R^T, left to right (ray_utils.py:96-97)
    const float* R = cam.c2w + 4 * r;
    d[r] = fadd(fadd(fmul(x, R[0]), fmul(y, R[1])), fmul(1.0f, R[2]));
  }
  const float mag = fsqrt(fadd(fadd(fmul(d[0], d[0]), fmul(d[1], d[1])), fmul(d[2], d[2])));
  const float den = fadd(mag, 1e-8f);  // ray_utils.py:100-101
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    rd[3 * i + r] = fdiv(d[r], den);
    ro[3 * i + r] = cam.c2w[4 * r + 3];
  }
}

extern "C" int32_t nerf_get_rays(int32_t H, int32_t W, const double* K, const double* c2w,
                                 float* rays_o, float* rays_d, void* stream) {
  NERF_CHECK_ARG(H > 0 && W > 0, "nerf_get_rays: H=%d W=%d must be positive", H, W);
  NERF_CHECK_ARG(K && c2w && rays_o && rays_d, "nerf_get_rays: null pointer");
  Cam64 cam;
  memcpy(cam.K, K, sizeof(cam.K));
  memcpy(cam.c2w, c2w, sizeof(cam.c2w));
  const int64_t n = (int64_t)H * W;
  get_rays_f64_kernel<<<(unsigned)((n + 255) / 256), 256, 0, as_stream(stream)>>>(H, W, cam, rays_o,
                                                                                rays_d);
  NERF_LAUNCH_CHECK("nerf_get_rays");
  return NERF_OK;
}

extern "C" int32_t nerf_get_rays_tf(int32_t H, int32_t W, const float* K, const float* c2w,
                                    float* rays_o, float* rays_d, void* stream) {
  NERF_CHECK_ARG(H > 0 && W > 0, "nerf_get_rays_tf: H=%d W=%d must be positive", H, W);
  NERF_CHECK_ARG(K && c2w && rays_o && rays_d, "nerf_get_rays_tf: null pointer");
  Cam32 cam;
  memcpy(cam.K, K, sizeof(cam.K));
  memcpy(cam.c2w, c2w, sizeof(cam.c2w));
  const int64_t n = (int64_t)H * W;
  get_rays_f32_kernel<<<(unsigned)((n + 255) / 256), 256, 0, as_stream(stream)>>>(H, W, cam, rays_o,
                                                                                rays_d);
  NERF_LAUNCH_CHECK("nerf_get_rays_tf");
  return NERF_OK;
}

// ---------------------------------------------------------------------------
// Stratified coarse sampling: one thread per (ray, bin).
// ---------------------------------------------------------------------------
// Edge k of tf.linspace(start, stop, n+1) (TF linspace_nd): endpoints copied,
// interior start + delta * k with delta = (stop - start) / n.
__device__ __forceinline__ float linspace_at(float start, float stop, float delta, int k, int n) {
  if (k == 0) return start;
  if (k == n) return stop;
  return fadd(start, fmul(delta, (float)k));
}

__global__ void coarse_kernel(int64_t n_rays, int nc, int inv, int perturb,
                              const float* __restrict__ ro, const float* __restrict__ rd,
                              const float* __restrict__ near, const float* __restrict__ far,
                              const float* __restrict__ u_vals, float* __restrict__ edges,
                              float* __restrict__ widths, float* __restrict__ t_vals,
                              float* __restrict__ xyz, float* __restrict__ dirs) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= n_rays * nc) return;
  const int64_t r = gid / nc;
  const int k = (int)(gid % nc);
  const float n0 = near[r], f0 = far[r];
  float start = n0, stop = f0;
  if (inv) {  // 1 / near, 1 / far                          (ray_utils.py:200-203)
    start = fdiv(1.0f, n0);
    stop = fdiv(1.0f, f0);
  }
  const float delta = fdiv(fsub(stop, start), (float)nc);
  float left = linspace_at(start, stop, delta, k, nc);
  float right = linspace_at(start, stop, delta, k + 1, nc);
  if (inv) {
    left = fdiv(1.0f, left);
    right = fdiv(1.0f, right);
  }
  const float width = fsub(right, left);  // ray_utils.py:226
  float t;
  if (perturb)
    t = fadd(left, fmul(u_vals[gid], width));  // ray_utils.py:234
  else
    t = fmul(0.5f, fadd(left, right));  // ray_utils.py:243
  edges[r * (nc + 1) + k] = left;
  if (k == nc - 1) edges[r * (nc + 1) + nc] = right;
  widths[gid] = width;
  t_vals[gid] = t;
  if (xyz) {  // o + t * d                                     (ray_utils.py:251-258)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float d = rd[3 * r + c];
      xyz[3 * gid + c] = fadd(ro[3 * r + c], fmul(t, d));
      dirs[3 * gid + c] = d;
    }
  }
}

extern "C" int32_t nerf_create_input_batch_coarse(int64_t n_rays, int32_t n_coarse,
                                                  int32_t lin_inv_depth, int32_t perturb,
                                                  const float* rays_o, const float* rays_d,
                                                  const float* near, const float* far,
                                                  const float* u_vals, float* bin_edges,
                                                  float* bin_widths, float* t_vals,
                                                  float* xyz_inputs, float* dir_inputs,
                                                  void* stream) {
  NERF_CHECK_ARG(n_rays >= 0 && n_coarse > 0, "nerf_create_input_batch_coarse: bad sizes");
  if (n_rays == 0) return NERF_OK;  // empty batch
  NERF_CHECK_ARG(near && far && bin_edges && bin_widths && t_vals,
                 "nerf_create_input_batch_coarse: null pointer");
  NERF_CHECK_ARG(!perturb || u_vals, "nerf_create_input_batch_coarse: perturb needs u_vals");
  NERF_CHECK_ARG((xyz_inputs == nullptr) == (dir_inputs == nullptr),
                 "nerf_create_input_batch_coarse: xyz_inputs and dir_inputs go together");
  NERF_CHECK_ARG(!xyz_inputs || (rays_o && rays_d),
                 "nerf_create_input_batch_coarse: xyz_inputs needs rays");
  const int64_t n = n_rays * n_coarse;
  if (n == 0) return NERF_OK;
  coarse_kernel<<<(unsigned)((n + 255) / 256), 256, 0, as_stream(stream)>>>(
      n_rays, n_coarse, lin_inv_depth, perturb, rays_o, rays_d, near, far, u_vals, bin_edges,
      bin_widths, t_vals, xyz_inputs, dir_inputs);
  NERF_LAUNCH_CHECK("nerf_create_input_batch_coarse");
  return NERF_OK;
}

// ---------------------------------------------------------------------------
// Training ray batches from images resident on the GPU (the sample mode of
// base_dataset.py:555-621): ray b is pixel pix[b] of image img[b]; its ray is
// get_rays_tf's (ray_utils.py:53-106) for that pixel, its colour the uint8
// pixel / 255 (base_dataset.py:595-598), its bounds the image's (near, far).
// ---------------------------------------------------------------------------
__global__ void sample_batch_kernel(int64_t batch, int H, int W, const uint8_t* __restrict__ images,
                                    const float* __restrict__ poses, const float* __restrict__ Ks,
                                    const float* __restrict__ bounds, const int32_t* __restrict__ img,
                                    const int64_t* __restrict__ pix, float* __restrict__ ro,
                                    float* __restrict__ rd, float* __restrict__ near,
                                    float* __restrict__ far, float* __restrict__ rgb) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  const int i = img[b];
  const int64_t p = pix[b];
  const int v = (int)(p / W), u = (int)(p % W);
  const float* K = Ks + 9 * i;
  const float* c2w = poses + 16 * i;
  const float x = fdiv(fsub((float)u, K[2]), K[0]);
  const float y = fdiv(fsub((float)v, K[5]), K[4]);
  float d[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const float* R = c2w + 4 * r;
    d[r] = fadd(fadd(fmul(x, R[0]), fmul(y, R[1])), fmul(1.0f, R[2]));
  }
  const float mag = fsqrt(fadd(fadd(fmul(d[0], d[0]), fmul(d[1], d[1])), fmul(d[2], d[2])));
  const float den = fadd(mag, 1e-8f);
  const uint8_t* px = images + (((int64_t)i * H + v) * W + u) * 3;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    rd[3 * b + r] = fdiv(d[r], den);
    ro[3 * b + r] = c2w[4 * r + 3];
    rgb[3 * b + r] = fdiv((float)px[r], 255.0f);
  }
  near[b] = bounds[2 * i];
  far[b] = bounds[2 * i + 1];
}

extern "C" int32_t nerf_sample_batch(int64_t batch, int32_t n_images, int32_t H, int32_t W,
                                     const uint8_t* images, const float* poses, const float* Ks,
                                     const float* bounds, const int32_t* image_idx,
                                     const int64_t* pixel_idx, float* rays_o, float* rays_d,
                                     float* near, float* far, float* rgb, void* stream) {
  NERF_CHECK_ARG(batch >= 0 && n_images > 0 && H > 0 && W > 0, "nerf_sample_batch: bad sizes");
  if (batch == 0) return NERF_OK;
  NERF_CHECK_ARG(images && poses && Ks && bounds && image_idx && pixel_idx && rays_o && rays_d && near &&
                     far && rgb,
                 "nerf_sample_batch: null pointer");
  if (batch == 0) return NERF_OK;
  sample_batch_kernel<<<(unsigned)((batch + 255) / 256), 256, 0, as_stream(stream)>>>(
      batch, H, W, images, poses, Ks, bounds, image_idx, pixel_idx, rays_o, rays_d, near, far, rgb);
  NERF_LAUNCH_CHECK("nerf_sample_batch");
  return NERF_OK;
}
