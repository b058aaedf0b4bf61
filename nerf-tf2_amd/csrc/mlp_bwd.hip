// Fused MLP backward and weight gradients (placeholder).
#include "common.h"
#include "mlp_kernels.h"

extern "C" int32_t nerf_mlp_backward(const nerf_mlp_desc* desc, int32_t precision, const void* packed,
                                     int64_t M, const float* d_rgb, const float* d_sigma,
                                     const float* rgb, const float* sigma, const void* stash, void* dz,
                                     void* stream) {
  nerf_set_error("nerf_mlp_backward: not built yet");
  return NERF_ERR_UNSUPPORTED;
}
extern "C" int64_t nerf_mlp_wgrad_workspace_bytes(const nerf_mlp_desc* desc, int64_t M) { return 0; }
extern "C" int32_t nerf_mlp_weight_grads(const nerf_mlp_desc* desc, int32_t precision, int64_t M,
                                         const void* stash, const void* dz, float* grad_flat,
                                         int32_t accumulate, void* workspace, void* stream) {
  nerf_set_error("nerf_mlp_weight_grads: not built yet");
  return NERF_ERR_UNSUPPORTED;
}
