/*
 * nerf_hip.h — C ABI of the MI355X (gfx950) NeRF ray-marching hot path.
 *
 * The reference (thatbrguy/nerf-tf2) has no FFI: its hot path is a set of
 * Python functions over TensorFlow ops.  Each entry point below replaces one
 * of those functions (cited as file:line of the reference) and is what a
 * ctypes / cffi binding of that function binds.  The Python package
 * nerf-tf2_amd keeps the reference's function names, argument order, dict
 * keys and shapes on top of these entry points.
 *
 * Conventions (all entry points):
 *   - Every pointer argument is a device pointer unless stated "host".
 *   - Tensors are float32, row-major, contiguous, unless stated otherwise.
 *   - Sizes are int64; the caller allocates every output and workspace; the
 *     library allocates and frees no memory and keeps no pointers after return.
 *   - Work is enqueued on `stream` (a hipStream_t, NULL = default stream) and
 *     is asynchronous; no entry point synchronises the host.
 *   - Return 0 on success, NERF_ERR_INVALID_ARG for a bad argument (nothing is
 *     launched), NERF_ERR_LAUNCH if a kernel launch fails.  nerf_last_error()
 *     returns the calling thread's last message.
 *   - Stateless and re-entrant: calls on different streams may overlap.
 */
#ifndef NERF_HIP_H_
#define NERF_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NERF_OK 0
#define NERF_ERR_INVALID_ARG (-1)
#define NERF_ERR_LAUNCH (-2)
#define NERF_ERR_UNSUPPORTED (-3)

#define NERF_PREC_F32 0  /* MLP matmuls on v_mfma_f32_32x32x2_f32 */
#define NERF_PREC_BF16 1 /* MLP matmuls on v_mfma_f32_32x32x16_bf16, fp32 accumulate */

/* Thread-local message of the last failed call (never NULL). */
const char* nerf_last_error(void);
/* ABI version of this header; bumped on any signature change. */
int32_t nerf_abi_version(void);

/* ------------------------------------------------------------------------
 * Rays.
 * ---------------------------------------------------------------------- */

/* get_rays (nerf/utils/ray_utils.py:6-51) followed by the caller's cast to
 * float32 (nerf/core/base_dataset.py:849-852): pixel (u,v) integer coords,
 * computed in float64.  K (3x3) and c2w (4x4) are HOST row-major float64.
 * rays_o, rays_d: (H*W, 3). */
int32_t nerf_get_rays(int32_t H, int32_t W, const double* K, const double* c2w,
                      float* rays_o, float* rays_d, void* stream);

/* get_rays_tf (nerf/utils/ray_utils.py:53-106): the float32 TF form.
 * K (3x3) and c2w (4x4) are HOST row-major float32. */
int32_t nerf_get_rays_tf(int32_t H, int32_t W, const float* K, const float* c2w,
                         float* rays_o, float* rays_d, void* stream);

/* Training ray batch from images resident on the GPU: the sample mode of
 * base_dataset.py:555-621 without the host tf.data step.  images uint8
 * (n_images, H, W, 3); poses (n_images, 4, 4), Ks (n_images, 3, 3), bounds
 * (n_images, 2) float32 [near, far]; ray b is pixel pixel_idx[b] (int64,
 * row-major v*W + u) of image image_idx[b] (int32), drawn by the caller
 * (base_dataset.py:607-610 draws them with tf.random.uniform).  Outputs
 * rays_o, rays_d (batch, 3) exactly as get_rays_tf for that pixel, near, far
 * (batch, 1) and rgb (batch, 3) = pixel / 255 (base_dataset.py:595-598). */
int32_t nerf_sample_batch(int64_t batch, int32_t n_images, int32_t H, int32_t W,
                          const uint8_t* images, const float* poses, const float* Ks,
                          const float* bounds, const int32_t* image_idx, const int64_t* pixel_idx,
                          float* rays_o, float* rays_d, float* near, float* far, float* rgb,
                          void* stream);

/* ------------------------------------------------------------------------
 * create_input_batch_coarse_model (nerf/utils/ray_utils.py:137-274).
 *   near, far: (N,1); u_vals: (N, n_coarse) draws in [0,1) (ignored when
 *   perturb == 0); outputs bin_edges (N, n_coarse+1), bin_widths
 *   (N, n_coarse), t_vals (N, n_coarse); xyz_inputs / dir_inputs
 *   (N*n_coarse, 3) may be NULL (the fused MLP reads rays + t_vals).
 *   perturb == 0 gives bin mid-points and bin_widths (ray_utils.py:236-247).
 * ---------------------------------------------------------------------- */
int32_t nerf_create_input_batch_coarse(int64_t n_rays, int32_t n_coarse, int32_t lin_inv_depth,
                                       int32_t perturb, const float* rays_o, const float* rays_d,
                                       const float* near, const float* far, const float* u_vals,
                                       float* bin_edges, float* bin_widths, float* t_vals,
                                       float* xyz_inputs, float* dir_inputs, void* stream);

/* ------------------------------------------------------------------------
 * create_input_batch_fine_model (nerf/utils/ray_utils.py:276-406).
 *   bin_weights (N, n_coarse) coarse weights; bin_edges / bin_widths from
 *   nerf_create_input_batch_coarse; t_vals_coarse (N, n_coarse) sorted;
 *   u_vals (N, n_fine).  Outputs: t_vals (N, n_coarse+n_fine) sorted;
 *   optional (may be NULL): piece_idxs int32 (N, n_fine) (ray_utils.py:363),
 *   t_vals_fine (N, n_fine) in draw order (ray_utils.py:376),
 *   xyz_inputs / dir_inputs (N*(n_coarse+n_fine), 3),
 *   cdf (N, n_coarse+1): the cdf the draws are inverted against
 *   (ray_utils.py:342-345, leading 0 included).
 *   cdf_mode: NERF_CDF_TF_ORDER computes the denominator and the cdf in the
 *   reference's float32 operation order (bin indices bit-exact with it);
 *   NERF_CDF_COMPENSATED (the BF16 and learned-proposal forms) computes them
 *   with compensated summation: each cdf value within about 1 ulp of the exact
 *   prefix ratio.
 *   n_coarse <= 64, n_fine <= 256.
 * ---------------------------------------------------------------------- */
#define NERF_CDF_TF_ORDER 0
#define NERF_CDF_COMPENSATED 1
int32_t nerf_create_input_batch_fine(int64_t n_rays, int32_t n_coarse, int32_t n_fine,
                                     const float* rays_o, const float* rays_d,
                                     const float* bin_weights, const float* bin_edges,
                                     const float* bin_widths, const float* t_vals_coarse,
                                     const float* u_vals, float* t_vals, int32_t* piece_idxs,
                                     float* t_vals_fine, float* xyz_inputs, float* dir_inputs,
                                     int32_t cdf_mode, float* cdf, void* stream);

/* ------------------------------------------------------------------------
 * post_process_model_output (nerf/utils/ray_utils.py:484-551) with
 * compute_weights (:426-482) and sigma_to_alpha (:408-424).
 *   sigma (N*S,1), sample_rgb (N*S,3) (NULL: weights only, pred_rgb = white
 *   background or 0), t_vals (N,S) -> weights (N,S),
 *   pred_rgb (N,3), pred_depth (N,), acc_map (N,).  Transmittance is an
 *   exclusive, compensated log-space scan over one wave per ray.
 *   The _f64 variant takes and returns float64 (gradient checks).
 * ---------------------------------------------------------------------- */
int32_t nerf_composite_forward(int64_t n_rays, int32_t S, int32_t white_bg, const float* sigma,
                               const float* sample_rgb, const float* t_vals, float* weights,
                               float* pred_rgb, float* pred_depth, float* acc_map, void* stream);
int32_t nerf_composite_forward_f64(int64_t n_rays, int32_t S, int32_t white_bg, const double* sigma,
                                   const double* sample_rgb, const double* t_vals, double* weights,
                                   double* pred_rgb, double* pred_depth, double* acc_map,
                                   void* stream);

/* Backward of nerf_composite_forward (the gradient TF's GradientTape takes
 * through ray_utils.py:408-551).  Upstream gradients g_rgb (N,3) required;
 * g_depth (N,), g_acc (N,), g_weights (N,S) may be NULL (= zero).
 * Outputs d_sigma (N*S,1), d_rgb (N*S,3). */
int32_t nerf_composite_backward(int64_t n_rays, int32_t S, int32_t white_bg, const float* sigma,
                                const float* sample_rgb, const float* t_vals, const float* g_rgb,
                                const float* g_depth, const float* g_acc, const float* g_weights,
                                float* d_sigma, float* d_rgb, void* stream);
int32_t nerf_composite_backward_f64(int64_t n_rays, int32_t S, int32_t white_bg,
                                    const double* sigma, const double* sample_rgb,
                                    const double* t_vals, const double* g_rgb,
                                    const double* g_depth, const double* g_acc,
                                    const double* g_weights, double* d_sigma, double* d_rgb,
                                    void* stream);

/* ------------------------------------------------------------------------
 * The coarse / fine MLP: get_coarse_or_fine_model (nerf/core/model.py:334-394)
 * with PositionalEncoder (model.py:289-332) fused into the first layer.
 * ---------------------------------------------------------------------- */
typedef struct nerf_mlp_desc {
  int32_t depth;  /* trunk Dense layers, ReLU (reference: 8)                     */
  int32_t width;  /* trunk width, multiple of 32 (reference: 256)                */
  int32_t skip;   /* re-concatenate enc_xyz after trunk layer `skip` (4); -1 none */
  int32_t l_xyz;  /* xyz encoding frequencies, even (10)                         */
  int32_t l_dir;  /* direction encoding frequencies, even (4); 0 = density-only  */
                  /* net (no bottleneck / rgb branch) with a softplus density    */
} nerf_mlp_desc;

/* Number of float32 parameters in the flat "Keras" layout: for each layer in
 * Keras' trainable-variable order of the functional model (model.py:389-392:
 * dense_0 .. dense_{D-1}, dense_D, dense_{D+1}, rgb, sigma; a density-only net:
 * dense_0 .. dense_{D-1}, sigma) its kernel (fan_in x fan_out, row-major) then
 * its bias. */
int64_t nerf_mlp_param_count(const nerf_mlp_desc* desc);
/* Bytes of the packed device weights for `precision`. */
int64_t nerf_mlp_packed_bytes(const nerf_mlp_desc* desc, int32_t precision);
/* Elements of the training stash per sample (layer inputs saved by the
 * forward): float32 elements for NERF_PREC_F32, bf16 for NERF_PREC_BF16. */
int64_t nerf_mlp_stash_floats_per_sample(const nerf_mlp_desc* desc);
/* Elements of the per-sample gradient buffer (dz) used by backward /
 * weight_grads: float32 for NERF_PREC_F32, bf16 for NERF_PREC_BF16. */
int64_t nerf_mlp_dz_floats_per_sample(const nerf_mlp_desc* desc);
/* Samples per block of the stash and dz buffers for `precision` (-1 for an
 * unsupported description): both hold M * (elements per sample) elements, cut
 * into blocks of this many consecutive samples, each block stored chunk-major
 * (1: plain rows, for float32; 32 for bf16, so that the kernels write and
 * read whole 512-byte stretches).  A batch whose weight gradients are
 * summed over several nerf_mlp_weight_grads calls (accumulate) is split at
 * multiples of it.  No counterpart in the reference (TF keeps activations in
 * its own tensors). */
int32_t nerf_mlp_sample_block(const nerf_mlp_desc* desc, int32_t precision);

/* Pack flat float32 weights into the MFMA operand layout (forward and
 * transposed-for-backward copies, in fp32 or bf16). */
int32_t nerf_mlp_pack(const nerf_mlp_desc* desc, int32_t precision, const float* flat, void* packed,
                      void* stream);

/* Forward of the model call `model((xyz, dirs)) -> (rgb, sigma)`
 * (model.py:94-96, 114-116) for M samples.  Inputs are either
 *   xyz, dirs (M,3) each                               (rays_o == NULL), or
 *   rays_o, rays_d (N,3), t_vals (N, samples_per_ray)   with M = N*samples_per_ray,
 *   the sample point being rays_o + t*rays_d (ray_utils.py:251, 389).
 * Outputs rgb (M,3) (NULL allowed for a density-only net) and sigma (M,1).
 * stash: NULL, or M * nerf_mlp_stash_floats_per_sample() elements (float32,
 * or bf16 for NERF_PREC_BF16) that the backward reads. */
int32_t nerf_mlp_forward(const nerf_mlp_desc* desc, int32_t precision, const void* packed,
                         int64_t M, const float* xyz, const float* dirs, const float* rays_o,
                         const float* rays_d, const float* t_vals, int32_t samples_per_ray,
                         float* rgb, float* sigma, void* stash, void* stream);

/* Backward through the MLP for M samples: d_rgb (M,3) (NULL for density-only),
 * d_sigma (M,1), the forward's rgb and sigma outputs and stash -> dz, the
 * per-layer pre-activation gradients (M * nerf_mlp_dz_floats_per_sample()
 * elements, float32 or, for NERF_PREC_BF16, bf16). */
int32_t nerf_mlp_backward(const nerf_mlp_desc* desc, int32_t precision, const void* packed,
                          int64_t M, const float* d_rgb, const float* d_sigma, const float* rgb,
                          const float* sigma, const void* stash, void* dz, void* stream);

/* Weight gradients: grad_flat (param_count floats, flat layout) = or +=
 * (accumulate != 0) sum over the M samples of X^T dz for every layer, and the
 * bias sums.  stash and dz hold M samples as the forward and backward wrote
 * them; a sub-range of a larger batch starts and ends on multiples of
 * nerf_mlp_sample_block() (or at the batch's end).  workspace: a device buffer of
 * workspace_bytes bytes: nerf_mlp_wgrad_workspace_bytes(desc, M) (a bound for
 * both precisions) lets a bf16 call keep every layer group's split-K partial
 * tiles side by side and sum them in one launch; a smaller one down to what one
 * group uses works too (each group summed before the next, the same result);
 * a buffer smaller than that is rejected with NERF_ERR_INVALID_ARG before
 * anything is launched.
 * The summation order is fixed (deterministic, no float atomics).
 * M == 0 (an empty batch): grad_flat is zeroed (or left as is when
 * accumulating); stash, dz and workspace may then be NULL.  Every other entry
 * point returns NERF_OK without reading its pointers when its ray / sample
 * count is 0. */
int64_t nerf_mlp_wgrad_workspace_bytes(const nerf_mlp_desc* desc, int64_t M);
int32_t nerf_mlp_weight_grads(const nerf_mlp_desc* desc, int32_t precision, int64_t M,
                              const void* stash, const void* dz, float* grad_flat,
                              int32_t accumulate, void* workspace, int64_t workspace_bytes,
                              void* stream);

/* nerf_mlp_weight_grads without a workspace and without allocating: no
 * counterpart in the reference; for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The partial tiles live in a static device array of the library,
 * NERF_WGRAD_STATIC_SCRATCH_BYTES on every GPU it runs on, held until the process
 * ends; its address is looked up once per device on the first call there.  A call
 * whose partial tiles need more is rejected with NERF_ERR_INVALID_ARG before
 * anything is launched (the 8x256 reference net fits every batch up to 10^7
 * samples in either precision).  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (other streams or threads) race.  Nothing in the
 * array needs clearing (each call writes every partial it reads), and the result
 * is bitwise that of nerf_mlp_weight_grads (the same partial tiles, summed in the
 * same order). */
#define NERF_WGRAD_STATIC_SCRATCH_BYTES (256LL * (256 * 256 + 256) * 4) /* 64.25 MiB */
int32_t nerf_mlp_weight_grads_noalloc(const nerf_mlp_desc* desc, int32_t precision, int64_t M,
                                      const void* stash, const void* dz, float* grad_flat,
                                      int32_t accumulate, void* stream);

/* ------------------------------------------------------------------------
 * Losses and the optimizer step (model.py:32, 148-171, 413-418).
 * ---------------------------------------------------------------------- */

/* MeanSquaredError (SUM_OVER_BATCH_SIZE) gradient: g = scale * 2 (pred-true)/(N*3).
 * loss_out (1 float) receives sum((pred-true)^2)/(N*3) (not scaled). */
int32_t nerf_mse_grad(int64_t n_rays, const float* pred_rgb, const float* true_rgb, float scale,
                      float* g_rgb, float* loss_out, void* stream);

/* Keras Adam with the TF ApplyAdam update on a flat float32 buffer of n
 * values: m += (g-m)*one_minus_b1; v += (g*g-v)*one_minus_b2;
 * var -= (m*lr_t)/(sqrt(v)+eps).  The gradient is first multiplied by
 * *inv_scale (device scalar, NULL = 1).  If skip_flag (device int32, may be
 * NULL) is nonzero the step is skipped (loss-scaler overflow). */
int32_t nerf_adam_tf(int64_t n, float* var, float* m, float* v, const float* grad, float lr_t,
                     float one_minus_b1, float one_minus_b2, float eps, const float* inv_scale,
                     const int32_t* skip_flag, void* stream);

/* Sets *flag (device int32) to 1 if any of the n values is inf or NaN
 * (dynamic loss scaling), else leaves it unchanged. */
int32_t nerf_check_finite(int64_t n, const float* x, int32_t* flag, void* stream);

/* ------------------------------------------------------------------------
 * Device-resident state of a loss-scaled training step, so that a BF16 train
 * step never waits on the host: TF's LossScaleOptimizer (dynamic scale,
 * skip on overflow, the skipped step does not advance `iterations`) around
 * Keras Adam + ExponentialDecay (nerf/core/model.py:413-418).  The reference
 * trains in FP32 and has no loss scaler; this is the state of the BF16 form.
 * 64 bytes, caller-allocated in device memory, zero-filled, then `scale` set.
 *   nerf_step_begin: overflow = 0, inv_scale = 1/scale, and the Adam
 *     coefficients of step `iterations` (lr = lr0 * rate^(it/decay_steps),
 *     lr_t = lr * sqrt(1 - b2^(it+1)) / (1 - b1^(it+1)), float32 operands,
 *     powers evaluated in float64 and rounded once).
 *   nerf_check_finite(n, grad, &state->overflow, ...) between the two.
 *   nerf_adam_tf_state: nerf_adam_tf with lr_t, 1-b1, 1-b2, inv_scale and the
 *     skip flag taken from the state.
 *   nerf_step_end: on overflow scale = max(scale/factor, min_scale), good = 0,
 *     skipped += 1; else iterations += 1, good += 1 and, after `interval`
 *     clean steps, scale *= factor, good = 0.
 * ---------------------------------------------------------------------- */
typedef struct nerf_step_state {
  float scale;
  float inv_scale;
  int32_t overflow;
  int32_t good;
  int32_t iterations;
  int32_t skipped;
  float lr_t;
  float one_minus_b1;
  float one_minus_b2;
  int32_t reserved[7];
} nerf_step_state;

int32_t nerf_step_begin(nerf_step_state* state, float lr0, int32_t decay_steps, float decay_rate,
                        float b1, float b2, void* stream);
int32_t nerf_adam_tf_state(int64_t n, float* var, float* m, float* v, const float* grad, float eps,
                           const nerf_step_state* state, void* stream);
int32_t nerf_step_end(nerf_step_state* state, int32_t interval, float factor, float min_scale,
                      void* stream);

/* ------------------------------------------------------------------------
 * Float64 layer kernels for the finite-difference check of the hand-written
 * backward (the FP64 gradient check; no counterpart in the reference, whose
 * gradients come from TF's GradientTape over model.py:355-392).  Row-major
 * float64; kernels are Keras (fan_in, fan_out).
 *   encode:      out (M, 3+6L) = PositionalEncoder (model.py:319-330) in double
 *   dense fwd:   Y (M,N) = act([X1 (M,K1), X2 (M,K2)] @ kernel + bias),
 *                activation 0 identity, 1 ReLU, 2 sigmoid; X2 may be NULL (K2 = 0)
 *   dense dX:    dX (M,K) = dZ (M,N) @ kernel[:K]^T, times [X_hidden > 0] if relu_mask
 *   dense dW:    d_kernel[row0:row0+K] = X^T dZ; d_bias = sum_m dZ (NULL: skipped)
 * ---------------------------------------------------------------------- */
int32_t nerf_f64_encode(int64_t M, int32_t L, const double* x, double* out, void* stream);
int32_t nerf_f64_dense_forward(int64_t M, int32_t K1, const double* X1, int32_t K2,
                               const double* X2, int32_t N, const double* kernel,
                               const double* bias, int32_t activation, double* Y, void* stream);
int32_t nerf_f64_dense_backward_dx(int64_t M, int32_t N, const double* dZ, int32_t K,
                                   const double* kernel, const double* X_hidden, int32_t relu_mask,
                                   double* dX, void* stream);
int32_t nerf_f64_dense_wgrad(int64_t M, int32_t K, const double* X, int32_t N, const double* dZ,
                             int32_t row0, double* d_kernel, double* d_bias, void* stream);

/* ------------------------------------------------------------------------
 * Learned proposal (an addition with no counterpart in the reference: it
 * replaces the coarse pass of NeRF.forward, model.py:88-103, for placing the
 * fine samples).  Histogram upper-bound loss of mip-NeRF 360 between the
 * proposal weights on its bins (prop_edges (N, n_prop+1), prop_weights
 * (N, n_prop)) and the fine weights fine_weights (N, n_fine) at the sorted
 * sample positions t_fine (N, n_fine); the last fine interval ends at far (N,1).
 * The bound of a fine interval sums the proposal bins overlapping it, widened
 * by `dilate` bins on each side (0: the plain outer measure).
 * Outputs (each may be NULL): missed (N,) missed weight mass per ray, loss (N,)
 * per-ray loss, g_prop_weights (N, n_prop) = grad_scale * d loss / d prop_weights.
 * ---------------------------------------------------------------------- */
int32_t nerf_proposal_interlevel(int64_t n_rays, int32_t n_prop, int32_t n_fine,
                                 const float* prop_edges, const float* prop_weights,
                                 const float* t_fine, const float* fine_weights, const float* far,
                                 float* missed, float* loss, float* g_prop_weights,
                                 float grad_scale, int32_t dilate, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* NERF_HIP_H_ */
