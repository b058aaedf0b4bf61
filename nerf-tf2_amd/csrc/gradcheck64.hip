// Float64 layer kernels for the finite-difference gradient check of the
// hand-written MLP backward.  They implement, layer by layer and in double,
// the same algebra as the fused kernels (encoding; dense + bias + ReLU /
// sigmoid / identity; dX = dZ W^T over the hidden rows; dW = X^T dZ, db =
// sum dZ) so a central difference with h = 1e-6 can resolve them.  One thread
// per output element, sequential sums: small sizes only.
#include "common.h"

namespace {

__global__ void encode64_kernel(int64_t M, int L, const double* __restrict__ x, double* __restrict__ out) {
  const int F = 3 + 6 * L;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= M * F) return;
  const int64_t m = e / F;
  const int f = (int)(e % F);
  if (f < 3) {
    out[e] = x[3 * m + f];
    return;
  }
  const int c = (f - 3) / (2 * L), l = ((f - 3) % (2 * L)) / 2, sc = (f - 3) % 2;
  const double a = x[3 * m + c] * (double)(1 << l) * 3.14159265358979323846;
  out[e] = sc ? cos(a) : sin(a);  // model.py:323-330 feature order
}

// Y[m][n] = act(sum_k X1[m][k] W[k][n] + sum_k X2[m][k] W[K1 + k][n] + b[n])
__global__ void dense64_fwd_kernel(int64_t M, int K1, const double* __restrict__ X1, int K2,
                                   const double* __restrict__ X2, int N, const double* __restrict__ Wt,
                                   const double* __restrict__ b, int act, double* __restrict__ Y) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= M * N) return;
  const int64_t m = e / N;
  const int n = (int)(e % N);
  double s = 0.0;
  for (int k = 0; k < K1; ++k) s += X1[m * K1 + k] * Wt[(int64_t)k * N + n];
  for (int k = 0; k < K2; ++k) s += X2[m * K2 + k] * Wt[(int64_t)(K1 + k) * N + n];
  s += b[n];
  if (act == 1) s = s > 0.0 ? s : 0.0;
  if (act == 2) s = 1.0 / (1.0 + exp(-s));
  Y[e] = s;
}

// dX[m][k] = (sum_n dZ[m][n] W[k][n]) * mask(k), mask = [Xh[m][k] > 0] if relu_mask
__global__ void dense64_dx_kernel(int64_t M, int N, const double* __restrict__ dZ, int K,
                                  const double* __restrict__ Wt, const double* __restrict__ Xh,
                                  int relu_mask, double* __restrict__ dX) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= M * K) return;
  const int64_t m = e / K;
  const int k = (int)(e % K);
  double s = 0.0;
  for (int n = 0; n < N; ++n) s += dZ[m * N + n] * Wt[(int64_t)k * N + n];
  if (relu_mask && !(Xh[e] > 0.0)) s = 0.0;
  dX[e] = s;
}

// dW[row0 + k][n] = sum_m X[m][k] dZ[m][n] ; db[n] = sum_m dZ[m][n] (if db)
__global__ void dense64_wgrad_kernel(int64_t M, int K, const double* __restrict__ X, int N,
                                     const double* __restrict__ dZ, int row0, double* __restrict__ dW,
                                     double* __restrict__ db) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < (int64_t)K * N) {
    const int k = (int)(e / N), n = (int)(e % N);
    double s = 0.0;
    for (int64_t m = 0; m < M; ++m) s += X[m * K + k] * dZ[m * N + n];
    dW[(int64_t)(row0 + k) * N + n] = s;
  } else if (db && e < (int64_t)K * N + N) {
    const int n = (int)(e - (int64_t)K * N);
    double s = 0.0;
    for (int64_t m = 0; m < M; ++m) s += dZ[m * N + n];
    db[n] = s;
  }
}

inline unsigned nblocks(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

extern "C" int32_t nerf_f64_encode(int64_t M, int32_t L, const double* x, double* out, void* stream) {
  NERF_CHECK_ARG(M >= 0 && L >= 1 && L <= 16 && x && out, "nerf_f64_encode: bad arguments");
  if (M == 0) return NERF_OK;
  encode64_kernel<<<nblocks(M * (3 + 6 * L)), 256, 0, as_stream(stream)>>>(M, L, x, out);
  NERF_LAUNCH_CHECK("nerf_f64_encode");
  return NERF_OK;
}

extern "C" int32_t nerf_f64_dense_forward(int64_t M, int32_t K1, const double* X1, int32_t K2,
                                          const double* X2, int32_t N, const double* kernel,
                                          const double* bias, int32_t activation, double* Y,
                                          void* stream) {
  NERF_CHECK_ARG(M >= 0 && K1 > 0 && K2 >= 0 && N > 0 && X1 && (K2 == 0 || X2) && kernel && bias && Y &&
                     activation >= 0 && activation <= 2,
                 "nerf_f64_dense_forward: bad arguments");
  if (M == 0) return NERF_OK;
  dense64_fwd_kernel<<<nblocks(M * N), 256, 0, as_stream(stream)>>>(M, K1, X1, K2, X2, N, kernel, bias,
                                                                   activation, Y);
  NERF_LAUNCH_CHECK("nerf_f64_dense_forward");
  return NERF_OK;
}

extern "C" int32_t nerf_f64_dense_backward_dx(int64_t M, int32_t N, const double* dZ, int32_t K,
                                              const double* kernel, const double* X_hidden,
                                              int32_t relu_mask, double* dX, void* stream) {
  NERF_CHECK_ARG(M >= 0 && N > 0 && K > 0 && dZ && kernel && dX && (!relu_mask || X_hidden),
                 "nerf_f64_dense_backward_dx: bad arguments");
  if (M == 0) return NERF_OK;
  dense64_dx_kernel<<<nblocks(M * K), 256, 0, as_stream(stream)>>>(M, N, dZ, K, kernel, X_hidden, relu_mask,
                                                                  dX);
  NERF_LAUNCH_CHECK("nerf_f64_dense_backward_dx");
  return NERF_OK;
}

extern "C" int32_t nerf_f64_dense_wgrad(int64_t M, int32_t K, const double* X, int32_t N,
                                        const double* dZ, int32_t row0, double* d_kernel,
                                        double* d_bias, void* stream) {
  NERF_CHECK_ARG(M >= 0 && K > 0 && N > 0 && X && dZ && d_kernel && row0 >= 0,
                 "nerf_f64_dense_wgrad: bad arguments");
  dense64_wgrad_kernel<<<nblocks((int64_t)K * N + N), 256, 0, as_stream(stream)>>>(M, K, X, N, dZ, row0,
                                                                                  d_kernel, d_bias);
  NERF_LAUNCH_CHECK("nerf_f64_dense_wgrad");
  return NERF_OK;
}
