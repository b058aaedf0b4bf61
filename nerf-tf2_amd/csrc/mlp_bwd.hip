// Hand-written backward of the fused NeRF MLP (the gradient TF's GradientTape
// takes through model.py:355-392) and its weight gradients.
//
// nerf_mlp_backward: one wave per 32 samples walks the layers top-down.
//   rgb = sigmoid(z):  dz_rgb = d_rgb * rgb (1 - rgb)
//   sigma = relu(z):   dz_sigma = d_sigma * [sigma > 0]      (softplus: * (1 - e^-sigma))
//   dh = W dz (MFMA with the transposed packed weights; the accumulators of one
//   layer are the B operand of the next, as in the forward), dz = dh * [h > 0]
//   using the forward's stashed activations for the ReLU masks.
//   Every pre-activation gradient dz_l is written row-major for the weight GEMMs.
// nerf_mlp_weight_grads: dW_l = X_l^T dz_l and db_l = sum dz_l over the samples,
//   as a split-K MFMA GEMM whose partial tiles are summed in a fixed order.
#include "common.h"
#include "mlp_kernels.h"

namespace {

// ---------------------------------------------------------------------------
// dX chain
// ---------------------------------------------------------------------------
// acc[nb][i] *= [h > 0] for the row-major activation row h (natural order).
template <int NBo>
__device__ __forceinline__ void relu_mask_rows_f32(floatx16 (&acc)[NBo], const float* row, int g) {
#pragma unroll
  for (int nb = 0; nb < NBo; ++nb)
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const floatx4 h = *reinterpret_cast<const floatx4*>(row + 32 * nb + 8 * a + 4 * g);
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[nb][4 * a + r] = h[r] > 0.f ? acc[nb][4 * a + r] : 0.f;
    }
}

template <int W, int D, int SKIP, int LX, int LD>
__global__ __launch_bounds__(256) void mlp_bwd_f32_kernel(
    MlpPlan P, const char* __restrict__ pk, int64_t M, const float* __restrict__ d_rgb,
    const float* __restrict__ d_sigma, const float* __restrict__ rgb,
    const float* __restrict__ sigma, const float* __restrict__ stash, float* __restrict__ dz) {
  using C = NetC<W, D, SKIP, LX, LD>;
  __shared__ floatx4 s_act[4 * C::QH * 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, j = lane & 31, g = lane >> 5;
  floatx4* act = s_act + wv * C::QH * 64;
  const int64_t blk = (int64_t)blockIdx.x * 4 + wv;
  const int64_t nblk = (M + 31) / 32;
  if (blk >= nblk) return;
  const int64_t m = blk * 32 + j;
  const bool valid = m < M;
  const int64_t mm = valid ? m : M - 1;
  const float* srow = stash + mm * P.stash_per_sample;
  // rows of invalid lanes go to a dummy write: masked below
  float* zrow = dz + mm * P.dz_per_sample;

  const float sg = sigma[mm];
  const float ds = valid ? d_sigma[mm] : 0.f;
  const float dzs = C::RGB ? (sg > 0.f ? ds : 0.f) : ds * (1.f - expf(-sg));
  floatx16 acc[C::NB];
  if constexpr (C::RGB) {
    float dzr[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float y = rgb[3 * mm + c];
      const float dy = valid ? d_rgb[3 * mm + c] : 0.f;
      dzr[c] = dy * y * (1.f - y);  // sigmoid'
    }
    // rgb head: dh9[f] = sum_c W_rgb[f][c] dz_rgb[c]
    floatx16 accr[C::NBR];
    const floatx4* hr4 = reinterpret_cast<const floatx4*>(pk + P.p_head_rgb) + g * 16 * C::NBR;
#pragma unroll
    for (int nb = 0; nb < C::NBR; ++nb)
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const floatx4 w = hr4[16 * nb + i];
        accr[nb][i] = dzr[0] * w[0] + dzr[1] * w[1] + dzr[2] * w[2];
      }
    relu_mask_rows_f32(accr, srow + P.st_rbr, g);
    if (valid) {
      store_rows_f32(accr, zrow + P.dz_rbr, g);
      if (g == 0) {
        zrow[P.dz_rgb + 0] = dzr[0];
        zrow[P.dz_rgb + 1] = dzr[1];
        zrow[P.dz_rgb + 2] = dzr[2];
      }
    }
    // dh8 = W9[:W] dz9; the bottleneck is linear so dz8 = dh8
    wave_sync();
    store_act_f32(accr, act, lane);
    zero_acc(acc);
    wave_sync();
    mfma_lds_f32<C::NB, 4 * C::NBR>(acc, reinterpret_cast<const floatx4*>(pk + P.p_bwd[D + 1]), act,
                                    lane);
    if (valid) store_rows_f32(acc, zrow + P.dz_bott, g);
    // dh7 = W8 dz8
    wave_sync();
    store_act_f32(acc, act, lane);
    zero_acc(acc);
    wave_sync();
    mfma_lds_f32<C::NB, C::QH>(acc, reinterpret_cast<const floatx4*>(pk + P.p_bwd[D]), act, lane);
  } else {
    zero_acc(acc);
  }
  // sigma head: dh7[f] += W_sigma[f] dz_sigma
  {
    const float* hs = reinterpret_cast<const float*>(pk + P.p_head_sigma) + g * C::KH;
#pragma unroll
    for (int nb = 0; nb < C::NB; ++nb)
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const floatx4 w = *reinterpret_cast<const floatx4*>(hs + 16 * nb + 4 * a);
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[nb][4 * a + r] = acc[nb][4 * a + r] + w[r] * dzs;
      }
  }
  if (valid && g == 0) zrow[P.dz_sigma] = dzs;
  relu_mask_rows_f32(acc, srow + P.st_h[D - 1], g);
  if (valid) store_rows_f32(acc, zrow + P.dz_h[D - 1], g);
  // trunk, top-down: dh_{l-1} = W_l[:W] dz_l, dz_{l-1} = dh_{l-1} [h_{l-1} > 0]
#pragma unroll 1
  for (int l = D - 1; l >= 1; --l) {
    wave_sync();
    store_act_f32(acc, act, lane);
    zero_acc(acc);
    wave_sync();
    mfma_lds_f32<C::NB, C::QH>(acc, reinterpret_cast<const floatx4*>(pk + P.p_bwd[l]), act, lane);
    relu_mask_rows_f32(acc, srow + P.st_h[l - 1], g);
    if (valid) store_rows_f32(acc, zrow + P.dz_h[l - 1], g);
  }
}

template <int NBo>
__device__ __forceinline__ void relu_mask_rows_bf16(floatx16 (&acc)[NBo], const __bf16* row, int g) {
#pragma unroll
  for (int nb = 0; nb < NBo; ++nb)
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const bf16x4 h = *reinterpret_cast<const bf16x4*>(row + 32 * nb + 8 * a + 4 * g);
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[nb][4 * a + r] = (float)h[r] > 0.f ? acc[nb][4 * a + r] : 0.f;
    }
}

// bf16 twin of mlp_bwd_f32_kernel: dz is rounded to bf16 only as the B operand of
// the next MFMA; the rows written for the weight GEMMs stay fp32.
template <int W, int D, int SKIP, int LX, int LD>
__global__ __launch_bounds__(256) void mlp_bwd_bf16_kernel(
    MlpPlan P, const char* __restrict__ pk, int64_t M, const float* __restrict__ d_rgb,
    const float* __restrict__ d_sigma, const float* __restrict__ rgb,
    const float* __restrict__ sigma, const __bf16* __restrict__ stash, float* __restrict__ dz) {
  using C = NetC<W, D, SKIP, LX, LD>;
  constexpr int TH = W / 16;
  __shared__ bf16x8 s_act[4 * TH * 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, j = lane & 31, g = lane >> 5;
  bf16x8* act = s_act + wv * TH * 64;
  const int64_t blk = (int64_t)blockIdx.x * 4 + wv;
  const int64_t nblk = (M + 31) / 32;
  if (blk >= nblk) return;
  const int64_t m = blk * 32 + j;
  const bool valid = m < M;
  const int64_t mm = valid ? m : M - 1;
  const __bf16* srow = stash + mm * P.stash_per_sample;
  float* zrow = dz + mm * P.dz_per_sample;
  const float sg = sigma[mm];
  const float ds = valid ? d_sigma[mm] : 0.f;
  const float dzs = C::RGB ? (sg > 0.f ? ds : 0.f) : ds * (1.f - expf(-sg));
  floatx16 acc[C::NB];
  if constexpr (C::RGB) {
    float dzr[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float y = rgb[3 * mm + c];
      const float dy = valid ? d_rgb[3 * mm + c] : 0.f;
      dzr[c] = dy * y * (1.f - y);
    }
    floatx16 accr[C::NBR];
    const floatx4* hr4 = reinterpret_cast<const floatx4*>(pk + P.p_head_rgb) + g * 16 * C::NBR;
#pragma unroll
    for (int nb = 0; nb < C::NBR; ++nb)
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const floatx4 w = hr4[16 * nb + i];
        accr[nb][i] = dzr[0] * w[0] + dzr[1] * w[1] + dzr[2] * w[2];
      }
    relu_mask_rows_bf16(accr, srow + P.st_rbr, g);
    if (valid) {
      store_rows_f32(accr, zrow + P.dz_rbr, g);
      if (g == 0) {
        zrow[P.dz_rgb + 0] = dzr[0];
        zrow[P.dz_rgb + 1] = dzr[1];
        zrow[P.dz_rgb + 2] = dzr[2];
      }
    }
    wave_sync();
    store_act_bf16(accr, act, lane);
    zero_acc(acc);
    wave_sync();
    mfma_lds_bf16<C::NB, 2 * C::NBR>(acc, reinterpret_cast<const bf16x8*>(pk + P.p_bwd[D + 1]), act,
                                     lane);
    if (valid) store_rows_f32(acc, zrow + P.dz_bott, g);
    wave_sync();
    store_act_bf16(acc, act, lane);
    zero_acc(acc);
    wave_sync();
    mfma_lds_bf16<C::NB, TH>(acc, reinterpret_cast<const bf16x8*>(pk + P.p_bwd[D]), act, lane);
  } else {
    zero_acc(acc);
  }
  {
    const float* hs = reinterpret_cast<const float*>(pk + P.p_head_sigma) + g * C::KH;
#pragma unroll
    for (int nb = 0; nb < C::NB; ++nb)
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const floatx4 w = *reinterpret_cast<const floatx4*>(hs + 16 * nb + 4 * a);
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[nb][4 * a + r] = acc[nb][4 * a + r] + w[r] * dzs;
      }
  }
  if (valid && g == 0) zrow[P.dz_sigma] = dzs;
  relu_mask_rows_bf16(acc, srow + P.st_h[D - 1], g);
  if (valid) store_rows_f32(acc, zrow + P.dz_h[D - 1], g);
#pragma unroll 1
  for (int l = D - 1; l >= 1; --l) {
    wave_sync();
    store_act_bf16(acc, act, lane);
    zero_acc(acc);
    wave_sync();
    mfma_lds_bf16<C::NB, TH>(acc, reinterpret_cast<const bf16x8*>(pk + P.p_bwd[l]), act, lane);
    relu_mask_rows_bf16(acc, srow + P.st_h[l - 1], g);
    if (valid) store_rows_f32(acc, zrow + P.dz_h[l - 1], g);
  }
}

// ---------------------------------------------------------------------------
// Weight-gradient GEMM: part[split][k][n] = sum over the split's samples of
// X[m][xoff + k] * Z[m][zoff + n], for k < K, n < N.  128x128 tile per block,
// 4 waves of 64x64 (2x2 blocks of v_mfma_f32_32x32x2_f32), 32-sample LDS chunks.
// The block with tile_k == 0 also sums the columns of Z (bias gradient).
// ---------------------------------------------------------------------------
constexpr int kTile = 128, kChunk = 32, kPad = 4;

struct GemmArgs {
  const void* X;   // float or bf16 rows (x_bf16)
  int x_bf16;
  int64_t ldx, xoff;
  const float* Z;
  int64_t ldz, zoff;
  int64_t M, m_per_split;
  int K, N, tiles_k, tiles_n, splits;
  float* part;   // [splits][K][N]
  float* bpart;  // [splits][N] or null
};

__global__ __launch_bounds__(256) void wgrad_gemm_kernel(GemmArgs a) {
  __shared__ float Xs[kChunk][kTile + kPad];
  __shared__ float Zs[kChunk][kTile + kPad];
  // XCD-aware order: workgroup b runs on XCD b % 8, so all tiles of one sample
  // split (which read the same X / dz rows) share b % 8 and hence one L2.
  const int T = a.tiles_k * a.tiles_n;
  const int xcd = blockIdx.x % 8, yy = blockIdx.x / 8;
  const int tile = yy % T;
  const int split = (yy / T) * 8 + xcd;
  if (split >= a.splits) return;  // padding of the split count to a multiple of 8
  const int tk = tile / a.tiles_n, tn = tile % a.tiles_n;
  const int k0 = tk * kTile, n0 = tn * kTile;
  const int64_t mb = (int64_t)split * a.m_per_split;
  const int64_t me = mb + a.m_per_split < a.M ? mb + a.m_per_split : a.M;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wk = (wv >> 1) * 64, wn = (wv & 1) * 64;  // wave quadrant
  const int r = lane & 31, q = lane >> 5;
  floatx16 acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[x][y][i] = 0.f;
  float bsum = 0.f;
  const bool do_bias = a.bpart && tk == 0 && tid < kTile;
  // loader mapping: 256 threads x 16 floats = 32 rows x 128 cols
  const int lr = tid >> 3, lc = (tid & 7) * 16;
  for (int64_t c0 = mb; c0 < me; c0 += kChunk) {
    const int64_t mrow = c0 + lr;
    const bool rok = mrow < me;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int col = lc + 4 * v;
      floatx4 xv = {0.f, 0.f, 0.f, 0.f}, zv = {0.f, 0.f, 0.f, 0.f};
      if (rok) {
        const int64_t xo = mrow * a.ldx + a.xoff + k0 + col;
        if (a.x_bf16) {
          const __bf16* xb = reinterpret_cast<const __bf16*>(a.X);
          if (k0 + col + 3 < a.K) {
            const bf16x4 v = *reinterpret_cast<const bf16x4*>(xb + xo);
#pragma unroll
            for (int e = 0; e < 4; ++e) xv[e] = (float)v[e];
          } else {
#pragma unroll
            for (int e = 0; e < 4; ++e)
              if (k0 + col + e < a.K) xv[e] = (float)xb[xo + e];
          }
        } else {
          const float* xf = reinterpret_cast<const float*>(a.X);
          if (k0 + col + 3 < a.K) {
            xv = *reinterpret_cast<const floatx4*>(xf + xo);
          } else {
#pragma unroll
            for (int e = 0; e < 4; ++e)
              if (k0 + col + e < a.K) xv[e] = xf[xo + e];
          }
        }
        if (n0 + col + 3 < a.N) {
          zv = *reinterpret_cast<const floatx4*>(a.Z + mrow * a.ldz + a.zoff + n0 + col);
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e)
            if (n0 + col + e < a.N) zv[e] = a.Z[mrow * a.ldz + a.zoff + n0 + col + e];
        }
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        Xs[lr][col + e] = xv[e];
        Zs[lr][col + e] = zv[e];
      }
    }
    __syncthreads();
#pragma unroll 4
    for (int s = 0; s < kChunk / 2; ++s) {
      const int row = 2 * s + q;
      float av[2], bv[2];
#pragma unroll
      for (int x = 0; x < 2; ++x) av[x] = Xs[row][wk + 32 * x + r];
#pragma unroll
      for (int y = 0; y < 2; ++y) bv[y] = Zs[row][wn + 32 * y + r];
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = mfma_f32(av[x], bv[y], acc[x][y]);
    }
    if (do_bias) {
#pragma unroll 8
      for (int rr = 0; rr < kChunk; ++rr) bsum += Zs[rr][tid];
    }
    __syncthreads();
  }
  float* out = a.part + (int64_t)split * a.K * a.N;
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int kk = k0 + wk + 32 * x + 8 * (i / 4) + 4 * q + (i % 4);
        const int nn = n0 + wn + 32 * y + r;
        if (kk < a.K && nn < a.N) out[(int64_t)kk * a.N + nn] = acc[x][y][i];
      }
  if (do_bias && n0 + tid < a.N) a.bpart[(int64_t)split * a.N + n0 + tid] = bsum;
}

// bf16 variant for the bf16 stash: X (bf16) and dz (fp32, rounded to bf16 here)
// are staged transposed, [feature][sample], into LDS as packed sample pairs so
// each lane's MFMA operand (8 consecutive samples of one feature) is one
// ds_read_b128; v_mfma_f32_32x32x16_bf16 with fp32 accumulation.  Bias sums use
// the fp32 dz (16-lane shuffle reduction per feature group).
constexpr int kPitch16 = kChunk + 8;  // bf16 per LDS row (16 B aligned, staggered banks)

__global__ __launch_bounds__(256) void wgrad_gemm_bf16_kernel(GemmArgs a) {
  __shared__ __attribute__((aligned(16))) __bf16 Xt[kTile][kPitch16];
  __shared__ __attribute__((aligned(16))) __bf16 Zt[kTile][kPitch16];
  // XCD-aware order: workgroup b runs on XCD b % 8, so all tiles of one sample
  // split (which read the same X / dz rows) share b % 8 and hence one L2.
  const int T = a.tiles_k * a.tiles_n;
  const int xcd = blockIdx.x % 8, yy = blockIdx.x / 8;
  const int tile = yy % T;
  const int split = (yy / T) * 8 + xcd;
  if (split >= a.splits) return;  // padding of the split count to a multiple of 8
  const int tk = tile / a.tiles_n, tn = tile % a.tiles_n;
  const int k0 = tk * kTile, n0 = tn * kTile;
  const int64_t mb = (int64_t)split * a.m_per_split;
  const int64_t me = mb + a.m_per_split < a.M ? mb + a.m_per_split : a.M;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wk = (wv >> 1) * 64, wn = (wv & 1) * 64;
  const int r = lane & 31, q = lane >> 5;
  floatx16 acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[x][y][i] = 0.f;
  // staging: thread -> sample pair sp (rows 2sp, 2sp+1), feature group fg (8 features)
  const int sp = tid & 15, fg = tid >> 4;
  const bool do_bias = a.bpart && tk == 0;
  float bsum[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) bsum[e] = 0.f;
  const __bf16* Xb = reinterpret_cast<const __bf16*>(a.X);
  for (int64_t c0 = mb; c0 < me; c0 += kChunk) {
    float xv[2][8], zv[2][8];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int64_t m = c0 + 2 * sp + h;
      const bool ok = m < me;
      const int kf = k0 + 8 * fg, nf = n0 + 8 * fg;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        xv[h][e] = 0.f;
        zv[h][e] = 0.f;
      }
      if (ok) {
        const __bf16* xr = Xb + m * a.ldx + a.xoff;
        if (kf + 7 < a.K) {
          const bf16x8 v = *reinterpret_cast<const bf16x8*>(xr + kf);
#pragma unroll
          for (int e = 0; e < 8; ++e) xv[h][e] = (float)v[e];
        } else {
#pragma unroll
          for (int e = 0; e < 8; ++e)
            if (kf + e < a.K) xv[h][e] = (float)xr[kf + e];
        }
        const float* zr = a.Z + m * a.ldz + a.zoff;
        if (nf + 7 < a.N) {
          const floatx4 v0 = *reinterpret_cast<const floatx4*>(zr + nf);
          const floatx4 v1 = *reinterpret_cast<const floatx4*>(zr + nf + 4);
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            zv[h][e] = v0[e];
            zv[h][4 + e] = v1[e];
          }
        } else {
#pragma unroll
          for (int e = 0; e < 8; ++e)
            if (nf + e < a.N) zv[h][e] = zr[nf + e];
        }
      }
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
      typedef float floatx2 __attribute__((ext_vector_type(2)));
      const floatx2 xp = {xv[0][e], xv[1][e]};
      const floatx2 zp = {zv[0][e], zv[1][e]};
      *reinterpret_cast<bf16x2*>(&Xt[8 * fg + e][2 * sp]) = __builtin_convertvector(xp, bf16x2);
      *reinterpret_cast<bf16x2*>(&Zt[8 * fg + e][2 * sp]) = __builtin_convertvector(zp, bf16x2);
      if (do_bias) bsum[e] += zv[0][e] + zv[1][e];
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kChunk / 16; ++s) {
      bf16x8 av[2], bv[2];
#pragma unroll
      for (int x = 0; x < 2; ++x) av[x] = *reinterpret_cast<const bf16x8*>(&Xt[wk + 32 * x + r][16 * s + 8 * q]);
#pragma unroll
      for (int y = 0; y < 2; ++y) bv[y] = *reinterpret_cast<const bf16x8*>(&Zt[wn + 32 * y + r][16 * s + 8 * q]);
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = mfma_bf16(av[x], bv[y], acc[x][y]);
    }
    __syncthreads();
  }
  float* out = a.part + (int64_t)split * a.K * a.N;
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int kk = k0 + wk + 32 * x + 8 * (i / 4) + 4 * q + (i % 4);
        const int nn = n0 + wn + 32 * y + r;
        if (kk < a.K && nn < a.N) out[(int64_t)kk * a.N + nn] = acc[x][y][i];
      }
  if (do_bias) {
    // reduce over the 16 sample-pair lanes of this feature group (lanes fg*16 .. fg*16+15)
#pragma unroll
    for (int e = 0; e < 8; ++e) {
#pragma unroll
      for (int msk = 1; msk < 16; msk <<= 1) bsum[e] += shfl_xor(bsum[e], msk);
    }
    if (sp == 0) {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int n = n0 + 8 * fg + e;
        if (n < a.N) a.bpart[(int64_t)split * a.N + n] = bsum[e];
      }
    }
  }
}

// grad[k_off + row(k) * fan_out + n] (+)= sum_s part[s][k][n] for n < n_real;
// row(k) = row_base + k for hidden inputs, or the encoding slot map (-1 = pad).
struct ReduceArgs {
  const float* part;
  const float* bpart;
  int K, N, n_real, splits;
  int enc_L;       // 0: identity rows; else slot map with KS slots per half
  int enc_KS;
  int row_base;
  int64_t k_off, b_off;
  int fan_out;
  int accumulate;
  float* grad;
};

__global__ void wgrad_reduce_kernel(ReduceArgs a) {
  const int64_t n_el = (int64_t)a.K * a.N;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n_el + a.N;
       e += (int64_t)gridDim.x * blockDim.x) {
    if (e < n_el) {
      const int k = (int)(e / a.N), n = (int)(e % a.N);
      if (n >= a.n_real) continue;
      int row = k;
      if (a.enc_L) {
        const int g = k / a.enc_KS, s = k % a.enc_KS;
        row = enc_feat(g, s, a.enc_L);
        if (row < 0) continue;
      }
      float s = 0.f;
      for (int sp = 0; sp < a.splits; ++sp) s += a.part[(int64_t)sp * n_el + e];
      float* dst = a.grad + a.k_off + (int64_t)(a.row_base + row) * a.fan_out + n;
      *dst = a.accumulate ? *dst + s : s;
    } else if (a.bpart) {
      const int n = (int)(e - n_el);
      if (n >= a.n_real) continue;
      float s = 0.f;
      for (int sp = 0; sp < a.splits; ++sp) s += a.bpart[(int64_t)sp * a.N + n];
      float* dst = a.grad + a.b_off + n;
      *dst = a.accumulate ? *dst + s : s;
    }
  }
}

// One (X-segment, Z) GEMM + reduction into the flat gradient.
struct WgradJob {
  int64_t xoff;  // X column offset in the stash row
  int K;         // X columns
  int64_t zoff;  // Z column offset in the dz row
  int N;         // Z columns (padded)
  int n_real;    // real output columns
  int enc_L, enc_KS, row_base;
  int keras;     // Keras layer index
  int bias;      // also write the bias gradient
};

int wgrad_jobs(const MlpPlan& P, WgradJob* jobs) {
  int n = 0;
  const int W = P.W;
  auto add = [&](int64_t xoff, int K, int64_t zoff, int N, int n_real, int encL, int encKS,
                 int row_base, int keras, int bias) {
    jobs[n++] = WgradJob{xoff, K, zoff, N, n_real, encL, encKS, row_base, keras, bias};
  };
  // dense_0: X = xyz encoding slots
  add(P.st_encx, 2 * P.KX, P.dz_h[0], W, W, P.LX, P.KX, 0, 0, 1);
  for (int i = 1; i < P.D; ++i) {
    add(P.st_h[i - 1], W, P.dz_h[i], W, W, 0, 0, 0, i, 1);
    if (i - 1 == P.skip) add(P.st_encx, 2 * P.KX, P.dz_h[i], W, W, P.LX, P.KX, W, i, 0);
  }
  add(P.st_h[P.D - 1], W, P.dz_sigma, 1, 1, 0, 0, 0, P.D, 1);  // sigma
  if (P.rgb) {
    add(P.st_h[P.D - 1], W, P.dz_bott, W, W, 0, 0, 0, P.D + 1, 1);
    add(P.st_bott, W, P.dz_rbr, 32 * P.NBR, P.WR, 0, 0, 0, P.D + 2, 1);
    add(P.st_encd, 2 * P.KD, P.dz_rbr, 32 * P.NBR, P.WR, P.LD, P.KD, W, P.D + 2, 0);
    add(P.st_rbr, P.WR, P.dz_rgb, 3, 3, 0, 0, 0, P.D + 3, 1);
  }
  return n;
}

int64_t splits_for(int64_t M) {
  int64_t s = (M + 4095) / 4096;  // >= 4096 samples per split
  if (s > 128) s = 128;
  return s < 1 ? 1 : s;
}

template <int W, int D, int SKIP, int LX, int LD>
int launch_bwd_f32(const MlpPlan& P, const void* packed, int64_t M, const float* d_rgb,
                   const float* d_sigma, const float* rgb, const float* sigma, const void* stash,
                   void* dz, hipStream_t s) {
  const int64_t nblk = (M + 31) / 32;
  mlp_bwd_f32_kernel<W, D, SKIP, LX, LD><<<(unsigned)((nblk + 3) / 4), 256, 0, s>>>(
      P, reinterpret_cast<const char*>(packed), M, d_rgb, d_sigma, rgb, sigma,
      reinterpret_cast<const float*>(stash), reinterpret_cast<float*>(dz));
  return 0;
}

template <int W, int D, int SKIP, int LX, int LD>
int launch_bwd_bf16(const MlpPlan& P, const void* packed, int64_t M, const float* d_rgb,
                    const float* d_sigma, const float* rgb, const float* sigma, const void* stash,
                    void* dz, hipStream_t s) {
  const int64_t nblk = (M + 31) / 32;
  mlp_bwd_bf16_kernel<W, D, SKIP, LX, LD><<<(unsigned)((nblk + 3) / 4), 256, 0, s>>>(
      P, reinterpret_cast<const char*>(packed), M, d_rgb, d_sigma, rgb, sigma,
      reinterpret_cast<const __bf16*>(stash), reinterpret_cast<float*>(dz));
  return 0;
}

int dispatch_bwd(const MlpPlan& P, const void* packed, int64_t M, const float* d_rgb,
                 const float* d_sigma, const float* rgb, const float* sigma, const void* stash,
                 void* dz, hipStream_t s) {
#define NERF_BWD_INST(W_, D_, S_, LX_, LD_)                                                   \
  if (P.W == W_ && P.D == D_ && P.skip == S_ && P.LX == LX_ && P.LD == LD_) {                \
    if (P.prec == NERF_PREC_F32)                                                              \
      return launch_bwd_f32<W_, D_, S_, LX_, LD_>(P, packed, M, d_rgb, d_sigma, rgb, sigma,   \
                                                  stash, dz, s);                              \
    return launch_bwd_bf16<W_, D_, S_, LX_, LD_>(P, packed, M, d_rgb, d_sigma, rgb, sigma,    \
                                                 stash, dz, s);                               \
  }
  NERF_MLP_INSTANCES(NERF_BWD_INST)
#undef NERF_BWD_INST
  return -1;
}

}  // namespace

extern "C" int32_t nerf_mlp_backward(const nerf_mlp_desc* desc, int32_t precision,
                                     const void* packed, int64_t M, const float* d_rgb,
                                     const float* d_sigma, const float* rgb, const float* sigma,
                                     const void* stash, void* dz, void* stream) {
  MlpPlan P;
  NERF_CHECK_ARG(nerf_plan_from_desc(desc, precision, &P) == 0,
                 "nerf_mlp_backward: unsupported MLP description or precision");
  NERF_CHECK_ARG(M >= 0 && packed && d_sigma && sigma && stash && dz,
                 "nerf_mlp_backward: null pointer");
  NERF_CHECK_ARG(!P.rgb || (d_rgb && rgb), "nerf_mlp_backward: d_rgb and rgb required");
  if (M == 0) return NERF_OK;
  const int rc = dispatch_bwd(P, packed, M, d_rgb, d_sigma, rgb, sigma, stash, dz, as_stream(stream));
  NERF_CHECK_ARG(rc == 0, "nerf_mlp_backward: no kernel for this topology/precision");
  NERF_LAUNCH_CHECK("nerf_mlp_backward");
  return NERF_OK;
}

extern "C" int64_t nerf_mlp_wgrad_workspace_bytes(const nerf_mlp_desc* desc, int64_t M) {
  MlpPlan P;
  if (nerf_plan_from_desc(desc, 0, &P) || M < 0) return -1;
  const int64_t s = splits_for(M);
  const int64_t kmax = (P.W > 2 * P.KX ? P.W : 2 * P.KX);
  return s * (kmax * P.W + P.W) * 4 + 256;
}

extern "C" int32_t nerf_mlp_weight_grads(const nerf_mlp_desc* desc, int32_t precision, int64_t M,
                                         const void* stash, const void* dz, float* grad_flat,
                                         int32_t accumulate, void* workspace, void* stream) {
  MlpPlan P;
  NERF_CHECK_ARG(nerf_plan_from_desc(desc, precision, &P) == 0,
                 "nerf_mlp_weight_grads: unsupported MLP description or precision");
  NERF_CHECK_ARG(M > 0 && stash && dz && grad_flat && workspace,
                 "nerf_mlp_weight_grads: bad arguments");
  hipStream_t s = as_stream(stream);
  WgradJob jobs[2 * NERF_MAX_LAYERS];
  const int nj = wgrad_jobs(P, jobs);
  const int64_t splits = splits_for(M);
  const int64_t mps = (M + splits - 1) / splits;
  float* part = reinterpret_cast<float*>(workspace);
  for (int t = 0; t < nj; ++t) {
    const WgradJob& J = jobs[t];
    const int64_t kmax = (P.W > 2 * P.KX ? P.W : 2 * P.KX);
    float* bpart = J.bias ? part + splits * kmax * P.W : nullptr;
    GemmArgs a;
    a.X = stash;
    a.x_bf16 = precision == NERF_PREC_BF16;
    a.ldx = P.stash_per_sample;
    a.xoff = J.xoff;
    a.Z = reinterpret_cast<const float*>(dz);
    a.ldz = P.dz_per_sample;
    a.zoff = J.zoff;
    a.M = M;
    a.m_per_split = mps;
    a.K = J.K;
    a.N = J.N;
    a.tiles_k = (J.K + kTile - 1) / kTile;
    a.tiles_n = (J.N + kTile - 1) / kTile;
    a.splits = (int)splits;
    a.part = part;
    a.bpart = bpart;
    const int64_t nb = (int64_t)a.tiles_k * a.tiles_n * ((splits + 7) / 8) * 8;
    if (a.x_bf16)
      wgrad_gemm_bf16_kernel<<<(unsigned)nb, 256, 0, s>>>(a);
    else
      wgrad_gemm_kernel<<<(unsigned)nb, 256, 0, s>>>(a);
    ReduceArgs r;
    r.part = part;
    r.bpart = bpart;
    r.K = J.K;
    r.N = J.N;
    r.n_real = J.n_real;
    r.splits = (int)splits;
    r.enc_L = J.enc_L;
    r.enc_KS = J.enc_KS;
    r.row_base = J.row_base;
    r.k_off = P.k_off[J.keras];
    r.b_off = P.b_off[J.keras];
    r.fan_out = P.fan_out[J.keras];
    r.accumulate = accumulate;
    r.grad = grad_flat;
    const int64_t tot = (int64_t)J.K * J.N + J.N;
    int64_t rb = (tot + 255) / 256;
    if (rb > 1024) rb = 1024;
    wgrad_reduce_kernel<<<(unsigned)rb, 256, 0, s>>>(r);
  }
  NERF_LAUNCH_CHECK("nerf_mlp_weight_grads");
  return NERF_OK;
}
