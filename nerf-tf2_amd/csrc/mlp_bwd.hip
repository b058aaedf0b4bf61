// Hand-written backward of the fused NeRF MLP (the gradient TF's GradientTape
// takes through model.py:355-392) and its weight gradients.
//
// nerf_mlp_backward: one wave per 32 samples walks the layers top-down (the
// reference topology runs the register-resident twin of mlp_bwd_f32r.h in fp32
// and the eight-wave kernel of mlp_bwd_bf16ws8.h in bf16, same rows in and out).
//   rgb = sigmoid(z):  dz_rgb = d_rgb * rgb (1 - rgb)
//   sigma = relu(z):   dz_sigma = d_sigma * [sigma > 0]      (softplus: * (1 - e^-sigma))
//   dh = W dz (MFMA with the transposed packed weights; the accumulators of one
//   layer are the B operand of the next, as in the forward), dz = dh * [h > 0]
//   using the forward's stashed activations (or its ReLU bits) for the masks.
//   Every pre-activation gradient dz_l is written for the weight GEMMs: fp32
//   rows, or bf16 blocked by 32 samples (MlpPlan::sblk).
// nerf_mlp_weight_grads: dW_l = X_l^T dz_l and db_l = sum dz_l over the samples,
//   as a split-K MFMA GEMM whose partial tiles are summed in a fixed order.
#include <atomic>
#include <mutex>

#include "common.h"
#include "mlp_kernels.h"
#include "mlp_bwd_bf16ws.h"
#include "mlp_bwd_bf16ws8.h"
#include "mlp_bwd_f32r.h"

namespace {

typedef unsigned int u32x4_ __attribute__((ext_vector_type(4)));

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
__device__ __forceinline__ void relu_mask_rows_bf16(floatx16 (&acc)[NBo], BlkRow<const __bf16> row, int64_t off,
                                                    int g) {
#pragma unroll
  for (int nb = 0; nb < NBo; ++nb)
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const bf16x4 h = *reinterpret_cast<const bf16x4*>(row.at(off + 32 * nb + 8 * a + 4 * g));
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[nb][4 * a + r] = (float)h[r] > 0.f ? acc[nb][4 * a + r] : 0.f;
    }
}

// bf16 twin of mlp_bwd_f32_kernel: dz is rounded to bf16 as the B operand of the
// next MFMA and in the rows written for the weight GEMMs (bf16 dz rows).
template <int W, int D, int SKIP, int LX, int LD>
__global__ __launch_bounds__(256) void mlp_bwd_bf16_kernel(
    MlpPlan P, const char* __restrict__ pk, int64_t M, const float* __restrict__ d_rgb,
    const float* __restrict__ d_sigma, const float* __restrict__ rgb,
    const float* __restrict__ sigma, const __bf16* __restrict__ stash, __bf16* __restrict__ dz) {
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
  // this sample's rows of the stash and dz, blocked by 32 samples (the wave's block;
  // lanes past M read the block's last row)
  const int nrow = (int)(M - blk * 32 < 32 ? M - blk * 32 : 32);
  const BlkRow<const __bf16> srow{stash + blk * 32 * P.stash_per_sample, nrow, (int)(mm - blk * 32)};
  const BlkRow<__bf16> zrow{dz + blk * 32 * P.dz_per_sample, nrow, (int)(mm - blk * 32)};
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
    relu_mask_rows_bf16(accr, srow, P.st_rbr, g);
    if (valid) {
      store_rows_bf16(accr, zrow, P.dz_rbr, g);
      if (g == 0) {
        *zrow.at(P.dz_rgb + 0) = (__bf16)dzr[0];
        *zrow.at(P.dz_rgb + 1) = (__bf16)dzr[1];
        *zrow.at(P.dz_rgb + 2) = (__bf16)dzr[2];
      }
    }
    wave_sync();
    store_act_bf16(accr, act, lane);
    zero_acc(acc);
    wave_sync();
    mfma_lds_bf16<C::NB, 2 * C::NBR>(acc, reinterpret_cast<const bf16x8*>(pk + P.p_bwd[D + 1]), act,
                                     lane);
    if (valid) store_rows_bf16(acc, zrow, P.dz_bott, g);
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
  if (valid && g == 0) *zrow.at(P.dz_sigma) = (__bf16)dzs;
  relu_mask_rows_bf16(acc, srow, P.st_h[D - 1], g);
  if (valid) store_rows_bf16(acc, zrow, P.dz_h[D - 1], g);
#pragma unroll 1
  for (int l = D - 1; l >= 1; --l) {
    wave_sync();
    store_act_bf16(acc, act, lane);
    zero_acc(acc);
    wave_sync();
    mfma_lds_bf16<C::NB, TH>(acc, reinterpret_cast<const bf16x8*>(pk + P.p_bwd[l]), act, lane);
    relu_mask_rows_bf16(acc, srow, P.st_h[l - 1], g);
    if (valid) store_rows_bf16(acc, zrow, P.dz_h[l - 1], g);
  }
}

// ---------------------------------------------------------------------------
// fp32 weight-gradient GEMM: part[split][k][n] = sum over the split's samples of
// X[m][xoff + k] * Z[m][zoff + n], for k < K, n < N.  128x128 tile per block,
// 4 waves of 64x64 (2x2 blocks of v_mfma_f32_32x32x2_f32), 32-sample LDS chunks.
// The block with tile_k == 0 also sums the columns of Z (bias gradient).
// ---------------------------------------------------------------------------
constexpr int kTile = 128, kChunk = 32, kPad = 4;

struct GemmArgs {
  const float* X;
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
  // bias sums from the loaded dz values: thread t adds its 16 columns of every
  // row it loads; the 32 threads of a column group meet at the end
  float bs[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) bs[e] = 0.f;
  const bool do_bias = a.bpart && tk == 0;
  // loader mapping: 256 threads x 16 floats = 32 rows x 128 cols; the next chunk is
  // loaded into registers while the current one is multiplied
  const int lr = tid >> 3, lc = (tid & 7) * 16;
  floatx4 xv[4], zv[4];
  auto load = [&](int64_t c0) {
    const int64_t mrow = c0 + lr;
    const bool rok = mrow < me;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int col = lc + 4 * v;
      xv[v] = floatx4{0.f, 0.f, 0.f, 0.f};
      zv[v] = floatx4{0.f, 0.f, 0.f, 0.f};
      if (rok) {
        const int64_t xo = mrow * a.ldx + a.xoff + k0 + col;
        if (k0 + col + 3 < a.K) {
          xv[v] = *reinterpret_cast<const floatx4*>(a.X + xo);
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e)
            if (k0 + col + e < a.K) xv[v][e] = a.X[xo + e];
        }
        if (n0 + col + 3 < a.N) {
          zv[v] = *reinterpret_cast<const floatx4*>(a.Z + mrow * a.ldz + a.zoff + n0 + col);
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e)
            if (n0 + col + e < a.N) zv[v][e] = a.Z[mrow * a.ldz + a.zoff + n0 + col + e];
        }
      }
    }
  };
  if (mb < me) load(mb);
  for (int64_t c0 = mb; c0 < me; c0 += kChunk) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int col = lc + 4 * v;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        Xs[lr][col + e] = xv[v][e];
        Zs[lr][col + e] = zv[v][e];
        bs[4 * v + e] += zv[v][e];
      }
    }
    __syncthreads();
    if (c0 + kChunk < me) load(c0 + kChunk);  // wave-uniform
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
    // threads lr = 0..31 of column group t & 7 hold partial sums: through LDS, fixed order
    float* red = &Xs[0][0];  // [32 rows][128 columns] fits the 32 x 132 tile
#pragma unroll
    for (int e = 0; e < 16; ++e) red[lr * 128 + lc + e] = bs[e];
    __syncthreads();
    if (tid < kTile) {
      float v = 0.f;
#pragma unroll 8
      for (int rr = 0; rr < kChunk; ++rr) v += red[rr * 128 + tid];
      if (n0 + tid < a.N) a.bpart[(int64_t)split * a.N + n0 + tid] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// bf16 weight-gradient GEMM, one pass.  Workgroup s owns the sample range of
// split s and computes the whole K x N (<= 256 x 256) partial tile, so every X
// (bf16 stash) and dz (bf16) row is read once per layer.  64-sample chunks are
// staged transposed, [feature][sample], into an LDS tile as packed sample
// pairs, so each lane's MFMA operand (8 consecutive samples of one feature) is
// one ds_read_b128; v_mfma_f32_32x32x16_bf16, fp32 accumulation.  While chunk c
// is multiplied, chunk c+1 goes from registers into the other LDS buffer (or,
// with one buffer, into the same one after a barrier) and chunk c+2 (and c+3,
// with two register sets) is in flight; no branch in the loop.  Loads go through
// buffer resources whose range ends at the split's last sample, so samples past
// it read as zero.  Tile split over the 4 waves of wgrad_bf16_kernel: with
// NB >= 4 column blocks, wave w owns column blocks w, w + 4 and every row
// block; otherwise wave w owns row blocks w, w + 4.  Rows k >= K and columns
// n >= N of the tile are not written out (they hold the neighbouring features
// up to the next 8-feature chunk, zeros past it: the generic kernel does not
// load whole chunks beyond K or N).  Bias sums (sum over samples of dz) are
// taken from the staged bf16 dz.
// ---------------------------------------------------------------------------
constexpr int kWC = 64;       // samples per chunk
constexpr int kWP = kWC + 8;  // bf16 per LDS row (16-byte aligned rows, staggered banks)
constexpr int kWF = 256;      // max features per tile side

// Several layers' GEMMs of the same tile shape go in one launch: workgroup b
// runs job b / splits on sample split b % splits, so fewer splits per job keep
// the machine full and the partial tiles (and their reduction) small.
constexpr int kMaxJobs = 8;
struct Wgrad16Job {
  int64_t xoff, zoff;  // X / dz column offsets of the layer
  int K, N;
  float* part;   // [splits][K][N]
  float* bpart;  // [splits][N] or null
};
struct Wgrad16Args {
  const __bf16* X;
  int64_t ldx;
  const __bf16* Z;
  int64_t ldz;
  int64_t M, m_per_split;  // m_per_split: a multiple of the sample block
  int splits, njobs;
  int sb_shift;            // log2 of the stash / dz sample block (MlpPlan::sblk)
  Wgrad16Job job[kMaxJobs];
};

// Staging of blocked buffers (sample block 32): a wave owns CW 8-feature chunks of
// the tile for the whole launch; one load covers 64 / CW consecutive rows of each
// (runs of 128-512 contiguous bytes), lane l row l % (64 / CW) of chunk l / (64 / CW).
// Lanes l and l ^ 1 then hold rows 2q and 2q + 1 of one chunk: they trade halves
// (DPP), and each writes the sample pairs {2q, 2q + 1} of four features (the even
// lane features 0-3, the odd lane 4-7) into the transposed tile.
__device__ __forceinline__ uint32_t dpp_swap1(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);  // quad_perm [1, 0, 3, 2]
}
// d: this lane's 16 bytes (8 bf16 features of its row); result: bf16x2 dwords
// {row 2q, row 2q + 1} of features 4 odd + i, i = 0..3.  LANE_SEL: lane-dependent
// byte selectors put the even row's half low (own bytes first on even lanes, the
// partner's on odd; fewer VALU, two more registers); otherwise the even and odd
// rows' dwords are selected first and the selectors are constants.
template <bool LANE_SEL>
__device__ __forceinline__ u32x4_ pair_rows(u32x4_ d, int odd) {
  const uint32_t d0 = d[0], d1 = d[1], d2 = d[2], d3 = d[3];  // copied out of the vector first
  const uint32_t x0 = dpp_swap1(odd ? d0 : d2), x1 = dpp_swap1(odd ? d1 : d3);  // the partner's half
  const uint32_t o0 = odd ? d2 : d0, o1 = odd ? d3 : d1;                        // this lane's half
  u32x4_ r;
  if constexpr (LANE_SEL) {
    const uint32_t slo = odd ? 0x01000504u : 0x05040100u, shi = odd ? 0x03020706u : 0x07060302u;
    r[0] = __builtin_amdgcn_perm(x0, o0, slo);  // low halves: feature 4 odd + 0
    r[1] = __builtin_amdgcn_perm(x0, o0, shi);  // high halves: + 1
    r[2] = __builtin_amdgcn_perm(x1, o1, slo);
    r[3] = __builtin_amdgcn_perm(x1, o1, shi);
  } else {
    const uint32_t e0 = odd ? x0 : o0, e1 = odd ? x1 : o1;  // the even row's dwords
    const uint32_t f0 = odd ? o0 : x0, f1 = odd ? o1 : x1;  // the odd row's
    r[0] = __builtin_amdgcn_perm(f0, e0, 0x05040100u);
    r[1] = __builtin_amdgcn_perm(f0, e0, 0x07060302u);
    r[2] = __builtin_amdgcn_perm(f1, e1, 0x05040100u);
    r[3] = __builtin_amdgcn_perm(f1, e1, 0x07060302u);
  }
  return r;
}

// Byte offset of 8-element chunk ch of row m of a split (rows [0, rows) from its
// first sample; splits start on 32-sample block boundaries) in a buffer blocked
// by 32 samples with ld elements per sample, in 32-bit arithmetic (a split's
// buffer stays under 2^31 bytes); rows past the split get an offset outside its
// buffer resource (they read zeros).
__device__ __forceinline__ uint32_t blk32_chunk_off(uint32_t m, uint32_t ch, uint32_t ld, uint32_t rows) {
  const uint32_t b = m & ~31u;
  const uint32_t n = rows - b < 32u ? rows - b : 32u;
  return m < rows ? (b * ld + 8u * (ch * n + (m - b))) * 2u : 0x80000000u;
}

// blk_chunk_off for tile-chunks that end inside the split's whole blocks: lane
// base ((e / 8) 32 + ri) 16 (chunk at element e, row ri), tile-chunk c (64 rows,
// two blocks of 32), load k (rows k RI .. k RI + RI - 1, inside one block as RI
// divides 32); ld elements per sample.
template <int RI>
__device__ __forceinline__ uint32_t blk_fast_off(uint32_t lane_base, int c, int k, int64_t ld) {
  const uint32_t blk_bytes = (uint32_t)(64 * ld);  // 32 rows of ld bf16
  return lane_base + (uint32_t)(2 * c + ((k * RI) >> 5)) * blk_bytes + (uint32_t)(((k * RI) & 31) * 16);
}
// The 8 bf16 features of d as floats, added to s.
__device__ __forceinline__ void add_bf16x8(float (&s)[8], u32x4_ d) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t w = d[k];
    s[2 * k] += __builtin_bit_cast(float, w << 16);
    s[2 * k + 1] += __builtin_bit_cast(float, w & 0xFFFF0000u);
  }
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rows_rsrc(const __bf16* base, int64_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<__bf16*>(base), 0, (int)bytes, 0x00020000);
}

// The operands are blocked by 32 samples and staged as described at pair_rows:
// wave w owns NFG / 4 8-feature chunks of each operand (NFG = the operand's
// chunks), and the transposed LDS tile stores sample pair p of feature row R at
// pair position p ^ 4 ((R >> 4) & 7), which keeps every 8-sample MFMA operand
// contiguous and spreads the pair writes of a wave over the banks.
// The LDS tiles hold 8 NFG feature rows of each operand (NFG = the operand's
// 8-feature chunks), double-buffered except for the 64 x 256, 256 x 64 and
// 256 x 128 tiles: their registers allow two workgroups per CU, which two
// buffers (92-110 KB) would not; these jobs stream their rows at one chunk of
// loads in flight per thread, and the second workgroup doubles the bytes in
// flight (the 64 x 256 encoding-slot jobs of the 8x256 net: 220 -> 150 us for
// 720,896 samples) for one more barrier per chunk.
__host__ __device__ constexpr int wg16_nfg(int b) { return b >= 8 ? 32 : b >= 4 ? 16 : 8; }
__host__ __device__ constexpr int wg16_nbuf(int kb, int nb) {
  return (kb * nb == 16 && kb != 4) || (kb == 8 && nb == 4) ? 1 : 2;
}
// Waves per SIMD the register allocation must allow: two for the 256 x 128 tile
// (128 accumulators a wave; 247 registers then, 288 otherwise, which ran one
// workgroup per CU).  The 128 x 256 tile would spill under the same bound.
__host__ __device__ constexpr int wg16_min_waves(int kb, int nb) { return kb == 8 && nb == 4 ? 2 : 1; }
// Register sets of loads in flight: two where one set is at most 10 16-byte
// loads and the tile at most 8 blocks (the registers then stay within 256);
// the larger tiles would spill with two.
__host__ __device__ constexpr int wg16_nset(int kb, int nb) {
  return kb * nb <= 8 && wg16_nfg(kb) / 4 + wg16_nfg(nb) / 4 <= 10 ? 2 : 1;
}

template <int KB, int NB>
__global__ __launch_bounds__(256, wg16_min_waves(KB, NB)) void wgrad_bf16_kernel(Wgrad16Args a) {
  constexpr bool kModeN = NB >= 4;
  constexpr int RB = kModeN ? KB : (KB >= 8 ? 2 : 1);  // row blocks per wave
  constexpr int CB = kModeN ? NB / 4 : NB;             // column blocks per wave
  constexpr int NFGX = wg16_nfg(KB), NFGZ = wg16_nfg(NB);
  constexpr int CWX = NFGX / 4, CWZ = NFGZ / 4;  // chunks per wave
  constexpr int RIX = 64 / CWX, RIZ = 64 / CWZ;  // rows per load
  constexpr int NBUF = wg16_nbuf(KB, NB);
  constexpr int NSET = wg16_nset(KB, NB);
  __shared__ __attribute__((aligned(16))) __bf16 Xt[NBUF][8 * NFGX][kWP];
  __shared__ __attribute__((aligned(16))) __bf16 Zt[NBUF][8 * NFGZ][kWP];
  const int split = blockIdx.x % a.splits;
  const Wgrad16Job J = a.job[blockIdx.x / a.splits];
  const int64_t mb = (int64_t)split * a.m_per_split;
  const int64_t me = mb + a.m_per_split < a.M ? mb + a.m_per_split : a.M;
  const int64_t rows = me > mb ? me - mb : 0;
  const int t = threadIdx.x, lane = t & 63, r = lane & 31, q = lane >> 5;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  // this lane's chunk (fixed) and row of each operand
  const int chx = wv * CWX + lane / RIX, rix = lane % RIX, chz = wv * CWZ + lane / RIZ, riz = lane % RIZ;
  const int swx = 4 * ((chx >> 1) & 7), swz = 4 * ((chz >> 1) & 7);
  const int qx = q ^ (r >> 4);  // operand reads: swizzled 8-sample group of this lane
  const __amdgpu_buffer_rsrc_t rx = rows_rsrc(a.X + mb * a.ldx, rows * a.ldx * 2);
  const __amdgpu_buffer_rsrc_t rz = rows_rsrc(a.Z + mb * a.ldz, rows * a.ldz * 2);

  floatx16 acc[RB][CB];
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[i][c][e] = 0.f;
  float bsum[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) bsum[e] = 0.f;

  u32x4_ xr[NSET][CWX], zr[NSET][CWZ];  // NSET chunks of loads in flight
  const uint32_t bxf = (uint32_t)((((J.xoff >> 3) + chx) * 32 + rix) * 16);  // fast-path lane bases
  const uint32_t bzf = (uint32_t)((((J.zoff >> 3) + chz) * 32 + riz) * 16);
  const uint32_t chx32 = (uint32_t)((J.xoff >> 3) + chx), chz32 = (uint32_t)((J.zoff >> 3) + chz);
  const uint32_t ldx32 = (uint32_t)a.ldx, ldz32 = (uint32_t)a.ldz, rows32 = (uint32_t)rows;
  // chunks past the job's K (N) features: offsets outside the buffer, so that the
  // padding of a narrow tile reads zeros instead of the neighbouring features
  const uint32_t xoob = chx < ((J.K + 7) >> 3) ? 0u : 0x80000000u;
  const uint32_t zoob = chz < ((J.N + 7) >> 3) ? 0u : 0x80000000u;
  auto load_chunk = [&](auto set, int c) {
    constexpr int S = decltype(set)::value;
    if ((int64_t)(c + 1) * kWC <= (rows & ~(int64_t)31)) {  // whole blocks (wave-uniform)
#pragma unroll
      for (int k = 0; k < CWX; ++k)
        xr[S][k] = __builtin_amdgcn_raw_buffer_load_b128(rx, blk_fast_off<RIX>(bxf, c, k, a.ldx) | xoob, 0, kNT);
#pragma unroll
      for (int k = 0; k < CWZ; ++k)
        zr[S][k] = __builtin_amdgcn_raw_buffer_load_b128(rz, blk_fast_off<RIZ>(bzf, c, k, a.ldz) | zoob, 0, kNT);
    } else {
#pragma unroll
      for (int k = 0; k < CWX; ++k)
        xr[S][k] = __builtin_amdgcn_raw_buffer_load_b128(
            rx, blk32_chunk_off((uint32_t)(c * kWC + k * RIX + rix), chx32, ldx32, rows32) | xoob, 0, kNT);
#pragma unroll
      for (int k = 0; k < CWZ; ++k)
        zr[S][k] = __builtin_amdgcn_raw_buffer_load_b128(
            rz, blk32_chunk_off((uint32_t)(c * kWC + k * RIZ + riz), chz32, ldz32, rows32) | zoob, 0, kNT);
    }
  };
  auto store_chunk = [&](auto set, int buf) {
    constexpr int S = decltype(set)::value;
#pragma unroll
    for (int k = 0; k < CWX; ++k) {
      const int col = 2 * (((k * RIX + rix) >> 1) ^ swx);
      const u32x4_ px = pair_rows<true>(xr[S][k], rix & 1);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const uint32_t v = px[i];
        *reinterpret_cast<uint32_t*>(&Xt[buf][8 * chx + 4 * (rix & 1) + i][col]) = v;
      }
    }
#pragma unroll
    for (int k = 0; k < CWZ; ++k) {
      const int col = 2 * (((k * RIZ + riz) >> 1) ^ swz);
      const u32x4_ pz = pair_rows<true>(zr[S][k], riz & 1);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const uint32_t v = pz[i];
        *reinterpret_cast<uint32_t*>(&Zt[buf][8 * chz + 4 * (riz & 1) + i][col]) = v;
      }
      add_bf16x8(bsum, zr[S][k]);
    }
  };
  auto row_block = [&](int i) { return kModeN ? i : (wv + 4 * i < KB ? wv + 4 * i : KB - 1); };
  auto col_block = [&](int c) { return kModeN ? wv + 4 * c : c; };
  // 8 samples 16 s + 8 q .. of feature row 32 blk + r, at its swizzled position
  auto operand = [&](const __bf16 (*tile)[kWP], int blk, int s) {
    return *reinterpret_cast<const bf16x8*>(&tile[32 * blk + r][16 * (s ^ (blk & 3)) + 8 * qx]);
  };
  auto compute = [&](int buf) {
#pragma unroll
    for (int s = 0; s < kWC / 16; ++s) {
      bf16x8 bz[CB];
#pragma unroll
      for (int c = 0; c < CB; ++c) bz[c] = operand(Zt[buf], col_block(c), s);
#pragma unroll
      for (int i = 0; i < RB; ++i) {
        const bf16x8 ax = operand(Xt[buf], row_block(i), s);
#pragma unroll
        for (int c = 0; c < CB; ++c) acc[i][c] = mfma_bf16(ax, bz[c], acc[i][c]);
      }
    }
  };

  // chunk k goes to LDS buffer k % 2 (or, after a barrier, the one buffer) from
  // register set k % NSET, loaded NSET chunks before
  const int nchunks = (int)((rows + kWC - 1) / kWC);
  const std::integral_constant<int, 0> set0;
  if constexpr (NSET == 2) {
    const std::integral_constant<int, NSET - 1> set1;
    constexpr int B1 = NBUF == 2 ? 1 : 0;
    load_chunk(set0, 0);
    load_chunk(set1, 1);
    store_chunk(set0, 0);
    load_chunk(set0, 2);
    __syncthreads();
    for (int c = 0; c < nchunks; c += 2) {
      compute(0);  // chunk c
      if constexpr (NBUF == 1) __syncthreads();
      store_chunk(set1, B1);  // chunk c+1 (zeros past the end; never read then)
      load_chunk(set1, c + 3);
      __syncthreads();
      if (c + 1 >= nchunks) break;
      compute(B1);  // chunk c+1
      if constexpr (NBUF == 1) __syncthreads();
      store_chunk(set0, 0);  // chunk c+2
      load_chunk(set0, c + 4);
      __syncthreads();
    }
  } else {
    load_chunk(set0, 0);
    store_chunk(set0, 0);
    load_chunk(set0, 1);
    __syncthreads();
    for (int c = 0; c < nchunks; ++c) {
      if constexpr (NBUF == 2) {
        compute(c & 1);
        store_chunk(set0, (c + 1) & 1);  // chunk c+1 (zeros past the end; never read then)
      } else {
        compute(0);
        __syncthreads();
        store_chunk(set0, 0);
      }
      load_chunk(set0, c + 2);
      __syncthreads();
    }
  }

  float* out = J.part + (int64_t)split * J.K * J.N;
#pragma unroll
  for (int i = 0; i < RB; ++i) {
    const int kb = kModeN ? i : wv + 4 * i;
    if (kb >= KB) continue;
#pragma unroll
    for (int c = 0; c < CB; ++c) {
      const int nn = 32 * col_block(c) + r;
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int kk = 32 * kb + 8 * (e / 4) + 4 * q + (e % 4);
        if (kk < J.K && nn < J.N) out[(int64_t)kk * J.N + nn] = acc[i][c][e];
      }
    }
  }
  if (J.bpart) {
    // the RIZ lanes of a chunk (all in this wave), in a fixed order
#pragma unroll
    for (int e = 0; e < 8; ++e) {
#pragma unroll
      for (int msk = 1; msk < RIZ; msk <<= 1) bsum[e] += shfl_xor(bsum[e], msk);
    }
    if (riz == 0) {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int n = 8 * chz + e;
        if (n < J.N) J.bpart[(int64_t)split * J.N + n] = bsum[e];
      }
    }
  }
}

// The 256 x 256 tile with 8 waves (two per SIMD): wave w owns the 2 x 4 blocks
// of output column blocks 2 (w % 4) + {0, 1} and row blocks 4 (w / 4) + {0..3}
// (128 accumulators), so a k-step reads 6 operands from LDS for 8 MFMAs, and
// registers remain for two chunks of loads in flight per thread (chunk k is
// loaded into register set k % 2 two iterations before it goes to LDS).  Same
// chunking, staging (pair_rows) and transposed XOR-swizzled LDS tiles as
// wgrad_bf16_kernel: wave w owns feature chunks 4 w .. 4 w + 3, a load covers 16
// rows of each.
__global__ __launch_bounds__(512) void wgrad_bf16_w8_kernel(Wgrad16Args a) {
  constexpr int NFG = 32, CW = NFG / 8, RI = 64 / CW;  // feature chunks, chunks per wave, rows per load
  typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
  __shared__ __attribute__((aligned(16))) __bf16 Xt[2][kWF][kWP];
  __shared__ __attribute__((aligned(16))) __bf16 Zt[2][kWF][kWP];
  const int split = blockIdx.x % a.splits;
  const Wgrad16Job J = a.job[blockIdx.x / a.splits];
  const int64_t mb = (int64_t)split * a.m_per_split;
  const int64_t me = mb + a.m_per_split < a.M ? mb + a.m_per_split : a.M;
  const int64_t rows = me > mb ? me - mb : 0;
  const int t = threadIdx.x, lane = t & 63, r = lane & 31, q = lane >> 5;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  const int chunk = wv * CW + lane / RI, rin = lane % RI, odd = rin & 1;  // this lane's chunk (fixed) and row
  const int sw = 4 * ((chunk >> 1) & 7);
  const int qx = q ^ (r >> 4);
  const __amdgpu_buffer_rsrc_t rx = rows_rsrc(a.X + mb * a.ldx, rows * a.ldx * 2);
  const __amdgpu_buffer_rsrc_t rz = rows_rsrc(a.Z + mb * a.ldz, rows * a.ldz * 2);

  floatx16 acc[8];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[i][e] = 0.f;
  float bsum[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) bsum[e] = 0.f;

  u32x4_ xr[2][CW], zr[2][CW];
  const uint32_t chx32 = (uint32_t)((J.xoff >> 3) + chunk), chz32 = (uint32_t)((J.zoff >> 3) + chunk);
  const uint32_t ldx32 = (uint32_t)a.ldx, ldz32 = (uint32_t)a.ldz, rows32 = (uint32_t)rows;
  // per-load byte offsets in the VGPR offset, which the buffer range check covers
  // (rows past the split read zeros), not in the unchecked scalar offset
  auto load_chunk = [&](auto set, int c) {
    constexpr int S = decltype(set)::value;
#pragma unroll
    for (int k = 0; k < CW; ++k) {
      const uint32_t row = (uint32_t)(c * kWC + k * RI + rin);
      xr[S][k] = __builtin_amdgcn_raw_buffer_load_b128(rx, blk32_chunk_off(row, chx32, ldx32, rows32), 0, kNT);
      zr[S][k] = __builtin_amdgcn_raw_buffer_load_b128(rz, blk32_chunk_off(row, chz32, ldz32, rows32), 0, kNT);
    }
  };
  auto store_chunk = [&](auto set, int buf) {
    constexpr int S = decltype(set)::value;
#pragma unroll
    for (int k = 0; k < CW; ++k) {
      const int col = 2 * (((k * RI + rin) >> 1) ^ sw);  // the swizzled position of this lane's sample pair
      const u32x4_ px = pair_rows<false>(xr[S][k], odd), pz = pair_rows<false>(zr[S][k], odd);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const uint32_t vx = px[i], vz = pz[i];
        *reinterpret_cast<uint32_t*>(&Xt[buf][8 * chunk + 4 * odd + i][col]) = vx;
        *reinterpret_cast<uint32_t*>(&Zt[buf][8 * chunk + 4 * odd + i][col]) = vz;
      }
      add_bf16x8(bsum, zr[S][k]);
    }
  };
  auto operand = [&](const __bf16 (*tile)[kWP], int blk, int s) {
    return *reinterpret_cast<const bf16x8*>(&tile[32 * blk + r][16 * (s ^ (blk & 3)) + 8 * qx]);
  };
  const int cb0 = 2 * (wv & 3), rb0 = 4 * (wv >> 2);  // acc[2 i + c]: row block rb0 + i, column block cb0 + c
  auto compute = [&](int buf) {
#pragma unroll
    for (int s = 0; s < kWC / 16; ++s) {
      const bf16x8 bz0 = operand(Zt[buf], cb0, s), bz1 = operand(Zt[buf], cb0 + 1, s);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const bf16x8 ax = operand(Xt[buf], rb0 + i, s);
        acc[2 * i] = mfma_bf16(ax, bz0, acc[2 * i]);
        acc[2 * i + 1] = mfma_bf16(ax, bz1, acc[2 * i + 1]);
      }
      __builtin_amdgcn_sched_barrier(0);  // at most one sub-step of operands live
    }
  };

  const int nchunks = (int)((rows + kWC - 1) / kWC);
  const std::integral_constant<int, 0> set0;
  const std::integral_constant<int, 1> set1;
  load_chunk(set0, 0);
  load_chunk(set1, 1);
  store_chunk(set0, 0);
  load_chunk(set0, 2);
  __syncthreads();
  for (int c = 0; c < nchunks; c += 2) {
    compute(0);                // chunk c
    store_chunk(set1, 1);      // chunk c+1 (zeros past the end; never read then)
    load_chunk(set1, c + 3);
    __syncthreads();
    if (c + 1 >= nchunks) break;
    compute(1);                // chunk c+1
    store_chunk(set0, 0);      // chunk c+2
    load_chunk(set0, c + 4);
    __syncthreads();
  }

  // partial tile rows kk = 32 (rb0 + i) + 8 (e / 4) + 4 q + e % 4, column
  // nn = 32 (cb0 + c) + r, through a buffer resource over the tile: the VGPR offset
  // is range-checked, so rows kk >= K are dropped
  const __amdgpu_buffer_rsrc_t ro = __builtin_amdgcn_make_buffer_rsrc(
      J.part + (int64_t)split * J.K * J.N, 0, J.K * J.N * 4, 0x00020000);
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int nn = 32 * (cb0 + c) + r;
    const uint32_t vo = (uint32_t)((4 * q * J.N + nn) * 4);
    if (nn < J.N) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          // copied out first: a bit_cast of a vector element subscript reads element 0
          const float v = acc[2 * i + c][e];
          __builtin_amdgcn_raw_buffer_store_b32(
              __builtin_bit_cast(uint32_t, v), ro,
              vo + (uint32_t)((32 * (rb0 + i) + 8 * (e / 4) + (e % 4)) * J.N * 4), 0, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
  if (J.bpart) {
    // the RI lanes of a chunk (all in this wave), in a fixed order
#pragma unroll
    for (int e = 0; e < 8; ++e) {
#pragma unroll
      for (int msk = 1; msk < RI; msk <<= 1) bsum[e] += shfl_xor(bsum[e], msk);
    }
    if (rin == 0) {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int n = 8 * chunk + e;
        if (n < J.N) J.bpart[(int64_t)split * J.N + n] = bsum[e];
      }
    }
  }
}

int wg16_blocks(int n) { return n <= 1 ? 1 : n <= 2 ? 2 : n <= 4 ? 4 : 8; }

// Workgroups of the generic kernel for tile shape (kb, nb) that a CU holds at
// once (registers and LDS, from the runtime's occupancy query; 1 if it fails),
// at most 3: the splits of a narrow group scale with it.
int wg16_per_cu(int kb, int nb) {
  static std::atomic<int> memo[4][4];  // 0: not asked yet
  const int i = kb >= 8 ? 3 : kb >= 4 ? 2 : kb - 1, j = nb >= 8 ? 3 : nb >= 4 ? 2 : nb - 1;
  int v = memo[i][j].load(std::memory_order_relaxed);
  if (v) return v;
  const void* fn = nullptr;
#define NERF_WG16_FN(KB_, NB_) \
  if (kb == KB_ && nb == NB_) fn = reinterpret_cast<const void*>(&wgrad_bf16_kernel<KB_, NB_>);
#define NERF_WG16_FN_ROW(KB_) NERF_WG16_FN(KB_, 1) NERF_WG16_FN(KB_, 2) NERF_WG16_FN(KB_, 4) NERF_WG16_FN(KB_, 8)
  NERF_WG16_FN_ROW(1) NERF_WG16_FN_ROW(2) NERF_WG16_FN_ROW(4)
  NERF_WG16_FN(8, 1) NERF_WG16_FN(8, 2) NERF_WG16_FN(8, 4)
#undef NERF_WG16_FN_ROW
#undef NERF_WG16_FN
  int n = 1;
  if (!fn || hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, 256, 0) != hipSuccess) {
    (void)hipGetLastError();  // not left for the launch checks that follow
    n = 1;
  }
  if (n < 1) n = 1;
  v = n < 3 ? n : 3;
  memo[i][j].store(v, std::memory_order_relaxed);
  return v;
}

// All jobs of `a` share the tile shape (kb, nb) = wg16_blocks of their 32-feature blocks.
int launch_wgrad_bf16(const Wgrad16Args& a, int kb, int nb, hipStream_t s) {
  const unsigned grid = (unsigned)(a.splits * a.njobs);
  if (a.sb_shift != 5) return -1;  // the kernels stage buffers blocked by 32 samples
  if (kb == 8 && nb == 8) {
    wgrad_bf16_w8_kernel<<<grid, 512, 0, s>>>(a);
    return 0;
  }
#define NERF_WG16(KB_, NB_)                            \
  if (kb == KB_ && nb == NB_) {                        \
    wgrad_bf16_kernel<KB_, NB_><<<grid, 256, 0, s>>>(a); \
    return 0;                                          \
  }
#define NERF_WG16_ROW(KB_) NERF_WG16(KB_, 1) NERF_WG16(KB_, 2) NERF_WG16(KB_, 4) NERF_WG16(KB_, 8)
  NERF_WG16_ROW(1) NERF_WG16_ROW(2) NERF_WG16_ROW(4)
  NERF_WG16(8, 1) NERF_WG16(8, 2) NERF_WG16(8, 4)  // (8, 8): wgrad_bf16_w8_kernel above
#undef NERF_WG16_ROW
#undef NERF_WG16
  return -1;
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

__global__ void zero_kernel(float* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.f;
}

// Sum of n partials p[0], p[stride], ... in sequential order; the loads are
// issued 64 (then 16) at a time so that they are in flight together: a tile with
// few outputs has one thread per output, whose loads are the whole launch's.
__device__ __forceinline__ float sum_splits(const float* __restrict__ p, int64_t stride, int n) {
  float s = 0.f;
  int sp = 0;
  for (; sp + 64 <= n; sp += 64) {
    float v[64];
#pragma unroll
    for (int j = 0; j < 64; ++j) v[j] = p[(int64_t)(sp + j) * stride];
#pragma unroll
    for (int j = 0; j < 64; ++j) s += v[j];
  }
  for (; sp + 16 <= n; sp += 16) {
    float v[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = p[(int64_t)(sp + j) * stride];
#pragma unroll
    for (int j = 0; j < 16; ++j) s += v[j];
  }
  for (; sp < n; ++sp) s += p[(int64_t)sp * stride];
  return s;
}

// Blocks blk (of nblk) of the reduction of one job: element e of the K x N tile
// (and of the bias row) summed over the splits into the flat gradient.
__device__ __forceinline__ void reduce_job(const ReduceArgs& a, int64_t blk, int64_t nblk) {
  const int64_t n_el = (int64_t)a.K * a.N;
  for (int64_t e = blk * blockDim.x + threadIdx.x; e < n_el + a.N; e += nblk * blockDim.x) {
    if (e < n_el) {
      const int k = (int)(e / a.N), n = (int)(e % a.N);
      if (n >= a.n_real) continue;
      int row = k;
      if (a.enc_L) {
        const int g = k / a.enc_KS, s = k % a.enc_KS;
        row = enc_feat(g, s, a.enc_L);
        if (row < 0) continue;
      }
      const float s = sum_splits(a.part + e, n_el, a.splits);
      float* dst = a.grad + a.k_off + (int64_t)(a.row_base + row) * a.fan_out + n;
      *dst = a.accumulate ? *dst + s : s;
    } else if (a.bpart) {
      const int n = (int)(e - n_el);
      if (n >= a.n_real) continue;
      const float s = sum_splits(a.bpart + n, a.N, a.splits);
      float* dst = a.grad + a.b_off + n;
      *dst = a.accumulate ? *dst + s : s;
    }
  }
}

__global__ void wgrad_reduce_kernel(ReduceArgs a) { reduce_job(a, blockIdx.x, gridDim.x); }

// The reductions of several jobs (a group, or every job of the call) in one
// launch: job j owns blocks [first[j], first[j + 1]).
struct ReduceBatch {
  ReduceArgs job[2 * NERF_MAX_LAYERS];
  int first[2 * NERF_MAX_LAYERS + 1];
  int njobs;
};

__global__ void wgrad_reduce_batch_kernel(ReduceBatch b) {
  int j = 0;
  while (j + 1 < b.njobs && (int)blockIdx.x >= b.first[j + 1]) ++j;
  reduce_job(b.job[j], (int64_t)blockIdx.x - b.first[j], b.first[j + 1] - b.first[j]);
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
  add(P.st_h[P.D - 1], W, P.dz_sigma, 1, 1, 0, 0, 0, P.kl_sigma, 1);
  if (P.rgb) {
    add(P.st_h[P.D - 1], W, P.dz_bott, W, W, 0, 0, 0, P.kl_bott, 1);
    add(P.st_bott, W, P.dz_rbr, 32 * P.NBR, P.WR, 0, 0, 0, P.kl_rbr, 1);
    add(P.st_encd, 2 * P.KD, P.dz_rbr, 32 * P.NBR, P.WR, P.LD, P.KD, W, P.kl_rbr, 0);
    add(P.st_rbr, P.WR, P.dz_rgb, 3, 3, 0, 0, 0, P.kl_rgb, 1);
  }
  return n;
}

// Sample splits of the weight-gradient GEMMs: >= 2048 samples per split, at most
// 256 splits (one workgroup per CU for the single-pass bf16 kernel).
int64_t splits_for(int64_t M) {
  int64_t s = (M + 2047) / 2048;
  if (s > 256) s = 256;
  return s < 1 ? 1 : s;
}

// Splits per job of a group of n same-shape bf16 weight-gradient jobs (one
// launch): 256 workgroups in all, one per CU (each takes most of a CU's LDS),
// but no split under 256 samples.
int64_t wg16_splits(int64_t M, int n) {
  int64_t s = (256 + n - 1) / n;
  const int64_t cap = (M + 255) / 256;
  if (s > cap) s = cap;
  return s < 1 ? 1 : s;
}

// Fewest splits that keep one split's rows (and the 256 rows of zero padding a
// chunk may read past them) under 2^31 bytes: a split's buffer resources
// address its rows with 32-bit offsets.
int64_t wg16_min_splits(int64_t M, int64_t row_bytes) {
  const int64_t lim = (((int64_t)1 << 31) - 1) / row_bytes - 256;
  return (M + lim - 1) / lim;
}

template <int W, int D, int SKIP, int LX, int LD>
int launch_bwd_f32(const MlpPlan& P, const void* packed, int64_t M, const float* d_rgb,
                   const float* d_sigma, const float* rgb, const float* sigma, const void* stash,
                   void* dz, hipStream_t s) {
  if constexpr (W == 256 && LD > 0 && SKIP >= 0 && SKIP + 1 < D) {
    // the reference topology: the register-resident chain (16 samples per wave)
    launch_bwd_f32r<W, D, SKIP, LX, LD>(P, packed, M, d_rgb, d_sigma, rgb, sigma, stash, dz, s);
  } else {
    const int64_t nblk = (M + 31) / 32;
    mlp_bwd_f32_kernel<W, D, SKIP, LX, LD><<<(unsigned)((nblk + 3) / 4), 256, 0, s>>>(
        P, reinterpret_cast<const char*>(packed), M, d_rgb, d_sigma, rgb, sigma,
        reinterpret_cast<const float*>(stash), reinterpret_cast<float*>(dz));
  }
  return 0;
}

template <int W, int D, int SKIP, int LX, int LD>
int launch_bwd_bf16(const MlpPlan& P, const void* packed, int64_t M, const float* d_rgb,
                    const float* d_sigma, const float* rgb, const float* sigma, const void* stash,
                    void* dz, hipStream_t s) {
  const int64_t nblk = (M + 31) / 32;
  mlp_bwd_bf16_kernel<W, D, SKIP, LX, LD><<<(unsigned)((nblk + 3) / 4), 256, 0, s>>>(
      P, reinterpret_cast<const char*>(packed), M, d_rgb, d_sigma, rgb, sigma,
      reinterpret_cast<const __bf16*>(stash), reinterpret_cast<__bf16*>(dz));
  return 0;
}

// bf16: the weight-stationary chain for the reference topology, the per-wave one otherwise.
template <int W, int D, int SKIP, int LX, int LD>
int launch_bwd_bf16_any(const MlpPlan& P, const void* packed, int64_t M, const float* d_rgb,
                        const float* d_sigma, const float* rgb, const float* sigma, const void* stash,
                        void* dz, hipStream_t s) {
  // the stash / dz layout (MlpPlan::sblk) must be the one the kernel reads and writes
  if constexpr (W == 256 && LX == 10 && LD == 4 && SKIP >= 0 && SKIP + 1 < D)
    return P.sblk != 32 ? -1
                        : launch_bwd_bf16ws8<W, D, SKIP, LX, LD>(P, packed, M, d_rgb, d_sigma, rgb, sigma, stash, dz, s);
  else
    return P.sblk != 32 ? -1
                        : launch_bwd_bf16<W, D, SKIP, LX, LD>(P, packed, M, d_rgb, d_sigma, rgb, sigma, stash, dz, s);
}

int dispatch_bwd(const MlpPlan& P, const void* packed, int64_t M, const float* d_rgb,
                 const float* d_sigma, const float* rgb, const float* sigma, const void* stash,
                 void* dz, hipStream_t s) {
#define NERF_BWD_INST(W_, D_, S_, LX_, LD_)                                                   \
  if (P.W == W_ && P.D == D_ && P.skip == S_ && P.LX == LX_ && P.LD == LD_) {                \
    if (P.prec == NERF_PREC_F32)                                                              \
      return P.sblk != 1 ? -1 : launch_bwd_f32<W_, D_, S_, LX_, LD_>(P, packed, M, d_rgb,     \
                                                  d_sigma, rgb, sigma, stash, dz, s);         \
    return launch_bwd_bf16_any<W_, D_, S_, LX_, LD_>(P, packed, M, d_rgb, d_sigma, rgb, sigma,\
                                                     stash, dz, s);                           \
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
  NERF_CHECK_ARG(M >= 0, "nerf_mlp_backward: M < 0");
  if (M == 0) return NERF_OK;  // empty batch
  NERF_CHECK_ARG(packed && d_sigma && sigma && stash && dz,
                 "nerf_mlp_backward: null pointer");
  NERF_CHECK_ARG(!P.rgb || (d_rgb && rgb), "nerf_mlp_backward: d_rgb and rgb required");
  if (M == 0) return NERF_OK;
  const int rc = dispatch_bwd(P, packed, M, d_rgb, d_sigma, rgb, sigma, stash, dz, as_stream(stream));
  NERF_CHECK_ARG(rc == 0, "nerf_mlp_backward: no kernel for this topology/precision");
  NERF_LAUNCH_CHECK("nerf_mlp_backward");
  return NERF_OK;
}

// Partial-tile floats of the weight-gradient workspace: room for `s` tiles of the
// widest layer (kmax x W) and their bias rows, s = the larger of the float32
// split count and the bf16 bound: a group of n <= kMaxJobs same-shape jobs uses
// n x wg16_splits(M, n) <= min(256 + kMaxJobs - 1, kMaxJobs x ceil(M / 256)) tiles.
// This is the size callers are told to allocate; wgrad_used_floats is what one
// call's launches actually touch (never more).
static int64_t wgrad_workspace_floats(const MlpPlan& P, int64_t M) {
  int64_t b16 = kMaxJobs * ((M + 255) / 256);
  if (b16 > 256 + kMaxJobs - 1) b16 = 256 + kMaxJobs - 1;
  if (b16 < kMaxJobs) b16 = kMaxJobs;
  const int64_t row = 2 * (P.stash_per_sample > P.dz_per_sample ? P.stash_per_sample : P.dz_per_sample);
  if (b16 < kMaxJobs * wg16_min_splits(M, row)) b16 = kMaxJobs * wg16_min_splits(M, row);
  const int64_t s = splits_for(M) > b16 ? splits_for(M) : b16;
  const int64_t kmax = (P.W > 2 * P.KX ? P.W : 2 * P.KX);
  return s * (kmax * P.W + P.W);
}

// The bf16 launches: jobs grouped by tile shape (up to kMaxJobs per launch),
// each group with sg splits per job.  A group's partial tiles start at the
// workspace base and are summed before the next group's GEMM overwrites them,
// or, where the workspace holds every group's tiles side by side, all groups'
// GEMMs run first and one launch sums every split (the same sums, fewer
// launches).
struct Wg16Group {
  int members[kMaxJobs];
  int n, kb, nb;
  int64_t sg;
};

static int64_t wg16_group_floats(const Wg16Group& g, const WgradJob* jobs) {
  int64_t f = 0;
  for (int i = 0; i < g.n; ++i) {
    const WgradJob& J = jobs[g.members[i]];
    f += g.sg * J.K * J.N + (J.bias ? g.sg * J.N : 0);
  }
  return f;
}

// bound: assume the most workgroups per CU any tile shape gets (3) instead of
// asking the runtime, for a size that needs no device.
static int wg16_groups(const MlpPlan& P, int64_t M, const WgradJob* jobs, int nj, Wg16Group* groups,
                       bool bound = false) {
  bool done[2 * NERF_MAX_LAYERS] = {};
  const int64_t row_bytes = 2 * (P.stash_per_sample > P.dz_per_sample ? P.stash_per_sample : P.dz_per_sample);
  int ng = 0;
  for (int t0 = 0; t0 < nj; ++t0) {
    if (done[t0]) continue;
    if (jobs[t0].K > 256 || jobs[t0].N > 256) return -1;
    Wg16Group& g = groups[ng++];
    g.kb = wg16_blocks((jobs[t0].K + 31) / 32);
    g.nb = wg16_blocks((jobs[t0].N + 31) / 32);
    g.n = 0;
    for (int t = t0; t < nj && g.n < kMaxJobs; ++t)
      if (!done[t] && wg16_blocks((jobs[t].K + 31) / 32) == g.kb && wg16_blocks((jobs[t].N + 31) / 32) == g.nb) {
        g.members[g.n++] = t;
        done[t] = true;
      }
    g.sg = wg16_splits(M, g.n);
    if (g.sg < wg16_min_splits(M, row_bytes)) g.sg = wg16_min_splits(M, row_bytes);
  }
  // Where the generic kernel fits several workgroups per CU, a group gets that
  // many times the splits, as far as its partial tiles stay within the largest
  // group's or the allocation-free entry point's static scratch, whichever is
  // larger (so that entry point takes every batch it took before), and its tiles
  // are large enough (16,384 floats per split) for the split sums to run wide:
  // the sums over the splits of a narrow tile (one thread per output) cost more
  // than the extra workgroups gain.
  int64_t most = (int64_t)NERF_WGRAD_STATIC_SCRATCH_BYTES / 4;
  for (int i = 0; i < ng; ++i) {
    const int64_t f = wg16_group_floats(groups[i], jobs);
    if (f > most) most = f;
  }
  const int64_t cap = (M + 255) / 256;
  for (int i = 0; i < ng; ++i) {
    Wg16Group& g = groups[i];
    if (g.kb == 8 && g.nb == 8) continue;
    int64_t per = 0;  // floats per split
    for (int k = 0; k < g.n; ++k) per += jobs[g.members[k]].K * (int64_t)jobs[g.members[k]].N + jobs[g.members[k]].N;
    if (per < 16384) continue;
    const int pc = bound ? 3 : wg16_per_cu(g.kb, g.nb);
    int64_t sg = (256 * pc + g.n - 1) / g.n;
    if (sg > cap) sg = cap;
    if (sg > most / per) sg = most / per;
    if (sg > g.sg) g.sg = sg;
  }
  return ng;
}

// Workspace floats the launches of one call use (M > 0), or -1 for a layer the
// bf16 GEMM cannot take.
static int64_t wgrad_used_floats(const MlpPlan& P, int32_t precision, int64_t M) {
  WgradJob jobs[2 * NERF_MAX_LAYERS];
  const int nj = wgrad_jobs(P, jobs);
  if (precision == NERF_PREC_BF16) {
    Wg16Group groups[2 * NERF_MAX_LAYERS];
    const int ng = wg16_groups(P, M, jobs, nj, groups);
    if (ng < 0) return -1;
    int64_t need = 0;
    for (int i = 0; i < ng; ++i) {
      const int64_t f = wg16_group_floats(groups[i], jobs);
      if (f > need) need = f;
    }
    return need;
  }
  const int64_t kmax = (P.W > 2 * P.KX ? P.W : 2 * P.KX);
  return splits_for(M) * (kmax * P.W + P.W);
}

// Floats of every bf16 group's partial tiles side by side (see Wg16Group).
static int64_t wg16_stacked_floats(const MlpPlan& P, int64_t M, bool bound) {
  WgradJob jobs[2 * NERF_MAX_LAYERS];
  const int nj = wgrad_jobs(P, jobs);
  Wg16Group groups[2 * NERF_MAX_LAYERS];
  const int ng = wg16_groups(P, M, jobs, nj, groups, bound);
  int64_t f = 0;
  for (int i = 0; i < ng; ++i) f += wg16_group_floats(groups[i], jobs);
  return f;
}

// The workspace callers are told to allocate: room for either precision, in
// bf16 for every group's partial tiles side by side.
static int64_t wgrad_workspace_bytes_for(const nerf_mlp_desc* desc, int64_t M) {
  MlpPlan P, P16;
  if (nerf_plan_from_desc(desc, NERF_PREC_F32, &P) || M < 0) return -1;
  int64_t f = wgrad_workspace_floats(P, M);
  if (nerf_plan_from_desc(desc, NERF_PREC_BF16, &P16) == 0) {
    const int64_t st = wg16_stacked_floats(P16, M, true);
    if (st > f) f = st;
  }
  return f * 4 + 256;
}

extern "C" int64_t nerf_mlp_wgrad_workspace_bytes(const nerf_mlp_desc* desc, int64_t M) {
  return wgrad_workspace_bytes_for(desc, M);
}

// The launches of one weight-gradient call (M > 0) into the partial-tile buffer
// `part` of part_floats >= wgrad_used_floats(P, precision, M) floats.
static int32_t wgrad_run(const MlpPlan& P, int32_t precision, int64_t M, const void* stash,
                         const void* dz, float* grad_flat, int32_t accumulate, float* part,
                         int64_t part_floats, hipStream_t s) {
  WgradJob jobs[2 * NERF_MAX_LAYERS];
  const int nj = wgrad_jobs(P, jobs);
  const int64_t splits = splits_for(M);
  auto reduce_args = [&](const WgradJob& J, const float* jp, const float* jb, int nsplit) {
    ReduceArgs r;
    r.part = jp;
    r.bpart = jb;
    r.K = J.K;
    r.N = J.N;
    r.n_real = J.n_real;
    r.splits = nsplit;
    r.enc_L = J.enc_L;
    r.enc_KS = J.enc_KS;
    r.row_base = J.row_base;
    r.k_off = P.k_off[J.keras];
    r.b_off = P.b_off[J.keras];
    r.fan_out = P.fan_out[J.keras];
    r.accumulate = accumulate;
    r.grad = grad_flat;
    return r;
  };
  auto reduce_blocks = [](const WgradJob& J) {
    const int64_t tot = (int64_t)J.K * J.N + J.N;
    const int64_t rb = (tot + 255) / 256;
    return (int)(rb > 1024 ? 1024 : rb);
  };
  if (precision == NERF_PREC_BF16) {
    Wg16Group groups[2 * NERF_MAX_LAYERS];
    const int ng = wg16_groups(P, M, jobs, nj, groups);
    NERF_CHECK_ARG(ng > 0, "nerf_mlp_weight_grads: layer too wide for the bf16 weight GEMM");
    int64_t stacked = 0;
    for (int gi = 0; gi < ng; ++gi) stacked += wg16_group_floats(groups[gi], jobs);
    const bool stack = stacked <= part_floats;  // every group's tiles side by side
    ReduceBatch all;
    all.njobs = 0;
    all.first[0] = 0;
    float* base = part;
    for (int gi = 0; gi < ng; ++gi) {
      const Wg16Group& G = groups[gi];
      const int n = G.n;
      const int64_t sg = G.sg;
      Wgrad16Args b;
      b.X = reinterpret_cast<const __bf16*>(stash);
      b.ldx = P.stash_per_sample;
      b.Z = reinterpret_cast<const __bf16*>(dz);
      b.ldz = P.dz_per_sample;
      b.M = M;
      // splits start on sample-block boundaries of the stash / dz layout
      b.m_per_split = ((M + sg - 1) / sg + P.sblk - 1) / P.sblk * P.sblk;
      b.splits = (int)sg;
      b.sb_shift = P.sblk == 32 ? 5 : 0;
      b.njobs = n;
      float* wp = base;
      for (int i = 0; i < n; ++i) {
        const WgradJob& J = jobs[G.members[i]];
        b.job[i].xoff = J.xoff;
        b.job[i].zoff = J.zoff;
        b.job[i].K = J.K;
        b.job[i].N = J.N;
        b.job[i].part = wp;
        wp += sg * J.K * J.N;
        b.job[i].bpart = J.bias ? wp : nullptr;
        if (J.bias) wp += sg * J.N;
      }
      NERF_CHECK_ARG(launch_wgrad_bf16(b, G.kb, G.nb, s) == 0,
                     "nerf_mlp_weight_grads: no bf16 weight GEMM for this tile shape and stash layout");
      if (stack) base = wp;
      ReduceBatch one;
      ReduceBatch& rbt = stack ? all : one;
      if (!stack) {
        rbt.njobs = 0;
        rbt.first[0] = 0;
      }
      for (int i = 0; i < n; ++i) {
        const int k = rbt.njobs++;
        rbt.job[k] = reduce_args(jobs[G.members[i]], b.job[i].part, b.job[i].bpart, (int)sg);
        rbt.first[k + 1] = rbt.first[k] + reduce_blocks(jobs[G.members[i]]);
      }
      if (!stack) wgrad_reduce_batch_kernel<<<(unsigned)rbt.first[rbt.njobs], 256, 0, s>>>(rbt);
    }
    if (stack) wgrad_reduce_batch_kernel<<<(unsigned)all.first[all.njobs], 256, 0, s>>>(all);
  } else {
    for (int t = 0; t < nj; ++t) {
      const WgradJob& J = jobs[t];
      const int64_t kmax = (P.W > 2 * P.KX ? P.W : 2 * P.KX);
      float* bpart = J.bias ? part + splits * kmax * P.W : nullptr;
      GemmArgs a;
      a.X = reinterpret_cast<const float*>(stash);
      a.ldx = P.stash_per_sample;
      a.xoff = J.xoff;
      a.Z = reinterpret_cast<const float*>(dz);
      a.ldz = P.dz_per_sample;
      a.zoff = J.zoff;
      a.M = M;
      a.K = J.K;
      a.N = J.N;
      a.tiles_k = (J.K + kTile - 1) / kTile;
      a.tiles_n = (J.N + kTile - 1) / kTile;
      // three workgroups fit per CU: a four-tile layer takes 3/4 of the splits, so that
      // its 4 x 192 workgroups run as one round on 256 CUs
      const int64_t sj = a.tiles_k * a.tiles_n >= 4 && splits >= 4 ? splits * 3 / 4 : splits;
      a.m_per_split = (M + sj - 1) / sj;
      a.splits = (int)sj;
      a.part = part;
      a.bpart = bpart;
      const int64_t nb = (int64_t)a.tiles_k * a.tiles_n * ((sj + 7) / 8) * 8;
      wgrad_gemm_kernel<<<(unsigned)nb, 256, 0, s>>>(a);
      wgrad_reduce_kernel<<<(unsigned)reduce_blocks(J), 256, 0, s>>>(reduce_args(J, part, bpart, (int)sj));
    }
  }
  NERF_LAUNCH_CHECK("nerf_mlp_weight_grads");
  return NERF_OK;
}

// M == 0: the gradient of an empty batch is zero (a fill kernel, no memset node
// in a graph), or left as is when accumulating.
static int32_t wgrad_empty(const MlpPlan& P, float* grad_flat, int32_t accumulate, hipStream_t s) {
  if (!accumulate)
    zero_kernel<<<(unsigned)((P.n_params + 255) / 256 < 4096 ? (P.n_params + 255) / 256 : 4096), 256, 0, s>>>(
        grad_flat, P.n_params);
  NERF_LAUNCH_CHECK("nerf_mlp_weight_grads");
  return NERF_OK;
}

// Partial tiles of nerf_mlp_weight_grads_noalloc: one array per GPU, held until
// the process ends.  Host code never passes this name to a kernel (on the host
// it is a stand-in, not a device address): it uses the address
// hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) float scratch[NERF_WGRAD_STATIC_SCRATCH_BYTES / 4];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// This device's copy of `scratch`, looked up once per device (with the device current).
static float* scratch_on_current_device() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
    g_scratch_ptr[dev] = reinterpret_cast<float*>(p);
  }
  return g_scratch_ptr[dev];
}

extern "C" int32_t nerf_mlp_weight_grads_noalloc(const nerf_mlp_desc* desc, int32_t precision,
                                                 int64_t M, const void* stash, const void* dz,
                                                 float* grad_flat, int32_t accumulate,
                                                 void* stream) {
  MlpPlan P;
  NERF_CHECK_ARG(nerf_plan_from_desc(desc, precision, &P) == 0,
                 "nerf_mlp_weight_grads_noalloc: unsupported MLP description or precision");
  NERF_CHECK_ARG(M >= 0 && grad_flat, "nerf_mlp_weight_grads_noalloc: bad arguments");
  hipStream_t s = as_stream(stream);
  if (M == 0) return wgrad_empty(P, grad_flat, accumulate, s);
  NERF_CHECK_ARG(stash && dz, "nerf_mlp_weight_grads_noalloc: bad arguments");
  const int64_t need = wgrad_used_floats(P, precision, M);
  NERF_CHECK_ARG(need >= 0, "nerf_mlp_weight_grads_noalloc: layer too wide for the bf16 weight GEMM");
  NERF_CHECK_ARG(need * 4 <= (int64_t)NERF_WGRAD_STATIC_SCRATCH_BYTES,
                 "nerf_mlp_weight_grads_noalloc: M = %lld needs %lld bytes of partial tiles, the "
                 "static scratch holds %lld; use nerf_mlp_weight_grads with a workspace",
                 (long long)M, (long long)(need * 4), (long long)NERF_WGRAD_STATIC_SCRATCH_BYTES);
  float* part = scratch_on_current_device();
  if (!part) {
    nerf_set_error("nerf_mlp_weight_grads_noalloc: no device address for the static scratch");
    return NERF_ERR_LAUNCH;
  }
  return wgrad_run(P, precision, M, stash, dz, grad_flat, accumulate, part,
                   (int64_t)NERF_WGRAD_STATIC_SCRATCH_BYTES / 4, s);
}

extern "C" int32_t nerf_mlp_weight_grads(const nerf_mlp_desc* desc, int32_t precision, int64_t M,
                                         const void* stash, const void* dz, float* grad_flat,
                                         int32_t accumulate, void* workspace,
                                         int64_t workspace_bytes, void* stream) {
  MlpPlan P;
  NERF_CHECK_ARG(nerf_plan_from_desc(desc, precision, &P) == 0,
                 "nerf_mlp_weight_grads: unsupported MLP description or precision");
  NERF_CHECK_ARG(M >= 0 && grad_flat, "nerf_mlp_weight_grads: bad arguments");
  hipStream_t s = as_stream(stream);
  if (M == 0) return wgrad_empty(P, grad_flat, accumulate, s);
  NERF_CHECK_ARG(stash && dz && workspace, "nerf_mlp_weight_grads: bad arguments");
  const int64_t need = wgrad_used_floats(P, precision, M);
  NERF_CHECK_ARG(need >= 0, "nerf_mlp_weight_grads: layer too wide for the bf16 weight GEMM");
  NERF_CHECK_ARG(workspace_bytes >= need * 4,
                 "nerf_mlp_weight_grads: workspace of %lld bytes, this call needs %lld "
                 "(nerf_mlp_wgrad_workspace_bytes gives %lld for M = %lld)",
                 (long long)workspace_bytes, (long long)(need * 4),
                 (long long)wgrad_workspace_bytes_for(desc, M), (long long)M);
  return wgrad_run(P, precision, M, stash, dz, grad_flat, accumulate,
                   reinterpret_cast<float*>(workspace), workspace_bytes / 4, s);
}
