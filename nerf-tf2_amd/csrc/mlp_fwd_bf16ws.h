// Weight-stationary BF16 fused forward for the reference topology (W = 256).
//
// A workgroup of 4 waves carries 128 samples (4 blocks of 32) through the MLP.
// For each layer, wave w holds output blocks {2w, 2w+1} of the layer's packed
// bf16 weights in VGPRs (16 k-steps x 2 blocks x 16 B per lane) and streams
// the B operands of all 4 sample blocks from an LDS activation buffer, so each
// weight byte fetched from L2 feeds 128 samples and each LDS byte feeds 64
// output features.  The next layer's weight slice is prefetched into registers
// while the current layer runs.  Outputs (fp32 accumulators + bias +
// activation) are written as bf16 into the other LDS buffer in the B-operand
// layout; one barrier per layer.  The encodings live in LDS too; the sigma
// and rgb heads are fp32 dot products over the bf16 activations.
#pragma once
#include "mlp_kernels.h"

namespace ws {

constexpr int kWaves = 4;
constexpr int kBlocks = 4;       // sample blocks (of 32) per workgroup
constexpr int kT = 16;           // bf16 k-steps over a 256-wide hidden input
constexpr int kTX = 4;           // k-steps over the xyz encoding slots (32 per half)
constexpr int kTD = 2;           // k-steps over the direction encoding slots (16 per half)
constexpr int kSlab = kT * 64;   // bf16x8 per sample block in one activation buffer

// Load the first TL k-steps of this wave's weight slice (output blocks nb0,
// nb0+1): w[t][i] = A operand of k-step t, output block nb0 + i.
template <int TL, int TT>
__device__ __forceinline__ void load_slice(bf16x8 (&w)[TT][2], const bf16x8* __restrict__ wq, int NBo,
                                           int nb0, int lane) {
#pragma unroll
  for (int t = 0; t < TL; ++t)
#pragma unroll
    for (int i = 0; i < 2; ++i) w[t][i] = wq[(t * NBo + nb0 + i) * 64 + lane];
}

}  // namespace ws

template <int W, int D, int SKIP, int LX, int LD>
__global__ __launch_bounds__(256) void mlp_fwd_bf16ws_kernel(
    MlpPlan P, const char* __restrict__ pk, int64_t M, const float* __restrict__ xyz,
    const float* __restrict__ dirs, const float* __restrict__ ro, const float* __restrict__ rd,
    const float* __restrict__ tv, int spr, float* __restrict__ rgb_out,
    float* __restrict__ sig_out, __bf16* __restrict__ stash) {
  using namespace ws;
  using C = NetC<W, D, SKIP, LX, LD>;
  static_assert(W == 256 && LX == 10 && LD == 4, "weight-stationary kernel is for the 8x256 NeRF MLP");
  __shared__ bf16x8 s_act[2][kBlocks * kSlab];          // 2 x 64 KB
  __shared__ bf16x8 s_ex[kBlocks * kTX * 64];           // 16 KB
  __shared__ bf16x8 s_ed[kBlocks * kTD * 64];           // 8 KB
  __shared__ float s_red[kWaves][3][kBlocks * 32];     // rgb-head partials, 6 KB
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, j = lane & 31, g = lane >> 5;
  const int64_t base = (int64_t)blockIdx.x * (kBlocks * 32);

  // ---- encodings of this wave's own sample block (wave w encodes block w)
  {
    const int64_t m = base + wv * 32 + j;
    const int64_t mm = m < M ? m : M - 1;
    float x[3], d[3];
    load_sample(mm, xyz, dirs, ro, rd, tv, spr, x, d);
    float ex[C::KX], ed[C::KD];
    encode_slots<LX, C::KX>(x, g, ex);
    encode_slots<LD, C::KD>(d, g, ed);
    bf16x8 exb[kTX], edb[kTD];
    slots_to_bf16(ex, exb);
    slots_to_bf16(ed, edb);
#pragma unroll
    for (int t = 0; t < kTX; ++t) s_ex[(wv * kTX + t) * 64 + lane] = exb[t];
#pragma unroll
    for (int t = 0; t < kTD; ++t) s_ed[(wv * kTD + t) * 64 + lane] = edb[t];
    if (stash && m < M) {
      store_slots_bf16<C::KX>(ex, stash + m * P.stash_per_sample + P.st_encx + g * C::KX);
      store_slots_bf16<C::KD>(ed, stash + m * P.stash_per_sample + P.st_encd + g * C::KD);
    }
  }
  const int nb0 = 2 * wv;  // this wave's output blocks of a 256-wide layer

  // Runs one MFMA layer for all sample blocks.  `in` (or the encoding for layer
  // 0) provides T_h hidden k-steps, followed by T_e encoding k-steps from `enc`.
  // Result: bias + activation, bf16 into `out` (B-operand layout) and the stash.
  auto run_layer = [&](auto w_slice, auto t_h_c, auto t_e_c, const bf16x8* in, const bf16x8* enc,
                       int enc_T, bf16x8* out, const float* bias, int relu, int64_t st_off) {
    constexpr int T_h = decltype(t_h_c)::value;
    constexpr int T_e = decltype(t_e_c)::value;
    auto& w = *w_slice;
#pragma unroll 1
    for (int sb = 0; sb < kBlocks; ++sb) {
      floatx16 acc[2];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[i][r] = 0.f;
#pragma unroll
      for (int t = 0; t < T_h; ++t) {
        const bf16x8 b = in[sb * kSlab + t * 64 + lane];
        acc[0] = mfma_bf16(w[t][0], b, acc[0]);
        acc[1] = mfma_bf16(w[t][1], b, acc[1]);
      }
#pragma unroll
      for (int t = 0; t < T_e; ++t) {
        const bf16x8 b = enc[(sb * enc_T + t) * 64 + lane];
        acc[0] = mfma_bf16(w[T_h + t][0], b, acc[0]);
        acc[1] = mfma_bf16(w[T_h + t][1], b, acc[1]);
      }
      // bias + activation on the two output blocks nb0, nb0+1
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const floatx4 bv = *reinterpret_cast<const floatx4*>(bias + 32 * (nb0 + i) + 8 * a + 4 * g);
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            float v = acc[i][4 * a + r] + bv[r];
            if (relu) v = v > 0.f ? v : 0.f;
            acc[i][4 * a + r] = v;
          }
        }
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        out[sb * kSlab + (2 * (nb0 + i)) * 64 + lane] = acc_to_bf16(acc[i], 0);
        out[sb * kSlab + (2 * (nb0 + i) + 1) * 64 + lane] = acc_to_bf16(acc[i], 1);
      }
      if (stash && st_off >= 0) {
        const int64_t m = base + sb * 32 + j;
        if (m < M) {
          __bf16* row = stash + m * P.stash_per_sample + st_off + 32 * nb0;
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int a = 0; a < 4; ++a) {
              const floatx4_ v = {acc[i][4 * a], acc[i][4 * a + 1], acc[i][4 * a + 2], acc[i][4 * a + 3]};
              *reinterpret_cast<bf16x4*>(row + 32 * i + 8 * a + 4 * g) = __builtin_convertvector(v, bf16x4);
            }
        }
      }
    }
  };

  using I0 = std::integral_constant<int, 0>;
  using IT = std::integral_constant<int, kT>;
  using ITX = std::integral_constant<int, kTX>;
  using ITD = std::integral_constant<int, kTD>;
  auto wq = [&](int m) { return reinterpret_cast<const bf16x8*>(pk + P.p_fwd[m]); };
  auto bias = [&](int m) { return reinterpret_cast<const float*>(pk + P.p_bias[m]); };

  // layer 0 (encoding only) — its slice is small; load it directly
  {
    bf16x8 w0[kTX][2];
    ws::load_slice<kTX, kTX>(w0, wq(0), C::NB, nb0, lane);
    __syncthreads();  // encodings visible
    run_layer(&w0, I0{}, ITX{}, nullptr, s_ex, kTX, s_act[0], bias(0), 1, P.st_h[0]);
  }
  // hidden layers 1..D-1 and the bottleneck (layer D), weights prefetched one layer ahead
  bf16x8 wA[kT + kTX][2], wB[kT + kTX][2];
  if (SKIP == 0)
    ws::load_slice<kT + kTX, kT + kTX>(wA, wq(1), C::NB, nb0, lane);  // layer 1 with its enc rows
  else
    ws::load_slice<kT, kT + kTX>(wA, wq(1), C::NB, nb0, lane);
  int cur = 0;
#pragma unroll
  for (int l = 1; l <= D; ++l) {  // l == D: bottleneck (linear)
    auto& wcur = (l % 2) ? wA : wB;
    auto& wnext = (l % 2) ? wB : wA;
    if (l + 1 <= D) {
      if (l == SKIP)  // layer l+1 re-reads the encoding
        ws::load_slice<kT + kTX, kT + kTX>(wnext, wq(l + 1), C::NB, nb0, lane);
      else
        ws::load_slice<kT, kT + kTX>(wnext, wq(l + 1), C::NB, nb0, lane);
    }
    __syncthreads();  // layer l-1 outputs complete
    const int64_t st_off = l < D ? P.st_h[l] : P.st_bott;
    if (l - 1 == SKIP)
      run_layer(&wcur, IT{}, ITX{}, s_act[cur], s_ex, kTX, s_act[cur ^ 1], bias(l), 1, st_off);
    else
      run_layer(&wcur, IT{}, I0{}, s_act[cur], s_ex, kTX, s_act[cur ^ 1], bias(l), l < D, st_off);
    cur ^= 1;
    if (l == D - 1) {
      // sigma head on h_{D-1} (now in s_act[cur]): wave w handles sample block w
      __syncthreads();
      const float* hs = reinterpret_cast<const float*>(pk + P.p_head_sigma) + g * C::KH;
      float zs = 0.f;
#pragma unroll
      for (int t = 0; t < kT; ++t) {
        const bf16x8 hv = s_act[cur][wv * kSlab + t * 64 + lane];
        const int nb = t >> 1, h = t & 1;
#pragma unroll
        for (int r = 0; r < 8; ++r) zs += (float)hv[r] * hs[16 * nb + 8 * h + r];
      }
      zs = zs + shfl_xor(zs, 32);
      zs = zs + reinterpret_cast<const float*>(pk + P.p_head_sigma)[2 * C::KH];
      const int64_t m = base + wv * 32 + j;
      if (m < M && g == 0) sig_out[m] = relu(zs);
    }
  }
  // rgb branch dense_{D+1}: 4 output blocks, one per wave; input = bottleneck (s_act[cur]) + dir enc
  {
    const int mr = D + 1;
    const bf16x8* wr = reinterpret_cast<const bf16x8*>(pk + P.p_fwd[mr]);
    bf16x8 w1[kT + kTD];
#pragma unroll
    for (int t = 0; t < kT + kTD; ++t) w1[t] = wr[(t * C::NBR + wv) * 64 + lane];
    __syncthreads();
    const float* br = reinterpret_cast<const float*>(pk + P.p_bias[mr]);
    const floatx4* hr4 = reinterpret_cast<const floatx4*>(pk + P.p_head_rgb);
#pragma unroll 1
    for (int sb = 0; sb < kBlocks; ++sb) {
      floatx16 acc;
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[r] = 0.f;
#pragma unroll
      for (int t = 0; t < kT; ++t) acc = mfma_bf16(w1[t], s_act[cur][sb * kSlab + t * 64 + lane], acc);
#pragma unroll
      for (int t = 0; t < kTD; ++t) acc = mfma_bf16(w1[kT + t], s_ed[(sb * kTD + t) * 64 + lane], acc);
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const floatx4 bv = *reinterpret_cast<const floatx4*>(br + 32 * wv + 8 * a + 4 * g);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float v = acc[4 * a + r] + bv[r];
          acc[4 * a + r] = (float)(__bf16)(v > 0.f ? v : 0.f);
        }
      }
      if (stash) {
        const int64_t m = base + sb * 32 + j;
        if (m < M) {
          __bf16* row = stash + m * P.stash_per_sample + P.st_rbr + 32 * wv;
#pragma unroll
          for (int a = 0; a < 4; ++a) {
            const floatx4_ v = {acc[4 * a], acc[4 * a + 1], acc[4 * a + 2], acc[4 * a + 3]};
            *reinterpret_cast<bf16x4*>(row + 8 * a + 4 * g) = __builtin_convertvector(v, bf16x4);
          }
        }
      }
      // partial rgb dot over this wave's 32 features of the rgb branch
      float p0 = 0.f, p1 = 0.f, p2 = 0.f;
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const floatx4 w = hr4[g * 16 * C::NBR + 16 * wv + i];
        p0 += acc[i] * w[0];
        p1 += acc[i] * w[1];
        p2 += acc[i] * w[2];
      }
      p0 += shfl_xor(p0, 32);
      p1 += shfl_xor(p1, 32);
      p2 += shfl_xor(p2, 32);
      if (g == 0) {
        s_red[wv][0][sb * 32 + j] = p0;
        s_red[wv][1][sb * 32 + j] = p1;
        s_red[wv][2][sb * 32 + j] = p2;
      }
    }
    __syncthreads();
    // wave w finishes the colours of sample block w
    const float* bias3 = reinterpret_cast<const float*>(hr4 + 2 * 16 * C::NBR);
    const int64_t m = base + wv * 32 + j;
    if (g == 0 && m < M) {
      const int q = wv * 32 + j;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float z = ((s_red[0][c][q] + s_red[1][c][q]) + (s_red[2][c][q] + s_red[3][c][q])) + bias3[c];
        rgb_out[3 * m + c] = sigmoidf(z);
      }
    }
  }
}

template <int W, int D, int SKIP, int LX, int LD>
int launch_fwd_bf16ws(const MlpPlan& P, const void* packed, int64_t M, const float* xyz,
                      const float* dirs, const float* ro, const float* rd, const float* tv, int spr,
                      float* rgb, float* sigma, void* stash, hipStream_t s) {
  const int64_t nwg = (M + 127) / 128;
  mlp_fwd_bf16ws_kernel<W, D, SKIP, LX, LD><<<(unsigned)nwg, 256, 0, s>>>(
      P, reinterpret_cast<const char*>(packed), M, xyz, dirs, ro, rd, tv, spr, rgb, sigma,
      reinterpret_cast<__bf16*>(stash));
  return 0;
}
