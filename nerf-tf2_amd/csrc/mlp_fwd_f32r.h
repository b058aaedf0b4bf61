// Register-resident FP32 forward of the reference topology (W = 256 with the
// rgb branch), with or without the training stash.
//
// A wave owns 16 samples and carries them through every layer on
// v_mfma_f32_16x16x4f32 (lane l = j + 16 q: sample j, lane quarter q).  The
// accumulators of a layer's 16 output blocks hold, on lane (j, q), features
// 16 nb + 4 q + r of sample j; with the weights packed in that k order, they
// are the next layer's B operands as they stand, so the layer input stays in
// 64 VGPRs and the output in 64 accumulators, and no activation goes through
// LDS.  That keeps a wave under 256 registers, two waves share each SIMD, and
// one wave's epilogue, encoding and memory waits overlap the other's MFMAs.
// Weights stream from L2 as 16-byte buffer loads (4 k-steps of one output
// block per load), one half-group (8 output blocks) ahead of their MFMAs.
// The encodings are written once to the wave's LDS slots (each lane reads
// back only what it wrote) and re-read as B operands by the first layer, the
// skip layer and the rgb branch.  The sigma head, the rgb head and the
// per-sample reductions over the four lane quarters are VALU dot products and
// two cross-lane shuffles.  The training stash gets the same rows as from the
// 32-sample kernel (hidden rows in natural feature order, the encodings in the
// lane-half slot layout), written by buffer stores over the wave's 16 rows, and
// each ReLU layer's branch bits (64 per lane: act > 0), which the dX chain
// reads instead of the activation rows.
#pragma once
#include "mlp_kernels.h"
#include "mlp_fwd_bf16ws.h"  // ws::packed_rsrc, ws::ldf4, ws::static_for

namespace f32r {

constexpr int kWaves = 4;  // waves per workgroup (64 samples)

__device__ __forceinline__ floatx4 mfma16(float a, float b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Encoding of 3 components with L frequencies into the S slots of lane
// quarter q (enc_feat4 layout).  Argument = fl32(x_c * (2^l * fl32(pi))), as in
// encode_slots.
template <int L, int S>
__device__ __forceinline__ void encode4(const float (&x)[3], int q, float (&out)[S]) {
  constexpr int P0 = S / 2 - 2;
  constexpr float kPi = 3.14159265358979323846f;
  float x0 = x[0], x1 = x[1], x2 = x[2];
  asm volatile("" : "+v"(x0), "+v"(x1), "+v"(x2));
#pragma unroll
  for (int k = 0; k < S / 2; ++k) {
    const int p = q == 0 ? k - 2 : P0 + (q - 1) * (S / 2) + k;
    const bool is_pair = (q != 0 || k >= 2) && p < 3 * L;
    float s = 0.f, c = 0.f;
    if (is_pair) {
      const int comp = p / L, l = p - comp * L;
      const float xc = comp == 0 ? x0 : (comp == 1 ? x1 : x2);
      const float arg = fmul(xc, ldexpf(kPi, l));
      sincosf(arg, &s, &c);
    }
    float v0 = s, v1 = c;
    if (!is_pair) {
      v0 = (q == 0 && k == 0) ? x0 : ((q == 0 && k == 1) ? x2 : 0.f);
      v1 = (q == 0 && k == 0) ? x1 : 0.f;
    }
    out[2 * k] = v0;
    out[2 * k + 1] = v1;
  }
}

typedef __attribute__((address_space(3))) floatx4 lds_f4;

// acc[0 .. NBO) = sum over the layer's k-step groups of A * B.  Groups 0 .. NE-1
// read their B operand from the encoding slots (enc[grp * 64]), groups NE ..
// NE+NH-1 from act[grp - NE].  Step u runs group u / 2 on output blocks
// (u % 2) * NBO/2 .. + NBO/2 - 1; the weights of step u + 1 are loaded at the
// start of step u.  The loads are unconditional, so every wait counts only the
// loads of its own step.
template <int NBO, int NE, int NH>
__device__ __forceinline__ void layer(floatx4 (&acc)[16], const floatx4 (&act)[16], const lds_f4* enc,
                                      __amdgpu_buffer_rsrc_t rs, uint32_t lane_off, int64_t p_layer) {
  constexpr int NG = NE + NH, HB = NBO / 2;
  floatx4 w[2][HB];
  auto load_step = [&](int u, floatx4 (&dst)[HB]) {
    const int grp = u >> 1, h = u & 1;
#pragma unroll
    for (int i = 0; i < HB; ++i) dst[i] = ws::ldf4(rs, lane_off, p_layer + (int64_t)(grp * NBO + h * HB + i) * 1024);
  };
  load_step(0, w[0]);
  const floatx4 zero = {0.f, 0.f, 0.f, 0.f};
  ws::static_for<0, 2 * NG>([&](auto U) {
    constexpr int u = decltype(U)::value;
    constexpr int grp = u >> 1, h = u & 1;
    if constexpr (u + 1 < 2 * NG) load_step(u + 1, w[(u + 1) & 1]);
    floatx4 bv;
    if constexpr (grp < NE)
      bv = enc[grp * 64];
    else
      bv = act[grp - NE];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int i = 0; i < HB; ++i)
        acc[h * HB + i] = mfma16(w[u & 1][i][r], bv[r], (grp == 0 && r == 0) ? zero : acc[h * HB + i]);
    __builtin_amdgcn_sched_barrier(0);
  });
}

// act[nb] = act_fn(acc[nb] + bias) for output blocks 0 .. NBO-1 (bias in natural
// feature order: lane quarter q reads features 16 nb + 4 q .. + 3).
template <int NBO, bool RELU>
__device__ __forceinline__ void epilogue(floatx4 (&act)[16], const floatx4 (&acc)[16], __amdgpu_buffer_rsrc_t rs,
                                         uint32_t q_off, int64_t p_bias) {
#pragma unroll
  for (int nb = 0; nb < NBO; ++nb) {
    const floatx4 b = ws::ldf4(rs, q_off, p_bias + 64 * nb);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float v = acc[nb][r] + b[r];
      act[nb][r] = RELU ? (v > 0.f ? v : 0.f) : v;
    }
  }
}

// Rows r of this wave's stash: 16-byte stores of act[nb] at features 16 nb + 4 q.
template <int NBO>
__device__ __forceinline__ void store_rows(const floatx4 (&act)[16], __amdgpu_buffer_rsrc_t rst, uint32_t row_q,
                                           int64_t st_off) {
#pragma unroll
  for (int nb = 0; nb < NBO; ++nb)
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4_cvt, act[nb]), rst,
                                           row_q + (uint32_t)((st_off + 16 * nb) * 4), 0, 0);
}

// ReLU branch bits of act[0 .. NBO): bit 4 (nb % 8) + r of word nb / 8 is act[nb][r] > 0.
template <int NBO>
__device__ __forceinline__ void relu_bits(const floatx4 (&act)[16], uint32_t (&w)[2]) {
  w[0] = w[1] = 0u;
#pragma unroll
  for (int nb = 0; nb < NBO; ++nb)
#pragma unroll
    for (int r = 0; r < 4; ++r) w[nb >> 3] |= (act[nb][r] > 0.f ? 1u : 0u) << (4 * (nb & 7) + r);
}

// The encoding of this lane's quarter slots into the lane-half slot layout of the
// stash (K slots per half, all 2 K written: pads as zeros by quarter 0).
template <int L, int S, int K>
__device__ __forceinline__ void store_enc(const float (&v)[S], int q, __amdgpu_buffer_rsrc_t rst, uint32_t row,
                                          int64_t st_off) {
#pragma unroll
  for (int e = 0; e < S; ++e) {
    const int f = enc_feat4(q, e, L);
    if (f >= 0)
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v[e]), rst,
                                            row + (uint32_t)((st_off + enc_slot_of(f, L, K)) * 4), 0, 0);
  }
  if (q == 0) {
    constexpr int H = 2 + 3 * L;
#pragma unroll
    for (int t = 0; t < 2 * K; ++t) {
      const int g = t / K, e = t % K;
      if ((g == 0 && e == 3) || e >= H)
        __builtin_amdgcn_raw_buffer_store_b32(0u, rst, row + (uint32_t)((st_off + t) * 4), 0, 0);
    }
  }
}

}  // namespace f32r

template <int W, int D, int SKIP, int LX, int LD, bool STASH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void mlp_fwd_f32r_kernel(
    MlpPlan P, const char* __restrict__ pk, int64_t M, const float* __restrict__ xyz,
    const float* __restrict__ dirs, const float* __restrict__ ro, const float* __restrict__ rd,
    const float* __restrict__ tv, int spr, float* __restrict__ rgb_out, float* __restrict__ sig_out,
    float* __restrict__ stash) {
  using namespace f32r;
  static_assert(W == 256 && LD > 0 && SKIP >= 0 && SKIP + 1 < D, "register-resident kernel is for the 8x256 NeRF MLP");
  constexpr int SX = enc_slots4(LX), SD = enc_slots4(LD);
  constexpr int GX = SX / 4, GD = SD / 4;  // k-step groups of the encodings
  constexpr int NH = W / 16;               // k-step groups of a hidden input
  __shared__ floatx4 s_enc[kWaves][GX + GD][64];
  const int lane = threadIdx.x & 63, j = lane & 15, q = lane >> 4;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t m0 = ((int64_t)blockIdx.x * kWaves + wv) * 16;
  if (m0 >= M) return;  // no workgroup barrier in this kernel
  const int64_t m = m0 + j;
  const bool valid = m < M;
  const int64_t mm = valid ? m : M - 1;
  const __amdgpu_buffer_rsrc_t rs = ws::packed_rsrc(pk, P.packed_bytes);
  const uint32_t lane_off = 16u * (uint32_t)lane, q_off = 16u * (uint32_t)q;
  // the stash rows of this wave as a buffer resource (rows past M fall outside it)
  const int64_t w_rows = M - m0 < 16 ? M - m0 : 16;
  const __amdgpu_buffer_rsrc_t rst =
      STASH ? __builtin_amdgcn_make_buffer_rsrc(stash + m0 * P.stash_per_sample, 0,
                                                (int)(w_rows * P.stash_per_sample * 4), 0x00020000)
            : rs;
  const uint32_t row = (uint32_t)(j * P.stash_per_sample * 4), row_q = row + q_off;

  // ---- encodings into this lane's LDS slots (B layout)
  {
    float x[3], d[3];
    load_sample(mm, xyz, dirs, ro, rd, tv, spr, x, d);
    float ex[SX], ed[SD];
    encode4<LX, SX>(x, q, ex);
    encode4<LD, SD>(d, q, ed);
#pragma unroll
    for (int g = 0; g < GX; ++g) s_enc[wv][g][lane] = floatx4{ex[4 * g], ex[4 * g + 1], ex[4 * g + 2], ex[4 * g + 3]};
#pragma unroll
    for (int g = 0; g < GD; ++g)
      s_enc[wv][GX + g][lane] = floatx4{ed[4 * g], ed[4 * g + 1], ed[4 * g + 2], ed[4 * g + 3]};
    if constexpr (STASH) {
      store_enc<LX, SX, NetC<W, D, SKIP, LX, LD>::KX>(ex, q, rst, row, P.st_encx);
      store_enc<LD, SD, NetC<W, D, SKIP, LX, LD>::KD>(ed, q, rst, row, P.st_encd);
    }
  }
  const lds_f4* encx = (const lds_f4*)&s_enc[wv][0][lane];
  const lds_f4* encd = encx + GX * 64;

  floatx4 act[16], acc[16];
  // dense_0 on the xyz encoding
  layer<16, GX, 0>(acc, act, encx, rs, lane_off, P.p_fwd4[0]);
  // ReLU bits of trunk layer l: two words at element st_mask + 8 l + 2 q of the row
  auto store_bits = [&](int l) {
    uint32_t w[2];
    relu_bits<16>(act, w);
    const ws::u32x2 v = {w[0], w[1]};
    __builtin_amdgcn_raw_buffer_store_b64(v, rst, row + (uint32_t)((P.st_mask + 8 * l + 2 * q) * 4), 0, 0);
  };
  epilogue<16, true>(act, acc, rs, q_off, P.p_bias[0]);
  if constexpr (STASH) {
    store_rows<16>(act, rst, row_q, P.st_h[0]);
    store_bits(0);
  }
  // dense_1 .. dense_{D-1}; the encoding joins the input after dense_SKIP
  ws::static_for<1, D>([&](auto Lc) {
    constexpr int l = decltype(Lc)::value;
    if constexpr (l - 1 == SKIP)
      layer<16, GX, NH>(acc, act, encx, rs, lane_off, P.p_fwd4[l]);
    else
      layer<16, 0, NH>(acc, act, encx, rs, lane_off, P.p_fwd4[l]);
    epilogue<16, true>(act, acc, rs, q_off, P.p_bias[l]);
    if constexpr (STASH) {
      store_rows<16>(act, rst, row_q, P.st_h[l]);
      store_bits(l);
    }
  });
  // sigma head on h_{D-1} (model.py:374-375)
  float zs = 0.f;
#pragma unroll
  for (int nb = 0; nb < 16; ++nb) {
    const floatx4 w = ws::ldf4(rs, q_off, P.p_hs_nat + 64 * nb);
#pragma unroll
    for (int r = 0; r < 4; ++r) zs = zs + act[nb][r] * w[r];
  }
  zs = zs + shfl_xor(zs, 16);
  zs = zs + shfl_xor(zs, 32);
  zs = zs + reinterpret_cast<const float*>(pk + P.p_hs_nat)[W];
  const float sigma = relu(zs);
  // bottleneck dense_D, linear (model.py:377-378)
  layer<16, 0, NH>(acc, act, encx, rs, lane_off, P.p_fwd4[D]);
  epilogue<16, false>(act, acc, rs, q_off, P.p_bias[D]);
  if constexpr (STASH) store_rows<16>(act, rst, row_q, P.st_bott);
  // dense_{D+1} on concat(bottleneck, direction encoding) (model.py:380-384): 8 blocks of 16
  layer<8, GD, NH>(acc, act, encd, rs, lane_off, P.p_fwd4[D + 1]);
  epilogue<8, true>(act, acc, rs, q_off, P.p_bias[D + 1]);
  if constexpr (STASH) {
    store_rows<8>(act, rst, row_q, P.st_rbr);
    uint32_t w[2];
    relu_bits<8>(act, w);
    __builtin_amdgcn_raw_buffer_store_b32(w[0], rst, row + (uint32_t)((P.st_mask + 8 * D + q) * 4), 0, 0);
  }
  // rgb head, sigmoid (model.py:386-387): weights [feature][4]
  float p0 = 0.f, p1 = 0.f, p2 = 0.f;
#pragma unroll
  for (int nb = 0; nb < 8; ++nb)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      // feature 16 nb + 4 q + r: byte 16 (16 nb + r) + 64 q
      const floatx4 w = ws::ldf4(rs, 4u * q_off, P.p_hr_nat + 16 * (16 * nb + r));
      p0 = p0 + act[nb][r] * w[0];
      p1 = p1 + act[nb][r] * w[1];
      p2 = p2 + act[nb][r] * w[2];
    }
  p0 = p0 + shfl_xor(p0, 16);
  p1 = p1 + shfl_xor(p1, 16);
  p2 = p2 + shfl_xor(p2, 16);
  p0 = p0 + shfl_xor(p0, 32);
  p1 = p1 + shfl_xor(p1, 32);
  p2 = p2 + shfl_xor(p2, 32);
  if (valid && q == 0) {
    const float* hb = reinterpret_cast<const float*>(pk + P.p_hr_nat) + 4 * (W / 2);
    sig_out[m] = sigma;
    rgb_out[3 * m + 0] = sigmoidf(p0 + hb[0]);
    rgb_out[3 * m + 1] = sigmoidf(p1 + hb[1]);
    rgb_out[3 * m + 2] = sigmoidf(p2 + hb[2]);
  }
}

template <int W, int D, int SKIP, int LX, int LD>
void launch_fwd_f32r(const MlpPlan& P, const void* packed, int64_t M, const float* xyz, const float* dirs,
                     const float* ro, const float* rd, const float* tv, int spr, float* rgb, float* sigma,
                     void* stash, hipStream_t s) {
  const int64_t nwg = (M + 16 * f32r::kWaves - 1) / (16 * f32r::kWaves);
  if (stash)
    mlp_fwd_f32r_kernel<W, D, SKIP, LX, LD, true><<<(unsigned)nwg, 64 * f32r::kWaves, 0, s>>>(
        P, reinterpret_cast<const char*>(packed), M, xyz, dirs, ro, rd, tv, spr, rgb, sigma,
        reinterpret_cast<float*>(stash));
  else
    mlp_fwd_f32r_kernel<W, D, SKIP, LX, LD, false><<<(unsigned)nwg, 64 * f32r::kWaves, 0, s>>>(
        P, reinterpret_cast<const char*>(packed), M, xyz, dirs, ro, rd, tv, spr, rgb, sigma, nullptr);
}
