// Weight-stationary BF16 forward with eight waves per workgroup (W = 256), with
// or without the training stash.
//
// The layout is that of mlp_fwd_bf16ws.h (128 samples per workgroup, the
// activations of a layer in a double-buffered LDS tile in the MFMA B layout,
// weights in VGPRs as two halves of a layer's k-steps, accumulators starting
// from the bias), but wave w owns one 32-feature output block (w) instead of
// two.  A wave then holds 80 weight registers and 64 accumulators, about 200
// registers in all, so two waves share each SIMD and one wave's LDS and L2
// waits hide behind the other's MFMAs.  Each B operand read from LDS feeds one
// MFMA instead of two: the tile's read traffic doubles, to at most half of what
// the LDS delivers to two waves per SIMD.  The first half of a layer's k-steps
// runs the four sample blocks round-robin (four independent accumulators in
// flight); the second half runs them as two pairs, so the first pair's
// epilogue overlaps the second pair's MFMAs.  The sigma head, the encodings,
// the rgb branch (four 32-feature blocks) and the colours are done by waves
// 0-3.  Measured against a four-wave kernel (two blocks per wave, one wave per
// SIMD) on the same box: 7.20 -> 6.54 ms per 32,768-ray render launch.
//
// Training stash: the encodings and every layer's output rows (bf16, natural
// feature order), and the ReLU bits of the trunk layers: per 32-feature block
// nb, lane half g and layer l one 16-bit word (element st_mask + (2 nb + g) D +
// l of the row); a wave keeps its block's words for four layers and writes them
// as one 8-byte store per sample block.  The stash is blocked by 32 samples
// (MlpPlan::sblk), so the two 16-byte row chunks a wave's lanes (j, g) write for
// one sample block are 1 KB of contiguous memory, and each workgroup writes a
// 64 KB stretch per layer.  Stores go through a buffer resource over the
// workgroup's rows; lanes past M get an offset outside it and are dropped.  (At this
// register budget the compiler issues each B operand read two or three MFMAs
// ahead rather than five; pinning the ring with scheduling-group barriers
// changed nothing measurable, the other wave on the SIMD covers the wait.)
#pragma once
#include "mlp_fwd_bf16ws.h"

namespace ws8 {

using ws::kBlocks;
using ws::kHalf;
using ws::kRing;
using ws::kSlab;
using ws::kT;
using ws::kTD;
using ws::kTX;
using ws::lds_b8;
constexpr int kWaves = 8;

// k-steps OFF .. OFF+N-1 of output block nb of the packed layer at p_layer into w[0 .. N).
template <int N, int OFF>
__device__ __forceinline__ void load_w1(bf16x8 (&w)[kHalf], __amdgpu_buffer_rsrc_t r, int64_t p_layer, int NBo,
                                        int nb, uint32_t lane_off) {
#pragma unroll
  for (int t = 0; t < N; ++t) w[t] = ws::ldw(r, lane_off, p_layer + (int64_t)((OFF + t) * NBo + nb) * 1024);
}

// acc[sb] (+)= sum over k-steps T0 .. T0+H-1 of w[k - T0] * B(sb, k) for the four
// sample blocks.  Without an epilogue step s runs block s % 4, k-step s / 4;
// with one, blocks {0, 1} run interleaved over all H k-steps, then {2, 3}, and
// epi(sb) follows each block's last MFMA.
template <int T0, int H, bool INIT, bool EPI, typename Epi>
__device__ __forceinline__ void run_steps(floatx16 (&acc)[kBlocks], const bf16x8 (&w)[kHalf], const floatx16& cb,
                                          const lds_b8* in, const lds_b8* ex, Epi&& epi) {
  constexpr int S = kBlocks * H;
  auto blk = [](int s) { return EPI ? 2 * (s / (2 * H)) + (s & 1) : s % kBlocks; };
  auto kst = [](int s) { return EPI ? (s >> 1) % H : s / kBlocks; };
  constexpr int R = kRing;
  bf16x8 b[R];
#pragma unroll
  for (int s = 0; s < R - 1; ++s) b[s] = ws::b_operand(in, ex, blk(s), T0 + kst(s));
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (s + R - 1 < S) b[(s + R - 1) % R] = ws::b_operand(in, ex, blk(s + R - 1), T0 + kst(s + R - 1));
    const int sb = blk(s), k = kst(s);
    acc[sb] = mfma_bf16(w[k], b[s % R], (INIT && k == 0) ? cb : acc[sb]);
    if (EPI && k == H - 1) epi(sb);
  }
}

}  // namespace ws8

template <int W, int D, int SKIP, int LX, int LD, bool STASH>
__global__ __launch_bounds__(512) void mlp_fwd_bf16ws8_kernel(
    MlpPlan P, const char* __restrict__ pk, int64_t M, const float* __restrict__ xyz,
    const float* __restrict__ dirs, const float* __restrict__ ro, const float* __restrict__ rd,
    const float* __restrict__ tv, int spr, float* __restrict__ rgb_out, float* __restrict__ sig_out,
    __bf16* __restrict__ stash) {
  using namespace ws8;
  using C = NetC<W, D, SKIP, LX, LD>;
  static_assert(W == 256 && LX == 10 && LD == 4 && SKIP >= 0 && SKIP + 1 < D,
                "weight-stationary kernel is for the 8x256 NeRF MLP");
  static_assert(D == 8, "ReLU bit words: (2 nb + g) D is one chunk per lane half, four layers per store");
  __shared__ bf16x8 s_act[2][kBlocks * kSlab];  // 2 x 64 KB
  __shared__ bf16x8 s_ex[kBlocks * kTX * 64];   // 16 KB
  __shared__ bf16x8 s_ed[kBlocks * kTD * 64];   // 8 KB
  __shared__ float s_red[4][3][kBlocks * 32];   // rgb-head partials of waves 0-3, 6 KB
  const int lane = threadIdx.x & 63, j = lane & 31, g = lane >> 5;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform, 0 .. 7
  const int64_t base = (int64_t)blockIdx.x * (kBlocks * 32);
  const int nb0 = wv;  // this wave's output block of a 256-wide layer
  const uint32_t lane_off = 16u * (uint32_t)lane, g_off = 16u * (uint32_t)g;
  const __amdgpu_buffer_rsrc_t rs = ws::packed_rsrc(pk, P.packed_bytes);
  // the stash rows of this workgroup as a buffer resource (rows past M fall outside)
  const int64_t wg_rows = M - base < kBlocks * 32 ? M - base : kBlocks * 32;
  const __amdgpu_buffer_rsrc_t rst =
      STASH ? __builtin_amdgcn_make_buffer_rsrc(stash + base * P.stash_per_sample, 0,
                                                (int)(wg_rows * P.stash_per_sample * 2), 0x00020000)
            : rs;
  // Stash offsets (blocked by 32 samples): element u + 8 ch + el of this lane's
  // sample in sample block sb of the workgroup, u wave-uniform with u % 8 + el +
  // the store's elements <= 8: a lane part (out of range for rows past M, which
  // drops the store) plus a uniform part, both in the VGPR offset; the scalar
  // offset of these 8- and 16-byte stores stays 0.
  const int sps = (int)P.stash_per_sample;
  int nrow[kBlocks];  // rows of each sample block (wave-uniform)
#pragma unroll
  for (int sb = 0; sb < kBlocks; ++sb) {
    const int64_t left = M - base - 32 * sb;
    nrow[sb] = __builtin_amdgcn_readfirstlane((int)(left < 32 ? (left > 0 ? left : 0) : 32));
  }
  auto rows_in = [&](int sb) { return nrow[sb]; };
  uint32_t lbase[kBlocks];  // this lane's row start in each sample block; out of range past M
#pragma unroll
  for (int sb = 0; sb < kBlocks; ++sb) lbase[sb] = j < nrow[sb] ? (uint32_t)(64 * sb * sps + 16 * j) : 0x80000000u;
  auto voff = [&](int sb, int ch, int el) -> uint32_t {
    return lbase[sb] + (uint32_t)(16 * rows_in(sb) * ch + 2 * el);
  };
  auto uoff = [&](int sb, int64_t u) { return (int)(16 * rows_in(sb) * (int)(u >> 3) + 2 * (int)(u & 7)); };
  auto st16 = [&](bf16x8 v, uint32_t vo, int so) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(ws::u32x4, v), rst, vo + (uint32_t)so, 0, kNT);
  };
  auto load_half = [&](bf16x8 (&w)[kHalf], int m, bool enc, bool second) {
    if (enc) {
      constexpr int H = (kT + kTX) / 2;
      if (second)
        load_w1<H, H>(w, rs, P.p_fwd[m], C::NB, nb0, lane_off);
      else
        load_w1<H, 0>(w, rs, P.p_fwd[m], C::NB, nb0, lane_off);
    } else {
      constexpr int H = kT / 2;
      if (second)
        load_w1<H, H>(w, rs, P.p_fwd[m], C::NB, nb0, lane_off);
      else
        load_w1<H, 0>(w, rs, P.p_fwd[m], C::NB, nb0, lane_off);
    }
  };

  bf16x8 wlo[kHalf], whi[kHalf];
  floatx16 cb;
  floatx16 acc[kBlocks];

  // layer 0 weights (4 encoding k-steps) into whi, layer 1's first half into wlo
  load_w1<kTX, 0>(whi, rs, P.p_fwd[0], C::NB, nb0, lane_off);
  load_half(wlo, 1, SKIP == 0, false);
  cb = ws::bias_acc(rs, P.p_bias[0], nb0, g_off);

  // ---- encodings: wave w < 4 encodes sample block w
  if (wv < kBlocks) {
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
    if (STASH) {
      // half g's slots start K / 8 chunks in (K a multiple of 8)
      static_assert(C::KX % 8 == 0 && C::KD % 8 == 0, "encoding slots in whole chunks");
      const uint32_t vx = voff(wv, g * (C::KX / 8), 0), vd = voff(wv, g * (C::KD / 8), 0);
#pragma unroll
      for (int e = 0; e < C::KX; e += 4)
        __builtin_amdgcn_raw_buffer_store_b64(ws::bf16x4_bits(ex[e], ex[e + 1], ex[e + 2], ex[e + 3]), rst,
                                              vx + (uint32_t)uoff(wv, P.st_encx + e), 0, kNT);
#pragma unroll
      for (int e = 0; e < C::KD; e += 4)
        __builtin_amdgcn_raw_buffer_store_b64(ws::bf16x4_bits(ed[e], ed[e + 1], ed[e + 2], ed[e + 3]), rst,
                                              vd + (uint32_t)uoff(wv, P.st_encd + e), 0, kNT);
    }
  }

  // ReLU bit words of this wave's block for four trunk layers per sample block:
  // layer 4 q + t in half t % 2 of mw[sb][t / 2]
  uint32_t mw[kBlocks][2];
  auto store_masks = [&](int q) {
#pragma unroll
    for (int sb = 0; sb < kBlocks; ++sb) {
      const ws::u32x2 v = {mw[sb][0], mw[sb][1]};
      // (2 nb0 + g) D = 16 nb0 + 8 g: lane half g one chunk on
      __builtin_amdgcn_raw_buffer_store_b64(
          v, rst, voff(sb, g, 0) + (uint32_t)uoff(sb, P.st_mask + 2 * nb0 * D + 4 * q), 0, kNT);
    }
  };

  // Epilogue of sample block sb: bf16, ReLU, LDS (B layout: k-steps 2 nb0, 2 nb0 + 1),
  // and with the stash the row pieces at st_off and (trunk layers, MASK) the ReLU bits.
  auto epilogue = [&](lds_b8* out, bool relu, int64_t st_off, auto mask_c, int layer, int sb) {
    constexpr bool MASK = decltype(mask_c)::value;
    bf16x8 lo = acc_to_bf16(acc[sb], 0), hi = acc_to_bf16(acc[sb], 1);
    if (relu) {
      lo = ws::relu_bf16(lo);
      hi = ws::relu_bf16(hi);
    }
    out[sb * kSlab] = lo;
    out[sb * kSlab + 64] = hi;
    if constexpr (STASH) {
      // as row chunks: features 32 nb0 + 16 h + 8 g .. + 7 (lo: h = 0, hi: h = 1)
      const uint32_t vo = voff(sb, g, 0);
      st16(ws::swap_halves(lo), vo, uoff(sb, st_off + 32 * nb0));
      st16(ws::swap_halves(hi), vo, uoff(sb, st_off + 32 * nb0 + 16));
      if constexpr (MASK) {
        const uint32_t b = ws::nz_bits(lo, hi);  // bits 0..7 and 16..23
        const uint32_t v = (b & 0xFFu) | ((b >> 8) & 0xFF00u);
        uint32_t& w = mw[sb][(layer >> 1) & 1];
        w = (layer & 1) ? ((w & 0xFFFFu) | (v << 16)) : v;
      }
    }
  };

  __syncthreads();  // encodings visible
  // ---- layer 0 (encoding only, weights in whi)
  lds_b8* const ex_l = ws::lds_lane(s_ex, 0, lane);
  {
    lds_b8* const out_l = ws::lds_lane(s_act[0], 2 * nb0 * 64, lane);
    run_steps<kT, kTX, true, true>(acc, whi, cb, nullptr, ex_l,
                                   [&](int sb) { epilogue(out_l, true, P.st_h[0], std::true_type{}, 0, sb); });
  }
  load_half(whi, 1, SKIP == 0, true);  // layer 1's second half and bias
  cb = ws::bias_acc(rs, P.p_bias[1], nb0, g_off);
  __syncthreads();

  // ---- layers 1 .. D (D: the linear bottleneck), unrolled
  ws::static_for<1, D + 1>([&](auto Lc) {
    constexpr int l = decltype(Lc)::value;
    constexpr int cur = (l - 1) & 1;
    const bf16x8* in = s_act[cur];
    lds_b8* const in_l = ws::lds_lane(s_act[cur], 0, lane);
    lds_b8* const out_l = ws::lds_lane(s_act[cur ^ 1], 2 * nb0 * 64, lane);
    if constexpr (l == D) {
      // sigma head on h_{D-1} (the input of this layer): wave w < 4 handles sample block w
      if (wv < kBlocks) {
        const uint32_t hs_off = 4u * (uint32_t)(g * C::KH);
        float zs = 0.f;
#pragma unroll
        for (int t = 0; t < kT; ++t) {
          const bf16x8 hv = in[wv * kSlab + t * 64 + lane];
          const int nb = t >> 1, h = t & 1;
          const floatx4 w0 = ws::ldf4(rs, hs_off, P.p_head_sigma + 4 * (16 * nb + 8 * h));
          const floatx4 w1 = ws::ldf4(rs, hs_off, P.p_head_sigma + 4 * (16 * nb + 8 * h + 4));
#pragma unroll
          for (int r = 0; r < 4; ++r) zs += (float)hv[r] * w0[r];
#pragma unroll
          for (int r = 0; r < 4; ++r) zs += (float)hv[4 + r] * w1[r];
        }
        zs = zs + shfl_xor(zs, 32);
        zs = zs + reinterpret_cast<const float*>(pk + P.p_head_sigma)[2 * C::KH];
        const int64_t m = base + wv * 32 + j;
        if (m < M && g == 0) sig_out[m] = relu(zs);
      }
    }
    constexpr bool enc = l == SKIP + 1;
    constexpr int H = enc ? (kT + kTX) / 2 : kT / 2;
    run_steps<0, H, true, false>(acc, wlo, cb, in_l, ex_l, [](int) {});
    // wlo is free: the next layer's first half and bias (after the bottleneck: the
    // rgb branch's first 10 k-steps of output block w, waves 0-3)
    if constexpr (l < D) {
      load_half(wlo, l + 1, l + 1 == SKIP + 1, false);
      cb = ws::bias_acc(rs, P.p_bias[l + 1], nb0, g_off);
    } else {
      if (wv < 4) load_w1<kHalf, 0>(wlo, rs, P.p_fwd[D + 1], C::NBR, wv, lane_off);
    }
    const int64_t st_off = l < D ? P.st_h[l] : P.st_bott;
    run_steps<H, H, false, true>(acc, whi, cb, in_l, ex_l, [&](int sb) {
      epilogue(out_l, l < D, st_off, std::integral_constant<bool, (l < D)>{}, l, sb);
    });
    if constexpr (STASH && l < D && (l & 3) == 3) store_masks(l >> 2);
    if constexpr (l < D) {
      load_half(whi, l + 1, l + 1 == SKIP + 1, true);  // whi is free
    } else {
      if (wv < 4) {
        load_w1<kT + kTD - kHalf, kHalf>(whi, rs, P.p_fwd[D + 1], C::NBR, wv, lane_off);
        cb = ws::bias_acc(rs, P.p_bias[D + 1], wv, g_off);
      }
    }
    __syncthreads();  // layer l outputs complete
  });
  constexpr int fin = D & 1;  // buffer holding the bottleneck

  // ---- rgb branch dense_{D+1}: waves 0-3, output block w; input = bottleneck + dir enc
  const floatx4* hr4 = reinterpret_cast<const floatx4*>(pk + P.p_head_rgb);
  if (wv < 4) {
    lds_b8* const in_l = ws::lds_lane(s_act[fin], 0, lane);
    lds_b8* const ed_l = ws::lds_lane(s_ed, 0, lane);
    auto wk = [&](int t) { return t < kHalf ? wlo[t] : whi[t - kHalf]; };
#pragma unroll
    for (int sb = 0; sb < kBlocks; ++sb) {
      floatx16 a = cb;
#pragma unroll
      for (int t = 0; t < kT; ++t) a = mfma_bf16(wk(t), in_l[sb * kSlab + t * 64], a);
#pragma unroll
      for (int t = 0; t < kTD; ++t) a = mfma_bf16(wk(kT + t), ed_l[(sb * kTD + t) * 64], a);
      // ReLU, rounded to bf16 (the stashed activation), then the fp32 head partials
#pragma unroll
      for (int r = 0; r < 16; ++r) a[r] = (float)(__bf16)(a[r] > 0.f ? a[r] : 0.f);
      if (STASH) {
        const bf16x8 lo = acc_to_bf16(a, 0), hi = acc_to_bf16(a, 1);
        const uint32_t vo = voff(sb, g, 0);
        st16(ws::swap_halves(lo), vo, uoff(sb, P.st_rbr + 32 * wv));
        st16(ws::swap_halves(hi), vo, uoff(sb, P.st_rbr + 32 * wv + 16));
        // ReLU bits of rgb-branch block wv: bits 0..7 and 16..23 of nz_bits as 16 bits
        const uint32_t b = ws::nz_bits(lo, hi);
        __builtin_amdgcn_raw_buffer_store_b16((unsigned short)((b & 0xFFu) | ((b >> 8) & 0xFF00u)), rst,
                                              voff(sb, 0, g) + (uint32_t)uoff(sb, P.st_mask + 16 * D + 2 * wv), 0, kNT);
      }
      float p0 = 0.f, p1 = 0.f, p2 = 0.f;
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const floatx4 w = hr4[g * 16 * C::NBR + 16 * wv + i];
        p0 += a[i] * w[0];
        p1 += a[i] * w[1];
        p2 += a[i] * w[2];
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
  }
  __syncthreads();
  // wave w < 4 finishes the colours of sample block w
  if (wv < kBlocks) {
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
void launch_fwd_bf16ws8(const MlpPlan& P, const void* packed, int64_t M, const float* xyz, const float* dirs,
                        const float* ro, const float* rd, const float* tv, int spr, float* rgb, float* sigma,
                        void* stash, hipStream_t s) {
  const int64_t nwg = (M + 127) / 128;
  if (stash)
    mlp_fwd_bf16ws8_kernel<W, D, SKIP, LX, LD, true><<<(unsigned)nwg, 512, 0, s>>>(
        P, reinterpret_cast<const char*>(packed), M, xyz, dirs, ro, rd, tv, spr, rgb, sigma,
        reinterpret_cast<__bf16*>(stash));
  else
    mlp_fwd_bf16ws8_kernel<W, D, SKIP, LX, LD, false><<<(unsigned)nwg, 512, 0, s>>>(
        P, reinterpret_cast<const char*>(packed), M, xyz, dirs, ro, rd, tv, spr, rgb, sigma, nullptr);
}
