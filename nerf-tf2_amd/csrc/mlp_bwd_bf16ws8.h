// Weight-stationary BF16 dX chain with eight waves per workgroup (W = 256): the
// backward twin of mlp_fwd_bf16ws8_kernel, with the chain of mlp_bwd_bf16ws.h.
//
// A workgroup carries 128 samples (four blocks of 32); wave w owns one 32-feature
// output block (w) of every step and holds that slice of the transposed weights in
// VGPRs as two halves of the step's k-steps (the next step's first half is fetched
// while the current second half runs, and vice versa), with the four sample
// blocks' accumulators live: about 200 registers, so two waves share each SIMD.
// B operands are the previous step's bf16 dz in a double-buffered LDS tile (B
// layout), read through the register ring of ws8::run_steps.  Per sample the
// chain is (model.py:355-392 differentiated):
//   rgb head:    dz_rgb = dy * y (1 - y);  d h9 = W_rgb dz_rgb (fp32), masked by h9 > 0
//   rgb branch:  d bott = W_{D+1}[:bott] dz_rbr            (bottleneck: linear)
//   bottleneck:  d h_{D-1} = W_D dz_bott + W_sigma dz_sigma, masked by h_{D-1} > 0
//   trunk:       d h_{l-1} = W_l[:hidden] dz_l, masked by h_{l-1} > 0, l = D-1 .. 1
// The heads run on waves 0-3 (sample block w, as in the four-wave chain).  Each dz
// is written as bf16 row chunks into the blocked dz buffer (MlpPlan::sblk = 32; a
// store covers 1 KB of contiguous memory) and, rounded the same way, is the next
// step's B operand.  ReLU masks are the forward's 16-bit words of this wave's
// block, loaded per sample block four layers at a time.  Stash reads and dz
// writes go through buffer resources over the workgroup's rows; lanes past the
// last sample get a VGPR offset outside them (zeros read, stores dropped).
#pragma once
#include "mlp_bwd_bf16ws.h"
#include "mlp_fwd_bf16ws8.h"

template <int W, int D, int SKIP, int LX, int LD>
__global__ __launch_bounds__(512) void mlp_bwd_bf16ws8_kernel(
    MlpPlan P, const char* __restrict__ pk, int64_t M, const float* __restrict__ d_rgb,
    const float* __restrict__ d_sigma, const float* __restrict__ rgb, const float* __restrict__ sigma,
    const __bf16* __restrict__ stash, __bf16* __restrict__ dz) {
  using C = NetC<W, D, SKIP, LX, LD>;
  using ws::kBlocks;
  using ws::kHalf;
  using ws::kSlab;
  using ws::kT;
  using ws::lds_b8;
  static_assert(W == 256 && LX == 10 && LD == 4 && SKIP >= 0 && SKIP + 1 < D,
                "weight-stationary backward is for the 8x256 NeRF MLP");
  static_assert(C::NBR == 4, "rgb branch of 128 features");
  static_assert(D == 8, "ReLU bit words: (2 nb + g) D is one chunk per lane half, four layers per load");
  __shared__ bf16x8 s_act[2][kBlocks * kSlab];  // 2 x 64 KB
  __shared__ float s_dzs[kBlocks * 32];          // dz_sigma of the workgroup's samples
  const int lane = threadIdx.x & 63, j = lane & 31, g = lane >> 5;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // 0 .. 7
  const int nb0 = wv;  // this wave's output block of a 256-wide step
  const int64_t base = (int64_t)blockIdx.x * (kBlocks * 32);
  const uint32_t lane_off = 16u * (uint32_t)lane;
  const __amdgpu_buffer_rsrc_t rs = ws::packed_rsrc(pk, P.packed_bytes);
  const int64_t nrows = M - base < kBlocks * 32 ? M - base : kBlocks * 32;
  const int64_t sps = P.stash_per_sample, zps = P.dz_per_sample;
  const __amdgpu_buffer_rsrc_t rst = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<__bf16*>(stash + base * sps), 0, (int)(nrows * sps * 2), 0x00020000);
  const __amdgpu_buffer_rsrc_t rdz =
      __builtin_amdgcn_make_buffer_rsrc(dz + base * zps, 0, (int)(nrows * zps * 2), 0x00020000);
  // Offsets in the blocked rows: element u + 8 ch + el of this lane's sample in
  // sample block sb, u wave-uniform (u % 8 + el + the access's elements <= 8): a
  // lane part (out of range past the last sample) and a uniform part.
  int nrow[kBlocks];
  uint32_t sbase[kBlocks], zbase[kBlocks];
#pragma unroll
  for (int sb = 0; sb < kBlocks; ++sb) {
    const int64_t left = M - base - 32 * sb;
    nrow[sb] = __builtin_amdgcn_readfirstlane((int)(left < 32 ? (left > 0 ? left : 0) : 32));
    sbase[sb] = j < nrow[sb] ? (uint32_t)(64 * sb * sps + 16 * j) : 0x80000000u;
    zbase[sb] = j < nrow[sb] ? (uint32_t)(64 * sb * zps + 16 * j) : 0x80000000u;
  }
  auto sv = [&](int sb, int ch, int el) { return sbase[sb] + (uint32_t)(16 * nrow[sb] * ch + 2 * el); };
  auto zv = [&](int sb, int ch) { return zbase[sb] + (uint32_t)(16 * nrow[sb] * ch); };
  auto uo = [&](int sb, int64_t u) { return (int64_t)(16 * nrow[sb] * (int)(u >> 3) + 2 * (int)(u & 7)); };

  bf16x8 wlo[kHalf], whi[kHalf];
  floatx16 acc[kBlocks];
  floatx16 zero;
#pragma unroll
  for (int e = 0; e < 16; ++e) zero[e] = 0.f;

  // rgb-branch step (8 k-steps over the 128 rgb-branch features): halves of 4
  constexpr int KR = 2 * C::NBR;
  ws8::load_w1<KR / 2, 0>(wlo, rs, P.p_bwd[D + 1], C::NB, nb0, lane_off);
  ws8::load_w1<KR / 2, KR / 2>(whi, rs, P.p_bwd[D + 1], C::NB, nb0, lane_off);

  // ---- heads (waves 0-3: sample block w): dz_rgb, dz_sigma, d h9 -> dz_rbr
  if (wv < kBlocks) {
    const int64_t m = base + wv * 32 + j;
    const bool valid = m < M;
    const int64_t mm = valid ? m : M - 1;
    const float sg = sigma[mm];
    const float ds = valid ? d_sigma[mm] : 0.f;
    const float dzs = sg > 0.f ? ds : 0.f;
    float dzr[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float y = rgb[3 * mm + c];
      const float dy = valid ? d_rgb[3 * mm + c] : 0.f;
      dzr[c] = dy * y * (1.f - y);
    }
    if (g == 0) s_dzs[wv * 32 + j] = dzs;
    if (g == 0) {
      const bf16x4 hz = {(__bf16)dzs, (__bf16)0.f, (__bf16)0.f, (__bf16)0.f};
      const bf16x4 rz = {(__bf16)dzr[0], (__bf16)dzr[1], (__bf16)dzr[2], (__bf16)0.f};
      wsb::st4(hz, rdz, zv(wv, 0), uo(wv, P.dz_sigma));
      wsb::st4(rz, rdz, zv(wv, 0), uo(wv, P.dz_rgb));
    }
    const floatx4* hr4 = reinterpret_cast<const floatx4*>(pk + P.p_head_rgb) + g * 16 * C::NBR;
    lds_b8* const out = ws::lds_lane(s_act[0], wv * kSlab, lane);
#pragma unroll
    for (int nb = 0; nb < C::NBR; ++nb) {
      floatx16 a;
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const floatx4 w = hr4[16 * nb + i];
        a[i] = dzr[0] * w[0] + dzr[1] * w[1] + dzr[2] * w[2];
      }
      // the forward's 16 ReLU bits of rgb-branch block nb (lane halves in one dword)
      const uint32_t hw = wsb::ld_u32(rst, sv(wv, 0, 0), uo(wv, P.st_mask + 16 * D + 2 * nb));
      const uint32_t u = g ? hw >> 16 : hw & 0xFFFFu;
      const uint32_t bits = (u & 0xFFu) | ((u & 0xFF00u) << 8);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const bf16x8 v = wsb::mask_bits(acc_to_bf16(a, h), bits, 4 * h);
        out[(2 * nb + h) * 64] = v;
        wsb::st_group(v, rdz, zv(wv, g), uo(wv, P.dz_rbr + 32 * nb + 16 * h));
      }
    }
  }
  __syncthreads();

  // Epilogue of sample block sb: ReLU mask (MASKED) from the forward's bit word
  // `bits` of this wave's block, bf16 into LDS (B layout k-steps 2 nb0, 2 nb0 + 1;
  // TO_LDS) and the dz row chunks at dz_off.
  auto epilogue = [&](auto masked_c, auto lds_c, lds_b8* out, uint32_t bits, int64_t dz_off, int sb) {
    constexpr bool MASKED = decltype(masked_c)::value, TO_LDS = decltype(lds_c)::value;
    const uint32_t zr = zv(sb, g);  // lane half g: the second chunk of each 16-feature group
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      bf16x8 v = acc_to_bf16(acc[sb], h);
      if constexpr (MASKED) v = wsb::mask_bits(v, bits, 4 * h);
      if constexpr (TO_LDS) out[sb * kSlab + h * 64] = v;
      wsb::st_group(v, rdz, zr, uo(sb, dz_off + 32 * nb0 + 16 * h));
    }
  };

  // ---- rgb branch: d bott = W_{D+1} dz_rbr (linear); input buffer 0, output buffer 1
  {
    lds_b8* const in_l = ws::lds_lane(s_act[0], 0, lane);
    lds_b8* const out_l = ws::lds_lane(s_act[1], 2 * nb0 * 64, lane);
    ws8::run_steps<0, KR / 2, true, false>(acc, wlo, zero, in_l, in_l, [](int) {});
    ws8::load_w1<kT / 2, 0>(wlo, rs, P.p_bwd[D], C::NB, nb0, lane_off);
    ws8::run_steps<KR / 2, KR / 2, false, true>(acc, whi, zero, in_l, in_l, [&](int sb) {
      epilogue(std::false_type{}, std::true_type{}, out_l, 0u, P.dz_bott, sb);
    });
    ws8::load_w1<kT / 2, kT / 2>(whi, rs, P.p_bwd[D], C::NB, nb0, lane_off);
    __syncthreads();
  }

  // ---- bottleneck and trunk, unrolled: step s = 0 applies W_D (input dz_bott), then
  // W_{D-1} .. W_1.  The ReLU bit words of trunk layers 4q .. 4q+3 of this wave's block
  // are 8 contiguous bytes per lane, loaded per sample block every four steps.
  ws::u32x2 mq[kBlocks];
  ws::static_for<0, D>([&](auto Sc) {
    constexpr int s = decltype(Sc)::value;
    constexpr int layer = D - s;      // packed layer whose transpose is applied
    constexpr int cur = (s + 1) & 1;  // buffer holding this step's input dz
    lds_b8* const in_l = ws::lds_lane(s_act[cur], 0, lane);
    lds_b8* const out_l = ws::lds_lane(s_act[cur ^ 1], 2 * nb0 * 64, lane);  // unused for layer 1
    constexpr int ml = layer - 1;  // trunk layer whose ReLU masks this step's output
    if constexpr ((ml & 3) == 3) {
#pragma unroll
      for (int sb = 0; sb < kBlocks; ++sb)
        mq[sb] = __builtin_amdgcn_raw_buffer_load_b64(rst, sv(sb, g, 0),
                                                      (int)uo(sb, P.st_mask + 2 * nb0 * D + (ml & ~3)), 0);
    }
    uint32_t mbits[kBlocks];
#pragma unroll
    for (int sb = 0; sb < kBlocks; ++sb) {
      const ws::u32x2 q = mq[sb];  // copied out first: an element subscript of a captured vector array
      const uint32_t h = (q[(ml >> 1) & 1] >> (16 * (ml & 1))) & 0xFFFFu;
      mbits[sb] = (h & 0xFFu) | ((h & 0xFF00u) << 8);
    }
    constexpr int H = kT / 2;
    if constexpr (s == 0) {
      // accumulators start from the sigma-head term W_sigma dz_sigma
      const uint32_t hs_off = 4u * (uint32_t)(g * C::KH);
      floatx16 hsv;
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const floatx4 w = ws::ldf4(rs, hs_off, P.p_head_sigma + 4 * (16 * nb0 + 4 * a));
#pragma unroll
        for (int r = 0; r < 4; ++r) hsv[4 * a + r] = w[r];
      }
#pragma unroll
      for (int sb = 0; sb < kBlocks; ++sb) {
        const float dzs = s_dzs[sb * 32 + j];
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[sb][e] = hsv[e] * dzs;
      }
      ws8::run_steps<0, H, false, false>(acc, wlo, zero, in_l, in_l, [](int) {});
    } else {
      ws8::run_steps<0, H, true, false>(acc, wlo, zero, in_l, in_l, [](int) {});
    }
    if constexpr (layer > 1) ws8::load_w1<H, 0>(wlo, rs, P.p_bwd[layer - 1], C::NB, nb0, lane_off);
    ws8::run_steps<H, H, false, true>(acc, whi, zero, in_l, in_l, [&](int sb) {
      epilogue(std::true_type{}, std::integral_constant<bool, (layer > 1)>{}, out_l, mbits[sb], P.dz_h[layer - 1],
               sb);
    });
    if constexpr (layer > 1) {
      ws8::load_w1<H, H>(whi, rs, P.p_bwd[layer - 1], C::NB, nb0, lane_off);
      __syncthreads();
    }
  });
}

template <int W, int D, int SKIP, int LX, int LD>
int launch_bwd_bf16ws8(const MlpPlan& P, const void* packed, int64_t M, const float* d_rgb,
                       const float* d_sigma, const float* rgb, const float* sigma, const void* stash,
                       void* dz, hipStream_t s) {
  const int64_t nwg = (M + 127) / 128;
  mlp_bwd_bf16ws8_kernel<W, D, SKIP, LX, LD><<<(unsigned)nwg, 512, 0, s>>>(
      P, reinterpret_cast<const char*>(packed), M, d_rgb, d_sigma, rgb, sigma,
      reinterpret_cast<const __bf16*>(stash), reinterpret_cast<__bf16*>(dz));
  return 0;
}
