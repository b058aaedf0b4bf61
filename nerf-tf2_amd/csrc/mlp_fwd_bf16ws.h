// Weight-stationary BF16 fused forward for the reference topology (W = 256).
//
// A workgroup of 4 waves carries 128 samples (4 blocks of 32) through the MLP.
// For each layer, wave w owns output blocks {2w, 2w+1} (64 features) and holds
// that slice of the packed bf16 weights in VGPRs, split into two halves of the
// layer's k-steps (wlo, whi).  The layer runs k-half by k-half over all four
// sample blocks, with the accumulators of all four blocks live (128 AGPRs):
// once the first half has been used, the next layer's first half is fetched
// into wlo while the second half runs; once the second half is done, the next
// layer's second half is fetched into whi while the next layer's first half
// runs.  Every weight byte fetched from L2 feeds 128 samples, every LDS byte
// feeds 64 output features, and no more than one layer's slice is resident.
//
// B operands (activations, bf16, in the MFMA B layout) stream from an LDS
// buffer through a ring of six registers (five loads in flight; four measured
// 3-4 % slower, eight no better).  The
// accumulators start from the bias (the first MFMA's C operand is the bias in
// accumulator layout), so the epilogue of a sample block is: bf16 conversion,
// ReLU as a packed integer max on the bf16 bits (equal to ReLU before the
// rounding, which keeps the sign), two LDS stores and, in training, the stash
// rows (the ReLU bits stay in registers and leave as one 16-byte store per
// sample block every four layers).  Each block's epilogue is issued right after its last MFMA, so it
// overlaps the MFMAs of the next block.  Out-of-range samples of the last
// workgroup compute the values of the last sample; their stash stores go
// through a buffer resource over the workgroup's rows and fall outside its
// range, so the epilogue has no per-lane branch.  One barrier per
// layer; the sigma head and the rgb branch follow the reference's topology.
#pragma once
#include <type_traits>

#include "mlp_kernels.h"

namespace ws {

constexpr int kWaves = 4;
constexpr int kBlocks = 4;       // sample blocks (of 32) per workgroup
constexpr int kT = 16;           // bf16 k-steps over a 256-wide hidden input
constexpr int kTX = 4;           // k-steps over the xyz encoding slots (32 per half)
constexpr int kTD = 2;           // k-steps over the direction encoding slots (16 per half)
constexpr int kSlab = kT * 64;   // bf16x8 per sample block in one activation buffer
constexpr int kRing = 6;   // B operands in flight (LDS reads issued 5 steps ahead)
constexpr int kHalf = 10;        // register k-steps per half (20-step skip layer)

typedef short shortx8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// The packed weights as a buffer resource: every weight / bias load is a
// buffer load with one shared per-lane VGPR offset and a wave-uniform SGPR
// offset, so no per-lane 64-bit address is ever computed or kept live.  Loads
// past packed_bytes return zeros instead of faulting.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t packed_rsrc(const char* pk, int64_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<char*>(pk), 0, (int)bytes, 0x00020000);
}

__device__ __forceinline__ bf16x8 ldw(__amdgpu_buffer_rsrc_t r, uint32_t voff, int64_t soff) {
  return __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(r, voff, (int)soff, 0));
}

__device__ __forceinline__ floatx4 ldf4(__amdgpu_buffer_rsrc_t r, uint32_t voff, int64_t soff) {
  return __builtin_bit_cast(floatx4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, (int)soff, 0));
}

// k-steps OFF .. OFF+N-1 of output blocks nb0, nb0 + 1 of the packed layer at
// byte offset p_layer into w[0 .. N); lane_off = 16 * lane.
template <int N, int OFF>
__device__ __forceinline__ void load_w(bf16x8 (&w)[kHalf][2], __amdgpu_buffer_rsrc_t r, int64_t p_layer, int NBo,
                                       int nb0, uint32_t lane_off) {
#pragma unroll
  for (int t = 0; t < N; ++t)
#pragma unroll
    for (int i = 0; i < 2; ++i) w[t][i] = ldw(r, lane_off, p_layer + (int64_t)((OFF + t) * NBo + nb0 + i) * 1024);
}

// Bias of output block nb in accumulator layout: element 4a + r of lane
// (j, g) holds feature 32 nb + 8 a + 4 g + r (g_off = 16 * g bytes).
__device__ __forceinline__ floatx16 bias_acc(__amdgpu_buffer_rsrc_t rs, int64_t p_bias, int nb, uint32_t g_off) {
  floatx16 c;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const floatx4 v = ldf4(rs, g_off, p_bias + 4 * (32 * nb + 8 * a));
#pragma unroll
    for (int r = 0; r < 4; ++r) c[4 * a + r] = v[r];
  }
  return c;
}

// Between the accumulator (= B-operand) layout and row chunks.  Of a 16-feature
// group, lane (j, g) of the accumulator layout holds features {4g .. 4g+3,
// 8+4g .. 8+4g+3} of sample j (dwords d0 .. d3); a row chunk is the 8 consecutive
// features 8g .. 8g+7, one 16-byte store or load.  Two v_permlane32_swap trade
// the middle dwords between lanes j and j + 32; the map is its own inverse.
__device__ __forceinline__ bf16x8 swap_halves(bf16x8 v) {
  const u32x4 d = __builtin_bit_cast(u32x4, v);
  const auto a = __builtin_amdgcn_permlane32_swap(d[0], d[2], false, false);
  const auto b = __builtin_amdgcn_permlane32_swap(d[1], d[3], false, false);
  const u32x4 r = {a[0], b[0], a[1], b[1]};
  return __builtin_bit_cast(bf16x8, r);
}

typedef unsigned short ushortx2 __attribute__((ext_vector_type(2)));

// Nonzero bits of a block's 16 accumulator elements (lo: elements 0..7, hi:
// 8..15; non-negative bf16 after ReLU, so nonzero == positive): the pair
// (2k, 2k+1) goes to bits k and 16 + k, k = 0..7.
__device__ __forceinline__ uint32_t nz_bits(bf16x8 lo, bf16x8 hi) {
  const u32x4 a = __builtin_bit_cast(u32x4, lo), b = __builtin_bit_cast(u32x4, hi);
  const ushortx2 one = {1, 1};
  uint32_t w = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    // (the elements are copied out first: a bit_cast of a vector element
    // subscript reads element 0 with this compiler)
    const uint32_t ak = a[k], bk = b[k];
    w |= __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(ushortx2, ak), one)) << k;
    w |= __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(ushortx2, bk), one)) << (4 + k);
  }
  return w;
}

__device__ __forceinline__ bf16x8 relu_bf16(bf16x8 v) {
  const shortx8 z = {0, 0, 0, 0, 0, 0, 0, 0};
  return __builtin_bit_cast(bf16x8, __builtin_elementwise_max(__builtin_bit_cast(shortx8, v), z));
}

typedef __attribute__((address_space(3))) bf16x8 lds_b8;

// Per-lane LDS pointer to element `lane` (+ off) of a bf16x8 array in LDS,
// opaque to the optimiser, so that every access through it is this one VGPR
// plus an immediate offset (all offsets used below are < 64 KB).
__device__ __forceinline__ lds_b8* lds_lane(const bf16x8* arr, int off, int lane) {
  uint32_t a = (uint32_t)(uintptr_t)(const lds_b8*)arr + 16u * (uint32_t)(off + lane);
  asm volatile("" : "+v"(a));
  return (lds_b8*)(uintptr_t)a;
}

// B operand of k-step t (0 <= t < 20) of sample block sb: hidden features from
// `in` for t < 16, the xyz encoding for t >= 16 (per-lane pointers).
__device__ __forceinline__ bf16x8 b_operand(const lds_b8* in, const lds_b8* ex, int sb, int t) {
  return t < kT ? in[sb * kSlab + t * 64] : ex[(sb * kTX + (t - kT)) * 64];
}

// acc[sb][i] (+)= sum over k-steps T0 .. T0+H-1 of w[k - T0][i] * B(sb, k), for
// all four sample blocks.  INIT: the first k-step starts from the bias cb.
// After the last k-step of each block, epi(sb) runs (no-op for a first half).
struct NoEpilogue {
  __device__ void operator()(int) const {}
};

// Without an epilogue (a first half) step s runs sample block pair p = s / 2H,
// k-step (s / 2) % H, block 2p + s % 2: the two blocks of a pair alternate, so
// consecutive MFMAs go to four independent accumulators.  With an epilogue the
// blocks run one after the other, so each block's epilogue overlaps the next
// block's MFMAs.
template <int T0, int H, bool INIT, typename Epi>
__device__ __forceinline__ void run_steps(floatx16 (&acc)[kBlocks][2], const bf16x8 (&w)[kHalf][2],
                                          const floatx16 (&cb)[2], const lds_b8* in, const lds_b8* ex,
                                          Epi&& epi) {
  constexpr int S = kBlocks * H;
  constexpr bool PAIRS = std::is_same_v<std::decay_t<Epi>, NoEpilogue>;
  auto blk = [](int s) { return PAIRS ? 2 * (s / (2 * H)) + (s & 1) : s / H; };
  auto kst = [](int s) { return PAIRS ? (s >> 1) % H : s % H; };
  constexpr int R = kRing;
  bf16x8 b[R];
#pragma unroll
  for (int s = 0; s < R - 1; ++s) b[s] = b_operand(in, ex, blk(s), T0 + kst(s));
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (s + R - 1 < S) b[(s + R - 1) % R] = b_operand(in, ex, blk(s + R - 1), T0 + kst(s + R - 1));
    const int sb = blk(s), k = kst(s);
#pragma unroll
    for (int i = 0; i < 2; ++i)
      acc[sb][i] = mfma_bf16(w[k][i], b[s % R], (INIT && k == 0) ? cb[i] : acc[sb][i]);
    if (!PAIRS && k == H - 1) epi(sb);
  }
}

template <int A, int B, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (A < B) {
    f(std::integral_constant<int, A>{});
    static_for<A + 1, B>(f);
  }
}

}  // namespace ws

template <int W, int D, int SKIP, int LX, int LD, bool STASH>
__global__ __launch_bounds__(256) void mlp_fwd_bf16ws_kernel(
    MlpPlan P, const char* __restrict__ pk, int64_t M, const float* __restrict__ xyz,
    const float* __restrict__ dirs, const float* __restrict__ ro, const float* __restrict__ rd,
    const float* __restrict__ tv, int spr, float* __restrict__ rgb_out,
    float* __restrict__ sig_out, __bf16* __restrict__ stash) {
  using namespace ws;
  using C = NetC<W, D, SKIP, LX, LD>;
  static_assert(W == 256 && LX == 10 && LD == 4 && SKIP >= 0 && SKIP + 1 < D,
                "weight-stationary kernel is for the 8x256 NeRF MLP");
  static_assert(D % 4 == 0, "ReLU bit words are written four layers at a time");
  __shared__ bf16x8 s_act[2][kBlocks * kSlab];          // 2 x 64 KB
  __shared__ bf16x8 s_ex[kBlocks * kTX * 64];           // 16 KB
  __shared__ bf16x8 s_ed[kBlocks * kTD * 64];           // 8 KB
  __shared__ float s_red[kWaves][3][kBlocks * 32];     // rgb-head partials, 6 KB
  const int lane = threadIdx.x & 63, j = lane & 31, g = lane >> 5;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  const int64_t base = (int64_t)blockIdx.x * (kBlocks * 32);
  const int nb0 = 2 * wv;  // this wave's output blocks of a 256-wide layer
  const uint32_t lane_off = 16u * (uint32_t)lane, g_off = 16u * (uint32_t)g;
  const __amdgpu_buffer_rsrc_t rs = packed_rsrc(pk, P.packed_bytes);
  // The stash rows of this workgroup as a buffer resource: 32-bit per-lane byte
  // offsets (no 64-bit address per store), and rows past M are out of range, so
  // their stores are dropped.  row_off(sb): this lane's row of sample block sb.
  const int64_t wg_rows = M - base < kBlocks * 32 ? M - base : kBlocks * 32;
  const __amdgpu_buffer_rsrc_t rst =
      STASH ? __builtin_amdgcn_make_buffer_rsrc(stash + base * P.stash_per_sample, 0,
                                                (int)(wg_rows * P.stash_per_sample * 2), 0x00020000)
            : rs;
  auto row_off = [&](int sb) { return (uint32_t)((sb * 32 + j) * P.stash_per_sample * 2); };
  auto st16 = [&](bf16x8 v, uint32_t off) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), rst, off, 0, 0);
  };
  // first / second half of a layer's k-steps into wlo / whi
  auto load_half = [&](bf16x8 (&w)[kHalf][2], int m, bool enc, bool second) {
    if (enc) {
      constexpr int H = (kT + kTX) / 2;
      if (second)
        load_w<H, H>(w, rs, P.p_fwd[m], C::NB, nb0, lane_off);
      else
        load_w<H, 0>(w, rs, P.p_fwd[m], C::NB, nb0, lane_off);
    } else {
      constexpr int H = kT / 2;
      if (second)
        load_w<H, H>(w, rs, P.p_fwd[m], C::NB, nb0, lane_off);
      else
        load_w<H, 0>(w, rs, P.p_fwd[m], C::NB, nb0, lane_off);
    }
  };

  bf16x8 wlo[kHalf][2], whi[kHalf][2];
  floatx16 cb[2];
  floatx16 acc[kBlocks][2];

  // layer 0 weights (4 encoding k-steps) into whi, layer 1's first half into wlo
  load_w<kTX, 0>(whi, rs, P.p_fwd[0], C::NB, nb0, lane_off);
  load_half(wlo, 1, SKIP == 0, false);
  cb[0] = bias_acc(rs, P.p_bias[0], nb0, g_off);
  cb[1] = bias_acc(rs, P.p_bias[0], nb0 + 1, g_off);

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
    if (STASH) {
      __bf16* r = stash + mm * P.stash_per_sample;
      store_slots_bf16<C::KX>(ex, r + P.st_encx + g * C::KX);
      store_slots_bf16<C::KD>(ed, r + P.st_encd + g * C::KD);
    }
  }

  // ReLU bit words of the trunk layers, kept for four layers and then written as
  // one 16-byte chunk per sample block: the mask region of a stash row holds, per
  // wave and lane half, D consecutive words (one per trunk layer), so this lane's
  // words for layers 4q .. 4q+3 are 16 contiguous bytes.
  uint32_t mbits[kBlocks][4];
  auto store_masks = [&](int q) {
#pragma unroll
    for (int sb = 0; sb < kBlocks; ++sb) {
      const u32x4 v = {mbits[sb][0], mbits[sb][1], mbits[sb][2], mbits[sb][3]};
      __builtin_amdgcn_raw_buffer_store_b128(
          v, rst, row_off(sb) + (uint32_t)((P.st_mask + 2 * ((2 * wv + g) * D + 4 * q)) * 2), 0, 0);
    }
  };

  // Epilogue of sample block sb of a 256-wide layer: bf16, ReLU, LDS (B layout), stash.
  // out: per-lane pointer at k-step 2 nb0 of sample block 0 of the output buffer
  // MASK: also the ReLU bit word of the wave's 64 features (bits of output block
  // nb0 + i at 8 i + {0..7, 16..23}) of trunk layer `layer`, into mbits
  auto epilogue = [&](lds_b8* out, bool relu, int64_t st_off, auto mask_c, int layer, int sb) {
    constexpr bool MASK = decltype(mask_c)::value;
    uint32_t bits = 0;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      bf16x8 lo = acc_to_bf16(acc[sb][i], 0), hi = acc_to_bf16(acc[sb][i], 1);
      if (relu) {
        lo = relu_bf16(lo);
        hi = relu_bf16(hi);
      }
      out[sb * kSlab + (2 * i) * 64] = lo;
      out[sb * kSlab + (2 * i + 1) * 64] = hi;
      if (STASH) {
        // as row chunks: features 32 (nb0 + i) + 16 h + 8 g .. + 7 (lo: h = 0, hi: h = 1)
        const uint32_t o = row_off(sb) + (uint32_t)((st_off + 32 * (nb0 + i) + 8 * g) * 2);
        st16(swap_halves(lo), o);
        st16(swap_halves(hi), o + 32);
        bits |= nz_bits(lo, hi) << (8 * i);
      }
    }
    if constexpr (STASH && MASK) mbits[sb][layer & 3] = bits;
  };

  __syncthreads();  // encodings visible
  // ---- layer 0 (encoding only, weights in whi)
  lds_b8* const ex_l = lds_lane(s_ex, 0, lane);
  {
    lds_b8* const out_l = lds_lane(s_act[0], 2 * nb0 * 64, lane);
    run_steps<kT, kTX, true>(acc, whi, cb, nullptr, ex_l, [&](int sb) { epilogue(out_l, true, P.st_h[0], std::true_type{}, 0, sb); });
  }
  load_half(whi, 1, SKIP == 0, true);  // layer 1's second half and bias
  cb[0] = bias_acc(rs, P.p_bias[1], nb0, g_off);
  cb[1] = bias_acc(rs, P.p_bias[1], nb0 + 1, g_off);
  __syncthreads();

  // ---- layers 1 .. D (D: the linear bottleneck), unrolled
  static_for<1, D + 1>([&](auto Lc) {
    constexpr int l = decltype(Lc)::value;
    constexpr int cur = (l - 1) & 1;
    const bf16x8* in = s_act[cur];
    lds_b8* const in_l = lds_lane(s_act[cur], 0, lane);
    lds_b8* const out_l = lds_lane(s_act[cur ^ 1], 2 * nb0 * 64, lane);
    if constexpr (l == D) {
      // sigma head on h_{D-1} (the input of this layer): wave w handles sample block w
      const uint32_t hs_off = 4u * (uint32_t)(g * C::KH);  // this lane half's head weights
      float zs = 0.f;
#pragma unroll
      for (int t = 0; t < kT; ++t) {
        const bf16x8 hv = in[wv * kSlab + t * 64 + lane];
        const int nb = t >> 1, h = t & 1;
        const floatx4 w0 = ldf4(rs, hs_off, P.p_head_sigma + 4 * (16 * nb + 8 * h));
        const floatx4 w1 = ldf4(rs, hs_off, P.p_head_sigma + 4 * (16 * nb + 8 * h + 4));
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
    constexpr bool enc = l == SKIP + 1;
    constexpr int H = enc ? (kT + kTX) / 2 : kT / 2;
    run_steps<0, H, true>(acc, wlo, cb, in_l, ex_l, NoEpilogue{});
    // wlo is free: the next layer's first half and bias (after the bottleneck: the
    // rgb branch's 18 k-steps of output block wv and its bias)
    if constexpr (l < D) {
      load_half(wlo, l + 1, l + 1 == SKIP + 1, false);
      cb[0] = bias_acc(rs, P.p_bias[l + 1], nb0, g_off);
      cb[1] = bias_acc(rs, P.p_bias[l + 1], nb0 + 1, g_off);
    } else {
#pragma unroll
      for (int t = 0; t < kT + kTD; ++t)
        wlo[t >> 1][t & 1] = ldw(rs, lane_off, P.p_fwd[D + 1] + (int64_t)(t * C::NBR + wv) * 1024);
      cb[0] = bias_acc(rs, P.p_bias[D + 1], wv, g_off);
    }
    const int64_t st_off = l < D ? P.st_h[l] : P.st_bott;
    run_steps<H, H, false>(acc, whi, cb, in_l, ex_l, [&](int sb) { epilogue(out_l, l < D, st_off, std::integral_constant<bool, (l < D)>{}, l, sb); });
    if constexpr (STASH && l < D && (l & 3) == 3) store_masks(l >> 2);
    if constexpr (l < D) load_half(whi, l + 1, l + 1 == SKIP + 1, true);  // whi is free
    __syncthreads();  // layer l outputs complete
  });
  constexpr int fin = D & 1;  // buffer holding the bottleneck

  // ---- rgb branch dense_{D+1}: one output block per wave (wv), input = bottleneck + dir enc
  {
    lds_b8* const in_l = lds_lane(s_act[fin], 0, lane);
    lds_b8* const ed_l = lds_lane(s_ed, 0, lane);
    const floatx4* hr4 = reinterpret_cast<const floatx4*>(pk + P.p_head_rgb);
#pragma unroll
    for (int sb = 0; sb < kBlocks; ++sb) {
      floatx16 a = cb[0];
#pragma unroll
      for (int t = 0; t < kT; ++t) a = mfma_bf16(wlo[t >> 1][t & 1], in_l[sb * kSlab + t * 64], a);
#pragma unroll
      for (int t = 0; t < kTD; ++t)
        a = mfma_bf16(wlo[(kT + t) >> 1][(kT + t) & 1], ed_l[(sb * kTD + t) * 64], a);
      // ReLU, rounded to bf16 (the stashed activation), then the fp32 head partials
#pragma unroll
      for (int r = 0; r < 16; ++r) a[r] = (float)(__bf16)(a[r] > 0.f ? a[r] : 0.f);
      if (STASH) {
        const uint32_t o = row_off(sb) + (uint32_t)((P.st_rbr + 32 * wv + 8 * g) * 2);
        const bf16x8 lo = acc_to_bf16(a, 0), hi = acc_to_bf16(a, 1);
        st16(swap_halves(lo), o);
        st16(swap_halves(hi), o + 32);
        // ReLU bits of rgb-branch block wv: bits 0..7 and 16..23 of nz_bits as 16 bits
        const uint32_t b = nz_bits(lo, hi);
        __builtin_amdgcn_raw_buffer_store_b16((unsigned short)((b & 0xFFu) | ((b >> 8) & 0xFF00u)), rst,
                                              row_off(sb) + (uint32_t)((P.st_mask + 16 * D + 2 * wv + g) * 2), 0, 0);
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
  if (stash)
    mlp_fwd_bf16ws_kernel<W, D, SKIP, LX, LD, true><<<(unsigned)nwg, 256, 0, s>>>(
        P, reinterpret_cast<const char*>(packed), M, xyz, dirs, ro, rd, tv, spr, rgb, sigma,
        reinterpret_cast<__bf16*>(stash));
  else
    mlp_fwd_bf16ws_kernel<W, D, SKIP, LX, LD, false><<<(unsigned)nwg, 256, 0, s>>>(
        P, reinterpret_cast<const char*>(packed), M, xyz, dirs, ro, rd, tv, spr, rgb, sigma, nullptr);
  return 0;
}
