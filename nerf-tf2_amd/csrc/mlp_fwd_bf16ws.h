// Building blocks of the weight-stationary BF16 kernels of the reference
// topology (W = 256): the forward (mlp_fwd_bf16ws8.h) and the dX chain
// (mlp_bwd_bf16ws8.h), eight waves per workgroup, wave w owning output block w.
//
// A workgroup carries 128 samples (4 blocks of 32) through the MLP.  Each wave
// owns a 32-feature output block of every layer and holds that slice of the
// packed bf16 weights in VGPRs, split into two halves of the layer's k-steps
// (wlo, whi).  A layer runs k-half by k-half over all four sample blocks, with
// the accumulators of all four blocks live: once the first half has been used,
// the next layer's first half is fetched into wlo while the second half runs;
// once the second half is done, the next layer's second half is fetched into
// whi while the next layer's first half runs.  Every weight byte fetched from
// L2 feeds 128 samples, and no more than one layer's slice is resident.
//
// B operands (activations, bf16, in the MFMA B layout) stream from a double-
// buffered LDS tile through a ring of registers (ws8::run_steps).  Accumulators can
// start from the bias (the first MFMA's C operand is the bias in accumulator
// layout, bias_acc), and an epilogue converts a block to bf16 (acc_to_bf16),
// applies ReLU as a packed integer max on the bf16 bits (relu_bf16: equal to
// ReLU before the rounding, which keeps the sign) and stores it back to LDS;
// swap_halves turns accumulator-layout chunks into row chunks for the stash
// and the dz rows, and nz_bits extracts the ReLU bits.  All weight and bias
// loads are buffer loads with one shared per-lane VGPR offset.
#pragma once
#include <type_traits>

#include "mlp_kernels.h"

namespace ws {

constexpr int kBlocks = 4;       // sample blocks (of 32) per workgroup
constexpr int kT = 16;           // bf16 k-steps over a 256-wide hidden input
constexpr int kTX = 4;           // k-steps over the xyz encoding slots (32 per half)
constexpr int kTD = 2;           // k-steps over the direction encoding slots (16 per half)
constexpr int kSlab = kT * 64;   // bf16x8 per sample block in one activation buffer
constexpr int kRing = 6;   // B operands in flight (LDS reads issued 5 steps ahead)
constexpr int kHalf = 10;        // register k-steps per half (20-step skip layer)

typedef short shortx8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

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

// Four floats rounded to bf16 (as __builtin_convertvector) as the dwords of an 8-byte store.
__device__ __forceinline__ u32x2 bf16x4_bits(float a, float b, float c, float d) {
  const floatx4_ x = {a, b, c, d};
  return __builtin_bit_cast(u32x2, __builtin_convertvector(x, bf16x4));
}

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

template <int A, int B, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (A < B) {
    f(std::integral_constant<int, A>{});
    static_for<A + 1, B>(f);
  }
}

}  // namespace ws
