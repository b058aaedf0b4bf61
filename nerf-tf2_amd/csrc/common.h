// Shared helpers for the NeRF hot-path kernels (gfx950 / CDNA4, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/nerf_hip.h"

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short ushortx8 __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------------------
// Error state: one thread-local message per host thread (nerf_last_error()).
// ---------------------------------------------------------------------------
void nerf_set_error(const char* fmt, ...);

#define NERF_CHECK_ARG(cond, ...)          \
  do {                                     \
    if (!(cond)) {                         \
      nerf_set_error(__VA_ARGS__);         \
      return NERF_ERR_INVALID_ARG;         \
    }                                      \
  } while (0)

#define NERF_LAUNCH_CHECK(what)                                               \
  do {                                                                        \
    hipError_t e_ = hipGetLastError();                                        \
    if (e_ != hipSuccess) {                                                   \
      nerf_set_error("%s: launch failed: %s", what, hipGetErrorString(e_));   \
      return NERF_ERR_LAUNCH;                                                 \
    }                                                                         \
  } while (0)

static inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Cache policy of buffer loads and stores that stream a buffer no kernel reads
// again soon (the blocked BF16 stash and dz, written once and read once a kernel
// or two later, far beyond the L2): non-temporal (`nt` on gfx950), so that they
// do not displace the weights and tiles the kernels do reuse.  Only for stores of
// whole cache lines: the FP32 stash and dz rows, stored in 64-byte pieces that
// the L2 merges, take longer with it.
constexpr int kNT = 2;

// ---------------------------------------------------------------------------
// Exactly rounded float ops that the compiler may not contract or reorder.
// ---------------------------------------------------------------------------
// HIP's __f*_rn intrinsics are plain operators unless OCML_BASIC_ROUNDED_OPERATIONS
// is defined (and __fsqrt_rn is then the native, approximate sqrt), so these
// helpers use operators with contraction disabled locally: an fmul feeding an
// fadd is never fused into an FMA, whatever the global -ffp-contract setting.
// Division and sqrtf are correctly rounded (HIP's default
// -fhip-fp32-correctly-rounded-divide-sqrt).
#define NERF_EXACT_OP(name, T, op)                              \
  __device__ __forceinline__ T name(T a, T b) {                 \
    _Pragma("clang fp contract(off)") return a op b;            \
  }
NERF_EXACT_OP(fadd, float, +)
NERF_EXACT_OP(fsub, float, -)
NERF_EXACT_OP(fmul, float, *)
NERF_EXACT_OP(fdiv, float, /)
NERF_EXACT_OP(fadd, double, +)
NERF_EXACT_OP(fsub, double, -)
NERF_EXACT_OP(fmul, double, *)
NERF_EXACT_OP(fdiv, double, /)
#undef NERF_EXACT_OP
__device__ __forceinline__ float fsqrt(float a) { return sqrtf(a); }
__device__ __forceinline__ double fsqrt(double a) { return sqrt(a); }

// Error-free transformation a + b = s + e (Knuth TwoSum); no FMA, no reassociation.
template <typename T>
__device__ __forceinline__ void two_sum(T a, T b, T& s, T& e) {
  s = fadd(a, b);
  T bb = fsub(s, a);
  e = fadd(fsub(a, fsub(s, bb)), fsub(b, bb));
}

// Compensated ("double-T") add of two (hi, lo) pairs.
template <typename T>
__device__ __forceinline__ void dd_add(T ah, T al, T bh, T bl, T& rh, T& rl) {
  T s, e;
  two_sum(ah, bh, s, e);
  e = fadd(e, fadd(al, bl));
  rh = fadd(s, e);
  rl = fsub(e, fsub(rh, s));
}

// Wave64 lane shuffles (ds_bpermute based).
template <typename T>
__device__ __forceinline__ T shfl(T v, int src) { return __shfl(v, src, 64); }
template <typename T>
__device__ __forceinline__ T shfl_up(T v, int d) { return __shfl_up(v, d, 64); }
template <typename T>
__device__ __forceinline__ T shfl_down(T v, int d) { return __shfl_down(v, d, 64); }
template <typename T>
__device__ __forceinline__ T shfl_xor(T v, int m) { return __shfl_xor(v, m, 64); }

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += shfl_xor(v, m);
  return v;
}

__device__ __forceinline__ float relu(float x) { return x > 0.f ? x : 0.f; }

// fp32 -> bf16, round to nearest even (finite inputs).
__device__ __forceinline__ unsigned short f2bf(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float(((uint32_t)h) << 16); }
