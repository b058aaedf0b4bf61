// Device building blocks shared by the fused MLP forward and backward kernels.
#pragma once
#include "common.h"
#include "mlp_plan.h"

// (W, D, skip, L_xyz, L_dir) topologies with compiled kernels:
//   reference coarse/fine NeRF MLP, finite-difference check MLP, proposal MLP.
#define NERF_MLP_INSTANCES(X) \
  X(256, 8, 4, 10, 4)         \
  X(32, 4, 1, 10, 4)          \
  X(64, 4, -1, 8, 0)

template <int W, int D, int SKIP, int LX, int LD>
struct NetC {
  static constexpr int NB = W / 32;        // output blocks of a hidden layer
  static constexpr int KH = W / 2;         // fp32 k-steps over a hidden input
  static constexpr int QH = KH / 4;        // groups of 4 k-steps
  static constexpr int HX = 2 + 3 * LX;
  static constexpr int KX = (HX + 7) / 8 * 8;
  static constexpr bool RGB = LD > 0;
  static constexpr int HD = RGB ? 2 + 3 * LD : 0;
  static constexpr int KD = (HD + 7) / 8 * 8;
  static constexpr int KDA = RGB ? KD : 8;  // array size (unused without rgb branch)
  static constexpr int WR = W / 2;
  static constexpr int NBR = (WR + 31) / 32;
};

int nerf_plan_from_desc(const nerf_mlp_desc* d, int prec, MlpPlan* p);
int nerf_num_cus();
int nerf_dispatch_fwd(const MlpPlan& P, const void* packed, int64_t M, const float* xyz,
                      const float* dirs, const float* ro, const float* rd, const float* tv,
                      int spr, float* rgb, float* sigma, void* stash, hipStream_t s);
template <int W, int D, int SKIP, int LX, int LD>
int launch_fwd_bf16(const MlpPlan& P, const void* packed, int64_t M, const float* xyz,
                    const float* dirs, const float* ro, const float* rd, const float* tv, int spr,
                    float* rgb, float* sigma, void* stash, hipStream_t s);

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float softplus(float z) { return z > 20.f ? z : log1pf(expf(z)); }
__device__ __forceinline__ float sigmoidf(float z) { return 1.f / (1.f + expf(-z)); }

// Sample point: either explicit (xyz, dirs) rows or o + t*d of its ray
// (ray_utils.py:251, 389: one multiply, one add).
__device__ __forceinline__ void load_sample(int64_t m, const float* __restrict__ xyz,
                                            const float* __restrict__ dirs,
                                            const float* __restrict__ ro,
                                            const float* __restrict__ rd,
                                            const float* __restrict__ tv, int spr, float (&x)[3],
                                            float (&d)[3]) {
  // scalars first, then the arrays: per-component stores under the branches kept the
  // arrays in scratch memory
  float x0, x1, x2, d0 = 0.f, d1 = 0.f, d2 = 0.f;
  if (ro) {
    const int64_t r = m / spr;
    const float t = tv[m];
    d0 = rd[3 * r];
    d1 = rd[3 * r + 1];
    d2 = rd[3 * r + 2];
    x0 = fadd(ro[3 * r], fmul(t, d0));
    x1 = fadd(ro[3 * r + 1], fmul(t, d1));
    x2 = fadd(ro[3 * r + 2], fmul(t, d2));
  } else {
    x0 = xyz[3 * m];
    x1 = xyz[3 * m + 1];
    x2 = xyz[3 * m + 2];
    if (dirs) {
      d0 = dirs[3 * m];
      d1 = dirs[3 * m + 1];
      d2 = dirs[3 * m + 2];
    }
  }
  x[0] = x0;
  x[1] = x1;
  x[2] = x2;
  d[0] = d0;
  d[1] = d1;
  d[2] = d2;
}

// Positional encoding (model.py:309-330) directly in the lane-half slots of
// mlp_plan.h.  Argument = fl32(x_c * (2^l * fl32(pi))), exact power-of-two
// scaling of fl32(pi) as in the reference; sin/cos are full-range accurate.
template <int L, int K>
__device__ __forceinline__ void encode_slots(const float (&x)[3], int g, float (&out)[K]) {
  constexpr int P0 = (3 * L - 2) / 2;
  constexpr float kPi = 3.14159265358979323846f;
  // the components as register values: a select between elements of the array by a
  // lane-dependent index would otherwise be folded into an indexed load, which keeps
  // the array in scratch memory
  float x0 = x[0], x1 = x[1], x2 = x[2];
  asm volatile("" : "+v"(x0), "+v"(x1), "+v"(x2));
#pragma unroll
  for (int k = 0; k < K / 2; ++k) {
    const int p = g ? P0 + k : k - 2;
    const bool is_pair = g ? (P0 + k < 3 * L) : (k >= 2 && k - 2 < P0);
    float s = 0.f, c = 0.f;
    if (is_pair) {
      const int comp = p / L, l = p - comp * L;
      const float xc = comp == 0 ? x0 : (comp == 1 ? x1 : x2);
      const float arg = fmul(xc, ldexpf(kPi, l));
      sincosf(arg, &s, &c);
    }
    float v0 = s, v1 = c;
    if (!is_pair) {
      v0 = (g == 0 && k == 0) ? x0 : ((g == 0 && k == 1) ? x2 : 0.f);
      v1 = (g == 0 && k == 0) ? x1 : 0.f;
    }
    out[2 * k] = v0;
    out[2 * k + 1] = v1;
  }
}

template <int K>
__device__ __forceinline__ void store_slots_f32(const float (&v)[K], float* dst) {
#pragma unroll
  for (int e = 0; e < K; e += 4)
    *reinterpret_cast<floatx4*>(dst + e) = floatx4{v[e], v[e + 1], v[e + 2], v[e + 3]};
}

template <int NBo>
__device__ __forceinline__ void zero_acc(floatx16 (&acc)[NBo]) {
#pragma unroll
  for (int nb = 0; nb < NBo; ++nb)
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[nb][i] = 0.f;
}

__device__ __forceinline__ floatx16 mfma_f32(float a, float b, floatx16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
}

// acc[nb] += A(q-group, nb) * B for KR register k-steps (KR % 4 == 0).  Two
// weight groups in flight; scheduling barriers keep the compiler from hoisting
// every load of the unrolled loop (which would need KR*NBo VGPRs).
template <int NBo, int KR>
__device__ __forceinline__ void mfma_regs_f32(floatx16 (&acc)[NBo], const floatx4* __restrict__ wq,
                                              const float (&b)[KR], int lane) {
  constexpr int G = KR / 4;
  floatx4 wa[NBo], wb[NBo];
#pragma unroll
  for (int nb = 0; nb < NBo; ++nb) wa[nb] = wq[nb * 64 + lane];
#pragma unroll
  for (int q = 0; q < G; q += 2) {
    if (q + 1 < G) {
#pragma unroll
      for (int nb = 0; nb < NBo; ++nb) wb[nb] = wq[((q + 1) * NBo + nb) * 64 + lane];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int nb = 0; nb < NBo; ++nb) acc[nb] = mfma_f32(wa[nb][r], b[4 * q + r], acc[nb]);
    __builtin_amdgcn_sched_barrier(0);
    if (q + 2 < G) {
#pragma unroll
      for (int nb = 0; nb < NBo; ++nb) wa[nb] = wq[((q + 2) * NBo + nb) * 64 + lane];
    }
    if (q + 1 < G) {
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int nb = 0; nb < NBo; ++nb) acc[nb] = mfma_f32(wb[nb][r], b[4 * (q + 1) + r], acc[nb]);
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

// acc[nb] += A * B over Q groups of 4 k-steps whose B operand is in the wave's
// LDS slab act[q][lane]; weights and B operands are fetched one group ahead.
// The loop body has no branch: on the last pass the prefetch re-reads the
// last group (an in-bounds load whose value is unused), so the wait before
// each MFMA counts only the loads of its own group.  (With a conditional
// prefetch, the join of the two paths made the compiler wait for the newest
// weight loads within a few MFMAs of issuing them, and a B operand read just
// before its first MFMA stalled every group on the LDS latency.)
template <int NBo, int Q>
__device__ __forceinline__ void mfma_lds_f32(floatx16 (&acc)[NBo], const floatx4* __restrict__ wq,
                                             const floatx4* act, int lane) {
  static_assert(Q % 2 == 0, "Q must be even");
  floatx4 wa[NBo], wb[NBo];
#pragma unroll
  for (int nb = 0; nb < NBo; ++nb) wa[nb] = wq[nb * 64 + lane];
  floatx4 ba = act[lane];
  for (int q = 0; q < Q; q += 2) {
#pragma unroll
    for (int nb = 0; nb < NBo; ++nb) wb[nb] = wq[((q + 1) * NBo + nb) * 64 + lane];
    const floatx4 bb = act[(q + 1) * 64 + lane];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int nb = 0; nb < NBo; ++nb) acc[nb] = mfma_f32(wa[nb][r], ba[r], acc[nb]);
    __builtin_amdgcn_sched_barrier(0);
    const int qn = q + 2 < Q ? q + 2 : Q - 1;
#pragma unroll
    for (int nb = 0; nb < NBo; ++nb) wa[nb] = wq[(qn * NBo + nb) * 64 + lane];
    ba = act[qn * 64 + lane];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int nb = 0; nb < NBo; ++nb) acc[nb] = mfma_f32(wb[nb][r], bb[r], acc[nb]);
    __builtin_amdgcn_sched_barrier(0);
  }
}

// z = acc + bias, then ReLU (RELU = 1) or identity.
template <int NBo, int RELU>
__device__ __forceinline__ void bias_act(floatx16 (&acc)[NBo], const float* __restrict__ bias, int g) {
#pragma unroll
  for (int nb = 0; nb < NBo; ++nb)
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const floatx4 bv = *reinterpret_cast<const floatx4*>(bias + 32 * nb + 8 * a + 4 * g);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float v = acc[nb][4 * a + r] + bv[r];
        if (RELU) v = v > 0.f ? v : 0.f;
        acc[nb][4 * a + r] = v;
      }
    }
}

// Accumulators -> B-operand slab: k-step 16 nb + i lives at act[(16nb+i)/4][lane][i%4].
template <int NBo>
__device__ __forceinline__ void store_act_f32(const floatx16 (&acc)[NBo], floatx4* act, int lane) {
#pragma unroll
  for (int nb = 0; nb < NBo; ++nb)
#pragma unroll
    for (int a = 0; a < 4; ++a)
      act[(4 * nb + a) * 64 + lane] =
          floatx4{acc[nb][4 * a], acc[nb][4 * a + 1], acc[nb][4 * a + 2], acc[nb][4 * a + 3]};
}

// Accumulators -> a row-major feature row of this lane's sample (natural order).
template <int NBo>
__device__ __forceinline__ void store_rows_f32(const floatx16 (&acc)[NBo], float* row, int g) {
#pragma unroll
  for (int nb = 0; nb < NBo; ++nb)
#pragma unroll
    for (int a = 0; a < 4; ++a)
      *reinterpret_cast<floatx4*>(row + 32 * nb + 8 * a + 4 * g) =
          floatx4{acc[nb][4 * a], acc[nb][4 * a + 1], acc[nb][4 * a + 2], acc[nb][4 * a + 3]};
}

// Row-major feature row -> accumulator layout.
template <int NBo>
__device__ __forceinline__ void load_rows_f32(floatx16 (&acc)[NBo], const float* row, int g) {
#pragma unroll
  for (int nb = 0; nb < NBo; ++nb)
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const floatx4 v = *reinterpret_cast<const floatx4*>(row + 32 * nb + 8 * a + 4 * g);
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[nb][4 * a + r] = v[r];
    }
}

// Partial dot product of this lane's features with a head weight vector laid
// out by k-step (w[16 nb + i] for feature cfeat(nb, i, g)).
template <int NBo>
__device__ __forceinline__ float head_dot_f32(const floatx16 (&acc)[NBo], const float* __restrict__ w) {
  float s = 0.f;
#pragma unroll
  for (int nb = 0; nb < NBo; ++nb)
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const floatx4 wv = *reinterpret_cast<const floatx4*>(w + 16 * nb + 4 * a);
#pragma unroll
      for (int r = 0; r < 4; ++r) s = s + acc[nb][4 * a + r] * wv[r];
    }
  return s;
}

// rgb = sigmoid(h @ W_rgb + b) with W_rgb packed as [g][k-step][4].
template <int NBR>
__device__ __forceinline__ void rgb_head_f32(const floatx16 (&accr)[NBR], const float* __restrict__ hr,
                                             int g, float (&rgb)[3]) {
  constexpr int KS = 16 * NBR;
  float p0 = 0.f, p1 = 0.f, p2 = 0.f;
  const floatx4* w4 = reinterpret_cast<const floatx4*>(hr) + g * KS;
#pragma unroll
  for (int nb = 0; nb < NBR; ++nb)
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const floatx4 w = w4[16 * nb + i];
      const float h = accr[nb][i];
      p0 = p0 + h * w[0];
      p1 = p1 + h * w[1];
      p2 = p2 + h * w[2];
    }
  p0 = p0 + shfl_xor(p0, 32);
  p1 = p1 + shfl_xor(p1, 32);
  p2 = p2 + shfl_xor(p2, 32);
  const float* bias = hr + 2 * KS * 4;
  rgb[0] = sigmoidf(p0 + bias[0]);
  rgb[1] = sigmoidf(p1 + bias[1]);
  rgb[2] = sigmoidf(p2 + bias[2]);
}

// ---------------------------------------------------------------------------
// bf16 building blocks (v_mfma_f32_32x32x16_bf16: 8 bf16 of A and B per lane)
// ---------------------------------------------------------------------------
typedef float floatx8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ floatx16 mfma_bf16(bf16x8 a, bf16x8 b, floatx16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

typedef float floatx2_ __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_ __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4_cvt __attribute__((ext_vector_type(4)));

// Two floats -> one packed bf16 pair (round to nearest even): one v_cvt_pk_bf16_f32.
__device__ __forceinline__ uint32_t pack_bf16x2(float lo, float hi) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector((floatx2_){lo, hi}, bf16x2_));
}

// Converted pair by pair: a whole-vector conversion of values extracted from
// accumulators is split by the instruction selector into one conversion per
// element plus byte permutes.
__device__ __forceinline__ bf16x8 to_bf16x8(float a0, float a1, float a2, float a3, float a4,
                                            float a5, float a6, float a7) {
  const u32x4_cvt w = {pack_bf16x2(a0, a1), pack_bf16x2(a2, a3), pack_bf16x2(a4, a5), pack_bf16x2(a6, a7)};
  return __builtin_bit_cast(bf16x8, w);
}

// k-step t of a 32-feature block: accumulators 8h .. 8h+7 (h = t % 2).
__device__ __forceinline__ bf16x8 acc_to_bf16(const floatx16& a, int h) {
  return to_bf16x8(a[8 * h], a[8 * h + 1], a[8 * h + 2], a[8 * h + 3], a[8 * h + 4], a[8 * h + 5],
                   a[8 * h + 6], a[8 * h + 7]);
}

template <int K>
__device__ __forceinline__ void slots_to_bf16(const float (&v)[K], bf16x8 (&out)[K / 8]) {
#pragma unroll
  for (int t = 0; t < K / 8; ++t)
    out[t] = to_bf16x8(v[8 * t], v[8 * t + 1], v[8 * t + 2], v[8 * t + 3], v[8 * t + 4], v[8 * t + 5],
                       v[8 * t + 6], v[8 * t + 7]);
}

// acc[nb] += A * B over KT register k-steps (two weight k-steps in flight).
template <int NBo, int KT>
__device__ __forceinline__ void mfma_regs_bf16(floatx16 (&acc)[NBo], const bf16x8* __restrict__ wq,
                                               const bf16x8 (&b)[KT], int lane) {
  bf16x8 wa[NBo], wb[NBo];
#pragma unroll
  for (int nb = 0; nb < NBo; ++nb) wa[nb] = wq[nb * 64 + lane];
#pragma unroll
  for (int t = 0; t < KT; t += 2) {
    if (t + 1 < KT) {
#pragma unroll
      for (int nb = 0; nb < NBo; ++nb) wb[nb] = wq[((t + 1) * NBo + nb) * 64 + lane];
    }
#pragma unroll
    for (int nb = 0; nb < NBo; ++nb) acc[nb] = mfma_bf16(wa[nb], b[t], acc[nb]);
    __builtin_amdgcn_sched_barrier(0);
    if (t + 2 < KT) {
#pragma unroll
      for (int nb = 0; nb < NBo; ++nb) wa[nb] = wq[((t + 2) * NBo + nb) * 64 + lane];
    }
    if (t + 1 < KT) {
#pragma unroll
      for (int nb = 0; nb < NBo; ++nb) acc[nb] = mfma_bf16(wb[nb], b[t + 1], acc[nb]);
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

// acc[nb] += A * B over T k-steps whose B operand sits in the wave's LDS slab
// act[t][lane] (bf16x8); weights prefetched two k-steps ahead.
template <int NBo, int T>
__device__ __forceinline__ void mfma_lds_bf16(floatx16 (&acc)[NBo], const bf16x8* __restrict__ wq,
                                              const bf16x8* act, int lane) {
  static_assert(T % 2 == 0, "T must be even");
  bf16x8 wa[NBo], wb[NBo];
#pragma unroll
  for (int nb = 0; nb < NBo; ++nb) wa[nb] = wq[nb * 64 + lane];
  for (int t = 0; t < T; t += 2) {
#pragma unroll
    for (int nb = 0; nb < NBo; ++nb) wb[nb] = wq[((t + 1) * NBo + nb) * 64 + lane];
    bf16x8 b = act[t * 64 + lane];
#pragma unroll
    for (int nb = 0; nb < NBo; ++nb) acc[nb] = mfma_bf16(wa[nb], b, acc[nb]);
    __builtin_amdgcn_sched_barrier(0);
    if (t + 2 < T) {
#pragma unroll
      for (int nb = 0; nb < NBo; ++nb) wa[nb] = wq[((t + 2) * NBo + nb) * 64 + lane];
    }
    b = act[(t + 1) * 64 + lane];
#pragma unroll
    for (int nb = 0; nb < NBo; ++nb) acc[nb] = mfma_bf16(wb[nb], b, acc[nb]);
    __builtin_amdgcn_sched_barrier(0);
  }
}

// Accumulators -> bf16 B-operand slab: k-step 2 nb + h at act[(2nb+h)][lane].
template <int NBo>
__device__ __forceinline__ void store_act_bf16(const floatx16 (&acc)[NBo], bf16x8* act, int lane) {
#pragma unroll
  for (int nb = 0; nb < NBo; ++nb) {
    act[(2 * nb) * 64 + lane] = acc_to_bf16(acc[nb], 0);
    act[(2 * nb + 1) * 64 + lane] = acc_to_bf16(acc[nb], 1);
  }
}

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4_ __attribute__((ext_vector_type(4)));

// One sample's row of a bf16 stash or dz buffer blocked by 32 samples
// (MlpPlan::sblk): the sample is row j of the block of n rows starting at blk,
// whose 8-element chunk c of that row is at blk + 8 (c n + j).
template <typename T>
struct BlkRow {
  T* blk;
  int n, j;
  __device__ __forceinline__ T* at(int64_t e) const { return blk + 8 * ((e >> 3) * n + j) + (e & 7); }
};

// Accumulators -> features off .. of a blocked bf16 row (natural order), 4
// features per store; lane halves g = 0, 1 fill the two halves of each 8-feature
// chunk, so a store of the wave's 32 samples covers 512 contiguous bytes (whole
// cache lines: stored non-temporally, as the stash and dz are read back only
// after the L2 has turned over).
template <int NBo>
__device__ __forceinline__ void store_rows_bf16(const floatx16 (&acc)[NBo], BlkRow<__bf16> row, int64_t off, int g) {
#pragma unroll
  for (int nb = 0; nb < NBo; ++nb)
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const floatx4_ v = {acc[nb][4 * a], acc[nb][4 * a + 1], acc[nb][4 * a + 2], acc[nb][4 * a + 3]};
      __builtin_nontemporal_store(__builtin_convertvector(v, bf16x4),
                                  reinterpret_cast<bf16x4*>(row.at(off + 32 * nb + 8 * a + 4 * g)));
    }
}

// Encoding slots v (K a multiple of 4) -> elements off .. off + K - 1 of a blocked bf16 row.
template <int K>
__device__ __forceinline__ void store_slots_bf16(const float (&v)[K], BlkRow<__bf16> row, int64_t off) {
#pragma unroll
  for (int e = 0; e < K; e += 4) {
    const floatx4_ x = {v[e], v[e + 1], v[e + 2], v[e + 3]};
    __builtin_nontemporal_store(__builtin_convertvector(x, bf16x4), reinterpret_cast<bf16x4*>(row.at(off + e)));
  }
}

// Round the accumulators through bf16 (the values the next layer actually sees).
template <int NBo>
__device__ __forceinline__ void round_acc_bf16(floatx16 (&acc)[NBo]) {
#pragma unroll
  for (int nb = 0; nb < NBo; ++nb)
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[nb][i] = (float)(__bf16)acc[nb][i];
}
