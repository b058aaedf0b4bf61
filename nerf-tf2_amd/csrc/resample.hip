// Inverse-transform fine sampling: create_input_batch_fine_model
// (nerf/utils/ray_utils.py:276-406).  One wave64 per ray.
//
// Bit-exact with the oracle: every float op is an explicitly rounded
// intrinsic and the summation orders are fixed —
//   denom = reduce_sum(w' * width)  in the order of an 8-wide Eigen packet
//           reduction (lane j sums bins j, j+8, ...; then
//           ((p0+p4)+(p2+p6)) + ((p1+p5)+(p3+p7)), plus the n mod 8 tail);
//   agg   = cumsum(pdf * width)      sequential, each lane recomputing its own
//           prefix in left-to-right order.
// searchsorted(side='right') is an upper_bound binary search; the final
// tf.sort(concat(t_coarse, t_fine)) is a rank merge (coarse values first on
// ties), so equal values land in equal slots and the output is the sorted set.
//
// Compensated mode (NERF_CDF_COMPENSATED, the BF16 and learned forms): the
// products w' * width are split exactly into (hi, lo) pairs (fma residual),
// prefix-summed as compensated pairs (TwoSum), denom is the rounded total and
// cdf_k = S_k / denom with the division's remainder carried (fma), so every
// cdf value is within about one ulp of the exact prefix ratio instead of
// accumulating one rounding per bin.  pdf = w' / denom as in the reference.
#include "common.h"
#include <type_traits>

namespace {

constexpr int kWaves = 4;
constexpr int kMaxNc = 64;
constexpr int kMaxNf = 256;

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Compensated inclusive prefix of (ph, pl) pairs over lanes 0..63 (lanes past
// the bins hold zeros): Hillis-Steele steps with compensated pair adds.
__device__ __forceinline__ void comp_prefix(float& sh, float& sl, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const float oh = shfl_up(sh, d), ol = shfl_up(sl, d);
    if (lane >= d) dd_add(oh, ol, sh, sl, sh, sl);
  }
}

// (sh + sl) / denom rounded once up to the division's own error: the exact
// remainder of the first quotient is carried into a correction term.
__device__ __forceinline__ float comp_ratio(float sh, float sl, float denom) {
  const float q = fdiv(sh, denom);
  const float r = __builtin_fmaf(-q, denom, sh);
  return fadd(q, fdiv(fadd(r, sl), denom));
}

__global__ __launch_bounds__(256) void fine_kernel(
    int64_t n_rays, int nc, int nf, const float* __restrict__ ro, const float* __restrict__ rd,
    const float* __restrict__ bin_weights, const float* __restrict__ bin_edges,
    const float* __restrict__ bin_widths, const float* __restrict__ t_c,
    const float* __restrict__ u_vals, float* __restrict__ t_out, int32_t* __restrict__ idx_out,
    float* __restrict__ tf_out, float* __restrict__ xyz, float* __restrict__ dirs, int comp,
    float* __restrict__ cdf_out) {
  __shared__ float s_agg[kWaves][kMaxNc + 1];
  __shared__ float s_pdf[kWaves][kMaxNc];
  __shared__ float s_left[kWaves][kMaxNc];
  __shared__ float s_tc[kWaves][kMaxNc];
  __shared__ float s_tf[kWaves][kMaxNf];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int64_t ray = (int64_t)blockIdx.x * kWaves + w;
  if (ray >= n_rays) return;  // wave-uniform; no block-level barriers below
  const int S = nc + nf;

  // ---- pdf and cdf over the coarse bins (ray_utils.py:335-345)
  const bool on = lane < nc;
  const float wv = on ? bin_weights[ray * nc + lane] : 0.f;
  const float width = on ? bin_widths[ray * nc + lane] : 0.f;
  const float left = on ? bin_edges[ray * (nc + 1) + lane] : 0.f;
  const float wp = fadd(wv, 1e-5f);      // bin_weights + 1e-5
  const float prod = fmul(wp, width);    // bin_weights * bin_widths
  float pdf, agg;
  if (comp) {
    float sh = on ? prod : 0.f, sl = on ? __builtin_fmaf(wp, width, -prod) : 0.f;
    comp_prefix(sh, sl, lane);
    const float denom = shfl(sh, nc - 1);
    pdf = fdiv(wp, denom);
    agg = comp_ratio(sh, sl, denom);
  } else {
    const int nv = (nc / 8) * 8;
    float p = 0.f;
    for (int i = (lane & 7); i < nv; i += 8) p = fadd(p, shfl(prod, i));
    // every lane needs the same shuffles: loop bound is lane-independent
    float tail = 0.f;
    for (int i = nv; i < nc; ++i) tail = fadd(tail, shfl(prod, i));
    const float P0 = shfl(p, 0), P1 = shfl(p, 1), P2 = shfl(p, 2), P3 = shfl(p, 3);
    const float P4 = shfl(p, 4), P5 = shfl(p, 5), P6 = shfl(p, 6), P7 = shfl(p, 7);
    const float denom =
        fadd(tail, fadd(fadd(fadd(P0, P4), fadd(P2, P6)), fadd(fadd(P1, P5), fadd(P3, P7))));
    pdf = fdiv(wp, denom);               // ray_utils.py:339
    const float cw = fmul(pdf, width);   // pdf * bin_widths
    agg = 0.f;
    for (int k = 0; k < nc; ++k) {       // sequential inclusive cumsum (ray_utils.py:342)
      const float v = shfl(cw, k);
      if (k <= lane) agg = fadd(agg, v);
    }
  }
  if (cdf_out) {
    if (lane == 0) cdf_out[ray * (nc + 1)] = 0.f;
    if (on) cdf_out[ray * (nc + 1) + 1 + lane] = agg;
  }
  if (lane == 0) s_agg[w][0] = 0.f;      // concat 0 (ray_utils.py:345)
  if (on) {
    s_agg[w][lane + 1] = agg;
    s_pdf[w][lane] = pdf;
    s_left[w][lane] = left;
    s_tc[w][lane] = t_c[ray * nc + lane];
  }
  wave_sync();

  // ---- inverse transform (ray_utils.py:355-377)
  for (int q = lane; q < nf; q += 64) {
    const float u = u_vals[ray * nf + q];
    // searchsorted(agg[:, 1:-1], u, side='right'): first inner edge > u
    int lo = 0, hi = nc - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (s_agg[w][1 + mid] <= u)
        lo = mid + 1;
      else
        hi = mid;
    }
    const int idx = lo;
    const float agg_ = s_agg[w][idx];
    float pdf_ = s_pdf[w][idx];
    const float left_ = s_left[w][idx];
    const float mask = pdf_ < 1e-8f ? 0.f : 1.f;   // ray_utils.py:372
    pdf_ = fmaxf(pdf_, 1e-8f);                      // ray_utils.py:373
    const float tf = fadd(fmul(fdiv(fsub(u, agg_), pdf_), mask), left_);  // :376
    s_tf[w][q] = tf;
    if (idx_out) idx_out[ray * nf + q] = idx;
    if (tf_out) tf_out[ray * nf + q] = tf;
  }
  wave_sync();

  // ---- sort(concat(t_coarse, t_fine)) as a rank merge (ray_utils.py:380-385)
  float o3[3] = {0.f, 0.f, 0.f}, d3[3] = {0.f, 0.f, 0.f};
  if (xyz) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      o3[c] = ro[3 * ray + c];
      d3[c] = rd[3 * ray + c];
    }
  }
  float* trow = t_out + ray * S;
  auto emit = [&](int pos, float t) {
    trow[pos] = t;
    if (xyz) {  // o + t * d                                  (ray_utils.py:389-398)
      const int64_t m = ray * S + pos;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        xyz[3 * m + c] = fadd(o3[c], fmul(t, d3[c]));
        dirs[3 * m + c] = d3[c];
      }
    }
  };
  for (int q = lane; q < nf; q += 64) {
    const float v = s_tf[w][q];
    int rank = 0;
    for (int k = 0; k < nf; ++k) {
      const float o = s_tf[w][k];
      rank += (o < v) || (o == v && k < q);
    }
    int nle = 0;  // coarse values <= v (coarse first on ties)
    for (int k = 0; k < nc; ++k) nle += (s_tc[w][k] <= v);
    emit(rank + nle, v);
  }
  for (int j = lane; j < nc; j += 64) {
    const float v = s_tc[w][j];
    int nlt = 0;
    for (int k = 0; k < nf; ++k) nlt += (s_tf[w][k] < v);
    emit(j + nlt, v);
  }
}

// Value of `v` in lane `k` (a compile-time lane after unrolling): v_readlane
// into a scalar register, no LDS round trip.
__device__ __forceinline__ float lane_value(float v, int k) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), k));
}

// Lane i - 1's value in lane i, 0 in lane 0 (DPP wave_shr:1, bound_ctrl).
__device__ __forceinline__ float lane_shift_up(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xF, 0xF, true));
}

// Lane i ^ J's value (J < 64): DPP quad_perm for 1 and 2, row_ror:8 for 8,
// ds_swizzle (bit mode, within 32 lanes) for 4 and 16, ds_bpermute for 32.
template <int J>
__device__ __forceinline__ float lane_xor(float v) {
  const int x = __builtin_bit_cast(int, v);
  int r;
  if constexpr (J == 1) r = __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false);
  else if constexpr (J == 2) r = __builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false);
  else if constexpr (J == 8) r = __builtin_amdgcn_mov_dpp(x, 0x128, 0xF, 0xF, false);
  else if constexpr (J == 4 || J == 16) r = __builtin_amdgcn_ds_swizzle(x, 0x1F | (J << 10));
  else r = __builtin_bit_cast(int, shfl_xor(v, J));
  return __builtin_bit_cast(float, r);
}

// Count of the first n entries of the ascending row a that are <= x (LE) or
// < x (!LE): a branch-free binary search with a fixed trip count.
template <int N, bool LE>
__device__ __forceinline__ int count_below(const float* a, float x) {
  int c = 0;
#pragma unroll
  for (int step = 1 << (31 - __builtin_clz(N)); step > 0; step >>= 1) {
    if (c + step <= N) {
      const float e = a[c + step - 1];
      if (LE ? (e <= x) : (e < x)) c += step;
    }
  }
  return c;
}

// The same computation for the shapes of the reference configuration and the
// learned proposal (n_fine = 128, n_coarse <= 64, each a multiple of 8 or with
// its Eigen tail), organised for throughput: every cross-lane read of the pdf /
// cdf passes is a v_readlane of an unrolled, compile-time lane; the searchsorted
// is a fixed-trip binary search; the merge sorts the 128 fine values with a
// bitonic network in the wave (2 per lane), places each value by rank (its sorted
// index + the number of coarse values <= it; a coarse value: its index + the
// number of fine values < it — coarse first on ties, as the rank merge above), and
// writes the row from LDS with coalesced stores.  Equal floats are
// indistinguishable, so the output equals the sorted set bit for bit.
template <int NC, bool COMP>
__global__ __launch_bounds__(256) void fine128_kernel(
    int64_t n_rays, const float* __restrict__ ro, const float* __restrict__ rd,
    const float* __restrict__ bin_weights, const float* __restrict__ bin_edges,
    const float* __restrict__ bin_widths, const float* __restrict__ t_c,
    const float* __restrict__ u_vals, float* __restrict__ t_out, int32_t* __restrict__ idx_out,
    float* __restrict__ tf_out, float* __restrict__ xyz, float* __restrict__ dirs,
    float* __restrict__ cdf_out) {
  constexpr int NF = 128, S = NC + NF;
  static_assert(NC >= 2 && NC <= 64, "one coarse bin per lane");
  __shared__ float s_agg[kWaves][NC + 1];
  __shared__ float s_pdf[kWaves][NC];
  __shared__ float s_left[kWaves][NC];
  __shared__ float s_tc[kWaves][NC];
  __shared__ float s_sorted[kWaves][NF];
  __shared__ float s_row[kWaves][S];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int64_t ray = (int64_t)blockIdx.x * kWaves + w;
  if (ray >= n_rays) return;  // wave-uniform; no block-level barriers below

  // ---- pdf and cdf over the coarse bins (ray_utils.py:335-345)
  const bool on = lane < NC;
  const float wv = on ? bin_weights[ray * NC + lane] : 0.f;
  const float width = on ? bin_widths[ray * NC + lane] : 0.f;
  const float left = on ? bin_edges[ray * (NC + 1) + lane] : 0.f;
  const float tc = on ? t_c[ray * NC + lane] : 0.f;
  const float wp = fadd(wv, 1e-5f);
  const float prod = fmul(wp, width);
  float pdf, agg;
  if constexpr (COMP) {
    float sh = on ? prod : 0.f, sl = on ? __builtin_fmaf(wp, width, -prod) : 0.f;
    comp_prefix(sh, sl, lane);
    const float denom = lane_value(sh, NC - 1);
    pdf = fdiv(wp, denom);
    agg = comp_ratio(sh, sl, denom);
  } else {
    constexpr int NV = (NC / 8) * 8;
    // lane j < 8: packet sum P_j = prod_j + prod_{j+8} + ... in that order
    float pj = prod;
#pragma unroll
    for (int i = 1; 8 * i < NV; ++i) pj = fadd(pj, shfl_down(prod, 8 * i));
    float P[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) P[j] = lane_value(pj, j);
    float tail = 0.f;
#pragma unroll
    for (int i = NV; i < NC; ++i) tail = fadd(tail, lane_value(prod, i));
    const float denom =
        fadd(tail, fadd(fadd(fadd(P[0], P[4]), fadd(P[2], P[6])), fadd(fadd(P[1], P[5]), fadd(P[3], P[7]))));
    pdf = fdiv(wp, denom);
    const float cw = fmul(pdf, width);
    // sequential inclusive prefix: after step s lane i holds cw_{i-s} + ... + cw_i added
    // left to right (lane -1 reads 0), so after NC - 1 steps every lane < NC holds
    // ((cw_0 + cw_1) + ...) + cw_i exactly
    agg = cw;
#pragma unroll
    for (int step = 1; step < NC; ++step) agg = fadd(lane_shift_up(agg), cw);
  }
  if (cdf_out) {
    if (lane == 0) cdf_out[ray * (NC + 1)] = 0.f;
    if (on) cdf_out[ray * (NC + 1) + 1 + lane] = agg;
  }
  if (lane == 0) s_agg[w][0] = 0.f;
  if (on) {
    s_agg[w][lane + 1] = agg;
    s_pdf[w][lane] = pdf;
    s_left[w][lane] = left;
    s_tc[w][lane] = tc;
  }
  wave_sync();

  // ---- inverse transform (ray_utils.py:355-377), fine samples lane and lane + 64
  float v[2];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int q = lane + 64 * r;
    const float u = u_vals[ray * NF + q];
    const int idx = count_below<NC - 1, true>(&s_agg[w][1], u);  // searchsorted(agg[1:-1], u, 'right')
    const float agg_ = s_agg[w][idx];
    float pdf_ = s_pdf[w][idx];
    const float left_ = s_left[w][idx];
    const float mask = pdf_ < 1e-8f ? 0.f : 1.f;
    pdf_ = fmaxf(pdf_, 1e-8f);
    v[r] = fadd(fmul(fdiv(fsub(u, agg_), pdf_), mask), left_);
    if (idx_out) idx_out[ray * NF + q] = idx;
    if (tf_out) tf_out[ray * NF + q] = v[r];
  }

  // ---- bitonic sort of the 128 fine values; element e = lane + 64 r
  auto stage = [&](auto kc, auto jc) {
    constexpr int k = decltype(kc)::value, j = decltype(jc)::value;
    if constexpr (j == 64) {  // partner in the same lane (k == 128: ascending)
      const float a = v[0], b = v[1];
      v[0] = fminf(a, b);
      v[1] = fmaxf(a, b);
    } else {
      const bool lower = (lane & j) == 0;
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const int e = lane + 64 * r;
        const bool asc = (e & k) == 0;
        const float o = lane_xor<j>(v[r]);
        v[r] = (lower == asc) ? fminf(v[r], o) : fmaxf(v[r], o);
      }
    }
  };
  auto merge = [&](auto kc) {
    constexpr int k = decltype(kc)::value;
    if constexpr (k >= 128) stage(kc, std::integral_constant<int, 64>{});
    if constexpr (k >= 64) stage(kc, std::integral_constant<int, 32>{});
    if constexpr (k >= 32) stage(kc, std::integral_constant<int, 16>{});
    if constexpr (k >= 16) stage(kc, std::integral_constant<int, 8>{});
    if constexpr (k >= 8) stage(kc, std::integral_constant<int, 4>{});
    if constexpr (k >= 4) stage(kc, std::integral_constant<int, 2>{});
    stage(kc, std::integral_constant<int, 1>{});
  };
  merge(std::integral_constant<int, 2>{});
  merge(std::integral_constant<int, 4>{});
  merge(std::integral_constant<int, 8>{});
  merge(std::integral_constant<int, 16>{});
  merge(std::integral_constant<int, 32>{});
  merge(std::integral_constant<int, 64>{});
  merge(std::integral_constant<int, 128>{});
  s_sorted[w][lane] = v[0];
  s_sorted[w][lane + 64] = v[1];
  wave_sync();

  // ---- ranks -> row in LDS -> coalesced stores (ray_utils.py:380-398)
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int e = lane + 64 * r;
    s_row[w][e + count_below<NC, true>(&s_tc[w][0], v[r])] = v[r];
  }
  if (on) s_row[w][lane + count_below<NF, false>(&s_sorted[w][0], tc)] = tc;
  wave_sync();
  float o3[3] = {0.f, 0.f, 0.f}, d3[3] = {0.f, 0.f, 0.f};
  if (xyz) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      o3[c] = ro[3 * ray + c];
      d3[c] = rd[3 * ray + c];
    }
  }
  float* trow = t_out + ray * S;
  for (int i = lane; i < S; i += 64) {
    const float t = s_row[w][i];
    trow[i] = t;
    if (xyz) {
      const int64_t m = ray * S + i;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        xyz[3 * m + c] = fadd(o3[c], fmul(t, d3[c]));
        dirs[3 * m + c] = d3[c];
      }
    }
  }
}

}  // namespace

extern "C" int32_t nerf_create_input_batch_fine(int64_t n_rays, int32_t n_coarse, int32_t n_fine,
                                                const float* rays_o, const float* rays_d,
                                                const float* bin_weights, const float* bin_edges,
                                                const float* bin_widths,
                                                const float* t_vals_coarse, const float* u_vals,
                                                float* t_vals, int32_t* piece_idxs,
                                                float* t_vals_fine, float* xyz_inputs,
                                                float* dir_inputs, int32_t cdf_mode, float* cdf,
                                                void* stream) {
  NERF_CHECK_ARG(n_rays >= 0 && n_coarse >= 2 && n_coarse <= kMaxNc && n_fine >= 1 &&
                     n_fine <= kMaxNf,
                 "nerf_create_input_batch_fine: n_coarse=%d (2..%d) n_fine=%d (1..%d)", n_coarse,
                 kMaxNc, n_fine, kMaxNf);
  NERF_CHECK_ARG(cdf_mode == NERF_CDF_TF_ORDER || cdf_mode == NERF_CDF_COMPENSATED,
                 "nerf_create_input_batch_fine: cdf_mode=%d", cdf_mode);
  if (n_rays == 0) return NERF_OK;  // empty batch
  NERF_CHECK_ARG(bin_weights && bin_edges && bin_widths && t_vals_coarse && u_vals && t_vals,
                 "nerf_create_input_batch_fine: null pointer");
  NERF_CHECK_ARG((xyz_inputs == nullptr) == (dir_inputs == nullptr),
                 "nerf_create_input_batch_fine: xyz_inputs and dir_inputs go together");
  NERF_CHECK_ARG(!xyz_inputs || (rays_o && rays_d),
                 "nerf_create_input_batch_fine: xyz_inputs needs rays");
  const unsigned grid = (unsigned)((n_rays + kWaves - 1) / kWaves);
  hipStream_t s = as_stream(stream);
  const bool comp = cdf_mode == NERF_CDF_COMPENSATED;
#define NERF_FINE128(NC_, C_)                                                                    \
  fine128_kernel<NC_, C_><<<grid, 64 * kWaves, 0, s>>>(n_rays, rays_o, rays_d, bin_weights,      \
                                                      bin_edges, bin_widths, t_vals_coarse,      \
                                                      u_vals, t_vals, piece_idxs, t_vals_fine,   \
                                                      xyz_inputs, dir_inputs, cdf)
  if (n_fine == 128 && n_coarse == 64) {
    if (comp) NERF_FINE128(64, true); else NERF_FINE128(64, false);
  } else if (n_fine == 128 && n_coarse == 48) {
    if (comp) NERF_FINE128(48, true); else NERF_FINE128(48, false);
  } else {
    fine_kernel<<<grid, 64 * kWaves, 0, s>>>(
        n_rays, n_coarse, n_fine, rays_o, rays_d, bin_weights, bin_edges, bin_widths, t_vals_coarse,
        u_vals, t_vals, piece_idxs, t_vals_fine, xyz_inputs, dir_inputs, comp ? 1 : 0, cdf);
  }
#undef NERF_FINE128
  NERF_LAUNCH_CHECK("nerf_create_input_batch_fine");
  return NERF_OK;
}
