// Layout plan of the fused NeRF MLP: the flat "Keras" parameter layout, the
// packed MFMA operand layout, the training stash and the per-sample gradients.
//
// Register layout (v_mfma_f32_32x32x{2f32,16_bf16}): one wave owns 32 samples.
// Lane l = j + 32 g (j = sample in the block, g = lane half).  An MFMA output
// block nb (32 features) holds in accumulator i of lane l the feature
//   F(nb, i, g) = 32 nb + 8 (i/4) + 4 g + (i%4)
// of sample j.  Those accumulators are fed back as the B operand of the next
// layer without any transpose: fp32 k-step s = 16 nb + i uses F(nb, i, g) on
// lane half g; bf16 k-step t uses slots r = 0..7 of F(t/2, 8 (t%2) + r, g).
// The packed A operand (weights) is permuted to match, so the k order of each
// dot product is the only thing that differs from a plain matmul.
//
// Positional-encoding features are produced directly in lane slots: half g
// holds slots e = 0..H-1 (H = 2 + 3L).  Half 0: x, y, z, pad, then (sin, cos)
// pairs 0..P0-1; half 1: pairs P0..3L-1 (P0 = (3L-2)/2).  Pair p encodes
// component c = p / L at frequency l = p % L, i.e. reference features
// 3 + c 2L + 2l + {0: sin, 1: cos} (model.py:323-330).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define NERF_MAX_LAYERS 16

__host__ __device__ inline int ru(int x, int m) { return (x + m - 1) / m * m; }

// Reference feature index of encoding slot e on lane half g, or -1 for a pad slot.
__host__ __device__ inline int enc_feat(int g, int e, int L) {
  const int H = 2 + 3 * L, P0 = (3 * L - 2) / 2;
  if (e < 0 || e >= H) return -1;
  int p, sc;
  if (g == 0) {
    if (e < 3) return e;
    if (e == 3) return -1;
    p = (e - 4) / 2;
    sc = (e - 4) % 2;
  } else {
    p = P0 + e / 2;
    sc = e % 2;
  }
  const int c = p / L, l = p % L;
  return 3 + c * 2 * L + 2 * l + sc;
}

// Hidden feature held by accumulator i of output block nb on lane half g.
__host__ __device__ inline int cfeat(int nb, int i, int g) { return 32 * nb + 8 * (i / 4) + 4 * g + (i % 4); }

// Register-resident fp32 forward (v_mfma_f32_16x16x4f32): one wave owns 16
// samples, lane l = j + 16 q (j = sample, q = lane quarter).  Output block nb
// (16 features) holds in accumulator r of lane (j, q) the feature
//   G(nb, r, q) = 16 nb + 4 q + r
// which is also what k-step 4 nb + r of the next layer reads on quarter q, so
// a layer's output is the next layer's B operand in place.  An encoding of L
// frequencies takes S4(L) slots per quarter (a multiple of 4): quarter 0 holds
// x, y, z, pad, then (sin, cos) pairs 0 .. S4/2 - 3; quarter q >= 1 holds
// pairs S4/2 - 2 + (q - 1) S4/2 + e/2; pairs >= 3L are pads.
// Inverse of enc_feat: lane-half slot index g * K + e (K slots per half) of
// reference encoding feature f of an encoding with L frequencies.
__host__ __device__ inline int enc_slot_of(int f, int L, int K) {
  if (f < 3) return f;
  const int P0 = (3 * L - 2) / 2, p = (f - 3) / 2, sc = (f - 3) % 2;
  return p < P0 ? 4 + 2 * p + sc : K + 2 * (p - P0) + sc;
}

__host__ __device__ constexpr int enc_slots4(int L) { return ((3 * L + 3) / 2 + 3) / 4 * 4; }
__host__ __device__ inline int enc_feat4(int q, int e, int L) {
  const int S = enc_slots4(L), P0 = S / 2 - 2;
  if (e < 0 || e >= S) return -1;
  int p;
  if (q == 0) {
    if (e < 3) return e;
    if (e == 3) return -1;
    p = (e - 4) / 2;
  } else {
    p = P0 + (q - 1) * (S / 2) + e / 2;
  }
  if (p >= 3 * L) return -1;
  const int c = p / L, l = p % L;
  return 3 + c * 2 * L + 2 * l + e % 2;
}

struct MlpPlan {
  int D, W, skip, LX, LD, rgb;
  int NB, NBR, WR;      // hidden blocks, rgb-branch blocks, rgb-branch width
  int HX, HD;           // encoding slots per lane half
  int KX, KD;           // encoding slots per half padded to 8 (= fp32 k-steps)
  int KX16, KD16;       // bf16 k-steps for the encodings
  int FX, FD;           // encoding widths (3 + 6L)
  int n_keras;          // Keras layers in trainable-variable order
  int kl_sigma, kl_bott, kl_rbr, kl_rgb;  // their Keras indices (-1: absent)
  int fan_in[NERF_MAX_LAYERS], fan_out[NERF_MAX_LAYERS];
  int64_t k_off[NERF_MAX_LAYERS], b_off[NERF_MAX_LAYERS];
  int64_t n_params;
  // MFMA layers: m = 0..D-1 trunk, D bottleneck, D+1 rgb branch (index into arrays below)
  int n_mfma;
  int keras_of[NERF_MAX_LAYERS];   // Keras layer index of MFMA layer m
  int ks_hid[NERF_MAX_LAYERS];     // fp32 k-steps read from the hidden input (0 for m = 0)
  int ks_enc[NERF_MAX_LAYERS];     // fp32 k-steps from an encoding (KX / KD / 0)
  int enc_L[NERF_MAX_LAYERS];      // L of that encoding
  int nb_out[NERF_MAX_LAYERS];     // output blocks
  int relu[NERF_MAX_LAYERS];
  // packed byte offsets
  int64_t p_fwd[NERF_MAX_LAYERS], p_bwd[NERF_MAX_LAYERS], p_bias[NERF_MAX_LAYERS];
  int64_t p_head_sigma, p_head_rgb, packed_bytes;
  // fp32 only: A operands of the register-resident forward (per layer: the
  // encoding's k-step groups first, then the hidden input's; a group is 4
  // k-steps of every output block of 16), and both heads in natural feature
  // order (sigma: W weights + bias; rgb: [W/2][4] + bias[4]); -1 for bf16
  int64_t p_fwd4[NERF_MAX_LAYERS], p_hs_nat, p_hr_nat;
  // fp32 only: transposed A operands of the register-resident dX chain, layers
  // m >= 1 (k-step groups of 16 output features of layer m, each over the W/16
  // blocks of its hidden input; -1 otherwise)
  int64_t p_bwd4[NERF_MAX_LAYERS];
  int S4X, S4D;  // encoding slots per lane quarter in that forward
  int prec;
  // stash (elements per sample; fp32 or bf16 elements) and dz (floats per sample)
  int64_t st_encx, st_h[NERF_MAX_LAYERS], st_bott, st_rbr, st_encd, stash_per_sample;
  // ReLU bit masks (W = 256 with an rgb branch; -1 otherwise), written by the
  // forward for the dX chain.  bf16: per 32-feature block nb and lane half g, D
  // uint16 words, one per trunk layer (element (2 nb + g) D + l), 16 D elements
  // in all; then 8 elements for the rgb branch (one uint16 per block and lane
  // half).  fp32 (register-resident kernels, lane quarter q holding features
  // 16 nb + 4 q + r): per trunk layer l and quarter q two uint32 words (element
  // 8 l + 2 q + h, bit 4 (nb - 8 h) + r for nb = 8 h .. 8 h + 7), 8 D elements;
  // then one word per quarter for the rgb branch's 8 blocks (element 8 D + q)
  int64_t st_mask;
  int64_t dz_h[NERF_MAX_LAYERS], dz_bott, dz_rbr, dz_sigma, dz_rgb, dz_per_sample;
  // Samples per block of the stash and dz buffers.  A buffer of M samples with S
  // elements per sample is cut into blocks of sblk consecutive samples (the last
  // one may hold fewer); a block [b, b + n) stores its rows chunk-major: the
  // 8-element chunk c of sample m at element b S + 8 (c n + m - b).  sblk = 1 is
  // the plain row-major layout (fp32); 32 (one MFMA sample block) for bf16, whose
  // row chunks of 32 consecutive samples are then 512 contiguous bytes.
  int sblk;
};

// Element offset, in a blocked stash or dz buffer (MlpPlan::sblk = B, S elements
// per sample, M samples), of element e of sample m.
__host__ __device__ inline int64_t blk_elem(int64_t m, int64_t e, int64_t S, int B, int64_t M) {
  const int64_t b = m - m % B;
  const int64_t n = M - b < B ? M - b : B;
  return b * S + 8 * ((e >> 3) * n + (m - b)) + (e & 7);
}

// Returns 0 on success, -1 if the description is not supported.
inline int mlp_make_plan(int D, int W, int skip, int LX, int LD, int prec, MlpPlan* p) {
  // skip < D-1: the sigma head and the bottleneck read the last trunk layer only
  if (D < 2 || D > 10 || W < 32 || W > 256 || (W % 32) || skip < -1 || skip >= D - 1 || LX < 2 ||
      (LX % 2) || LX > 16 || LD < 0 || (LD % 2) || LD > 16 || (prec != 0 && prec != 1))
    return -1;
  MlpPlan& P = *p;
  P = MlpPlan{};
  P.D = D; P.W = W; P.skip = skip; P.LX = LX; P.LD = LD; P.rgb = LD > 0; P.prec = prec;
  // every bf16 stash and dz is blocked (the kernels' launchers check it)
  P.sblk = prec == 1 ? 32 : 1;
  P.NB = W / 32;
  P.WR = W / 2;
  P.NBR = (P.WR + 31) / 32;
  P.HX = 2 + 3 * LX;
  P.HD = LD ? 2 + 3 * LD : 0;
  P.KX = ru(P.HX, 8);
  P.KD = LD ? ru(P.HD, 8) : 0;
  P.KX16 = P.KX / 8;
  P.KD16 = P.KD / 8;
  P.FX = 3 + 6 * LX;
  P.FD = LD ? 3 + 6 * LD : 0;
  // Keras layers in trainable-variable order.  A functional model lists its
  // layers by decreasing depth from the outputs, ties in the order a depth-first
  // walk from the outputs [rgb, sigma] first meets them (model.py:389-392).  The
  // last trunk layer feeds the bottleneck (depth 3 below rgb), so it sits at
  // depth 4; sigma and rgb are both outputs (depth 0) and the walk meets rgb
  // first: dense_0 .. dense_{D-1}, dense_D, dense_{D+1}, rgb, sigma.  A
  // density-only net is dense_0 .. dense_{D-1}, sigma.
  int n = 0;
  const int trunk_out_in = W + (skip == D - 1 ? P.FX : 0);
  for (int i = 0; i < D; ++i) {
    P.fan_in[n] = i == 0 ? P.FX : W + (i - 1 == skip ? P.FX : 0);
    P.fan_out[n] = W;
    ++n;
  }
  P.kl_bott = P.kl_rbr = P.kl_rgb = -1;
  if (P.rgb) {
    P.kl_bott = n; P.fan_in[n] = trunk_out_in; P.fan_out[n] = W; ++n;   // dense_D (bottleneck)
    P.kl_rbr = n; P.fan_in[n] = W + P.FD; P.fan_out[n] = P.WR; ++n;     // dense_{D+1}
    P.kl_rgb = n; P.fan_in[n] = P.WR; P.fan_out[n] = 3; ++n;            // rgb
  }
  P.kl_sigma = n; P.fan_in[n] = trunk_out_in; P.fan_out[n] = 1; ++n;  // sigma
  P.n_keras = n;
  int64_t off = 0;
  for (int i = 0; i < n; ++i) {
    P.k_off[i] = off;
    off += (int64_t)P.fan_in[i] * P.fan_out[i];
    P.b_off[i] = off;
    off += P.fan_out[i];
  }
  P.n_params = off;
  // MFMA layers
  int m = 0;
  for (int i = 0; i < D; ++i) {
    P.keras_of[m] = i;
    P.ks_hid[m] = i == 0 ? 0 : W / 2;
    P.ks_enc[m] = i == 0 ? P.KX : (i - 1 == skip ? P.KX : 0);
    P.enc_L[m] = LX;
    P.nb_out[m] = P.NB;
    P.relu[m] = 1;
    ++m;
  }
  if (P.rgb) {
    P.keras_of[m] = P.kl_bott;
    P.ks_hid[m] = W / 2;
    P.ks_enc[m] = skip == D - 1 ? P.KX : 0;
    P.enc_L[m] = LX;
    P.nb_out[m] = P.NB;
    P.relu[m] = 0;
    ++m;
    P.keras_of[m] = P.kl_rbr;
    P.ks_hid[m] = W / 2;
    P.ks_enc[m] = P.KD;
    P.enc_L[m] = LD;
    P.nb_out[m] = P.NBR;
    P.relu[m] = 1;
    ++m;
  }
  P.n_mfma = m;
  // packed layout
  int64_t b = 0;
  auto align = [&](int64_t a) { b = (b + a - 1) / a * a; };
  for (int i = 0; i < P.n_mfma; ++i) {
    align(256);
    P.p_fwd[i] = b;
    if (prec == 0)
      b += (int64_t)(P.ks_hid[i] + P.ks_enc[i]) * P.nb_out[i] * 64 * 4;
    else {
      const int ke16 = P.ks_enc[i] == 0 ? 0 : (P.enc_L[i] == LX ? P.KX16 : P.KD16);
      b += (int64_t)(P.ks_hid[i] / 8 + ke16) * P.nb_out[i] * 64 * 16;
    }
  }
  for (int i = 0; i < P.n_mfma; ++i) {
    P.p_bwd[i] = -1;
    if (i == 0) continue;  // no dX for the encoding input
    align(256);
    P.p_bwd[i] = b;
    // A[n = input hidden feature][k = output feature]: k-steps over the outputs
    const int out_ks = P.nb_out[i] * 16;  // fp32 k-steps over 32*nb_out outputs
    if (prec == 0)
      b += (int64_t)out_ks * P.NB * 64 * 4;
    else
      b += (int64_t)(out_ks / 8) * P.NB * 64 * 16;
  }
  for (int i = 0; i < P.n_mfma; ++i) {
    align(256);
    P.p_bias[i] = b;
    b += (int64_t)P.nb_out[i] * 32 * 4;
  }
  align(256);
  P.p_head_sigma = b;  // [2][W/2] fp32 + bias (padded to 4)
  b += (int64_t)2 * (W / 2) * 4 + 16;
  align(256);
  P.p_head_rgb = b;  // [2][16*NBR][4] fp32 + bias[4]
  if (P.rgb) b += (int64_t)2 * 16 * P.NBR * 16 + 16;
  P.S4X = enc_slots4(LX);
  P.S4D = LD ? enc_slots4(LD) : 0;
  P.p_hs_nat = P.p_hr_nat = -1;
  for (int i = 0; i < P.n_mfma; ++i) {
    P.p_fwd4[i] = -1;
    if (prec != 0) continue;
    align(256);
    P.p_fwd4[i] = b;
    const int ke4 = P.ks_enc[i] == 0 ? 0 : (P.enc_L[i] == LX ? P.S4X : P.S4D);
    const int kh4 = P.ks_hid[i] == 0 ? 0 : W / 4;
    b += (int64_t)(ke4 + kh4) * (2 * P.nb_out[i]) * 64 * 4;
  }
  for (int i = 0; i < P.n_mfma; ++i) {
    P.p_bwd4[i] = -1;
    if (prec != 0 || i == 0) continue;
    align(256);
    P.p_bwd4[i] = b;
    b += (int64_t)(2 * P.nb_out[i]) * (W / 16) * 64 * 4 * 4;
  }
  if (prec == 0) {
    align(256);
    P.p_hs_nat = b;
    b += (int64_t)W * 4 + 16;
    align(256);
    P.p_hr_nat = b;
    if (P.rgb) b += (int64_t)P.WR * 16 + 16;
  }
  align(256);
  P.packed_bytes = b;
  // stash (elements per sample)
  int64_t s = 0;
  P.st_encx = s; s += 2 * P.KX;
  for (int i = 0; i < D; ++i) { P.st_h[i] = s; s += W; }
  P.st_bott = P.st_rbr = P.st_encd = -1;
  if (P.rgb) {
    P.st_bott = s; s += W;
    P.st_rbr = s; s += 32 * P.NBR;
    P.st_encd = s; s += 2 * P.KD;
  }
  s = ru((int)s, 8);
  P.st_mask = -1;
  if (W == 256 && P.rgb && P.NBR == 4) {
    P.st_mask = s;
    s += 16 * D + 8;
  }
  // rows padded to 64 elements: every stash row, and each 64-feature segment a
  // wave writes, starts on a 128-byte line (bf16), so the forward writes whole lines
  P.stash_per_sample = ru((int)s, 64);
  // dz (elements per sample: float32 for the fp32 form, bf16 for the bf16 form);
  // every segment starts on a multiple of 8 elements (16-byte aligned bf16 loads)
  int64_t z = 0;
  for (int i = 0; i < D; ++i) { P.dz_h[i] = z; z += W; }
  P.dz_bott = P.dz_rbr = P.dz_rgb = -1;
  if (P.rgb) {
    P.dz_bott = z; z += W;
    P.dz_rbr = z; z += 32 * P.NBR;
  }
  P.dz_sigma = z; z += 8;
  if (P.rgb) { P.dz_rgb = z; z += 8; }
  P.dz_per_sample = ru((int)z, 64);  // 128-byte rows (bf16), as the stash
  return 0;
}
