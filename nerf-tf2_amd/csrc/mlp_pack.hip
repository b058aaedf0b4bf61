// Packing of the flat "Keras" MLP parameters into the MFMA operand layout,
// and the size queries of the MLP entry points.
#include "common.h"
#include "mlp_plan.h"

namespace {

// Input row (of the Keras kernel) fed by fp32 k-step s on lane half g of MFMA layer m.
__device__ int fwd_row_f32(const MlpPlan& P, int m, int s, int g) {
  const int kh = P.ks_hid[m];
  if (s < kh) return cfeat(s / 16, s % 16, g);
  const int e = s - kh;
  const int ef = enc_feat(g, e, P.enc_L[m]);
  if (ef < 0) return -1;
  return (m == 0 ? 0 : P.W) + ef;
}

// Same for bf16 k-step t, slot r (0..7).
__device__ int fwd_row_bf16(const MlpPlan& P, int m, int t, int r, int g) {
  const int th = P.ks_hid[m] / 8;
  if (t < th) return cfeat(t / 2, 8 * (t % 2) + r, g);
  const int e = 8 * (t - th) + r;
  const int ef = enc_feat(g, e, P.enc_L[m]);
  if (ef < 0) return -1;
  return (m == 0 ? 0 : P.W) + ef;
}

// Input row fed by k-step s (0 .. ke4 + kh4 - 1: the encoding's first) on lane
// quarter q of MFMA layer m in the register-resident fp32 forward.
__device__ int fwd_row_f32x4(const MlpPlan& P, int m, int s, int q) {
  const int ke4 = P.ks_enc[m] == 0 ? 0 : (P.enc_L[m] == P.LX ? P.S4X : P.S4D);
  if (s < ke4) {
    const int ef = enc_feat4(q, s, P.enc_L[m]);
    return ef < 0 ? -1 : (m == 0 ? 0 : P.W) + ef;
  }
  const int t = s - ke4;  // hidden k-step 4 nb + r reads G(nb, r, q)
  return 16 * (t / 4) + 4 * q + t % 4;
}

__device__ float kernel_at(const MlpPlan& P, const float* flat, int kl, int row, int col) {
  if (row < 0 || row >= P.fan_in[kl] || col < 0 || col >= P.fan_out[kl]) return 0.f;
  return flat[P.k_off[kl] + (int64_t)row * P.fan_out[kl] + col];
}

// One launch packs everything: blockIdx.y selects the segment.
//   y in [0, n_mfma)            forward A operands of MFMA layer y
//   y in [n_mfma, 2 n_mfma)     backward (transposed) A operands
//   y in [2 n_mfma, 3 n_mfma)   biases
//   y == 3 n_mfma               sigma head, y == 3 n_mfma + 1 rgb head
//   y in [3 n_mfma + 2, 4 n_mfma + 2)  register-resident fp32 forward A operands
//   y == 4 n_mfma + 2           both heads in natural feature order (fp32)
//   y in [4 n_mfma + 3, 5 n_mfma + 3)  register-resident fp32 dX chain A operands
__global__ void pack_kernel(MlpPlan P, const float* __restrict__ flat, char* __restrict__ out) {
  const int seg = blockIdx.y;
  const int nm = P.n_mfma;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nthr = (int64_t)gridDim.x * blockDim.x;
  if (seg >= 4 * nm + 3) {
    // e = ((grp * NBi + nb) * 64 + lane) * 4 + r: A[n = input feature 16 nb + i][k = output
    // feature 16 grp + 4 q + r] of layer m (lane = i + 16 q)
    const int m = seg - (4 * nm + 3);
    if (P.prec != 0 || m == 0) return;
    const int kl = P.keras_of[m], NBi = P.W / 16, G = 2 * P.nb_out[m];
    float* o = reinterpret_cast<float*>(out + P.p_bwd4[m]);
    const int64_t n = (int64_t)G * NBi * 64 * 4;
    for (int64_t e = tid; e < n; e += nthr) {
      const int r = (int)(e % 4), lane = (int)((e / 4) % 64);
      const int nb = (int)((e / 256) % NBi), grp = (int)(e / (256 * NBi));
      const int i = lane & 15, q = lane >> 4;
      o[e] = kernel_at(P, flat, kl, 16 * nb + i, 16 * grp + 4 * q + r);
    }
    return;
  }
  if (seg >= 3 * nm + 2) {
    if (P.prec != 0) return;
    if (seg < 4 * nm + 2) {
      // e = ((grp * NBo4 + nb) * 64 + lane) * 4 + r; k-step 4 grp + r; lane = i + 16 q
      const int m = seg - (3 * nm + 2), kl = P.keras_of[m], NBo4 = 2 * P.nb_out[m];
      const int ke4 = P.ks_enc[m] == 0 ? 0 : (P.enc_L[m] == P.LX ? P.S4X : P.S4D);
      const int kh4 = P.ks_hid[m] == 0 ? 0 : P.W / 4;
      float* o = reinterpret_cast<float*>(out + P.p_fwd4[m]);
      const int64_t n = (int64_t)(ke4 + kh4) * NBo4 * 64;
      for (int64_t e = tid; e < n; e += nthr) {
        const int r = (int)(e % 4), lane = (int)((e / 4) % 64);
        const int nb = (int)((e / 256) % NBo4), grp = (int)(e / (256 * NBo4));
        const int i = lane & 15, q = lane >> 4;
        o[e] = kernel_at(P, flat, kl, fwd_row_f32x4(P, m, 4 * grp + r, q), 16 * nb + i);
      }
      return;
    }
    float* hs = reinterpret_cast<float*>(out + P.p_hs_nat);
    for (int64_t e = tid; e < P.W + 4; e += nthr)
      hs[e] = e < P.W ? kernel_at(P, flat, P.kl_sigma, (int)e, 0) : (e == P.W ? flat[P.b_off[P.kl_sigma]] : 0.f);
    if (!P.rgb) return;
    float* hr = reinterpret_cast<float*>(out + P.p_hr_nat);
    for (int64_t e = tid; e < (int64_t)P.WR * 4 + 4; e += nthr) {
      const int c = (int)(e % 4);
      if (e < (int64_t)P.WR * 4)
        hr[e] = c < 3 ? kernel_at(P, flat, P.kl_rgb, (int)(e / 4), c) : 0.f;
      else
        hr[e] = c < 3 ? flat[P.b_off[P.kl_rgb] + c] : 0.f;
    }
    return;
  }
  if (seg < nm) {
    const int m = seg, kl = P.keras_of[m], NBo = P.nb_out[m];
    if (P.prec == 0) {
      const int KS = P.ks_hid[m] + P.ks_enc[m];
      float* o = reinterpret_cast<float*>(out + P.p_fwd[m]);
      const int64_t n = (int64_t)KS * NBo * 64;
      for (int64_t e = tid; e < n; e += nthr) {  // e = ((q*NBo + nb)*64 + lane)*4 + r
        const int r = (int)(e % 4), lane = (int)((e / 4) % 64);
        const int nb = (int)((e / 256) % NBo), q = (int)(e / (256 * NBo));
        const int s = 4 * q + r, j = lane & 31, g = lane >> 5;
        o[e] = kernel_at(P, flat, kl, fwd_row_f32(P, m, s, g), 32 * nb + j);
      }
    } else {
      const int ke16 = P.ks_enc[m] == 0 ? 0 : (P.enc_L[m] == P.LX ? P.KX16 : P.KD16);
      const int KS = P.ks_hid[m] / 8 + ke16;
      unsigned short* o = reinterpret_cast<unsigned short*>(out + P.p_fwd[m]);
      const int64_t n = (int64_t)KS * NBo * 64 * 8;
      for (int64_t e = tid; e < n; e += nthr) {  // e = ((t*NBo + nb)*64 + lane)*8 + r
        const int r = (int)(e % 8), lane = (int)((e / 8) % 64);
        const int nb = (int)((e / 512) % NBo), t = (int)(e / (512 * NBo));
        const int j = lane & 31, g = lane >> 5;
        o[e] = f2bf(kernel_at(P, flat, kl, fwd_row_bf16(P, m, t, r, g), 32 * nb + j));
      }
    }
  } else if (seg < 2 * nm) {
    const int m = seg - nm;
    if (m == 0) return;
    const int kl = P.keras_of[m], NBo = P.nb_out[m];
    if (P.prec == 0) {
      const int KS = NBo * 16;  // k-steps over the layer's outputs
      float* o = reinterpret_cast<float*>(out + P.p_bwd[m]);
      const int64_t n = (int64_t)KS * P.NB * 64;
      for (int64_t e = tid; e < n; e += nthr) {
        const int r = (int)(e % 4), lane = (int)((e / 4) % 64);
        const int nb = (int)((e / 256) % P.NB), q = (int)(e / (256 * P.NB));
        const int s = 4 * q + r, j = lane & 31, g = lane >> 5;
        // A[n = input hidden feature 32nb+j][k = output feature cfeat(s)]
        o[e] = kernel_at(P, flat, kl, 32 * nb + j, cfeat(s / 16, s % 16, g));
      }
    } else {
      const int KS = NBo * 2;
      unsigned short* o = reinterpret_cast<unsigned short*>(out + P.p_bwd[m]);
      const int64_t n = (int64_t)KS * P.NB * 64 * 8;
      for (int64_t e = tid; e < n; e += nthr) {
        const int r = (int)(e % 8), lane = (int)((e / 8) % 64);
        const int nb = (int)((e / 512) % P.NB), t = (int)(e / (512 * P.NB));
        const int j = lane & 31, g = lane >> 5;
        o[e] = f2bf(kernel_at(P, flat, kl, 32 * nb + j, cfeat(t / 2, 8 * (t % 2) + r, g)));
      }
    }
  } else if (seg < 3 * nm) {
    const int m = seg - 2 * nm, kl = P.keras_of[m];
    float* o = reinterpret_cast<float*>(out + P.p_bias[m]);
    const int n = 32 * P.nb_out[m];
    for (int64_t e = tid; e < n; e += nthr)
      o[e] = e < P.fan_out[kl] ? flat[P.b_off[kl] + e] : 0.f;
  } else if (seg == 3 * nm) {
    const int kl = P.kl_sigma;
    float* o = reinterpret_cast<float*>(out + P.p_head_sigma);
    const int KH = P.W / 2;
    for (int64_t e = tid; e < 2 * KH + 4; e += nthr) {
      if (e < 2 * KH) {
        const int g = (int)(e / KH), s = (int)(e % KH);
        o[e] = kernel_at(P, flat, kl, cfeat(s / 16, s % 16, g), 0);
      } else {
        o[e] = e == 2 * KH ? flat[P.b_off[kl]] : 0.f;
      }
    }
  } else if (P.rgb) {
    const int kl = P.kl_rgb;
    float* o = reinterpret_cast<float*>(out + P.p_head_rgb);
    const int KS = 16 * P.NBR;
    for (int64_t e = tid; e < 2 * KS * 4 + 4; e += nthr) {
      if (e < 2 * KS * 4) {
        const int c = (int)(e % 4), s = (int)((e / 4) % KS), g = (int)(e / (4 * KS));
        o[e] = c < 3 ? kernel_at(P, flat, kl, cfeat(s / 16, s % 16, g), c) : 0.f;
      } else {
        const int c = (int)(e - 2 * KS * 4);
        o[e] = c < 3 ? flat[P.b_off[kl] + c] : 0.f;
      }
    }
  }
}

}  // namespace

int nerf_plan_from_desc(const nerf_mlp_desc* d, int prec, MlpPlan* p) {
  if (!d) return -1;
  return mlp_make_plan(d->depth, d->width, d->skip, d->l_xyz, d->l_dir, prec, p);
}

#define NERF_PLAN_OR_FAIL(desc, prec, P, name)                                                 \
  MlpPlan P;                                                                                   \
  NERF_CHECK_ARG(nerf_plan_from_desc(desc, prec, &P) == 0,                                     \
                 "%s: unsupported MLP description or precision", name)

extern "C" int64_t nerf_mlp_param_count(const nerf_mlp_desc* desc) {
  MlpPlan P;
  if (nerf_plan_from_desc(desc, 0, &P)) return -1;
  return P.n_params;
}
extern "C" int64_t nerf_mlp_packed_bytes(const nerf_mlp_desc* desc, int32_t precision) {
  MlpPlan P;
  if (nerf_plan_from_desc(desc, precision, &P)) return -1;
  return P.packed_bytes;
}
extern "C" int64_t nerf_mlp_stash_floats_per_sample(const nerf_mlp_desc* desc) {
  MlpPlan P;
  if (nerf_plan_from_desc(desc, 0, &P)) return -1;
  return P.stash_per_sample;
}
extern "C" int64_t nerf_mlp_dz_floats_per_sample(const nerf_mlp_desc* desc) {
  MlpPlan P;
  if (nerf_plan_from_desc(desc, 0, &P)) return -1;
  return P.dz_per_sample;
}
extern "C" int32_t nerf_mlp_sample_block(const nerf_mlp_desc* desc, int32_t precision) {
  MlpPlan P;
  if (nerf_plan_from_desc(desc, precision, &P)) return -1;
  return P.sblk;
}

extern "C" int32_t nerf_mlp_pack(const nerf_mlp_desc* desc, int32_t precision, const float* flat,
                                 void* packed, void* stream) {
  NERF_PLAN_OR_FAIL(desc, precision, P, "nerf_mlp_pack");
  NERF_CHECK_ARG(flat && packed, "nerf_mlp_pack: null pointer");
  dim3 grid(64, 5 * P.n_mfma + 3), block(256);
  pack_kernel<<<grid, block, 0, as_stream(stream)>>>(P, flat, reinterpret_cast<char*>(packed));
  NERF_LAUNCH_CHECK("nerf_mlp_pack");
  return NERF_OK;
}

// Compute units of the current device (cached per device).
int nerf_num_cus() {
  static int cache[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (!cache[dev]) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
      n = 256;
    cache[dev] = n;
  }
  return cache[dev];
}
