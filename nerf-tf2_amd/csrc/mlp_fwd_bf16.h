// BF16 fused forward (placeholder until the bf16 MFMA path lands).
#pragma once
#include "mlp_kernels.h"

template <int W, int D, int SKIP, int LX, int LD>
int launch_fwd_bf16(const MlpPlan&, const void*, int64_t, const float*, const float*, const float*,
                    const float*, const float*, int, float*, float*, void*, hipStream_t) {
  return -1;
}
