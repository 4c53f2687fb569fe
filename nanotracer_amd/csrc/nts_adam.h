// Adam arithmetic (adam.rs:70-90, adam_step.wgsl:24-36) and its bias
// corrections, shared by the optimiser kernels.
#pragma once
#include "nts_internal.h"

struct AdamK {
  float lr, b1, b2, eps, bc1, bc2;
};

__device__ __forceinline__ float adam_lane(float p, float g, float& m, float& v, const AdamK& c) {
  const float mn = c.b1 * m + (1.0f - c.b1) * g;
  const float vn = c.b2 * v + (1.0f - c.b2) * g * g;
  m = mn;
  v = vn;
  const float mh = mn / c.bc1;
  const float vh = vn / c.bc2;
  return p - c.lr * mh / (sqrtf(vh) + c.eps);
}

// apply_constraints.wgsl:24-41 on one splat whose rotation (q), log-scales
// (s) and logit (o) are in registers: quaternion renormalised, log-scales
// clamped to [smin, smax], logit to [-omax, omax]; written to the planes.
__device__ __forceinline__ void constrain_lane(float* __restrict__ P, size_t c, uint32_t i, const float q[4],
                                               const float s[3], float o, float smin, float smax, float omax) {
  const float n2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
  const float nrm = nts_maxf(sqrtf(n2), 1e-8f);
#pragma unroll
  for (int k = 0; k < 4; ++k) P[(P_ROT + k) * c + i] = q[k] / nrm;
#pragma unroll
  for (int k = 0; k < 3; ++k) P[(P_SCALE + k) * c + i] = nts_minf(nts_maxf(s[k], smin), smax);
  P[P_LOGIT * c + i] = nts_minf(nts_maxf(o, -omax), omax);
}

// compiler-rt __powisf2, which Rust's f32::powi lowers to (adam_gpu.rs:275-276).
static inline float powi_f32(float a, int32_t b) {
  const int recip = b < 0;
  float r = 1.0f;
  while (1) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return recip ? 1.0f / r : r;
}

static inline AdamK make_adam(const nts_adam_cfg* cfg, uint64_t t) {
  AdamK k;
  k.lr = cfg->lr;
  k.b1 = cfg->beta1;
  k.b2 = cfg->beta2;
  k.eps = cfg->eps;
  k.bc1 = 1.0f - powi_f32(cfg->beta1, (int32_t)t);
  k.bc2 = 1.0f - powi_f32(cfg->beta2, (int32_t)t);
  return k;
}

