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

