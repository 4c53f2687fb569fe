// Adam (adam_step.wgsl:24-36), the fused per-splat optimiser step
// (grad_acc += |dpos|_1 (accumulate_grad_pos.wgsl:17-23), Adam on all 59
// attribute planes, then the constraints of apply_constraints.wgsl:24-41),
// and each of those three on its own.
#include "nts_adam.h"

__global__ void __launch_bounds__(256) k_adam_flat(float* __restrict__ p, const float* __restrict__ g,
                                                   float* __restrict__ m, float* __restrict__ v, uint64_t n, AdamK c) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    float mm = m[i], vv = v[i];
    p[i] = adam_lane(p[i], g[i], mm, vv, c);
    m[i] = mm;
    v[i] = vv;
  }
}

// The scene's optimiser step over all 59 planes.  FUSED adds what the
// reference runs as separate dispatches around AdamGpu::step_all: the
// densify-gate accumulator before the step (train.rs:228) and the
// constraints after it; without FUSED it is step_all alone
// (adam_gpu.rs:334-427).
template <bool FUSED>
__global__ void __launch_bounds__(256) k_adam_scene(float* __restrict__ P, const float* __restrict__ G,
                                                    float* __restrict__ M, float* __restrict__ V,
                                                    float* __restrict__ grad_acc, uint32_t cap, uint32_t n,
                                                    AdamK cpos, AdamK cattr, float smin, float smax, float omax) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t c = cap;
  if (FUSED)
    grad_acc[i] = grad_acc[i] + fabsf(G[(P_POS + 0) * c + i]) + fabsf(G[(P_POS + 1) * c + i]) +
                  fabsf(G[(P_POS + 2) * c + i]);
  float q[4], s[3], o = 0.0f;
#pragma unroll 4
  for (int k = 0; k < P_COUNT; ++k) {
    const size_t j = (size_t)k * c + i;
    float mm = M[j], vv = V[j];
    const float np = adam_lane(P[j], G[j], mm, vv, k < P_ROT ? cpos : cattr);
    M[j] = mm;
    V[j] = vv;
    if (FUSED && k >= P_ROT && k < P_ROT + 4) q[k - P_ROT] = np;
    else if (FUSED && k >= P_SCALE && k < P_SCALE + 3) s[k - P_SCALE] = np;
    else if (FUSED && k == P_LOGIT) o = np;
    else P[j] = np;
  }
  if (FUSED) constrain_lane(P, c, i, q, s, o, smin, smax, omax);
}

// apply_constraints.wgsl:24-41 alone (AdamGpu::apply_constraints,
// adam_gpu.rs:431-489).
__global__ void __launch_bounds__(256) k_constraints(float* __restrict__ P, uint32_t cap, uint32_t n, float smin,
                                                     float smax, float omax) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t c = cap;
  float q[4], s[3];
#pragma unroll
  for (int k = 0; k < 4; ++k) q[k] = P[(P_ROT + k) * c + i];
#pragma unroll
  for (int k = 0; k < 3; ++k) s[k] = P[(P_SCALE + k) * c + i];
  constrain_lane(P, c, i, q, s, P[P_LOGIT * c + i], smin, smax, omax);
}

// accumulate_grad_pos.wgsl:17-23 alone (AdamGpu::accumulate_grad,
// adam_gpu.rs:492-541): grad_acc += |dx| + |dy| + |dz| of the scene's
// position gradient.
__global__ void __launch_bounds__(256) k_accumulate_grad(const float* __restrict__ G, float* __restrict__ grad_acc,
                                                         uint32_t cap, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t c = cap;
  grad_acc[i] = grad_acc[i] + fabsf(G[(P_POS + 0) * c + i]) + fabsf(G[(P_POS + 1) * c + i]) +
                fabsf(G[(P_POS + 2) * c + i]);
}

extern "C" int nts_adam_step_flat(nts_ctx* ctx, nts_buffer* params, const nts_buffer* grads, nts_buffer* m,
                                  nts_buffer* v, uint64_t n, const nts_adam_cfg* cfg, uint64_t t) {
  NTS_CHECK_ARG(ctx && params && grads && m && v && cfg && t > 0);
  NTS_CHECK_ARG(n * 4 <= params->bytes && n * 4 <= grads->bytes && n * 4 <= m->bytes && n * 4 <= v->bytes);
  if (n == 0) return NTS_OK;
  uint32_t grid = nts_div_up(n, 256);
  if (grid > 8192) grid = 8192;
  NTS_LAUNCH(ctx, "adam_flat", k_adam_flat, dim3(grid), dim3(256), 0, (float*)params->ptr,
             (const float*)grads->ptr, (float*)m->ptr, (float*)v->ptr, n, make_adam(cfg, t));
  return NTS_OK;
}

int nts_adam_scene_dev(nts_ctx* ctx, nts_scene* s, const nts_adam_cfg* cfg_pos, const nts_adam_cfg* cfg_attr,
                       uint64_t t, const nts_constraints* cons) {
  if (s->n == 0) return NTS_OK;
  NTS_LAUNCH(ctx, "adam_step", k_adam_scene<true>, dim3(nts_div_up(s->n, 256)), dim3(256), 0, s->params,
             (const float*)s->grads, s->m, s->v, s->grad_acc, s->cap, s->n, make_adam(cfg_pos, t),
             make_adam(cfg_attr, t), cons->log_scale_min, cons->log_scale_max, cons->logit_max);
  return NTS_OK;
}

extern "C" int nts_adam_step(nts_ctx* ctx, nts_scene* scene, const nts_adam_cfg* cfg_pos,
                             const nts_adam_cfg* cfg_attr, uint64_t t, const nts_constraints* cons) {
  NTS_CHECK_ARG(ctx && scene && cfg_pos && cfg_attr && cons && t > 0);
  NTS_NEED_TRAINABLE(scene);
  return nts_adam_scene_dev(ctx, scene, cfg_pos, cfg_attr, t, cons);
}

extern "C" int nts_adam_step_all(nts_ctx* ctx, nts_scene* scene, const nts_adam_cfg* cfg_pos,
                                 const nts_adam_cfg* cfg_attr, uint64_t t) {
  NTS_CHECK_ARG(ctx && scene && cfg_pos && cfg_attr && t > 0);
  NTS_NEED_TRAINABLE(scene);
  if (scene->n == 0) return NTS_OK;
  NTS_LAUNCH(ctx, "adam_step_all", k_adam_scene<false>, dim3(nts_div_up(scene->n, 256)), dim3(256), 0,
             scene->params, (const float*)scene->grads, scene->m, scene->v, scene->grad_acc, scene->cap, scene->n,
             make_adam(cfg_pos, t), make_adam(cfg_attr, t), 0.0f, 0.0f, 0.0f);
  return NTS_OK;
}

extern "C" int nts_apply_constraints(nts_ctx* ctx, nts_scene* scene, const nts_constraints* cons) {
  NTS_CHECK_ARG(ctx && scene && cons);
  if (scene->n == 0) return NTS_OK;
  NTS_LAUNCH(ctx, "apply_constraints", k_constraints, dim3(nts_div_up(scene->n, 256)), dim3(256), 0, scene->params,
             scene->cap, scene->n, cons->log_scale_min, cons->log_scale_max, cons->logit_max);
  return NTS_OK;
}

extern "C" int nts_accumulate_grad(nts_ctx* ctx, nts_scene* scene) {
  NTS_CHECK_ARG(ctx && scene);
  NTS_NEED_TRAINABLE(scene);
  if (scene->n == 0) return NTS_OK;
  NTS_LAUNCH(ctx, "accumulate_grad", k_accumulate_grad, dim3(nts_div_up(scene->n, 256)), dim3(256), 0,
             (const float*)scene->grads, scene->grad_acc, scene->cap, scene->n);
  return NTS_OK;
}
