// Projection (EWA covariance + order-3 SH) and its backward, one thread per
// splat.  Parameters are read from the scene's attribute planes (fully
// coalesced); projected records are 48-byte ProjectedSplat structs.
#include "nts_project.h"

__global__ void __launch_bounds__(256) k_project(const float* __restrict__ P, uint32_t cap, uint32_t n, DevCamera cam,
                                                 Projected* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  SplatP s;
  load_splat(P, cap, i, s);
  out[i] = project_p(s, cam);
}

template <bool FULL>
__global__ void __launch_bounds__(256) k_project_backward(const float* __restrict__ P, uint32_t cap, uint32_t n,
                                                          DevCamera cam, const ProjectedGrad* __restrict__ pgrad,
                                                          float* __restrict__ G) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  SplatP s;
  load_splat(P, cap, i, s);
  float g[P_COUNT];
  project_backward_p<FULL>(s, cam, pgrad[i], g);
#pragma unroll
  for (int k = 0; k < P_COUNT; ++k) G[(size_t)k * cap + i] = g[k];
}

extern "C" int nts_project(nts_ctx* ctx, const nts_scene* scene, const nts_camera* cam, nts_buffer* projected) {
  NTS_CHECK_ARG(ctx && scene && cam && projected);
  NTS_CHECK_ARG(projected->bytes >= (uint64_t)scene->n * sizeof(Projected));
  if (scene->n == 0) return NTS_OK;
  DevCamera dc;
  memcpy(&dc, cam, sizeof dc);
  NTS_LAUNCH(ctx, "project", k_project, dim3(nts_div_up(scene->n, 256)), dim3(256), 0, scene->params, scene->cap,
             scene->n, dc, (Projected*)projected->ptr);
  return NTS_OK;
}

extern "C" int nts_project_backward_ex(nts_ctx* ctx, nts_scene* scene, const nts_buffer* projected_grad,
                                       const nts_camera* cam, int mode) {
  NTS_CHECK_ARG(ctx && scene && cam && projected_grad);
  NTS_CHECK_ARG(mode == NTS_BWD_REF_PARITY || mode == NTS_BWD_RECOMPUTE);
  NTS_NEED_TRAINABLE(scene);
  NTS_CHECK_ARG(projected_grad->bytes >= (uint64_t)scene->n * sizeof(ProjectedGrad));
  if (scene->n == 0) return NTS_OK;
  DevCamera dc;
  memcpy(&dc, cam, sizeof dc);
  if (mode == NTS_BWD_RECOMPUTE)
    NTS_LAUNCH(ctx, "project_backward_full", k_project_backward<true>, dim3(nts_div_up(scene->n, 256)), dim3(256), 0,
               scene->params, scene->cap, scene->n, dc, (const ProjectedGrad*)projected_grad->ptr, scene->grads);
  else
    NTS_LAUNCH(ctx, "project_backward", k_project_backward<false>, dim3(nts_div_up(scene->n, 256)), dim3(256), 0,
               scene->params, scene->cap, scene->n, dc, (const ProjectedGrad*)projected_grad->ptr, scene->grads);
  return NTS_OK;
}

extern "C" int nts_project_backward(nts_ctx* ctx, nts_scene* scene, const nts_buffer* projected_grad,
                                    const nts_camera* cam) {
  return nts_project_backward_ex(ctx, scene, projected_grad, cam, NTS_BWD_REF_PARITY);
}
