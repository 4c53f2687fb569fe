// Host restatement of the nano-optimize differentiable splat rasteriser.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the smoke
// check and the CPU-baseline leg of bench.py load it to compare the gfx950
// path against.  No product code links or calls it.
//
// It restates, operation by operation, the WGSL kernels and Rust dispatch of
// crates/nano-optimize (each function cites the file:line it follows), in
// fp32 (bit parity with the device path: both sides use nts_math.h and
// -ffp-contract=off) and in fp64 (finite-difference checks).
//
// Layouts:
//   splats (host SoA, crates/nano-optimize/src/splat_store.rs:21-36):
//     pos[3n] xyz, rot[4n] wxyz, log_scale[3n], logit[n], sh_dc[3n],
//     sh_rest[45n] planar per splat (R[15] G[15] B[15]).
//   projected record (raster.rs:27-39): 12 values per splat
//     {mean.x, mean.y, depth, radius, conic.a, conic.b, conic.c, visible,
//      r, g, b, opacity}.
//   projected grad (raster.rs:46-57): 12 values per splat
//     {dmean.x, dmean.y, dopacity, pad, dconic.a, .b, .c, pad, dr, dg, db, pad}.
//   images: 4 floats per pixel, row-major.
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_camera {   // raster.rs:61-71 CameraUniform (96 bytes)
  float view[16];             // column-major 4x4 (glam to_cols_array)
  float cam_pos[3];
  float near_;
  float focal;
  float width;
  float height;
  uint32_t n_splats;
} orc_camera;

typedef struct orc_tiling {   // tile_binner.rs:21-42 TilingParams
  uint32_t width, height, tile_size;
  float depth_max;
} orc_tiling;

typedef struct orc_splats_f32 {
  const float *pos, *rot, *log_scale, *logit, *sh_dc, *sh_rest;
} orc_splats_f32;

typedef struct orc_splats_f64 {
  const double *pos, *rot, *log_scale, *logit, *sh_dc, *sh_rest;
} orc_splats_f64;

typedef struct orc_grads_f32 {
  float *pos, *rot, *log_scale, *logit, *sh_dc, *sh_rest;
} orc_grads_f32;

int orc_version(void);
int orc_num_threads(void);

// exp / sigmoid from the shared scalar header (pinned against float64 exp).
void orc_expf_array(const float* x, float* y, uint64_t n);
// logarithm / cos(2 pi u) / LCG / Box-Muller of the densify step, shared
// with the device (nts_math.h).
void orc_logf_array(const float* x, float* y, uint64_t n);
void orc_cos2pi_array(const float* x, float* y, uint64_t n);
uint32_t orc_lcg_jump(uint32_t state, uint64_t k);
void orc_box_muller_stream(uint32_t state, uint64_t n, float* out, float* uniforms);

void orc_camera_from_pose(const float eye[3], const float target[3], const float up[3],
                          float fov_y, uint32_t width, uint32_t height, uint32_t n_splats,
                          orc_camera* out);

void orc_project_f32(const orc_splats_f32* s, uint32_t n, const orc_camera* cam, float* projected);
void orc_project_f64(const orc_splats_f64* s, uint32_t n, const orc_camera* cam, double* projected);

// Binning over a single view; keys in the reference u32 format
// (tile_id << 16 | depth16).  orc_tile_count returns the total pair count.
uint64_t orc_tile_count(const float* projected, uint32_t n, const orc_tiling* t, uint32_t* counts);
void orc_exclusive_scan_u32(const uint32_t* in, uint64_t n, uint32_t* out);
void orc_tile_emit(const float* projected, uint32_t n, const uint32_t* offsets,
                   const orc_tiling* t, uint32_t* keys, uint32_t* payloads);
void orc_sort_pairs_u32(uint32_t* keys, uint32_t* vals, uint64_t n);
void orc_sort_pairs_u64(uint64_t* keys, uint32_t* vals, uint64_t n);
void orc_tile_ranges(const uint32_t* sorted_keys, uint64_t k, uint32_t num_tiles, uint32_t* ranges);
// Full bin (count, scan, emit, stable sort, ranges).  Call once with
// keys == NULL to get the pair count, then again with buffers of that size.
uint64_t orc_bin(const float* projected, uint32_t n, const orc_tiling* t,
                 uint32_t* keys, uint32_t* payloads, uint32_t* ranges);
// orc_bin with cull: only the tiles each record's alpha >= 1/255 ellipse can
// reach (nts_alpha_tile_rect), as the device's batched and serving paths bin.
uint64_t orc_bin_ex(const float* projected, uint32_t n, const orc_tiling* t, int cull, uint32_t* keys,
                    uint32_t* payloads, uint32_t* ranges);

// Forward alpha blend (rasterize.wgsl:48-118).  out: 4 floats/pixel
// (r, g, b, 1 - T).  final_t / last_idx may be NULL; last_idx is the number
// of entries of the tile's slice up to and including the last splat that
// contributed to the pixel.
void orc_composite(const float* projected, const uint32_t* payloads, const uint32_t* ranges,
                   const orc_tiling* t, float* out, float* final_t, uint32_t* last_idx);
// orc_composite's image and, per pixel, the number of splats blended into it.
void orc_composite_used(const float* projected, const uint32_t* payloads, const uint32_t* ranges,
                        const orc_tiling* t, float* out, uint32_t* n_used);
// float64 blend; last_idx / n_used (contributing splats per pixel) may be NULL.
void orc_composite_f64(const double* projected, const uint32_t* payloads, const uint32_t* ranges,
                       const orc_tiling* t, double* out, uint32_t* last_idx, uint32_t* n_used);

// Reference-parity backward (rasterize_backward.wgsl:73-210).  Accumulates
// (in float64, then rounds) into projected_grad[12n], which is overwritten.
void orc_composite_backward_ref(const float* projected, const uint32_t* payloads,
                                const uint32_t* ranges, const orc_tiling* t, uint32_t n,
                                const float* forward_out, const float* dl_dc, float* projected_grad);

// Recompute-mode backward: starts at each pixel's last contributing splat
// with the saved final T (outputs of orc_composite), no 1.0001 bail.
void orc_composite_backward_recompute(const float* projected, const uint32_t* payloads, const uint32_t* ranges,
                                      const orc_tiling* t, uint32_t n, const float* final_t,
                                      const uint32_t* last_idx, const float* dl_dc, float* projected_grad);

// project_backward.wgsl:79-304.  Writes every gradient of every splat.
void orc_project_backward(const orc_splats_f32* s, uint32_t n, const orc_camera* cam,
                          const float* projected_grad, orc_grads_f32* g);
// The exact chain rule of the forward (adds the colour clamp gate, the SH
// direction term and the Jacobian term that the reference drops).
void orc_project_backward_full(const orc_splats_f32* s, uint32_t n, const orc_camera* cam,
                               const float* projected_grad, orc_grads_f32* g);

// Adam (adam.rs:70-90) on a flat slab; bias corrections as adam_gpu.rs:275-276.
float orc_powi_f32(float a, int32_t b);
void orc_adam_step(float* params, const float* grads, float* m, float* v, uint64_t n,
                   float lr, float beta1, float beta2, float eps, uint64_t t);
void orc_apply_constraints(float* rot, float* log_scale, float* logit, uint32_t n,
                           float log_scale_min, float log_scale_max, float logit_max);
void orc_accumulate_grad(const float* dpos, float* acc, uint32_t n);

// dL/dC for a rendered image against a target (both 4 floats/pixel, w
// ignored).  kind 0 = L1 (mean |d| over 3WH, sign(0) = 0), 1 = MSE
// (loss.rs:163-179).  Returns the loss (float64 accumulation).
double orc_loss_grad(const float* pred, const float* target, uint32_t w, uint32_t h,
                     int kind, float* dl_dc);

// Central difference of the mean-square loss over 3WH with respect to one
// parameter (splat idx, attribute attr in pos/rot/log_scale/logit/sh_dc/
// sh_rest order, component comp), re-blending in float64 only the tiles of
// the splat's float32 binning.  proj64 / proj32: float64 / float32
// projections of every splat; payloads / ranges: the float32 binning;
// target: float4 per pixel.  Returns 0, or 1 when the +-h evaluations are
// discontinuous (tile set, depth bucket, visibility or any pixel's
// contribution pattern changes), 2 when the float64 contribution pattern
// differs from the float32 forward's on one of the splat's pixels, -1 on
// bad arguments.
// With weights (float4 per pixel) the differentiated function is instead
// the linear functional sum_pixels <w, C> (a vector-Jacobian product with
// fixed dL/dC), not divided by 3WH.
int orc_fd_local_mse(const orc_splats_f64* s, uint32_t n, const orc_camera* cam, const orc_tiling* t,
                     const double* proj64, const float* proj32, const uint32_t* payloads, const uint32_t* ranges,
                     const float* target, const float* weights, uint32_t idx, int attr, int comp, double h,
                     double* fd);
// The same for `count` parameters at once (threads over parameters).
void orc_fd_local_mse_batch(const orc_splats_f64* s, uint32_t n, const orc_camera* cam, const orc_tiling* t,
                            const double* proj64, const float* proj32, const uint32_t* payloads,
                            const uint32_t* ranges, const float* target, const float* weights, uint32_t count,
                            const uint32_t* idx, const int32_t* attr, const int32_t* comp, const double* h,
                            double* fd, int32_t* disc);

// SSIM (loss.rs:23-415): the normalised 11-tap kernel, one separable blur
// (or its adjoint) of a w x h plane, and the 3-channel SSIM loss
// 1 - mean(SSIM) with dL/dC (float4 images, w ignored / written 0).
void orc_ssim_kernel(float* k);
void orc_ssim_blur(const float* src, int w, int h, float* dst, int adjoint);
float orc_ssim_loss_grad(const float* pred, const float* target, int w, int h, float* grad);

// Prune / densify on the host SoA (train.rs:414-602).  Arrays must have
// capacity for the result.  moments: 12 slabs (m,v per attribute, in the
// attribute order pos, rot, scale, logit, dc, rest), may be NULL.
typedef struct orc_splats_mut {
  float *pos, *rot, *log_scale, *logit, *sh_dc, *sh_rest;
} orc_splats_mut;
uint32_t orc_prune(orc_splats_mut* s, uint32_t n, float threshold, float** moments, float* grad_acc);
uint32_t orc_densify(orc_splats_mut* s, uint32_t n, uint32_t grad_count, float grad_threshold,
                     float split_log_scale_threshold, uint32_t max_splats, uint32_t iter,
                     float** moments, float* grad_acc, uint32_t counts_out[3]);

// Reinhard + sRGB + rgba8 (tonemap.wgsl:27-41).
void orc_tonemap(const float* image, uint32_t w, uint32_t h, uint8_t* out);

// Timing driver of the CPU baseline (see nts_oracle.cpp): the reference step
// over n_views views of one scene, `reps` timed runs after `warmup`, per-view
// and Adam wall times in seconds.  threads > 0 sets the OpenMP thread count.
int orc_time_step_views(const orc_splats_f32* s, uint32_t n, const orc_camera* cams, const orc_tiling* tilings,
                        const float* const* targets, uint32_t n_views, int warmup, int reps, int threads,
                        double* view_secs, double* adam_secs);

#ifdef __cplusplus
}
#endif
