/* nts.h — C ABI of the MI355X (gfx950) differentiable Gaussian-splat
 * rasteriser.
 *
 * This is the drop-in boundary for the public Rust surface of
 * crates/nano-optimize (src/lib.rs:21-45).  Each declaration names the
 * reference item it replaces.  A thin Rust FFI crate binds these symbols
 * (extern "C" + safe wrappers); Python binds them with ctypes.
 *
 * Conventions
 *  - Every call returns int: 0 = ok, negative = NTS_E_*.  Nothing aborts
 *    or throws across the ABI; nts_last_error() gives the message of the
 *    last failure on the calling thread.
 *  - Work is asynchronous on the context's HIP stream (default priority).
 *    nts_sync() and the *_download calls are the blocking calls.
 *  - The library owns all device memory; the caller owns host arrays.
 *  - One nts_ctx per host thread; handles are usable only with the context
 *    that created them.
 */
#ifndef NTS_H
#define NTS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NTS_OK 0
#define NTS_E_INVALID (-1)      /* bad argument / size mismatch */
#define NTS_E_HIP (-2)          /* HIP runtime error */
#define NTS_E_NOMEM (-3)        /* device allocation or HBM budget exceeded */
#define NTS_E_CAPACITY (-4)     /* a device capacity was exceeded (pairs, splats) */
#define NTS_E_STATE (-5)        /* handle used in the wrong state */

typedef struct nts_ctx nts_ctx;
typedef struct nts_buffer nts_buffer;
typedef struct nts_scene nts_scene;
typedef struct nts_bins nts_bins;

/* Mirror of CameraUniform (crates/nano-optimize/src/raster.rs:61-71), 96 B:
 * column-major view matrix, camera position, near plane, focal length in
 * pixels, image size, splat count. */
typedef struct nts_camera {
  float view[16];
  float cam_pos[3];
  float near_;
  float focal;
  float width;
  float height;
  uint32_t n_splats;
} nts_camera;

/* Mirror of TilingParams (crates/nano-optimize/src/tile_binner.rs:21-42). */
typedef struct nts_tiling {
  uint32_t width, height, tile_size;
  float depth_max;
} nts_tiling;

/* Host splat arrays in the SplatBuffer layout
 * (crates/nano-optimize/src/splat_store.rs:21-36): pos xyz, rotation wxyz,
 * log-scale xyz, opacity logit, SH DC rgb, 45 SH-rest planar per splat. */
typedef struct nts_splats_host {
  uint32_t n;
  float* pos;        /* 3n */
  float* rot;        /* 4n */
  float* log_scale;  /* 3n */
  float* logit;      /* n  */
  float* sh_dc;      /* 3n */
  float* sh_rest;    /* 45n */
} nts_splats_host;

/* Mirror of AdamConfig (crates/nano-optimize/src/adam.rs:22-39). */
typedef struct nts_adam_cfg {
  float lr, beta1, beta2, eps;
} nts_adam_cfg;

/* Constraint bounds (crates/nano-optimize/src/train.rs:37-39). */
typedef struct nts_constraints {
  float log_scale_min, log_scale_max, logit_max;
} nts_constraints;

#define NTS_LOSS_L1 0
#define NTS_LOSS_MSE 1
#define NTS_LOSS_SSIM 2  /* (1 - lambda) MSE + lambda (1 - SSIM); batched steps */

#define NTS_BWD_REF_PARITY 0  /* rasterize_backward.wgsl semantics */
#define NTS_BWD_RECOMPUTE 1   /* final-T + last-index recompute, full chain rule */

/* ---- context (replaces WgpuCtx::new, gpu.rs:31-77) -------------------- */
int nts_ctx_create(int device, uint64_t hbm_budget_bytes, nts_ctx** out);
int nts_ctx_destroy(nts_ctx* ctx);
int nts_sync(nts_ctx* ctx);
const char* nts_last_error(void);
int nts_version(void);
/* Per-kernel HIP-event timing on the context stream (off by default). */
int nts_prof_enable(nts_ctx* ctx, int on);
int nts_prof_reset(nts_ctx* ctx);
/* Total ms and launch count of kernels whose name starts with `prefix`. */
int nts_prof_query(nts_ctx* ctx, const char* prefix, double* total_ms, uint64_t* launches);
/* Newline-separated "name launches total_ms" report into buf. */
int nts_prof_report(nts_ctx* ctx, char* buf, size_t cap);
/* Ceilings measured on the device, for roofline figures (no reference
 * counterpart): the bandwidth of a plain copy of copy_bytes (read + write
 * bytes over time, GB/s), the LDS bandwidth of conflict-free 128-bit reads
 * over all CUs (TB/s), and the shader clock during that LDS run (GHz, from
 * the cycle counter over the wall clock).  Allocates 2 x copy_bytes. */
int nts_probe_peaks(nts_ctx* ctx, uint64_t copy_bytes, double* hbm_copy_gbs, double* lds_read_tbs,
                    double* clock_ghz);
/* The FP32 VALU ceiling measured on the device (no reference counterpart):
 * independent v_fma_f32 chains at eight waves per SIMD, as TFLOP/s (2 flops
 * per lane FMA) and as wave instructions issued per SIMD per shader cycle
 * (the clock of the same run). */
int nts_probe_valu(nts_ctx* ctx, double* fma_tflops, double* fma_per_simd_cycle);
/* Instruction-mix ceiling of the recompute blend backward (no reference
 * counterpart): the (splat, wave) evaluations per second a launch of the
 * backward's own per-evaluation code reaches at the kernel's workgroup size,
 * LDS and occupancy, with every pixel taking part and no memory traffic or
 * barriers (tile_size 8, 16 or 32; bwd_batch 64, 128 or 256).  With the
 * evaluation count of a step (nts_batch_eval_counts) it bounds the
 * backward's time from below by its arithmetic alone. */
int nts_probe_blend_mix(nts_ctx* ctx, uint32_t tile_size, uint32_t bwd_batch, double* bwd_hits_per_s);

/* ---- device buffers (replaces WgpuCtx::storage_buffer[_zeroed] and
 *      readback, gpu.rs:79-152) ------------------------------------------ */
int nts_buffer_create(nts_ctx* ctx, uint64_t bytes, nts_buffer** out);
int nts_buffer_destroy(nts_ctx* ctx, nts_buffer* buf);
uint64_t nts_buffer_size(const nts_buffer* buf);
int nts_buffer_upload(nts_ctx* ctx, nts_buffer* buf, uint64_t offset, const void* host, uint64_t bytes);
int nts_buffer_download(nts_ctx* ctx, const nts_buffer* buf, uint64_t offset, void* host, uint64_t bytes);
int nts_buffer_zero(nts_ctx* ctx, nts_buffer* buf);

/* ---- scene: parameters + gradients + Adam moments + grad accumulator
 *      (replaces GpuSplatBuffer, GradSplatBuffers, AdamMomentBuffers,
 *      splat_gpu.rs:36-384) ---------------------------------------------- */
#define NTS_SCENE_TRAINABLE 1u /* also hold gradients, both Adam moments and grad_acc */
/* A scene of up to `capacity` splats.  With NTS_SCENE_TRAINABLE it holds
 * GpuSplatBuffer + GradSplatBuffers + AdamMomentBuffers + the grad_acc slab
 * (train.rs:94-96, 236 B x 4 + 4 B per splat); without it only the
 * parameters (236 B per splat), as a viewer's GpuSplatBuffer
 * (nano-view/src/lib.rs:642-652): calls that need gradients, moments or
 * grad_acc then fail with NTS_E_STATE. */
int nts_scene_create(nts_ctx* ctx, uint32_t capacity, uint32_t flags, nts_scene** out);
int nts_scene_destroy(nts_ctx* ctx, nts_scene* scene);
uint32_t nts_scene_count(const nts_scene* scene);
/* GpuSplatBuffer::upload (splat_gpu.rs:94-120): sets the splat count and the
 * parameters, and resets gradients, moments and grad_acc to zero (a fresh
 * upload has fresh optimiser state, train.rs:94-96). */
int nts_scene_upload(nts_ctx* ctx, nts_scene* scene, const nts_splats_host* host);
/* GpuSplatBuffer::sync_from (splat_gpu.rs:124-150): rewrites the parameters
 * of a scene holding host->n splats in place; gradients, moments and
 * grad_acc are kept.  NTS_E_INVALID when the count differs (the reference
 * reallocates on densify / prune instead). */
int nts_scene_sync_params(nts_ctx* ctx, nts_scene* scene, const nts_splats_host* host);
/* GpuSplatBuffer::readback (splat_gpu.rs:152-187); host->n is set. */
int nts_scene_download(nts_ctx* ctx, const nts_scene* scene, nts_splats_host* host);
/* GradSplatBuffers readback (train.rs:370-389 reads them for norms). */
int nts_scene_download_grads(nts_ctx* ctx, const nts_scene* scene, nts_splats_host* grads);
/* Overwrite gradients (tests feed synthetic gradients into Adam). */
int nts_scene_upload_grads(nts_ctx* ctx, nts_scene* scene, const nts_splats_host* grads);
/* GradSplatBuffers::zero (splat_gpu.rs:63-75). */
int nts_scene_zero_grads(nts_ctx* ctx, nts_scene* scene);
/* AdamMomentBuffers::{download_to_cpu, upload_from_cpu} (splat_gpu.rs:229-384). */
int nts_scene_download_moments(nts_ctx* ctx, const nts_scene* scene, nts_splats_host* m, nts_splats_host* v);
int nts_scene_upload_moments(nts_ctx* ctx, nts_scene* scene, const nts_splats_host* m, const nts_splats_host* v);
int nts_scene_download_grad_acc(nts_ctx* ctx, const nts_scene* scene, float* host);
int nts_scene_upload_grad_acc(nts_ctx* ctx, nts_scene* scene, const float* host);

/* ---- forward (Rasterizer::project, raster.rs:731-794) ----------------- */
/* CameraUniform::from_pose (raster.rs:79-99): glam look_at_rh, near 0.01. */
int nts_camera_from_pose(const float eye[3], const float target[3], const float up[3], float fov_y,
                         uint32_t width, uint32_t height, uint32_t n_splats, nts_camera* out);
/* projected: >= 48*n bytes of ProjectedSplat records (raster.rs:27-39). */
int nts_project(nts_ctx* ctx, const nts_scene* scene, const nts_camera* cam, nts_buffer* projected);

/* ---- binning (TileBinner::bin, tile_binner.rs:181-382) ---------------- */
/* Count, scan, emit (tile << 16 | depth16) keys, onesweep radix sort,
 * per-tile ranges.  Like the reference it reads the pair total back.  Keys
 * are u32, the reference's format, when the tile id fits 16 bits (tiles <=
 * 65,536, every view up to 4096x4096 at tile 16); u64 otherwise, where the
 * reference's key wraps (tile_emit.wgsl:60-61).  The bins also hold the
 * order in which composite and composite_backward take the tiles, most
 * pairs first (4 B per tile; grid order when there are no pairs): it
 * changes no image or gradient term, only the order workgroups start. */
int nts_bin(nts_ctx* ctx, const nts_buffer* projected, uint32_t n_splats, const nts_tiling* tiling,
            nts_bins** out);
int nts_bins_destroy(nts_ctx* ctx, nts_bins* bins);
int nts_bins_info(const nts_bins* bins, uint64_t* total_pairs, uint32_t* num_tiles);
/* Sorted keys in the reference's u32 format (tile_id << 16 | depth16);
 * NTS_E_CAPACITY when the tile id does not fit 16 bits. */
int nts_bins_download_keys_u32(nts_ctx* ctx, const nts_bins* bins, uint32_t* host);
int nts_bins_download_keys_u64(nts_ctx* ctx, const nts_bins* bins, uint64_t* host);
int nts_bins_download_payloads(nts_ctx* ctx, const nts_bins* bins, uint32_t* host);
/* ranges: 2 u32 per tile [begin, end), empty tiles (0, 0). */
int nts_bins_download_ranges(nts_ctx* ctx, const nts_bins* bins, uint32_t* host);

/* PrefixScan::scan (prefix_scan.rs:127-186): exclusive, wrapping u32, in place. */
int nts_scan_u32(nts_ctx* ctx, nts_buffer* data, uint64_t n);
/* RadixSort::sort (radix_sort.rs:144-303): stable sort of u32 keys with u32
 * payloads, in place.  key_bits limits the sorted low bits (32 = all). */
int nts_sort_pairs_u32(nts_ctx* ctx, nts_buffer* keys, nts_buffer* vals, uint64_t n, uint32_t key_bits);
int nts_sort_pairs_u64(nts_ctx* ctx, nts_buffer* keys, nts_buffer* vals, uint64_t n, uint32_t key_bits);
/* The binning half of TileBinner::bin (tile_binner.rs:181-382: RadixSort::sort
 * of the (tile_id << 16 | depth16) keys, then tile_ranges.wgsl) in one call:
 * sorts n pairs in place into the order nts_sort_pairs_u32/u64 give over the
 * low 16 + tile_bits bits, and writes ranges[t] = [begin, end) (2 u32) for
 * each tile t < num_tiles ((0, 0) for a tile with no pair).  Keys are below
 * 2^(16 + tile_bits) or all ones (a dropped pair: it sorts last and lies in
 * no range); with any all-ones key, every tile id is below 2^tile_bits - 1.
 * key_bytes 4 or 8.  Only the tile bits are sorted device-wide; the depth
 * order inside each group of tiles is made in LDS. */
int nts_sort_tile_pairs(nts_ctx* ctx, nts_buffer* keys, nts_buffer* vals, uint64_t n, uint32_t key_bytes,
                        uint32_t tile_bits, nts_buffer* ranges, uint32_t num_tiles);
/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  nts_sort_pairs_u32 with the same result and errors, but every
 * buffer beyond keys and vals (the ping-pong pairs, digit histograms and
 * offsets, look-back status words, counters) lives in one static device
 * array of 19.0 MB per GPU, held until the process ends: nothing is
 * allocated.  At most NTS_SORT_NOALLOC_MAX_PAIRS pairs (NTS_E_CAPACITY
 * beyond).  All calls on one GPU share that array, so calls that overlap on
 * one GPU (several contexts or threads) race; sequential calls are safe.
 * Like nts_sort_pairs_u32 it waits for the sort to finish (it reads the
 * look-back error word into the context's pinned word). */
#define NTS_SORT_NOALLOC_MAX_PAIRS (1u << 21)
int nts_sort_pairs_u32_noalloc(nts_ctx* ctx, nts_buffer* keys, nts_buffer* vals, uint64_t n, uint32_t key_bits);

/* ---- composite (Rasterizer::composite, raster.rs:468-529) ------------- */
/* image: 16 B per pixel (r, g, b, 1 - T).  Any tile size in {8, 16, 32};
 * the reference's kernel is fixed at 16. */
int nts_composite(nts_ctx* ctx, const nts_buffer* projected, const nts_bins* bins, nts_buffer* image,
                  const nts_tiling* tiling);
/* As nts_composite, also saving per pixel the final transmittance (f32)
 * and the slice-relative index + 1 of the last contributing splat (u32),
 * which the recompute backward starts from.  Either may be NULL.  No
 * reference counterpart: the reference stores only 1 - T (rasterize.wgsl:116). */
int nts_composite_ex(nts_ctx* ctx, const nts_buffer* projected, const nts_bins* bins, nts_buffer* image,
                     const nts_tiling* tiling, nts_buffer* final_t, nts_buffer* last_idx);

/* ---- loss (dL/dC contract, train.rs:180-196, loss.rs:163-179) ---------- */
/* dl_dc: 16 B per pixel (w = 0).  *loss receives the scalar loss. */
int nts_loss_grad(nts_ctx* ctx, const nts_buffer* image, const nts_buffer* target, nts_buffer* dl_dc,
                  uint32_t width, uint32_t height, int kind, double* loss);
/* combined_loss_grad (loss.rs:390-415): (1 - lambda) MSE + lambda (1 - SSIM)
 * with lambda clamped to [0, 1]; SSIM per channel over an 11-tap Gaussian
 * window (sigma 1.5, reflect-101 borders, C1 = 1e-4, C2 = 9e-4, loss.rs:23-338)
 * averaged over the three channels (ssim_loss_grad, loss.rs:346-383).
 * lambda = 0 is nts_loss_grad with NTS_LOSS_MSE.  *total, *mse and *ssim
 * (any may be NULL) receive the three components the reference returns. */
int nts_loss_grad_ssim(nts_ctx* ctx, const nts_buffer* image, const nts_buffer* target, nts_buffer* dl_dc,
                       uint32_t width, uint32_t height, float lambda, double* total, double* mse, double* ssim);

/* ---- backward (Rasterizer::composite_backward, raster.rs:653-724;
 *      project_backward, raster.rs:559-645) ---------------------------- */
/* projected_grad: 48 B per splat, accumulated into (zero it first, like
 * Rasterizer::zero_projected_grad).  Mode NTS_BWD_REF_PARITY only here. */
int nts_composite_backward(nts_ctx* ctx, const nts_buffer* projected, const nts_bins* bins,
                           const nts_buffer* forward_out, const nts_buffer* dl_dc,
                           nts_buffer* projected_grad, const nts_tiling* tiling);
/* Either mode: REF_PARITY reads forward_out (T = 1 - w, walk from the
 * slice end, bail at T > 1.0001); RECOMPUTE reads final_t / last_idx of
 * nts_composite_ex (walk from each pixel's last contributor, no bail). */
int nts_composite_backward_ex(nts_ctx* ctx, const nts_buffer* projected, const nts_bins* bins,
                              const nts_buffer* forward_out, const nts_buffer* final_t, const nts_buffer* last_idx,
                              const nts_buffer* dl_dc, nts_buffer* projected_grad, const nts_tiling* tiling, int mode);
/* Writes the scene's parameter gradients (reference chain rule). */
int nts_project_backward(nts_ctx* ctx, nts_scene* scene, const nts_buffer* projected_grad,
                         const nts_camera* cam);
/* mode NTS_BWD_RECOMPUTE: the exact chain rule, adding the colour clamp
 * gate, d(view dir)/d(position) and dJ/dt that project_backward.wgsl:17-24
 * drops. */
int nts_project_backward_ex(nts_ctx* ctx, nts_scene* scene, const nts_buffer* projected_grad,
                            const nts_camera* cam, int mode);

/* ---- optimiser (AdamGpu, adam_gpu.rs:334-541) ------------------------- */
/* AdamGpu::step on a flat slab of n floats (adam_step.wgsl). */
int nts_adam_step_flat(nts_ctx* ctx, nts_buffer* params, const nts_buffer* grads, nts_buffer* m,
                       nts_buffer* v, uint64_t n, const nts_adam_cfg* cfg, uint64_t t);
/* AdamGpu::step_all (adam_gpu.rs:334-427): Adam on every attribute plane
 * of the scene (positions with cfg_pos, the rest with cfg_attr, bias
 * corrections with step t).  Gradients come from the scene. */
int nts_adam_step_all(nts_ctx* ctx, nts_scene* scene, const nts_adam_cfg* cfg_pos, const nts_adam_cfg* cfg_attr,
                      uint64_t t);
/* AdamGpu::apply_constraints (adam_gpu.rs:431-489, apply_constraints.wgsl):
 * quaternion renormalisation (max(|q|, 1e-8)), log-scales clamped to
 * [log_scale_min, log_scale_max], logits to +-logit_max.  Works on
 * parameter-only scenes too. */
int nts_apply_constraints(nts_ctx* ctx, nts_scene* scene, const nts_constraints* cons);
/* AdamGpu::accumulate_grad (adam_gpu.rs:492-541, accumulate_grad_pos.wgsl):
 * grad_acc += |dx| + |dy| + |dz| of the scene's position gradient. */
int nts_accumulate_grad(nts_ctx* ctx, nts_scene* scene);
/* accumulate_grad + step_all + apply_constraints fused in one kernel:
 * grad_acc += |dpos|_1, Adam on every attribute (positions use cfg_pos,
 * the rest cfg_attr, bias corrections with step t), then quaternion
 * renormalisation and the log-scale / logit clamps. */
int nts_adam_step(nts_ctx* ctx, nts_scene* scene, const nts_adam_cfg* cfg_pos, const nts_adam_cfg* cfg_attr,
                  uint64_t t, const nts_constraints* cons);

/* ---- batched multi-view step and render (no reference counterpart: the
 *      reference renders one view per dispatch chain, train.rs:155-250) -- */
typedef struct nts_batch nts_batch;

/* One view of a batch: camera (CameraUniform layout), depth range of the
 * 16-bit depth keys, image size (must equal cam.width / cam.height). */
typedef struct nts_view_desc {
  nts_camera cam;
  float depth_max;
  uint32_t width, height;
  uint32_t pad;
} nts_view_desc;

#define NTS_STEP_GRAPH 1        /* capture the step once, replay it (same views) */
#define NTS_STEP_WRITE_GRADS 2  /* also store the summed parameter gradients */

/* Workspace for up to max_views views (<= 32) of a scene of <= max_splats
 * splats, <= max_pixels pixels in total and <= pair_capacity (splat, tile)
 * pairs.  All per-step memory comes from it: a step allocates nothing. */
int nts_batch_create(nts_ctx* ctx, uint32_t max_views, uint32_t max_splats, uint64_t max_pixels,
                     uint64_t pair_capacity, nts_batch** out);
int nts_batch_destroy(nts_ctx* ctx, nts_batch* batch);
/* One training iteration over n_views views in single launches per stage:
 * project + count, scan, emit, view-segmented sort, ranges, blend forward
 * with an L1 or MSE loss fused (targets: float4 per pixel, views
 * concatenated; NTS_LOSS_SSIM adds two tiled SSIM passes over all views
 * after the blend), blend backward, then per-splat project_backward summed
 * over views fused with accumulate_grad + Adam (step t) + constraints.
 * The per-view losses are reported by nts_batch_losses.
 * Nothing synchronises with the host; with NTS_STEP_GRAPH the step is
 * captured in a hipGraph on first use and replayed while the views and
 * settings stay the same and t advances by one.  Pairs beyond
 * pair_capacity are dropped and flagged (nts_batch_stats); a step whose
 * pairs overflowed, or whose sort look-back timed out, skips its update on
 * the device (parameters, moments, gradients and grad_acc keep their values;
 * the step number t is still used up) and latches the error for
 * nts_batch_stats. */
int nts_train_step_views(nts_ctx* ctx, nts_scene* scene, nts_batch* batch, const nts_view_desc* views,
                         uint32_t n_views, const nts_buffer* targets, uint32_t tile_size, int loss_kind,
                         int bwd_mode, const nts_adam_cfg* cfg_pos, const nts_adam_cfg* cfg_attr, uint64_t t,
                         const nts_constraints* cons, int flags);
/* Forward only (project, bin, blend) of n_views views into the batch image. */
int nts_render_views(nts_ctx* ctx, const nts_scene* scene, nts_batch* batch, const nts_view_desc* views,
                     uint32_t n_views, uint32_t tile_size);
/* Splats staged per LDS batch by the blend forward (64, 128, 256 or 512)
 * and backward (64, 128 or 256) of this batch's steps; 256 / 256 by
 * default.  No reference counterpart: the reference stages 256
 * (rasterize.wgsl:40). */
int nts_batch_set_lds_batch(nts_ctx* ctx, nts_batch* batch, uint32_t fwd_batch, uint32_t bwd_batch);
/* How the batch's blend kernels take their tiles.  Both take them most pairs
 * first (the order TileBinner-style ranges give, 4 buckets per octave).
 * NTS_SCHED_GRID (the default): one workgroup per (view, tile), the
 * reference's dispatch (raster.rs:526,721) in that order.
 * NTS_SCHED_PERSISTENT: as many workgroups as the device holds at once,
 * each taking the next tile from a device counter until none is left.
 * Images and gradient terms do not depend on the choice.  No reference
 * counterpart. */
#define NTS_SCHED_GRID 0
#define NTS_SCHED_PERSISTENT 1
int nts_batch_set_scheduler(nts_ctx* ctx, nts_batch* batch, int scheduler);
/* Binning rule of the batch's steps and renders.  NTS_BIN_CULLED (the
 * default): a splat gets pairs only for the tiles of its alpha >= 1/255
 * box, and a pair of such a tile that the alpha >= 1/255 ellipse cannot
 * reach is keyed all ones (sorted past every tile, in no range); no pair
 * that can change a pixel is dropped.  NTS_BIN_REFERENCE: the pairs of
 * TileBinner::bin (tile_binner.rs:181-382, tile_count.wgsl / tile_emit.wgsl),
 * so each view's keys, payloads and ranges (nts_batch_download_bins) are
 * the reference's.  Images are identical under both rules. */
#define NTS_BIN_CULLED 0
#define NTS_BIN_REFERENCE 1
int nts_batch_set_binning(nts_ctx* ctx, nts_batch* batch, int binning);
/* Evaluation counting (no reference counterpart; a measurement aid, off by
 * default).  While on, the batch's steps run the grid-scheduler blend
 * kernels in a form that also counts their work, and nts_batch_eval_counts
 * returns the last step's counts: [0] (splat, wave) evaluations of the
 * forward (a wave of 64 lanes testing one staged splat against its pixels),
 * [1] (pixel, splat) pairs the forward blended, [2] (splat, wave)
 * evaluations of the backward, [3] (pixel, splat) pairs that took part in
 * the backward.  Images and gradients are unchanged; the counting form is
 * slower and is not the one timed.  With NTS_SCHED_PERSISTENT nothing is
 * counted (the counts read 0). */
int nts_batch_set_eval_counting(nts_ctx* ctx, nts_batch* batch, int on);
int nts_batch_eval_counts(nts_ctx* ctx, const nts_batch* batch, uint64_t counts[4]);
/* Blend factor lambda of NTS_LOSS_SSIM steps, (1 - lambda) MSE + lambda
 * (1 - SSIM) (TrainConfig::ssim_lambda, train.rs:48-52; 0.2 by default,
 * main.rs:189-190), clamped to [0, 1].  The first lambda > 0 allocates the
 * SSIM planes (60 B per pixel of max_pixels), so the steps allocate
 * nothing; NTS_LOSS_SSIM steps fail until then.  lambda = 0 makes such
 * steps plain MSE (loss.rs:398-402).  nts_batch_losses then reports the
 * combined per-view loss. */
int nts_batch_set_ssim_lambda(nts_ctx* ctx, nts_batch* batch, float lambda);
int nts_batch_losses(nts_ctx* ctx, const nts_batch* batch, double* losses);
/* Pairs and visible (splat, view) records of the last call.  *overflow is
 * nonzero when any call since the previous nts_batch_stats overflowed the
 * pair capacity: the number of training steps whose update was skipped for
 * it (1 for a forward-only overflow).  A sort look-back timeout since then
 * is NTS_E_HIP.  Reading clears the latched errors. */
int nts_batch_stats(nts_ctx* ctx, const nts_batch* batch, uint64_t* total_pairs, uint32_t* visible,
                    uint32_t* overflow);
int nts_batch_download_image(nts_ctx* ctx, const nts_batch* batch, uint32_t view, float* host);
int nts_batch_download_state(nts_ctx* ctx, const nts_batch* batch, uint32_t view, float* final_t,
                             uint32_t* last_idx);
int nts_batch_download_projected_grad(nts_ctx* ctx, const nts_batch* batch, uint32_t view, uint32_t n,
                                      float* host);
/* Keys (reference u32 format, view-local tile id), payloads and ranges of
 * one view's segment; *count receives the view's pair count. */
int nts_batch_download_bins(nts_ctx* ctx, const nts_batch* batch, uint32_t view, uint64_t* count, uint32_t* keys,
                            uint32_t* payloads, uint32_t* ranges);

/* ---- render server (no reference counterpart: the viewer renders one view
 *      per frame, nano-view/src/lib.rs:761-796) --------------------------- */
/* Many scenes resident in HBM; each request is one camera of one scene and
 * is admitted into the batch being formed while the batch's views, pixels,
 * projected records, tiles and estimated (splat, tile) pairs fit a slot's
 * workspace.  `slots` batches are in flight at once, each on its own
 * stream; a dispatcher thread retires finished batches and refills free
 * slots from the FIFO of pending requests (continuous batching).  Every
 * result is bit-identical to rendering its request alone. */
typedef struct nts_server nts_server;
/* Admission by memory: each slot owns one device workspace of slot_bytes,
 * and a batch lays out its working set inside it (projected records, the
 * (splat, tile) pairs it expects, their sort workspace, its pixels).  A
 * request joins the batch being formed while that working set still fits
 * (pairs estimated from the scene's earlier renders) and the caps hold, so
 * a 4K view of a large scene takes most of a slot while many small views
 * share one; the rest wait for the next free slot.  With slot_bytes = 0 the
 * workspace is sized to hold every cap at once (then all caps must be set). */
typedef struct nts_server_cfg {
  uint64_t hbm_budget;     /* bytes for resident scenes + slot workspaces (0 = no limit) */
  uint32_t slots;          /* batches in flight (1..8) */
  uint32_t max_views;      /* requests per batch (<= 32) */
  uint64_t max_pixels;     /* per batch (sizes the pinned result staging) */
  uint64_t max_records;    /* per batch: sum of the requested scenes' splat counts (0 = no cap) */
  uint64_t pair_capacity;  /* per batch (splat, tile) pairs (0 = what the workspace holds) */
  uint32_t tile_size;      /* 8, 16 or 32 */
  uint32_t lds_batch;      /* blend forward LDS batch: 64, 128, 256 or 512 */
  uint64_t slot_bytes;     /* each slot's workspace (0 = sized from the caps) */
} nts_server_cfg;
typedef struct nts_serve_stats {
  uint64_t submitted, completed, failed;
  uint64_t batches, views, overflow_reruns;
  uint32_t max_in_flight, scenes, pending;
  uint32_t memory_limited; /* batches closed because the next request's working set did not fit */
  uint64_t hbm_in_use;
} nts_serve_stats;
typedef struct nts_request_info {
  double latency_ms;     /* submit to result in the caller's buffer */
  uint64_t pairs;        /* (splat, tile) pairs of the view */
  uint32_t batch_views;  /* views in the batch that rendered it */
  int status;            /* NTS_OK or the failure code */
} nts_request_info;
#define NTS_OUT_RGBA8 0  /* tonemapped (tonemap.wgsl), 4 B per pixel */
#define NTS_OUT_F32 1    /* (r, g, b, 1 - T) float, 16 B per pixel */
/* Free and total memory of a device (hipMemGetInfo), for sizing a server's
 * hbm_budget and slot_bytes (the reference's gpu-mem probe,
 * crates/gpu-mem/src/lib.rs:70, reads sysfs instead). */
int nts_device_memory(int device, uint64_t* free_bytes, uint64_t* total_bytes);
int nts_server_create(int device, const nts_server_cfg* cfg, nts_server** out);
int nts_server_destroy(nts_server* server);
/* Uploads a resident scene (parameters only); NTS_E_NOMEM over the budget. */
int nts_server_add_scene(nts_server* server, const nts_splats_host* splats, uint32_t* scene_id);
/* The same from a binary little-endian 3DGS PLY file (nts_ply_read): the
 * serving loader for scenes on disk. */
int nts_server_add_scene_ply(nts_server* server, const char* path, uint32_t* scene_id);
/* Queues a render; `out` (host, width * height * 4 or 16 bytes) must stay
 * valid until the request is polled done or waited for. */
int nts_server_submit(nts_server* server, uint32_t scene_id, const nts_view_desc* view, int format, void* out,
                      uint64_t* request_id);
/* Non-blocking: *done = 1 once finished (the request is then released). */
int nts_server_poll(nts_server* server, uint64_t request_id, int* done, nts_request_info* info);
int nts_server_wait(nts_server* server, uint64_t request_id, nts_request_info* info);
int nts_server_stats(nts_server* server, nts_serve_stats* out);

/* ---- prune / densify (train.rs:414-602, run on the device) -------------- */
/* AdamGpu::zero_grad_acc (adam_gpu.rs:492-541): grad_acc = 0. */
int nts_scene_zero_grad_acc(nts_ctx* ctx, nts_scene* scene);
/* prune_low_opacity (train.rs:572-602): removes splats with logit < threshold
 * (the reference uses -5.3) from parameters, both Adam moments and grad_acc,
 * leaving exactly the order of the reference's descending swap_remove.
 * *removed receives the count.  Synchronises with the host. */
int nts_prune(nts_ctx* ctx, nts_scene* scene, float logit_threshold, uint32_t* removed);
/* densify (train.rs:414-549): candidates grad_acc / max(grad_count, 1) >
 * grad_threshold in index order up to max_splats - n; clone when the largest
 * log-scale is <= split_log_scale (copy appended, both logits + ln 1/2),
 * else split (two children at parent + q (N(0,1) sigma), log-scales - ln 1.6,
 * the LCG stream seeded iter * 1664525 + 1013904223; parents removed by
 * descending swap_remove).  New splats get zero moments and grad_acc.
 * counts = {cloned, split children, parents removed}.  The scene needs room
 * for n + clones + 2 splits (NTS_E_CAPACITY otherwise).  grad_acc is left as
 * the reference's densify leaves it; the reference zeroes it next
 * (train.rs:327-331): nts_scene_zero_grad_acc.  Synchronises with the host. */
int nts_densify(nts_ctx* ctx, nts_scene* scene, uint32_t grad_count, float grad_threshold, float split_log_scale,
                uint32_t max_splats, uint32_t iter, uint32_t counts[3]);

/* ---- 3DGS PLY files (crates/nano-splat/src/ply.rs:17-256; host only) ---- */
/* Vertex count of a binary little-endian 3DGS PLY (read_ply's header). */
int nts_ply_count(const char* path, uint32_t* n);
/* read_ply into caller arrays holding out->n >= the count splats; out->n
 * is set to the count.  Properties are mapped by name; missing normals are
 * allowed, missing f_rest_i read as 0, other missing properties, non-float
 * properties and non-binary files are NTS_E_INVALID. */
int nts_ply_read(const char* path, nts_splats_host* out);
/* write_ply: the reference's header and record order, normals (0, 1, 0). */
int nts_ply_write(const char* path, const nts_splats_host* splats);

/* ---- viewer (Rasterizer::tonemap, raster.rs:405-455) ------------------- */
int nts_tonemap(nts_ctx* ctx, const nts_buffer* image, nts_buffer* rgba8, uint32_t width, uint32_t height);

#ifdef __cplusplus
}
#endif
#endif /* NTS_H */
