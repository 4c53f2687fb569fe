// Shared definitions of the tile alpha-blend kernels.
#pragma once
#include "nts_internal.h"

#define NTS_MAX_VIEWS 32

// One view of a launch: image size, tile grid, and where its tiles, its
// projected records and its pixels start in the launch-wide arrays.
struct ViewDesc {
  uint32_t width, height;
  uint32_t tiles_x, tiles_y;
  uint32_t tile_base;   // first global tile id of the view
  uint32_t proj_base;   // offset of the view's records in projected[] / projected_grad[]
  uint64_t pix_base;    // offset of the view's first pixel in the image arrays
  float inv_n;          // loss scale: 1/(3WH) for L1, 2/(3WH) for MSE
  uint32_t pad;
};

// Optional loss epilogue of the forward blend: dL/dC and the per-view loss
// sum (the reference computes these on the host, train.rs:180-196).
struct LossArgs {
  const float4* target;  // float4 per pixel (w ignored); nullptr = no loss
  float4* dl_dc;
  double* losses;        // per view, accumulated
  int kind;              // NTS_LOSS_L1 or NTS_LOSS_MSE
  // Optional buffer the launch clears to zero behind the blend, a slice per
  // workgroup after its last barrier (the training step's projected
  // gradients, which the backward then accumulates into); nullptr = none.
  float4* clear;
  uint64_t clear_n4;     // float4 words of `clear`
};

// Passed by value as a kernel argument (kernarg segment), so launches need
// no upload and stay capturable.
struct ViewTable {
  uint32_t n;
  uint32_t total_tiles;
  const uint32_t* order;  // global tile of each workgroup (nullptr: grid order)
  // Persistent tile scheduler (nullptr: one workgroup per tile): a grid of
  // resident workgroups takes tiles, in `order`, from these two counters
  // (forward, backward), which must be zero when the launch starts.
  uint32_t* tickets;
  // Evaluation counters (nullptr: none), added into by the counting form of
  // the blend kernels: [0] (splat, wave) evaluations of the forward, [1] its
  // blended (pixel, splat) pairs, [2] (splat, wave) evaluations of the
  // backward, [3] its contributing (pixel, splat) pairs.
  unsigned long long* counts;
  // Words per projected-gradient record the backward adds into:
  // NTS_GRAD_WORDS (the ProjectedGrad layout) or NTS_GRAD9_WORDS (compact).
  uint32_t grad_words;
  ViewDesc v[NTS_MAX_VIEWS];
};

// Forms of the blend kernels: one workgroup per tile, the persistent tile
// scheduler, or one workgroup per tile with the evaluation counters.
#define NTS_BLEND_GRID 0
#define NTS_BLEND_PERSIST 1
#define NTS_BLEND_COUNT 2

// View of global tile gt (tile_base ascending).
__device__ __forceinline__ uint32_t view_of_tile(const ViewTable& vt, uint32_t gt) {
  uint32_t lo = 0, hi = vt.n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (vt.v[mid].tile_base <= gt) lo = mid; else hi = mid;
  }
  return lo;
}

// e^x for x <= 0: the arithmetic of nts_expf with the branches that cannot
// fire for non-positive arguments removed (bit-identical on that domain).
__device__ __forceinline__ float nts_expf_nonpos(float x) {
  if (!(x > -104.0f)) return (x != x) ? x : 0.0f;
  const float n = __builtin_rintf(x * 1.44269504088896341f);
  float r = nts_fmaf(n, -0.693359375f, x);
  r = nts_fmaf(n, 2.12194440e-4f, r);
  float p = nts_exp_poly(r);
  int k = (int)n;
  if (k < -125) {
    p = p * nts_pow2i(-64);
    k += 64;
  }
  return p * nts_pow2i(k);
}

// e^x for the blend's power test, x <= 0: equal to nts_expf for
// x >= -80 (the same reduction and polynomial; the result there is a normal
// number, so the scaling is one ldexp).  Below -80 it returns e^-80, which
// like e^x gives alpha = opacity e^x < 1/255 for any opacity <= 1: every
// blend decision, and so every image, is the one nts_expf gives.  No range,
// NaN or subnormal branches.
__device__ __forceinline__ float nts_exp_blend(float x) {
  x = __builtin_amdgcn_fmed3f(x, -80.0f, 0.0f);  // one clamp instruction; x <= 0 on entry
  const float n = __builtin_rintf(x * 1.44269504088896341f);
  float r = nts_fmaf(n, -0.693359375f, x);
  r = nts_fmaf(n, 2.12194440e-4f, r);
  return __builtin_ldexpf(nts_exp_poly(r), (int)n);
}

// Splats staged through LDS per batch when the caller does not choose: the
// forward's is within a few percent of the best at tiles 16 and 32; a
// backward batch of 256 is the slowest there (its per-wave partial sums
// grow with the batch), 128 the best or near it.
#define NTS_DEFAULT_FWD_BATCH 256
#define NTS_DEFAULT_BWD_BATCH 128
// Where the tile size is known: 64 below tile 32 (4% faster than 128 on the
// single-view 1080p step at tile 16), 128 at tile 32.
static inline uint32_t nts_default_bwd_batch(uint32_t tile_size) { return tile_size >= 32 ? 128u : 64u; }

// Host-side launchers of the blend kernels (k_raster.hip).  tile_size in
// {8, 16, 32}; lds_batch in {64, 128, 256, 512} forward, {64, 128, 256}
// backward.
bool nts_blend_config_ok(uint32_t tile_size, uint32_t fwd_batch, uint32_t bwd_batch);
// tonemap.wgsl:27-41 over np pixels (rgba8 out).
int nts_tonemap_dev(nts_ctx* ctx, const float4* image, uint32_t* rgba8, uint64_t np);
// The blend order of n_tiles tiles (ViewTable::order): most pairs first,
// from the tile ranges.  With n_heavy set, the number of tiles of at least
// NTS_TILE_HEAVY pairs (the first *n_heavy of the order) goes to that device
// word.
#define NTS_TILE_HEAVY 2048u
int nts_tile_order(nts_ctx* ctx, const uint2* ranges, uint32_t n_tiles, uint32_t* order,
                   uint32_t* n_heavy = nullptr);
int nts_blend_forward(nts_ctx* ctx, uint32_t tile_size, uint32_t lds_batch, const Projected* proj,
                      const uint32_t* payloads, const uint2* ranges, const ViewTable& vt, float4* out, float* final_t,
                      uint32_t* last_idx, const LossArgs* loss);
// pgrad: records of vt.grad_words words per (view, splat), added into.
int nts_blend_backward(nts_ctx* ctx, uint32_t tile_size, uint32_t lds_batch, int mode, const Projected* proj,
                       const uint32_t* payloads, const uint2* ranges, const ViewTable& vt, const float4* fwd_out,
                       const float* final_t, const uint32_t* last_idx, const float4* dl_dc, float* pgrad);

// ---- combined MSE + SSIM loss (k_ssim.hip) ---------------------------------
// Normalised 11-tap Gaussian window (sigma 1.5), taps computed on the host.
struct SsimKernel {
  float k[11];
};
// Views of one SSIM launch (kernel argument): 64 x 16 pixel tiles per view,
// tile_base prefix over the views, pixel offsets into the concatenated
// images, and the MSE gradient scale 2 / (3WH) per view.
struct SsimViews {
  uint32_t n;
  uint32_t tile_base[NTS_MAX_VIEWS + 1];
  uint32_t tiles_x[NTS_MAX_VIEWS];
  uint32_t w[NTS_MAX_VIEWS], h[NTS_MAX_VIEWS];
  float inv_mse[NTS_MAX_VIEWS];
  uint64_t pix_base[NTS_MAX_VIEWS];
};
int nts_ssim_views(uint32_t n_views, const uint32_t* widths, const uint32_t* heights, const uint64_t* pix_base,
                   SsimViews* out);
// dL/dC of (1 - lam) MSE + lam (1 - SSIM) for every view (0 < lam <= 1).
// planes: 15 float planes of `plane` floats (plane >= the pixels of all
// views); sums: 4 doubles per view (SSIM r, g, b, squared error), cleared
// here.  losses (optional): the per-view total, written on the device.
int nts_ssim_launch(nts_ctx* ctx, const SsimViews& sv, const float4* pred, const float4* target, float* planes,
                    uint64_t plane, double* sums, float4* dl_dc, float lam, double* losses);
