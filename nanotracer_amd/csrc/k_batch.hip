// Multi-view training step and multi-view render over one scene.
//
// Every stage is one launch across all views of the batch:
//   project + tile count   one thread per splat, parameters read once, every
//                          view projected from registers;
//   scan                   over the (view, splat) counts;
//   emit                   (view-local tile << 16 | depth16) keys, one segment
//                          per view;
//   binning sort           onesweep over the tile bits, segmented by view,
//                          the tile ranges (global tile ids: view tile base +
//                          local tile), the tile order (by pair count,
//                          heaviest first), then each tile by depth in LDS;
//   blend forward          one workgroup per (view, tile) in that order, loss fused in the
//                          epilogue (dL/dC and per-view loss);
//   blend backward         per (view, tile), per-view projected gradients;
//   project_backward+Adam  one thread per splat: the views' parameter
//                          gradients are summed in registers and fed straight
//                          into Adam + constraints + grad_acc.
// All buffers live in a preallocated workspace and nothing synchronises with
// the host, so the whole step can be captured once in a hipGraph and replayed.
#include <cstring>
#include <vector>

#include "nts_adam.h"
#include "nts_bin.h"
#include "nts_project.h"
#include "nts_raster.h"

struct nts_batch {
  uint32_t max_views = 0, max_splats = 0;
  uint64_t max_pixels = 0, pair_cap = 0, max_tiles = 0;
  void* base = nullptr;
  uint64_t bytes = 0;
  DevCamera* d_cams = nullptr;
  Projected* proj = nullptr;
  float* pgrad = nullptr;  // compact per-(view, splat) gradient records (NTS_GRAD9_WORDS words)
  uint32_t* counts = nullptr;  // exclusive offsets after the scan
  uint64_t* d_total = nullptr;
  uint64_t* seg_begin = nullptr;
  uint32_t* tile_base = nullptr;
  uint32_t* flags = nullptr;    // [0] overflow, [1] visible count, [2..3] total copy
  // Errors latched since the last nts_batch_stats: bit 0 a pair overflow,
  // bit 1 a sort look-back timeout; [1] counts the steps whose Adam update
  // was skipped for either.  Only nts_batch_stats clears it.
  uint32_t* sticky = nullptr;
  uint64_t* d_step = nullptr;   // Adam step counter for captured replays
  void* keys_a = nullptr;
  void* keys_b = nullptr;
  uint32_t* vals_a = nullptr;
  uint32_t* vals_b = nullptr;
  uint2* ranges = nullptr;
  uint32_t* order = nullptr;  // the blend kernels' tile order (most pairs first)
  float4* image = nullptr;
  float4* dl_dc = nullptr;
  float* final_t = nullptr;
  uint32_t* last_idx = nullptr;
  double* losses = nullptr;
  void* scan_ws = nullptr;
  void* sort_ws = nullptr;
  uint32_t* sort_err = nullptr;
  DevCamera* h_cams = nullptr;  // pinned staging
  // last launch shape (for downloads)
  uint32_t n_views = 0;
  uint64_t pix_base[NTS_MAX_VIEWS + 1] = {};
  uint32_t width[NTS_MAX_VIEWS] = {}, height[NTS_MAX_VIEWS] = {};
  // captured step
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<uint8_t> graph_key;
  uint64_t graph_last_t = 0;
  bool last_alt = false;
  uint32_t last_key_bytes = 4;
  uint32_t last_tile_base[NTS_MAX_VIEWS + 1] = {};
  // splats staged per LDS batch by the blend kernels
  uint32_t fwd_batch = NTS_DEFAULT_FWD_BATCH, bwd_batch = NTS_DEFAULT_BWD_BATCH;
  // tile scheduler of the blend kernels (NTS_SCHED_*)
  int scheduler = NTS_SCHED_GRID;
  // blend kernels in their counting form (evaluation counters in flags[8..15])
  bool count_evals = false;
  // binning rule (NTS_BIN_*): the alpha >= 1/255 region, or the reference's box
  int binning = NTS_BIN_CULLED;
  // combined MSE + SSIM loss of NTS_LOSS_SSIM steps (planes allocated by the setter)
  float ssim_lambda = 0.2f;
  void* ssim_base = nullptr;
  uint64_t ssim_bytes = 0;
  bool last_loss_final = false;  // losses[] already hold the per-view loss
};

struct U32Table {
  uint32_t n;
  uint32_t v[NTS_MAX_VIEWS + 1];
};

__global__ void k_set_u32s(uint32_t* __restrict__ dst, U32Table t) {
  const uint32_t i = threadIdx.x;
  if (blockIdx.x == 0 && i < t.n) dst[i] = t.v[i];
}

// Latches this step's pair overflow (flags[0]) and sort look-back timeout
// (*sort_err) into the sticky error word; it runs after the sort, so both
// are final.
__global__ void k_latch_errors(const uint32_t* __restrict__ flags, const uint32_t* __restrict__ sort_err,
                               uint32_t* __restrict__ sticky) {
  const uint32_t i = threadIdx.x;
  if (blockIdx.x != 0) return;
  if (i == 0) {
    const uint32_t bits = (flags[0] ? 1u : 0u) | (*sort_err ? 2u : 0u);
    if (bits) atomicOr(sticky, bits);
  }
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// Held to 4 waves per SIMD (125 VGPRs, no spills; the compiler's own choice
// is 139 VGPRs and 3 waves): 6% faster over 1M splats x 16 views.
template <bool CULL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_project_views(const float* __restrict__ P, uint32_t cap, uint32_t n,
                                                       const DevCamera* __restrict__ cams, TileTable tt,
                                                       Projected* __restrict__ proj, uint32_t* __restrict__ counts,
                                                       uint32_t* __restrict__ visible) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t vis = 0;
  if (i < n) {
    SplatP s;
    load_splat(P, cap, i, s);
    for (uint32_t v = 0; v < tt.n; ++v) {
      const DevCamera cam = cams[v];
      const Projected p = project_p(s, cam);
      proj[(size_t)v * n + i] = p;
      uint32_t c = 0;
      int32_t x0, y0, x1, y1;
      const ViewTiles vt = tt.t[v];
      // CULL: only the tiles the splat's alpha >= 1/255 ellipse can reach;
      // otherwise the reference's tile rectangle (the emit uses the same rule)
      if (record_tiles<CULL>(p, (float)vt.tile_size, (int32_t)vt.tiles_x, (int32_t)vt.tiles_y, x0, y0, x1, y1)) {
        c = (uint32_t)((x1 - x0 + 1) * (y1 - y0 + 1));
        ++vis;
      }
      counts[(size_t)v * n + i] = c;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) vis += __shfl_xor(vis, off, 64);
  if ((threadIdx.x & 63u) == 0 && vis) atomicAdd(visible, vis);
}

// The order in which the blend kernels take their tiles: most pairs first.
// The hardware hands out a launch's workgroups in grid order, so a heavy
// tile late in the grid would run on alone at the end of the launch; taken
// first, the light tiles fill in behind it.  Tiles are bucketed by pair count,
// four buckets per octave, in descending bucket order (within a bucket the
// order is what the LDS atomics give; each tile's image and gradient terms
// do not depend on it).  One workgroup: count, exclusive scan from the top
// bucket down, scatter.
#define NTS_ORDER_BUCKETS 132
__device__ __forceinline__ uint32_t tile_order_bucket(uint32_t pairs) {
  if (pairs == 0) return 0;
  const uint32_t e = 31u - (uint32_t)__builtin_clz(pairs);  // floor(log2(pairs))
  const uint32_t f = e >= 2 ? (pairs >> (e - 2)) & 3u : (pairs << (2 - e)) & 3u;
  return 1u + 4u * e + f;  // <= 128
}

__global__ void __launch_bounds__(1024) k_tile_order(const uint2* __restrict__ ranges, uint32_t n_tiles,
                                                     uint32_t* __restrict__ order, uint32_t* __restrict__ n_heavy) {
  __shared__ uint32_t s_cnt[NTS_ORDER_BUCKETS];
  __shared__ uint32_t s_pos[NTS_ORDER_BUCKETS];
  for (uint32_t i = threadIdx.x; i < NTS_ORDER_BUCKETS; i += blockDim.x) s_cnt[i] = 0;
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < n_tiles; t += blockDim.x) {
    const uint2 r = ranges[t];
    atomicAdd(&s_cnt[tile_order_bucket(r.y - r.x)], 1u);
  }
  __syncthreads();
  if (threadIdx.x < NTS_ORDER_BUCKETS) {
    uint32_t above = 0;
    for (uint32_t j = threadIdx.x + 1; j < NTS_ORDER_BUCKETS; ++j) above += s_cnt[j];
    s_pos[threadIdx.x] = above;
    // tiles of NTS_TILE_HEAVY pairs or more: the buckets from its own up
    // (a power of two starts a bucket)
    if (n_heavy && threadIdx.x == tile_order_bucket(NTS_TILE_HEAVY)) *n_heavy = above + s_cnt[threadIdx.x];
  }
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < n_tiles; t += blockDim.x) {
    const uint2 r = ranges[t];
    order[atomicAdd(&s_pos[tile_order_bucket(r.y - r.x)], 1u)] = t;
  }
}

int nts_tile_order(nts_ctx* ctx, const uint2* ranges, uint32_t n_tiles, uint32_t* order, uint32_t* n_heavy) {
  if (n_tiles == 0) return n_heavy ? nts_fill_u32(ctx, n_heavy, 0, 1) : NTS_OK;
  NTS_LAUNCH(ctx, "tile_order", k_tile_order, dim3(1), dim3(1024), 0, ranges, n_tiles, order, n_heavy);
  return NTS_OK;
}

// Segment bounds per view from the scanned counts, clamped to the pair
// capacity (overflow is flagged, never written past).
__global__ void k_view_segments(const uint32_t* __restrict__ offsets, RecTable rt, const uint64_t* __restrict__ d_total,
                                uint64_t cap, uint64_t* __restrict__ seg_begin, uint32_t* __restrict__ flags) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint64_t total = *d_total;
  const uint32_t nrec = rt.base[rt.n];
  for (uint32_t v = 0; v < rt.n; ++v) {
    const uint64_t b = rt.base[v] < nrec ? (uint64_t)offsets[rt.base[v]] : total;
    seg_begin[v] = b < cap ? b : cap;
  }
  seg_begin[rt.n] = total < cap ? total : cap;
  flags[0] = total > cap ? 1u : 0u;
  flags[2] = (uint32_t)total;
  flags[3] = (uint32_t)(total >> 32);
}

__global__ void k_set_step(uint64_t* d_step, uint64_t t) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *d_step = t;
}
__global__ void k_inc_step(uint64_t* d_step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *d_step += 1;
}

// Bias correction 1 - beta^t with the host powi arithmetic (device side so
// that captured replays advance t without new kernel arguments).
__device__ __forceinline__ float dev_bias(float beta, uint64_t t) {
  int32_t b = (int32_t)t;
  float a = beta, r = 1.0f;
  while (1) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return 1.0f - r;
}

struct AdamCfgs {
  float lr_pos, lr_attr, b1_pos, b2_pos, b1_attr, b2_attr, eps_pos, eps_attr;
  float smin, smax, omax;
};

// Per splat: the V views' projected gradients chained back to the
// parameters and summed, then accumulate_grad + Adam + constraints.  The
// geometry (through the 2-D mean, conic and opacity) is summed over the
// views first, then the colour one channel at a time (its SH planes take
// their Adam step as soon as their sums are complete), then the position,
// rotation, scale and opacity planes: no more than 23 parameters and sums
// are live across a view loop, so several waves share each SIMD.
#define ADAM_LDS_VIEWS 16
// A compact gradient record (NTS_GRAD9_WORDS words) as the three float4s of
// the ProjectedGrad layout.
__device__ __forceinline__ ProjectedGrad load_grad9(const float* __restrict__ g) {
  ProjectedGrad r;
  r.a = make_float4(g[0], g[1], g[2], 0.0f);
  r.b = make_float4(g[3], g[4], g[5], 0.0f);
  r.c = make_float4(g[6], g[7], g[8], 0.0f);
  return r;
}
template <bool FULL>
__global__ void __launch_bounds__(256) k_backward_adam_views(float* __restrict__ P, float* __restrict__ G,
                                                             float* __restrict__ M, float* __restrict__ V,
                                                             float* __restrict__ grad_acc, uint32_t cap, uint32_t n,
                                                             uint32_t n_views, const DevCamera* __restrict__ cams,
                                                             const float* __restrict__ pgrad,
                                                             const uint64_t* __restrict__ d_step, AdamCfgs cfg,
                                                             int write_grads, const uint32_t* __restrict__ flags,
                                                             const uint32_t* __restrict__ sort_err,
                                                             uint32_t* __restrict__ sticky) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  // A step whose pairs overflowed the capacity or whose sort timed out has
  // truncated or misordered pairs: its update is skipped (parameters, moments,
  // gradients and grad_acc keep their values) and counted in the sticky word.
  if (flags[0] | *sort_err) {
    if (i == 0) atomicAdd(sticky + 1, 1u);
    return;
  }
  if (i >= n) return;
  const size_t c = cap;
  const uint64_t t = *d_step;
  AdamK kp, ka;
  kp.lr = cfg.lr_pos; kp.b1 = cfg.b1_pos; kp.b2 = cfg.b2_pos; kp.eps = cfg.eps_pos;
  kp.bc1 = dev_bias(cfg.b1_pos, t); kp.bc2 = dev_bias(cfg.b2_pos, t);
  ka.lr = cfg.lr_attr; ka.b1 = cfg.b1_attr; ka.b2 = cfg.b2_attr; ka.eps = cfg.eps_attr;
  ka.bc1 = dev_bias(cfg.b1_attr, t); ka.bc2 = dev_bias(cfg.b2_attr, t);
  // one plane's gradient into G and its Adam step; returns the new value
  auto step = [&](int k, float p, float gk) {
    const size_t j = (size_t)k * c + i;
    if (write_grads) G[j] = gk;
    float mm = M[j], vv = V[j];
    const float np = adam_lane(p, gk, mm, vv, k < P_ROT ? kp : ka);
    M[j] = mm;
    V[j] = vv;
    return np;
  };

  // ---- geometry over the views ----
  float q[11];  // pos, rot, log-scale, logit
#pragma unroll
  for (int k = 0; k < 3; ++k) q[k] = P[(size_t)(P_POS + k) * c + i];
#pragma unroll
  for (int k = 0; k < 4; ++k) q[3 + k] = P[(size_t)(P_ROT + k) * c + i];
#pragma unroll
  for (int k = 0; k < 3; ++k) q[7 + k] = P[(size_t)(P_SCALE + k) * c + i];
  q[10] = P[(size_t)P_LOGIT * c + i];
  float gq[11];
#pragma unroll
  for (int k = 0; k < 11; ++k) gq[k] = 0.0f;
  uint32_t kept = 0;  // views that keep the splat and passed it a gradient
  // Up to ADAM_LDS_VIEWS views, the colour terms read here are kept in LDS
  // for the colour pass (each thread reads back only its own words), which
  // otherwise reads them again once per channel.
  __shared__ float s_col[3][ADAM_LDS_VIEWS][256];
  const bool stash = n_views <= ADAM_LDS_VIEWS;
  {
    ProjectedGrad next = load_grad9(pgrad + (size_t)i * NTS_GRAD9_WORDS);  // the next view's record loads while this one is chained back
    for (uint32_t v = 0; v < n_views; ++v) {
      const ProjectedGrad pg = next;
      if (v + 1 < n_views) next = load_grad9(pgrad + ((size_t)(v + 1) * n + i) * NTS_GRAD9_WORDS);
      if (pg.a.x == 0.0f && pg.a.y == 0.0f && pg.a.z == 0.0f && pg.b.x == 0.0f && pg.b.y == 0.0f &&
          pg.b.z == 0.0f && pg.c.x == 0.0f && pg.c.y == 0.0f && pg.c.z == 0.0f)
        continue;  // the view gave this splat no gradient: its terms are all zero
      if (project_backward_geo<FULL>(q, cams[v], pg.a, pg.b, gq)) {
        kept |= 1u << v;
        if (stash) {
          s_col[0][v][threadIdx.x] = pg.c.x;
          s_col[1][v][threadIdx.x] = pg.c.y;
          s_col[2][v][threadIdx.x] = pg.c.z;
        }
      }
    }
  }

  // ---- colour, one channel at a time ----
  float3 psh = make_float3(0.0f, 0.0f, 0.0f);
#pragma unroll 1
  for (int ch = 0; ch < 3; ++ch) {
    const float dc = P[(size_t)(P_DC + ch) * c + i];
    float rest[15], grest[15];
#pragma unroll
    for (int k = 0; k < 15; ++k) {
      rest[k] = P[(size_t)(P_REST + 15 * ch + k) * c + i];
      grest[k] = 0.0f;
    }
    float gdc = 0.0f;
    for (uint32_t m = kept; m; m &= m - 1) {
      const uint32_t v = (uint32_t)__builtin_ctz(m);
      float dcol;
      if (stash) {
        dcol = s_col[ch][v][threadIdx.x];
      } else {
        dcol = pgrad[((size_t)v * n + i) * NTS_GRAD9_WORDS + 6 + ch];
      }
      project_backward_colour<FULL>(q[0], q[1], q[2], dc, rest, cams[v], dcol, gdc, grest, psh);
    }
    P[(size_t)(P_DC + ch) * c + i] = step(P_DC + ch, dc, gdc);
#pragma unroll
    for (int k = 0; k < 15; ++k) {
      const int pk = P_REST + 15 * ch + k;
      P[(size_t)pk * c + i] = step(pk, rest[k], grest[k]);
    }
  }

  // ---- position, rotation, scale, opacity ----
  if (FULL) {
    gq[0] = gq[0] - psh.x;
    gq[1] = gq[1] - psh.y;
    gq[2] = gq[2] - psh.z;
  }
  grad_acc[i] = grad_acc[i] + fabsf(gq[0]) + fabsf(gq[1]) + fabsf(gq[2]);
#pragma unroll
  for (int k = 0; k < 3; ++k) P[(size_t)(P_POS + k) * c + i] = step(P_POS + k, q[k], gq[k]);
  float nq[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) nq[k] = step(P_ROT + k, q[3 + k], gq[3 + k]);
  const float n2 = nq[0] * nq[0] + nq[1] * nq[1] + nq[2] * nq[2] + nq[3] * nq[3];
  const float nrm = nts_maxf(sqrtf(n2), 1e-8f);
#pragma unroll
  for (int k = 0; k < 4; ++k) P[(size_t)(P_ROT + k) * c + i] = nq[k] / nrm;
#pragma unroll
  for (int k = 0; k < 3; ++k)
    P[(size_t)(P_SCALE + k) * c + i] = nts_minf(nts_maxf(step(P_SCALE + k, q[7 + k], gq[7 + k]), cfg.smin), cfg.smax);
  P[(size_t)P_LOGIT * c + i] = nts_minf(nts_maxf(step(P_LOGIT, q[10], gq[10]), -cfg.omax), cfg.omax);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static uint64_t al256(uint64_t x) { return (x + 255) & ~255ull; }

extern "C" int nts_batch_create(nts_ctx* ctx, uint32_t max_views, uint32_t max_splats, uint64_t max_pixels,
                                uint64_t pair_capacity, nts_batch** out) {
  NTS_CHECK_ARG(ctx && out && max_views > 0 && max_views <= NTS_MAX_VIEWS && max_splats > 0 && max_pixels > 0);
  NTS_CHECK_ARG(pair_capacity > 0 && pair_capacity < (1ull << 30));
  *out = nullptr;
  nts_batch* b = new nts_batch();
  b->max_views = max_views;
  b->max_splats = max_splats;
  b->max_pixels = max_pixels;
  b->pair_cap = pair_capacity;
  // tiles at the smallest tile size (8): ceil per axis adds at most one row and column per view
  b->max_tiles = max_pixels / 64 + (uint64_t)max_views * 1024 + 64;
  const uint64_t vn = (uint64_t)max_views * max_splats;
  uint64_t off = 0;
  auto take = [&](uint64_t bytes) { const uint64_t o = off; off += al256(bytes); return o; };
  const uint64_t o_cams = take(sizeof(DevCamera) * max_views);
  const uint64_t o_proj = take(vn * sizeof(Projected));
  const uint64_t o_pgrad = take((vn * NTS_GRAD9_WORDS * 4 + 15) & ~15ull);  // whole float4s (the forward clears it)
  const uint64_t o_counts = take(vn * 4);
  const uint64_t o_total = take(8);
  const uint64_t o_seg = take(8 * (max_views + 1));
  const uint64_t o_tb = take(4 * (max_views + 1));
  const uint64_t o_flags = take(64);
  const uint64_t o_sticky = take(8);
  const uint64_t o_step = take(8);
  const uint64_t o_ka = take(pair_capacity * 8);
  const uint64_t o_kb = take(pair_capacity * 8);
  const uint64_t o_va = take(pair_capacity * 4);
  const uint64_t o_vb = take(pair_capacity * 4);
  const uint64_t o_rg = take(b->max_tiles * 8);
  const uint64_t o_ord = take(b->max_tiles * 4);
  const uint64_t o_img = take(max_pixels * 16);
  const uint64_t o_dl = take(max_pixels * 16);
  const uint64_t o_ft = take(max_pixels * 4);
  const uint64_t o_li = take(max_pixels * 4);
  const uint64_t o_loss = take(8 * max_views);
  const uint64_t o_scan = take(nts_scan_ws_bytes(vn));
  const uint64_t o_sort = take(nts_sort_ws_bytes(max_views, pair_capacity, 8));
  b->bytes = off;
  int rc = nts_dev_alloc(ctx, &b->base, b->bytes);
  if (rc) {
    delete b;
    return rc;
  }
  char* p = (char*)b->base;
  b->d_cams = (DevCamera*)(p + o_cams);
  b->proj = (Projected*)(p + o_proj);
  b->pgrad = (float*)(p + o_pgrad);
  b->counts = (uint32_t*)(p + o_counts);
  b->d_total = (uint64_t*)(p + o_total);
  b->seg_begin = (uint64_t*)(p + o_seg);
  b->tile_base = (uint32_t*)(p + o_tb);
  b->flags = (uint32_t*)(p + o_flags);
  b->sticky = (uint32_t*)(p + o_sticky);
  b->d_step = (uint64_t*)(p + o_step);
  b->keys_a = p + o_ka;
  b->keys_b = p + o_kb;
  b->vals_a = (uint32_t*)(p + o_va);
  b->vals_b = (uint32_t*)(p + o_vb);
  b->ranges = (uint2*)(p + o_rg);
  b->order = (uint32_t*)(p + o_ord);
  b->image = (float4*)(p + o_img);
  b->dl_dc = (float4*)(p + o_dl);
  b->final_t = (float*)(p + o_ft);
  b->last_idx = (uint32_t*)(p + o_li);
  b->losses = (double*)(p + o_loss);
  b->scan_ws = p + o_scan;
  b->sort_ws = p + o_sort;
  hipError_t e = hipHostMalloc((void**)&b->h_cams, sizeof(DevCamera) * NTS_MAX_VIEWS, hipHostMallocDefault);
  if (e != hipSuccess) {
    nts_dev_free(ctx, b->base, b->bytes);
    delete b;
    return nts_fail(NTS_E_HIP, "hipHostMalloc: %s", hipGetErrorString(e));
  }
  rc = nts_fill_u32(ctx, b->base, 0, b->bytes / 4);
  if (rc) return rc;
  *out = b;
  return NTS_OK;
}

static void drop_graph(nts_batch* b) {
  if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->graph_exec = nullptr;
  b->graph = nullptr;
  b->graph_key.clear();
}

extern "C" int nts_batch_destroy(nts_ctx* ctx, nts_batch* b) {
  if (!ctx || !b) return NTS_OK;
  (void)hipStreamSynchronize(ctx->stream);
  drop_graph(b);
  if (b->h_cams) (void)hipHostFree(b->h_cams);
  if (b->ssim_base) nts_dev_free(ctx, b->ssim_base, b->ssim_bytes);
  nts_dev_free(ctx, b->base, b->bytes);
  delete b;
  return NTS_OK;
}

struct BatchPlan {
  uint32_t n_views, tile_size, n;
  ViewTable vt;
  TileTable tt;
  uint32_t tile_bits, key_bits, key_bytes;
  uint64_t pixels;
};

static int plan_batch(nts_batch* b, const nts_scene* scene, const nts_view_desc* views, uint32_t n_views,
                      uint32_t tile_size, int loss_kind, BatchPlan& pl) {
  if (n_views == 0 || n_views > b->max_views) return nts_fail(NTS_E_INVALID, "n_views %u (max %u)", n_views, b->max_views);
  if (tile_size != 8 && tile_size != 16 && tile_size != 32)
    return nts_fail(NTS_E_INVALID, "tile size %u not in {8, 16, 32}", tile_size);
  if (scene->n > b->max_splats) return nts_fail(NTS_E_CAPACITY, "scene has %u splats, batch holds %u", scene->n,
                                                b->max_splats);
  memset(&pl, 0, sizeof pl);
  pl.n_views = n_views;
  pl.tile_size = tile_size;
  pl.n = scene->n;
  pl.vt.n = n_views;
  pl.vt.grad_words = NTS_GRAD9_WORDS;
  pl.tt.n = n_views;
  uint64_t pix = 0, tiles = 0;
  uint32_t max_tiles_view = 1;
  for (uint32_t v = 0; v < n_views; ++v) {
    const uint32_t w = views[v].width, h = views[v].height;
    if (w == 0 || h == 0) return nts_fail(NTS_E_INVALID, "view %u has an empty image", v);
    if ((float)w != views[v].cam.width || (float)h != views[v].cam.height)
      return nts_fail(NTS_E_INVALID, "view %u: camera size does not match %ux%u", v, w, h);
    const uint32_t tx = (w + tile_size - 1) / tile_size, ty = (h + tile_size - 1) / tile_size;
    ViewDesc& d = pl.vt.v[v];
    d.width = w;
    d.height = h;
    d.tiles_x = tx;
    d.tiles_y = ty;
    d.tile_base = (uint32_t)tiles;
    d.proj_base = v * scene->n;
    d.pix_base = pix;
    const float n_scalars = (float)w * (float)h * 3.0f;
    d.inv_n = loss_kind == NTS_LOSS_L1 ? 1.0f / n_scalars : 2.0f / n_scalars;
    pl.tt.t[v] = ViewTiles{tx, ty, tile_size, views[v].depth_max};
    pix += (uint64_t)w * h;
    tiles += (uint64_t)tx * ty;
    if (tx * ty > max_tiles_view) max_tiles_view = tx * ty;
  }
  if (pix > b->max_pixels) return nts_fail(NTS_E_CAPACITY, "%llu pixels exceed the batch's %llu",
                                           (unsigned long long)pix, (unsigned long long)b->max_pixels);
  if (tiles > b->max_tiles) return nts_fail(NTS_E_CAPACITY, "%llu tiles exceed the batch's %llu",
                                            (unsigned long long)tiles, (unsigned long long)b->max_tiles);
  if ((uint64_t)n_views * scene->n >= (1ull << 32)) return nts_fail(NTS_E_CAPACITY, "views x splats >= 2^32");
  pl.vt.total_tiles = (uint32_t)tiles;
  pl.vt.order = b->order;
  // the persistent scheduler's tile counters: flags[4] (forward) and
  // flags[5] (backward), zeroed with the flags at the start of every launch
  pl.vt.tickets = b->scheduler == NTS_SCHED_PERSISTENT ? b->flags + 4 : nullptr;
  // the evaluation counters: flags[8..15] as four 64-bit words (grid form only)
  pl.vt.counts = b->count_evals && !pl.vt.tickets ? reinterpret_cast<unsigned long long*>(b->flags + 8) : nullptr;
  pl.pixels = pix;
  // culled pairs carry the all-ones key: tile ids below 2^tb - 1
  pl.tile_bits = nts_tile_bits_strict(max_tiles_view);
  pl.key_bits = 16 + pl.tile_bits;
  pl.key_bytes = pl.key_bits <= 32 ? 4 : 8;
  return NTS_OK;
}

// Forward half of a step: project, bin (segmented by view), blend.
static int launch_forward(nts_ctx* ctx, const nts_scene* scene, nts_batch* b, const BatchPlan& pl,
                          const LossArgs* loss, bool keep_state, bool** alt_out) {
  const uint32_t n = pl.n, V = pl.n_views;
  const uint64_t vn = (uint64_t)V * n;
  int rc = nts_fill_u32(ctx, b->flags, 0, 16);
  if (rc) return rc;
  {
    const bool cull = b->binning == NTS_BIN_CULLED;
    if (n > 0 && cull)
      NTS_LAUNCH(ctx, "project_views", k_project_views<true>, dim3(nts_div_up(n, 256)), dim3(256), 0, scene->params,
                 scene->cap, n, (const DevCamera*)b->d_cams, pl.tt, b->proj, b->counts, b->flags + 1);
    else if (n > 0)
      NTS_LAUNCH(ctx, "project_views", k_project_views<false>, dim3(nts_div_up(n, 256)), dim3(256), 0, scene->params,
                 scene->cap, n, (const DevCamera*)b->d_cams, pl.tt, b->proj, b->counts, b->flags + 1);
  }
  rc = nts_scan_u32_ws(ctx, b->counts, b->counts, vn, b->d_total, b->scan_ws);
  if (rc) return rc;
  RecTable rt;
  memset(&rt, 0, sizeof rt);
  rt.n = V;
  for (uint32_t v = 0; v <= V; ++v) rt.base[v] = v * n;
  NTS_LAUNCH(ctx, "view_segments", k_view_segments, dim3(1), dim3(64), 0, (const uint32_t*)b->counts, rt,
             (const uint64_t*)b->d_total, b->pair_cap, b->seg_begin, b->flags);
  bool alt = false;
  if (n > 0) {
    const uint32_t nrec = (uint32_t)vn;
    const bool cull = b->binning == NTS_BIN_CULLED;
    if (pl.key_bytes == 4 && cull)
      NTS_LAUNCH(ctx, "emit_views", (k_emit_balanced<uint32_t, true>), dim3(nts_div_up(nrec, 256)), dim3(256), 0,
                 (const Projected*)b->proj, rt, pl.tt, (const uint32_t*)b->counts, (const uint64_t*)b->d_total,
                 b->pair_cap, (uint32_t*)b->keys_a, b->vals_a);
    else if (pl.key_bytes == 4)
      NTS_LAUNCH(ctx, "emit_views", (k_emit_balanced<uint32_t, false>), dim3(nts_div_up(nrec, 256)), dim3(256), 0,
                 (const Projected*)b->proj, rt, pl.tt, (const uint32_t*)b->counts, (const uint64_t*)b->d_total,
                 b->pair_cap, (uint32_t*)b->keys_a, b->vals_a);
    else if (cull)
      NTS_LAUNCH(ctx, "emit_views", (k_emit_balanced<uint64_t, true>), dim3(nts_div_up(nrec, 256)), dim3(256), 0,
                 (const Projected*)b->proj, rt, pl.tt, (const uint32_t*)b->counts, (const uint64_t*)b->d_total,
                 b->pair_cap, (uint64_t*)b->keys_a, b->vals_a);
    else
      NTS_LAUNCH(ctx, "emit_views", (k_emit_balanced<uint64_t, false>), dim3(nts_div_up(nrec, 256)), dim3(256), 0,
                 (const Projected*)b->proj, rt, pl.tt, (const uint32_t*)b->counts, (const uint64_t*)b->d_total,
                 b->pair_cap, (uint64_t*)b->keys_a, b->vals_a);
  }
  // tile ranges (zero for tiles with no pair) and tile bases before the sort,
  // which writes the ranges and the tiles' blend order
  rc = nts_fill_u32(ctx, b->ranges, 0, (uint64_t)pl.vt.total_tiles * 2);
  if (rc) return rc;
  U32Table tb;
  memset(&tb, 0, sizeof tb);
  tb.n = V + 1;
  for (uint32_t v = 0; v <= V; ++v) tb.v[v] = v < V ? pl.vt.v[v].tile_base : pl.vt.total_tiles;
  NTS_LAUNCH(ctx, "set_tile_base", k_set_u32s, dim3(1), dim3(64), 0, b->tile_base, tb);
  if (pl.key_bytes == 4)
    rc = nts_bin_sort_ws_launch32(ctx, (uint32_t*)b->keys_a, b->vals_a, (uint32_t*)b->keys_b, b->vals_b, b->seg_begin,
                                  V, b->pair_cap, pl.tile_bits, b->tile_base, pl.vt.total_tiles, b->ranges, b->order,
                                  b->sort_ws, &alt, &b->sort_err);
  else
    rc = nts_bin_sort_ws_launch64(ctx, (uint64_t*)b->keys_a, b->vals_a, (uint64_t*)b->keys_b, b->vals_b, b->seg_begin,
                                  V, b->pair_cap, pl.tile_bits, b->tile_base, pl.vt.total_tiles, b->ranges, b->order,
                                  b->sort_ws, &alt, &b->sort_err);
  if (rc) return rc;
  uint32_t* pay = alt ? b->vals_b : b->vals_a;
  NTS_LAUNCH(ctx, "sort_latch", k_latch_errors, dim3(1), dim3(64), 0, (const uint32_t*)b->flags,
             (const uint32_t*)b->sort_err, b->sticky);
  rc = nts_blend_forward(ctx, pl.tile_size, b->fwd_batch, b->proj, pay, b->ranges, pl.vt, b->image,
                         keep_state ? b->final_t : nullptr, keep_state ? b->last_idx : nullptr, loss);
  if (rc) return rc;
  if (alt_out) **alt_out = alt;
  b->last_alt = alt;
  b->last_key_bytes = pl.key_bytes;
  for (uint32_t v = 0; v < V; ++v) b->last_tile_base[v] = pl.vt.v[v].tile_base;
  b->last_tile_base[V] = pl.vt.total_tiles;
  return NTS_OK;
}

static void remember_shape(nts_batch* b, const BatchPlan& pl) {
  b->n_views = pl.n_views;
  for (uint32_t v = 0; v < pl.n_views; ++v) {
    b->pix_base[v] = pl.vt.v[v].pix_base;
    b->width[v] = pl.vt.v[v].width;
    b->height[v] = pl.vt.v[v].height;
  }
  b->pix_base[pl.n_views] = pl.pixels;
}

static int upload_cams(nts_ctx* ctx, nts_batch* b, const nts_view_desc* views, uint32_t V) {
  for (uint32_t v = 0; v < V; ++v) memcpy(&b->h_cams[v], &views[v].cam, sizeof(DevCamera));
  NTS_HIP_TRY(hipMemcpyAsync(b->d_cams, b->h_cams, sizeof(DevCamera) * V, hipMemcpyHostToDevice, ctx->stream));
  return NTS_OK;
}

static int record_step(nts_ctx* ctx, nts_scene* scene, nts_batch* b, const BatchPlan& pl, const nts_buffer* targets,
                       int loss_kind, int mode, const nts_adam_cfg* cp, const nts_adam_cfg* ca,
                       const nts_constraints* cons, int write_grads) {
  const uint32_t n = pl.n, V = pl.n_views;
  int rc = nts_fill_u32(ctx, b->losses, 0, 2 * (uint64_t)V);
  if (rc) return rc;
  // NTS_LOSS_SSIM at lambda 0 is the fused MSE (loss.rs:398-402)
  const bool ssim = loss_kind == NTS_LOSS_SSIM && b->ssim_lambda > 0.0f;
  LossArgs la{(const float4*)targets->ptr, b->dl_dc, b->losses, loss_kind == NTS_LOSS_L1 ? NTS_LOSS_L1 : NTS_LOSS_MSE,
              nullptr, 0};
  // The projected gradients start at zero: cleared by the forward blend behind
  // its tiles (its HBM is mostly idle), or by a fill where that launch has no
  // loss epilogue or no tiles.
  const uint64_t pgrad_n4 = ((uint64_t)V * n * NTS_GRAD9_WORDS + 3) / 4;  // within the allocation's whole float4s
  if (!ssim && pl.vt.total_tiles > 0) {
    la.clear = (float4*)b->pgrad;
    la.clear_n4 = pgrad_n4;
  } else {
    rc = nts_fill_u32(ctx, b->pgrad, 0, pgrad_n4 * 4);
    if (rc) return rc;
  }
  bool alt = false;
  bool* altp = &alt;
  rc = launch_forward(ctx, scene, b, pl, ssim ? nullptr : &la, mode == NTS_BWD_RECOMPUTE, &altp);
  if (rc) return rc;
  if (ssim) {
    uint32_t ws[NTS_MAX_VIEWS], hs[NTS_MAX_VIEWS];
    uint64_t pb[NTS_MAX_VIEWS];
    for (uint32_t v = 0; v < V; ++v) {
      ws[v] = pl.vt.v[v].width;
      hs[v] = pl.vt.v[v].height;
      pb[v] = pl.vt.v[v].pix_base;
    }
    SsimViews sv;
    rc = nts_ssim_views(V, ws, hs, pb, &sv);
    if (rc) return rc;
    double* sums = (double*)b->ssim_base;
    float* planes = (float*)((char*)b->ssim_base + 8 * 4 * NTS_MAX_VIEWS);
    rc = nts_ssim_launch(ctx, sv, b->image, (const float4*)targets->ptr, planes, b->max_pixels, sums, b->dl_dc,
                         b->ssim_lambda, b->losses);
    if (rc) return rc;
  }
  const uint32_t* pay = alt ? b->vals_b : b->vals_a;
  rc = nts_blend_backward(ctx, pl.tile_size, b->bwd_batch, mode, b->proj, pay, b->ranges, pl.vt, b->image,
                          b->final_t, b->last_idx, b->dl_dc, b->pgrad);
  if (rc) return rc;
  AdamCfgs cfg;
  cfg.lr_pos = cp->lr; cfg.b1_pos = cp->beta1; cfg.b2_pos = cp->beta2; cfg.eps_pos = cp->eps;
  cfg.lr_attr = ca->lr; cfg.b1_attr = ca->beta1; cfg.b2_attr = ca->beta2; cfg.eps_attr = ca->eps;
  cfg.smin = cons->log_scale_min; cfg.smax = cons->log_scale_max; cfg.omax = cons->logit_max;
  if (n > 0) {
    if (mode == NTS_BWD_RECOMPUTE)
      NTS_LAUNCH(ctx, "backward_adam_views", k_backward_adam_views<true>, dim3(nts_div_up(n, 256)), dim3(256), 0,
                 scene->params, scene->grads, scene->m, scene->v, scene->grad_acc, scene->cap, n, V,
                 (const DevCamera*)b->d_cams, (const float*)b->pgrad, (const uint64_t*)b->d_step, cfg,
                 write_grads, (const uint32_t*)b->flags, (const uint32_t*)b->sort_err, b->sticky);
    else
      NTS_LAUNCH(ctx, "backward_adam_views", k_backward_adam_views<false>, dim3(nts_div_up(n, 256)), dim3(256), 0,
                 scene->params, scene->grads, scene->m, scene->v, scene->grad_acc, scene->cap, n, V,
                 (const DevCamera*)b->d_cams, (const float*)b->pgrad, (const uint64_t*)b->d_step, cfg,
                 write_grads, (const uint32_t*)b->flags, (const uint32_t*)b->sort_err, b->sticky);
  }
  return NTS_OK;
}

extern "C" int nts_train_step_views(nts_ctx* ctx, nts_scene* scene, nts_batch* b, const nts_view_desc* views,
                                    uint32_t n_views, const nts_buffer* targets, uint32_t tile_size, int loss_kind,
                                    int bwd_mode, const nts_adam_cfg* cfg_pos, const nts_adam_cfg* cfg_attr,
                                    uint64_t t, const nts_constraints* cons, int flags) {
  NTS_CHECK_ARG(ctx && scene && b && views && targets && cfg_pos && cfg_attr && cons && t > 0);
  NTS_CHECK_ARG(loss_kind == NTS_LOSS_L1 || loss_kind == NTS_LOSS_MSE || loss_kind == NTS_LOSS_SSIM);
  NTS_CHECK_ARG(bwd_mode == NTS_BWD_REF_PARITY || bwd_mode == NTS_BWD_RECOMPUTE);
  NTS_NEED_TRAINABLE(scene);
  if (loss_kind == NTS_LOSS_SSIM && b->ssim_lambda > 0.0f && !b->ssim_base)
    return nts_fail(NTS_E_INVALID, "NTS_LOSS_SSIM needs nts_batch_set_ssim_lambda first (it allocates the planes)");
  BatchPlan pl;
  int rc = plan_batch(b, scene, views, n_views, tile_size, loss_kind, pl);
  if (rc) return rc;
  NTS_CHECK_ARG(targets->bytes >= pl.pixels * 16);
  remember_shape(b, pl);
  b->last_loss_final = loss_kind == NTS_LOSS_SSIM && b->ssim_lambda > 0.0f;
  const int write_grads = (flags & NTS_STEP_WRITE_GRADS) ? 1 : 0;
  const bool use_graph = (flags & NTS_STEP_GRAPH) != 0 && !ctx->prof;
  if (!use_graph) {
    b->graph_key.clear();  // cameras and step counter change under a captured graph
    rc = upload_cams(ctx, b, views, n_views);
    if (rc) return rc;
    NTS_LAUNCH(ctx, "set_step", k_set_step, dim3(1), dim3(64), 0, b->d_step, t);
    return record_step(ctx, scene, b, pl, targets, loss_kind, bwd_mode, cfg_pos, cfg_attr, cons, write_grads);
  }
  // Captured replay: the key holds everything baked into the graph.
  std::vector<uint8_t> key;
  auto put = [&](const void* p, size_t nbytes) {
    const uint8_t* c = (const uint8_t*)p;
    key.insert(key.end(), c, c + nbytes);
  };
  const void* ids[3] = {scene, targets->ptr, scene->params};
  put(ids, sizeof ids);
  // the plane strides baked into the kernels follow the capacity; the
  // creation number tells apart scenes that reuse freed addresses
  put(&scene->gen, sizeof scene->gen);
  put(&scene->cap, 4);
  put(&scene->n, 4);
  put(views, sizeof(nts_view_desc) * n_views);
  int lam_bits;
  memcpy(&lam_bits, &b->ssim_lambda, 4);
  const int meta[10] = {(int)tile_size, loss_kind, bwd_mode, write_grads, (int)b->fwd_batch, (int)b->bwd_batch,
                        loss_kind == NTS_LOSS_SSIM ? lam_bits : 0, b->scheduler, (int)b->count_evals, b->binning};
  put(meta, sizeof meta);
  put(cfg_pos, sizeof *cfg_pos);
  put(cfg_attr, sizeof *cfg_attr);
  put(cons, sizeof *cons);
  if (key != b->graph_key || t != b->graph_last_t + 1) {
    // (re)capture: cameras and the step counter are set outside the graph
    drop_graph(b);
    rc = upload_cams(ctx, b, views, n_views);
    if (rc) return rc;
    NTS_LAUNCH(ctx, "set_step", k_set_step, dim3(1), dim3(64), 0, b->d_step, t - 1);
    NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
    NTS_HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int rc2 = NTS_OK;
    {
      PendingEvent pe;
      (void)pe;
      hipLaunchKernelGGL(k_inc_step, dim3(1), dim3(64), 0, ctx->stream, b->d_step);
      rc2 = record_step(ctx, scene, b, pl, targets, loss_kind, bwd_mode, cfg_pos, cfg_attr, cons, write_grads);
    }
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc2) {
      if (g) (void)hipGraphDestroy(g);
      return rc2;
    }
    if (e != hipSuccess) return nts_fail(NTS_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    b->graph = g;
    e = hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      drop_graph(b);
      return nts_fail(NTS_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    b->graph_key = key;
  }
  NTS_HIP_TRY(hipGraphLaunch(b->graph_exec, ctx->stream));
  b->graph_last_t = t;
  return NTS_OK;
}

extern "C" int nts_render_views(nts_ctx* ctx, const nts_scene* scene, nts_batch* b, const nts_view_desc* views,
                                uint32_t n_views, uint32_t tile_size) {
  NTS_CHECK_ARG(ctx && scene && b && views);
  BatchPlan pl;
  int rc = plan_batch(b, scene, views, n_views, tile_size, NTS_LOSS_L1, pl);
  if (rc) return rc;
  remember_shape(b, pl);
  b->graph_key.clear();
  rc = upload_cams(ctx, b, views, n_views);
  if (rc) return rc;
  return launch_forward(ctx, scene, b, pl, nullptr, true, nullptr);
}

extern "C" int nts_batch_set_lds_batch(nts_ctx* ctx, nts_batch* b, uint32_t fwd_batch, uint32_t bwd_batch) {
  NTS_CHECK_ARG(ctx && b);
  if (!nts_blend_config_ok(16, fwd_batch, bwd_batch))
    return nts_fail(NTS_E_INVALID, "LDS batch %u / %u: forward in {64, 128, 256, 512}, backward in {64, 128, 256}",
                    fwd_batch, bwd_batch);
  b->fwd_batch = fwd_batch;
  b->bwd_batch = bwd_batch;
  return NTS_OK;
}

extern "C" int nts_batch_set_scheduler(nts_ctx* ctx, nts_batch* b, int scheduler) {
  NTS_CHECK_ARG(ctx && b && (scheduler == NTS_SCHED_GRID || scheduler == NTS_SCHED_PERSISTENT));
  b->scheduler = scheduler;
  return NTS_OK;
}

extern "C" int nts_batch_set_binning(nts_ctx* ctx, nts_batch* b, int binning) {
  NTS_CHECK_ARG(ctx && b && (binning == NTS_BIN_CULLED || binning == NTS_BIN_REFERENCE));
  b->binning = binning;
  return NTS_OK;
}

extern "C" int nts_batch_set_eval_counting(nts_ctx* ctx, nts_batch* b, int on) {
  NTS_CHECK_ARG(ctx && b);
  b->count_evals = on != 0;
  return NTS_OK;
}

extern "C" int nts_batch_eval_counts(nts_ctx* ctx, const nts_batch* b, uint64_t* counts) {
  NTS_CHECK_ARG(ctx && b && counts);
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  NTS_HIP_TRY(hipMemcpyAsync(ctx->host_word, b->flags + 8, 32, hipMemcpyDeviceToHost, ctx->stream));
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  memcpy(counts, ctx->host_word, 32);
  return NTS_OK;
}

extern "C" int nts_batch_set_ssim_lambda(nts_ctx* ctx, nts_batch* b, float lambda) {
  NTS_CHECK_ARG(ctx && b && lambda == lambda);
  const float lam = lambda < 0.0f ? 0.0f : (lambda > 1.0f ? 1.0f : lambda);
  if (lam > 0.0f && !b->ssim_base) {
    const uint64_t bytes = 8 * 4 * NTS_MAX_VIEWS + 15 * b->max_pixels * 4;
    const int rc = nts_dev_alloc(ctx, &b->ssim_base, bytes);
    if (rc) return rc;
    b->ssim_bytes = bytes;
  }
  b->ssim_lambda = lam;
  return NTS_OK;
}

extern "C" int nts_batch_losses(nts_ctx* ctx, const nts_batch* b, double* losses) {
  NTS_CHECK_ARG(ctx && b && losses);
  NTS_HIP_TRY(hipMemcpyAsync(losses, b->losses, 8 * (size_t)b->n_views, hipMemcpyDeviceToHost, ctx->stream));
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (uint32_t v = 0; v < b->n_views && !b->last_loss_final; ++v) {
    const uint64_t np = b->pix_base[v + 1] - b->pix_base[v];
    losses[v] = losses[v] / ((double)np * 3.0);
  }
  return NTS_OK;
}

extern "C" int nts_batch_stats(nts_ctx* ctx, const nts_batch* b, uint64_t* total_pairs, uint32_t* visible,
                               uint32_t* overflow) {
  NTS_CHECK_ARG(ctx && b);
  // Everything queued before this call has finished before the words are
  // read (into the context's pinned words), and the latch is cleared before
  // the call returns.
  uint32_t* hw = ctx->host_word;  // [0..3] flags, [4..5] sticky, [6] sort error
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  hw[6] = 0;
  NTS_HIP_TRY(hipMemcpyAsync(hw, b->flags, 16, hipMemcpyDeviceToHost, ctx->stream));
  NTS_HIP_TRY(hipMemcpyAsync(hw + 4, b->sticky, 8, hipMemcpyDeviceToHost, ctx->stream));
  if (b->sort_err) NTS_HIP_TRY(hipMemcpyAsync(hw + 6, b->sort_err, 4, hipMemcpyDeviceToHost, ctx->stream));
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  uint32_t f[4], st[2];
  memcpy(f, hw, 16);
  memcpy(st, hw + 4, 8);
  const uint32_t err = hw[6];
  // the latched errors are reported once
  int rc = nts_fill_u32(ctx, b->sticky, 0, 2);
  if (rc) return rc;
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (err || (st[0] & 2u))
    return nts_fail(NTS_E_HIP, "radix sort look-back timed out (%u step updates skipped)", st[1]);
  if (total_pairs) *total_pairs = (uint64_t)f[2] | ((uint64_t)f[3] << 32);
  if (visible) *visible = f[1];
  if (overflow) *overflow = (st[0] & 1u) ? (st[1] > 0 ? st[1] : 1u) : 0u;
  return NTS_OK;
}

extern "C" int nts_batch_download_image(nts_ctx* ctx, const nts_batch* b, uint32_t view, float* host) {
  NTS_CHECK_ARG(ctx && b && host && view < b->n_views);
  const uint64_t np = b->pix_base[view + 1] - b->pix_base[view];
  NTS_HIP_TRY(hipMemcpyAsync(host, b->image + b->pix_base[view], np * 16, hipMemcpyDeviceToHost, ctx->stream));
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return NTS_OK;
}

extern "C" int nts_batch_download_state(nts_ctx* ctx, const nts_batch* b, uint32_t view, float* final_t,
                                        uint32_t* last_idx) {
  NTS_CHECK_ARG(ctx && b && view < b->n_views);
  const uint64_t np = b->pix_base[view + 1] - b->pix_base[view];
  if (final_t)
    NTS_HIP_TRY(hipMemcpyAsync(final_t, b->final_t + b->pix_base[view], np * 4, hipMemcpyDeviceToHost, ctx->stream));
  if (last_idx)
    NTS_HIP_TRY(
        hipMemcpyAsync(last_idx, b->last_idx + b->pix_base[view], np * 4, hipMemcpyDeviceToHost, ctx->stream));
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return NTS_OK;
}

extern "C" int nts_batch_download_projected_grad(nts_ctx* ctx, const nts_batch* b, uint32_t view, uint32_t n,
                                                 float* host) {
  NTS_CHECK_ARG(ctx && b && host && view < b->n_views && n <= b->max_splats);
  // the compact records, spread into the ProjectedGrad layout (zero padding words)
  std::vector<float> g9((size_t)n * NTS_GRAD9_WORDS);
  NTS_HIP_TRY(hipMemcpyAsync(g9.data(), b->pgrad + (size_t)view * n * NTS_GRAD9_WORDS, g9.size() * 4,
                             hipMemcpyDeviceToHost, ctx->stream));
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (size_t i = 0; i < n; ++i) {
    const float* g = &g9[i * NTS_GRAD9_WORDS];
    float* o = host + i * NTS_GRAD_WORDS;
    const float rec[12] = {g[0], g[1], g[2], 0.0f, g[3], g[4], g[5], 0.0f, g[6], g[7], g[8], 0.0f};
    memcpy(o, rec, sizeof rec);
  }
  return NTS_OK;
}

extern "C" int nts_batch_download_bins(nts_ctx* ctx, const nts_batch* b, uint32_t view, uint64_t* count,
                                       uint32_t* keys, uint32_t* payloads, uint32_t* ranges) {
  NTS_CHECK_ARG(ctx && b && count && view < b->n_views);
  uint64_t sb[2];
  NTS_HIP_TRY(hipMemcpyAsync(sb, b->seg_begin + view, 16, hipMemcpyDeviceToHost, ctx->stream));
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  const uint64_t k = sb[1] - sb[0];
  *count = k;
  const void* kbase = b->last_alt ? b->keys_b : b->keys_a;
  const uint32_t* pbase = b->last_alt ? b->vals_b : b->vals_a;
  if (keys && k) {
    if (b->last_key_bytes == 4) {
      NTS_HIP_TRY(hipMemcpyAsync(keys, (const uint32_t*)kbase + sb[0], k * 4, hipMemcpyDeviceToHost, ctx->stream));
    } else {
      std::vector<uint64_t> tmp(k);
      NTS_HIP_TRY(hipMemcpyAsync(tmp.data(), (const uint64_t*)kbase + sb[0], k * 8, hipMemcpyDeviceToHost,
                                 ctx->stream));
      NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
      for (uint64_t i = 0; i < k; ++i) {
        if (tmp[i] >> 32) return nts_fail(NTS_E_CAPACITY, "tile id does not fit the 16-bit reference key");
        keys[i] = (uint32_t)tmp[i];
      }
    }
  }
  if (payloads && k)
    NTS_HIP_TRY(hipMemcpyAsync(payloads, pbase + sb[0], k * 4, hipMemcpyDeviceToHost, ctx->stream));
  const uint32_t t0 = b->last_tile_base[view], t1 = b->last_tile_base[view + 1];
  if (ranges)
    NTS_HIP_TRY(hipMemcpyAsync(ranges, b->ranges + t0, (uint64_t)(t1 - t0) * 8, hipMemcpyDeviceToHost, ctx->stream));
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (ranges)  // relative to the view's segment, as a single-view bin
    for (uint32_t t = 0; t < t1 - t0; ++t)
      if (ranges[2 * t + 1] > ranges[2 * t]) {
        ranges[2 * t] -= (uint32_t)sb[0];
        ranges[2 * t + 1] -= (uint32_t)sb[0];
      }
  return NTS_OK;
}
