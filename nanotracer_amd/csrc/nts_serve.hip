// Render server: many trained scenes resident in HBM, render requests
// (one camera of one scene each) admitted by workspace capacity and
// batched as they arrive, several batches in flight on separate streams.
//
// A dispatcher thread owns the GPU work.  Each of its slots is a complete
// forward workspace (projected records, pairs, sort, ranges, image) with
// its own HIP stream; a batch is up to max_views requests of any scenes and
// resolutions, rendered by one launch per stage exactly like the batched
// training forward (project, scan, view segments, emit, segmented sort,
// ranges, blend), then tonemapped and copied to pinned host memory.  The
// thread retires finished slots (event query), copies each result to the
// caller's buffer, and refills free slots from the FIFO of pending requests
// while the other slots' batches still run: continuous batching with
// overlap on the device.  The copies out of the pinned staging run on a pool
// of copier threads in pieces of 4 MB; a slot takes its next batch once its
// results are out.
//
// Admission: a request joins the batch being formed while the batch's
// views, pixels, projected records, tiles and estimated pairs fit the slot
// (pairs per pixel are learned per scene from its renders).  A batch whose
// pairs overflowed anyway is re-run one request at a time; a request that
// does not fit an empty slot fails with NTS_E_CAPACITY.  The forward is
// deterministic per view and a view's work never depends on the others in
// its batch, so every result is bit-identical alone and in any mix.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "nts_bin.h"
#include "nts_project.h"
#include "nts_raster.h"

namespace {

struct SceneTab {
  const float* params[NTS_MAX_VIEWS];
  uint32_t cap[NTS_MAX_VIEWS];
};

struct U32Pack {
  uint32_t n;
  uint32_t v[NTS_MAX_VIEWS + 1];
};

__global__ void __launch_bounds__(256) k_project_requests(SceneTab st, RecTable rt, const DevCamera* __restrict__ cams,
                                                          TileTable tt, Projected* __restrict__ proj,
                                                          uint32_t* __restrict__ counts) {
  const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= rt.base[rt.n]) return;
  const uint32_t v = rec_view(rt, f), i = f - rt.base[v];
  SplatP s;
  load_splat(st.params[v], st.cap[v], i, s);
  const Projected p = project_p(s, cams[v]);
  proj[f] = p;
  uint32_t c = 0;
  int32_t x0, y0, x1, y1;
  const ViewTiles vt = tt.t[v];
  if (record_tiles<true>(p, (float)vt.tile_size, (int32_t)vt.tiles_x, (int32_t)vt.tiles_y, x0, y0, x1, y1))
    c = (uint32_t)((x1 - x0 + 1) * (y1 - y0 + 1));
  counts[f] = c;
}

__global__ void k_request_segments(const uint32_t* __restrict__ offsets, RecTable rt,
                                   const uint64_t* __restrict__ d_total, uint64_t cap, uint64_t* __restrict__ seg_begin,
                                   uint32_t* __restrict__ tile_base, U32Pack tb, uint32_t* __restrict__ flags) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint64_t total = *d_total;
  const uint32_t nrec = rt.base[rt.n];
  for (uint32_t v = 0; v < rt.n; ++v) {
    const uint64_t b = rt.base[v] < nrec ? (uint64_t)offsets[rt.base[v]] : total;
    seg_begin[v] = b < cap ? b : cap;
  }
  seg_begin[rt.n] = total < cap ? total : cap;
  for (uint32_t v = 0; v <= rt.n; ++v) tile_base[v] = tb.v[v];
  flags[0] = total > cap ? 1u : 0u;
  flags[2] = (uint32_t)total;
  flags[3] = (uint32_t)(total >> 32);
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

enum ReqState { REQ_PENDING = 0, REQ_RUNNING = 1, REQ_DONE = 2, REQ_FAILED = 3 };

struct Request {
  uint64_t id = 0;
  uint32_t scene = 0;
  nts_view_desc view{};
  void* out = nullptr;
  int format = NTS_OUT_RGBA8;
  bool solo = false;
  int state = REQ_PENDING;
  int err = NTS_OK;
  double t_submit = 0.0, t_done = 0.0;
  uint32_t batch_views = 0;
  uint64_t pairs = 0;
  uint32_t chunks_left = 0;  // result copies still to do (copier threads)
};

struct Resident {
  nts_scene* scene = nullptr;
  double pairs_per_px = 0.0;  // learned from renders (0 = none yet)
};

struct Slot {
  nts_ctx* ctx = nullptr;
  void* base = nullptr;
  uint64_t bytes = 0;       // the slot's workspace; each batch lays out its working set inside it
  uint64_t pair_cap = 0;    // pairs the current batch has room for
  DevCamera* d_cams = nullptr;
  Projected* proj = nullptr;
  uint32_t* counts = nullptr;
  uint64_t* d_total = nullptr;
  uint64_t* seg_begin = nullptr;
  uint32_t* tile_base = nullptr;
  uint32_t* flags = nullptr;
  void *keys_a = nullptr, *keys_b = nullptr;
  uint32_t *vals_a = nullptr, *vals_b = nullptr;
  uint2* ranges = nullptr;
  uint32_t* order = nullptr;  // the blend order of the tiles (most pairs first)
  float4* image = nullptr;
  uint32_t* rgba = nullptr;
  void* scan_ws = nullptr;
  void* sort_ws = nullptr;
  uint32_t* sort_err = nullptr;
  // pinned host staging
  DevCamera* h_cams = nullptr;
  uint64_t* h_meta = nullptr;  // [0] flags 0-1, [1] sort error word, [2..] seg_begin
  uint8_t* h_out = nullptr;
  hipEvent_t done = nullptr;
  std::vector<Request*> reqs;
  // per request: pixel offset in the batch image, pixel count, and byte
  // offset of its result in h_out (results of both formats packed end to end)
  std::vector<uint64_t> out_off, pix, stage_off;
  bool busy = false;
  uint32_t copies_left = 0;  // result copies out of h_out still to do: the slot waits for them
};

// One piece of a result: copied from a slot's pinned staging into the
// caller's buffer by a copier thread.
struct CopyJob {
  Request* r;
  Slot* slot;
  const uint8_t* src;
  uint8_t* dst;
  size_t bytes;
};

}  // namespace

struct nts_server {
  int device = 0;
  nts_server_cfg cfg{};
  uint64_t max_tiles = 0;
  uint64_t slot_bytes = 0;   // each slot's workspace
  uint64_t max_records = 0;  // per batch (the caps of cfg, 0 = no cap)
  uint64_t pair_cap = 0;
  nts_ctx* ctx0 = nullptr;  // uploads
  std::vector<Resident> scenes;
  std::vector<Slot> slots;
  uint64_t used = 0;  // HBM bytes of scenes and slots
  std::deque<Request*> pending;
  std::unordered_map<uint64_t, Request*> reqs;
  uint64_t next_id = 1;
  std::mutex mu;
  std::condition_variable cv, done_cv;
  std::thread th;
  bool stop = false;
  // result copies, off the dispatcher thread
  std::vector<std::thread> copiers;
  std::deque<CopyJob> jobs;
  std::condition_variable job_cv;
  bool stop_copiers = false;
  nts_serve_stats stats{};
};

namespace {

uint64_t al256(uint64_t x) { return (x + 255) & ~255ull; }

// Working set of one batch of V views (rec projected records, px pixels,
// tiles tiles, keys of kb bytes, room for pair_cap pairs), laid out from the
// start of its slot's workspace.
struct BatchLayout {
  uint64_t cams, proj, counts, total, seg, tb, flags, ka, kb, va, vb, rg, order, img, rgba, scan, sort, end;
};

BatchLayout layout_batch(uint32_t V, uint64_t rec, uint64_t px, uint64_t tiles, uint32_t kb, uint64_t pair_cap) {
  BatchLayout L{};
  uint64_t off = 0;
  auto take = [&](uint64_t bytes) { const uint64_t o = off; off += al256(bytes); return o; };
  const uint64_t r1 = rec > 0 ? rec : 1, p1 = pair_cap > 0 ? pair_cap : 1;
  L.cams = take(sizeof(DevCamera) * V);
  L.proj = take(r1 * sizeof(Projected));
  L.counts = take(r1 * 4);
  L.total = take(8);
  L.seg = take(8 * (V + 1));
  L.tb = take(4 * (V + 1));
  L.flags = take(64);
  L.ka = take(p1 * kb);
  L.kb = take(p1 * kb);
  L.va = take(p1 * 4);
  L.vb = take(p1 * 4);
  L.rg = take((tiles > 0 ? tiles : 1) * 8);
  L.order = take((tiles > 0 ? tiles : 1) * 4);
  L.img = take(px * 16);
  L.rgba = take(px * 4);
  L.scan = take(nts_scan_ws_bytes(r1));
  L.sort = take(nts_sort_ws_bytes(V, p1, kb));
  L.end = off;
  return L;
}

// The most pairs (below 2^30) a batch of this shape can hold in `bytes`;
// 0 when not even its fixed part fits.
uint64_t pairs_that_fit(uint64_t bytes, uint32_t V, uint64_t rec, uint64_t px, uint64_t tiles, uint32_t kb) {
  if (layout_batch(V, rec, px, tiles, kb, 1).end > bytes) return 0;
  uint64_t lo = 1, hi = (1ull << 30) - 1;
  if (layout_batch(V, rec, px, tiles, kb, hi).end <= bytes) return hi;
  while (hi - lo > 1) {  // invariant: lo fits, hi does not
    const uint64_t mid = lo + (hi - lo) / 2;
    if (layout_batch(V, rec, px, tiles, kb, mid).end <= bytes) lo = mid; else hi = mid;
  }
  return lo;
}

// Key bytes of a batch whose largest view has max_tiles_view tiles.
uint32_t key_bytes_for(uint64_t max_tiles_view) {
  return 16 + nts_tile_bits_strict(max_tiles_view) <= 32 ? 4u : 8u;
}

int slot_create(nts_server* sv, Slot& sl) {
  const nts_server_cfg& c = sv->cfg;
  int rc = nts_ctx_create(sv->device, 0, &sl.ctx);
  if (rc) return rc;
  sl.bytes = sv->slot_bytes;
  if (sv->cfg.hbm_budget && sv->used + sl.bytes > sv->cfg.hbm_budget)
    return nts_fail(NTS_E_NOMEM, "server slot of %llu bytes exceeds the HBM budget", (unsigned long long)sl.bytes);
  rc = nts_dev_alloc(sl.ctx, &sl.base, sl.bytes);
  if (rc) return rc;
  sv->used += sl.bytes;
  const uint64_t V = c.max_views;
  NTS_HIP_TRY(hipHostMalloc((void**)&sl.h_cams, sizeof(DevCamera) * V, hipHostMallocDefault));
  NTS_HIP_TRY(hipHostMalloc((void**)&sl.h_meta, 8 * (V + 4), hipHostMallocDefault));
  NTS_HIP_TRY(hipHostMalloc((void**)&sl.h_out, c.max_pixels * 16, hipHostMallocDefault));
  NTS_HIP_TRY(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
  return nts_fill_u32(sl.ctx, sl.base, 0, sl.bytes / 4);
}

void slot_destroy(nts_server* sv, Slot& sl) {
  if (!sl.ctx) return;
  (void)hipStreamSynchronize(sl.ctx->stream);
  if (sl.done) (void)hipEventDestroy(sl.done);
  if (sl.h_cams) (void)hipHostFree(sl.h_cams);
  if (sl.h_meta) (void)hipHostFree(sl.h_meta);
  if (sl.h_out) (void)hipHostFree(sl.h_out);
  if (sl.base) {
    nts_dev_free(sl.ctx, sl.base, sl.bytes);
    sv->used -= sl.bytes;
  }
  nts_ctx_destroy(sl.ctx);
  sl.ctx = nullptr;
}

uint32_t tiles_of(const nts_view_desc& v, uint32_t ts) {
  return ((v.width + ts - 1) / ts) * ((v.height + ts - 1) / ts);
}

// Launch the forward of the slot's requests on its stream (no host sync).
int launch_batch(nts_server* sv, Slot& sl) {
  nts_ctx* ctx = sl.ctx;
  const uint32_t V = (uint32_t)sl.reqs.size(), ts = sv->cfg.tile_size;
  SceneTab st;
  RecTable rt;
  TileTable tt;
  ViewTable vt;
  U32Pack tb;
  memset(&st, 0, sizeof st);
  memset(&rt, 0, sizeof rt);
  memset(&tt, 0, sizeof tt);
  memset(&vt, 0, sizeof vt);
  vt.grad_words = NTS_GRAD_WORDS;
  memset(&tb, 0, sizeof tb);
  rt.n = tt.n = vt.n = V;
  tb.n = V + 1;
  uint32_t rec = 0, tiles = 0, max_tiles_view = 1;
  uint64_t pix = 0;
  sl.out_off.assign(V, 0);
  sl.pix.assign(V, 0);
  sl.stage_off.assign(V, 0);
  uint64_t stage = 0;
  for (uint32_t v = 0; v < V; ++v) {
    const Request& r = *sl.reqs[v];
    const nts_scene* sc = sv->scenes[r.scene].scene;
    st.params[v] = sc->params;
    st.cap[v] = sc->cap;
    rt.base[v] = rec;
    const uint32_t w = r.view.width, h = r.view.height;
    const uint32_t tx = (w + ts - 1) / ts, ty = (h + ts - 1) / ts;
    tt.t[v] = ViewTiles{tx, ty, ts, r.view.depth_max};
    ViewDesc& d = vt.v[v];
    d.width = w;
    d.height = h;
    d.tiles_x = tx;
    d.tiles_y = ty;
    d.tile_base = tiles;
    d.proj_base = rec;
    d.pix_base = pix;
    tb.v[v] = tiles;
    sl.out_off[v] = pix;
    sl.pix[v] = (uint64_t)w * h;
    sl.stage_off[v] = stage;
    stage += sl.pix[v] * (r.format == NTS_OUT_RGBA8 ? 4 : 16);
    rec += sc->n;
    tiles += tx * ty;
    pix += (uint64_t)w * h;
    if (tx * ty > max_tiles_view) max_tiles_view = tx * ty;
    memcpy(&sl.h_cams[v], &r.view.cam, sizeof(DevCamera));
  }
  rt.base[V] = rec;
  vt.total_tiles = tiles;
  tb.v[V] = tiles;
  // culled pairs carry the all-ones key: tile ids below 2^tbits - 1
  const uint32_t tbits = nts_tile_bits_strict(max_tiles_view);
  const uint32_t kb = key_bytes_for(max_tiles_view);
  // the batch's working set, laid out in the slot's workspace with every
  // byte left over given to the (splat, tile) pairs
  uint64_t cap = pairs_that_fit(sl.bytes, V, rec, pix, tiles, kb);
  if (cap > sv->pair_cap) cap = sv->pair_cap;
  if (cap == 0)
    return nts_fail(NTS_E_CAPACITY, "batch of %u views does not fit a slot's %llu bytes", V,
                    (unsigned long long)sl.bytes);
  {
    const BatchLayout L = layout_batch(V, rec, pix, tiles, kb, cap);
    char* p = (char*)sl.base;
    sl.d_cams = (DevCamera*)(p + L.cams);
    sl.proj = (Projected*)(p + L.proj);
    sl.counts = (uint32_t*)(p + L.counts);
    sl.d_total = (uint64_t*)(p + L.total);
    sl.seg_begin = (uint64_t*)(p + L.seg);
    sl.tile_base = (uint32_t*)(p + L.tb);
    sl.flags = (uint32_t*)(p + L.flags);
    sl.keys_a = p + L.ka;
    sl.keys_b = p + L.kb;
    sl.vals_a = (uint32_t*)(p + L.va);
    sl.vals_b = (uint32_t*)(p + L.vb);
    sl.ranges = (uint2*)(p + L.rg);
    sl.order = (uint32_t*)(p + L.order);
    sl.image = (float4*)(p + L.img);
    sl.rgba = (uint32_t*)(p + L.rgba);
    sl.scan_ws = p + L.scan;
    sl.sort_ws = p + L.sort;
    sl.pair_cap = cap;
  }
  NTS_HIP_TRY(hipMemcpyAsync(sl.d_cams, sl.h_cams, sizeof(DevCamera) * V, hipMemcpyHostToDevice, ctx->stream));
  int rc = nts_fill_u32(ctx, sl.flags, 0, 16);
  if (rc) return rc;
  rc = nts_fill_u32(ctx, sl.ranges, 0, (uint64_t)tiles * 2);
  if (rc) return rc;
  if (rec > 0)
    NTS_LAUNCH(ctx, "serve_project", k_project_requests, dim3(nts_div_up(rec, 256)), dim3(256), 0, st, rt,
               (const DevCamera*)sl.d_cams, tt, sl.proj, sl.counts);
  rc = nts_scan_u32_ws(ctx, sl.counts, sl.counts, rec > 0 ? rec : 1, sl.d_total, sl.scan_ws);
  if (rc) return rc;
  NTS_LAUNCH(ctx, "serve_segments", k_request_segments, dim3(1), dim3(64), 0, (const uint32_t*)sl.counts, rt,
             (const uint64_t*)sl.d_total, sl.pair_cap, sl.seg_begin, sl.tile_base, tb, sl.flags);
  bool alt = false;
  if (rec > 0) {
    if (kb == 4)
      NTS_LAUNCH(ctx, "serve_emit", (k_emit_balanced<uint32_t, true>), dim3(nts_div_up(rec, 256)), dim3(256), 0,
                 (const Projected*)sl.proj, rt, tt, (const uint32_t*)sl.counts, (const uint64_t*)sl.d_total,
                 sl.pair_cap, (uint32_t*)sl.keys_a, sl.vals_a);
    else
      NTS_LAUNCH(ctx, "serve_emit", (k_emit_balanced<uint64_t, true>), dim3(nts_div_up(rec, 256)), dim3(256), 0,
                 (const Projected*)sl.proj, rt, tt, (const uint32_t*)sl.counts, (const uint64_t*)sl.d_total,
                 sl.pair_cap, (uint64_t*)sl.keys_a, sl.vals_a);
  }
  // the sort writes the tile ranges (zero-filled above; tile bases set by
  // the segments kernel) and the tiles' blend order
  if (kb == 4)
    rc = nts_bin_sort_ws_launch32(ctx, (uint32_t*)sl.keys_a, sl.vals_a, (uint32_t*)sl.keys_b, sl.vals_b, sl.seg_begin,
                                  V, sl.pair_cap, tbits, sl.tile_base, tiles, sl.ranges, sl.order, sl.sort_ws, &alt,
                                  &sl.sort_err);
  else
    rc = nts_bin_sort_ws_launch64(ctx, (uint64_t*)sl.keys_a, sl.vals_a, (uint64_t*)sl.keys_b, sl.vals_b, sl.seg_begin,
                                  V, sl.pair_cap, tbits, sl.tile_base, tiles, sl.ranges, sl.order, sl.sort_ws, &alt,
                                  &sl.sort_err);
  if (rc) return rc;
  const uint32_t* pay = alt ? sl.vals_b : sl.vals_a;
  vt.order = sl.order;
  rc = nts_blend_forward(ctx, ts, sv->cfg.lds_batch, sl.proj, pay, sl.ranges, vt, sl.image, nullptr, nullptr, nullptr);
  if (rc) return rc;
  // outputs: rgba8 (tonemapped) for the requests that asked for it, float
  // images for the others, each copied to its offset in the pinned staging
  bool any8 = false;
  for (uint32_t v = 0; v < V; ++v) any8 = any8 || sl.reqs[v]->format == NTS_OUT_RGBA8;
  if (any8) {
    rc = nts_tonemap_dev(ctx, sl.image, sl.rgba, pix);
    if (rc) return rc;
  }
  for (uint32_t v = 0; v < V; ++v) {
    if (sl.reqs[v]->format == NTS_OUT_RGBA8)
      NTS_HIP_TRY(hipMemcpyAsync(sl.h_out + sl.stage_off[v], sl.rgba + sl.out_off[v], sl.pix[v] * 4,
                                 hipMemcpyDeviceToHost, ctx->stream));
    else
      NTS_HIP_TRY(hipMemcpyAsync(sl.h_out + sl.stage_off[v], sl.image + sl.out_off[v], sl.pix[v] * 16,
                                 hipMemcpyDeviceToHost, ctx->stream));
  }
  sl.h_meta[1] = 0;  // the slot is idle: no copy into h_meta is in flight
  NTS_HIP_TRY(hipMemcpyAsync(sl.h_meta, sl.flags, 8, hipMemcpyDeviceToHost, ctx->stream));
  NTS_HIP_TRY(hipMemcpyAsync(sl.h_meta + 1, sl.sort_err, 4, hipMemcpyDeviceToHost, ctx->stream));
  NTS_HIP_TRY(hipMemcpyAsync(sl.h_meta + 2, sl.seg_begin, 8 * (V + 1), hipMemcpyDeviceToHost, ctx->stream));
  NTS_HIP_TRY(hipEventRecord(sl.done, ctx->stream));
  return NTS_OK;
}

void finish(nts_server* sv, Request* r, int state, int err) {
  r->state = state;
  r->err = err;
  r->t_done = now_ms();
  if (state == REQ_DONE) {
    sv->stats.completed++;
  } else {
    sv->stats.failed++;
  }
}

// Estimated pairs of a request: learned pairs per pixel, else the prior
// of 8 tiles per splat.
double est_pairs(nts_server* sv, const Request& r) {
  const Resident& rs = sv->scenes[r.scene];
  const double px = (double)r.view.width * r.view.height;
  return rs.pairs_per_px > 0.0 ? rs.pairs_per_px * px * 1.25 : 8.0 * rs.scene->n;
}

bool fits_alone(nts_server* sv, const Request& r) {
  const uint64_t px = (uint64_t)r.view.width * r.view.height;
  const uint64_t n = sv->scenes[r.scene].scene->n;
  const uint64_t t = tiles_of(r.view, sv->cfg.tile_size);
  return px <= sv->cfg.max_pixels && n <= sv->max_records && tiles_of(r.view, 8) <= sv->max_tiles &&
         pairs_that_fit(sv->slot_bytes, 1, n, px, t, key_bytes_for(t)) > 0;
}

// Form a batch for a free slot from the FIFO head.  Caller holds the lock.
// A request joins while the batch's working set (records, estimated pairs,
// sort workspace, pixels: layout_batch) still fits the slot's workspace and
// the per-batch caps hold; the first request that does not fit waits for the
// next slot (FIFO).
void form_batch(nts_server* sv, Slot& sl) {
  sl.reqs.clear();
  uint64_t px = 0, rec = 0, tiles = 0, max_view_tiles = 1;
  double pairs = 0.0;
  while (!sv->pending.empty() && sl.reqs.size() < sv->cfg.max_views) {
    Request* r = sv->pending.front();
    if (!fits_alone(sv, *r)) {
      sv->pending.pop_front();
      finish(sv, r, REQ_FAILED, NTS_E_CAPACITY);
      continue;
    }
    if (r->solo && !sl.reqs.empty()) break;
    const uint64_t rpx = (uint64_t)r->view.width * r->view.height;
    const uint64_t rrec = sv->scenes[r->scene].scene->n;
    const uint64_t rt = tiles_of(r->view, sv->cfg.tile_size);
    const double rp = est_pairs(sv, *r);
    if (!sl.reqs.empty()) {
      const bool caps = px + rpx <= sv->cfg.max_pixels && rec + rrec <= sv->max_records &&
                        tiles + rt <= sv->max_tiles;
      const uint64_t mvt = rt > max_view_tiles ? rt : max_view_tiles;
      uint64_t room = pairs_that_fit(sv->slot_bytes, (uint32_t)sl.reqs.size() + 1, rec + rrec, px + rpx,
                                     tiles + rt, key_bytes_for(mvt));
      if (room > sv->pair_cap) room = sv->pair_cap;
      const bool memory = pairs + rp <= 0.9 * (double)room;
      if (!caps || !memory) {  // FIFO: the head waits for the next slot
        if (caps) sv->stats.memory_limited++;
        break;
      }
    }
    sv->pending.pop_front();
    r->state = REQ_RUNNING;
    sl.reqs.push_back(r);
    px += rpx;
    rec += rrec;
    tiles += rt;
    pairs += rp;
    if (rt > max_view_tiles) max_view_tiles = rt;
    if (r->solo) break;
  }
}

void retire(nts_server* sv, Slot& sl) {
  sl.busy = false;
  const uint32_t V = (uint32_t)sl.reqs.size();
  const bool overflow = (sl.h_meta[0] & 0xffffffffu) != 0;
  const uint32_t sort_err = (uint32_t)(sl.h_meta[1] & 0xffffffffu);
  sv->stats.batches++;
  sv->stats.views += V;
  if (overflow || sort_err) {
    sv->stats.overflow_reruns++;
    // re-run one at a time, ahead of everything else; a lone request that
    // overflows cannot be served
    for (uint32_t v = V; v-- > 0;) {
      Request* r = sl.reqs[v];
      if (r->solo || V == 1) {
        finish(sv, r, REQ_FAILED, NTS_E_CAPACITY);
      } else {
        r->solo = true;
        r->state = REQ_PENDING;
        sv->pending.push_front(r);
      }
    }
    sl.reqs.clear();
    return;
  }
  const uint64_t* seg = sl.h_meta + 2;
  for (uint32_t v = 0; v < V; ++v) {
    Request* r = sl.reqs[v];
    r->pairs = seg[v + 1] - seg[v];
    r->batch_views = V;
    Resident& rs = sv->scenes[r->scene];
    const double ppx = (double)r->pairs / ((double)r->view.width * r->view.height);
    rs.pairs_per_px = rs.pairs_per_px > 0.0 ? (ppx > 0.9 * rs.pairs_per_px ? ppx : 0.9 * rs.pairs_per_px) : ppx;
    // the result leaves the pinned staging in pieces, on the copier
    // threads; the request is done with its last piece, the slot free with
    // the batch's last
    const size_t bytes = sl.pix[v] * (r->format == NTS_OUT_RGBA8 ? 4 : 16);
    const size_t chunk = (size_t)4 << 20;
    const uint32_t pieces = (uint32_t)((bytes + chunk - 1) / chunk);
    r->chunks_left = pieces;
    sl.copies_left += pieces;
    for (uint32_t c = 0; c < pieces; ++c) {
      const size_t off = (size_t)c * chunk;
      sv->jobs.push_back(CopyJob{r, &sl, sl.h_out + sl.stage_off[v] + off, (uint8_t*)r->out + off,
                                 bytes - off < chunk ? bytes - off : chunk});
    }
  }
  sl.reqs.clear();
  sv->job_cv.notify_all();
}

// A copier thread: result pieces from the pinned staging to the callers'
// buffers, outside the lock.
void copier(nts_server* sv) {
  std::unique_lock<std::mutex> lk(sv->mu);
  for (;;) {
    sv->job_cv.wait(lk, [&] { return sv->stop_copiers || !sv->jobs.empty(); });
    if (sv->jobs.empty()) return;  // stopping, nothing left
    const CopyJob j = sv->jobs.front();
    sv->jobs.pop_front();
    lk.unlock();
    memcpy(j.dst, j.src, j.bytes);
    lk.lock();
    if (--j.r->chunks_left == 0) {
      finish(sv, j.r, REQ_DONE, NTS_OK);
      sv->done_cv.notify_all();
    }
    if (--j.slot->copies_left == 0) sv->cv.notify_one();  // the slot is free again
  }
}

void run(nts_server* sv) {
  (void)hipSetDevice(sv->device);
  std::unique_lock<std::mutex> lk(sv->mu);
  while (!sv->stop) {
    bool progress = false;
    uint32_t in_flight = 0;
    for (Slot& sl : sv->slots) {
      if (sl.busy && hipEventQuery(sl.done) == hipSuccess) {
        retire(sv, sl);
        progress = true;
        sv->done_cv.notify_all();
      }
    }
    for (Slot& sl : sv->slots) {
      if (sl.busy || sl.copies_left || sv->pending.empty()) continue;
      form_batch(sv, sl);
      if (sl.reqs.empty()) {
        sv->done_cv.notify_all();  // requests failed admission
        continue;
      }
      const int rc = launch_batch(sv, sl);
      if (rc) {
        for (Request* r : sl.reqs) finish(sv, r, REQ_FAILED, rc);
        sl.reqs.clear();
        sv->done_cv.notify_all();
        continue;
      }
      sl.busy = true;
      progress = true;
    }
    for (Slot& sl : sv->slots) in_flight += (sl.busy || sl.copies_left) ? 1u : 0u;
    if (in_flight > sv->stats.max_in_flight) sv->stats.max_in_flight = in_flight;
    if (!progress) sv->cv.wait_for(lk, std::chrono::microseconds(in_flight ? 20 : 200));
  }
}

}  // namespace

extern "C" int nts_server_create(int device, const nts_server_cfg* cfg, nts_server** out) {
  NTS_CHECK_ARG(cfg && out && cfg->slots >= 1 && cfg->slots <= 8);
  NTS_CHECK_ARG(cfg->max_views >= 1 && cfg->max_views <= NTS_MAX_VIEWS);
  NTS_CHECK_ARG(cfg->max_pixels > 0 && cfg->max_records < (1ull << 32) && cfg->pair_capacity < (1ull << 30));
  // without slot_bytes every cap sizes the workspace; with it, 0 = no cap
  NTS_CHECK_ARG(cfg->slot_bytes > 0 || (cfg->max_records > 0 && cfg->pair_capacity > 0));
  NTS_CHECK_ARG(nts_blend_config_ok(cfg->tile_size, cfg->lds_batch, 64));
  *out = nullptr;
  nts_server* sv = new nts_server();
  sv->device = device;
  sv->cfg = *cfg;
  sv->max_tiles = cfg->max_pixels / 64 + (uint64_t)cfg->max_views * 1024 + 64;
  sv->max_records = cfg->max_records ? cfg->max_records : 0xffffffffull;
  sv->pair_cap = cfg->pair_capacity ? cfg->pair_capacity : (1ull << 30) - 1;
  // a workspace that holds the caps at once (widest keys), unless given
  sv->slot_bytes = cfg->slot_bytes ? al256(cfg->slot_bytes)
                                   : layout_batch(cfg->max_views, cfg->max_records, cfg->max_pixels, sv->max_tiles,
                                                  8, cfg->pair_capacity).end;
  int rc = nts_ctx_create(device, 0, &sv->ctx0);
  if (rc == NTS_OK) {
    sv->slots.resize(cfg->slots);
    for (Slot& sl : sv->slots) {
      rc = slot_create(sv, sl);
      if (rc) break;
    }
  }
  if (rc) {
    for (Slot& sl : sv->slots) slot_destroy(sv, sl);
    if (sv->ctx0) nts_ctx_destroy(sv->ctx0);
    delete sv;
    return rc;
  }
  sv->th = std::thread(run, sv);
  // result copy-out: enough threads for a host's memory bandwidth, not more
  // than the slots can keep busy
  const unsigned hw = std::thread::hardware_concurrency();
  const unsigned nc = hw >= 16 ? 8u : (hw >= 4 ? hw / 2 : 1u);
  for (unsigned k = 0; k < nc; ++k) sv->copiers.emplace_back(copier, sv);
  *out = sv;
  return NTS_OK;
}

extern "C" int nts_server_destroy(nts_server* sv) {
  if (!sv) return NTS_OK;
  {
    std::lock_guard<std::mutex> g(sv->mu);
    sv->stop = true;
  }
  sv->cv.notify_all();
  if (sv->th.joinable()) sv->th.join();
  {
    std::lock_guard<std::mutex> g(sv->mu);
    sv->stop_copiers = true;  // the copiers finish the queued pieces first
  }
  sv->job_cv.notify_all();
  for (std::thread& t : sv->copiers)
    if (t.joinable()) t.join();
  for (Slot& sl : sv->slots) slot_destroy(sv, sl);
  for (Resident& r : sv->scenes) nts_scene_destroy(sv->ctx0, r.scene);
  for (auto& kv : sv->reqs) delete kv.second;
  nts_ctx_destroy(sv->ctx0);
  delete sv;
  return NTS_OK;
}

extern "C" int nts_server_add_scene(nts_server* sv, const nts_splats_host* host, uint32_t* scene_id) {
  NTS_CHECK_ARG(sv && host && scene_id && host->n > 0);
  if (host->n > sv->max_records) return nts_fail(NTS_E_CAPACITY, "scene of %u splats exceeds a slot's %llu records",
                                                 host->n, (unsigned long long)sv->max_records);
  const uint64_t need = (((uint64_t)host->n + 63) & ~63ull) * 4 * P_COUNT;
  std::lock_guard<std::mutex> g(sv->mu);
  if (sv->cfg.hbm_budget && sv->used + need > sv->cfg.hbm_budget)
    return nts_fail(NTS_E_NOMEM, "scene of %llu bytes exceeds the HBM budget (%llu of %llu in use)",
                    (unsigned long long)need, (unsigned long long)sv->used, (unsigned long long)sv->cfg.hbm_budget);
  nts_scene* s = nullptr;
  int rc = nts_scene_alloc_params(sv->ctx0, host->n, &s);
  if (rc) return rc;
  s->n = host->n;
  rc = nts_upload_planes(sv->ctx0, s, s->params, host);
  if (rc) {
    nts_scene_destroy(sv->ctx0, s);
    return rc;
  }
  sv->used += s->bytes;
  Resident r;
  r.scene = s;
  sv->scenes.push_back(r);
  *scene_id = (uint32_t)(sv->scenes.size() - 1);
  return NTS_OK;
}

// A resident scene straight from a 3DGS PLY file (nano-splat ply.rs:116-134).
extern "C" int nts_server_add_scene_ply(nts_server* sv, const char* path, uint32_t* scene_id) {
  NTS_CHECK_ARG(sv && path && scene_id);
  uint32_t n = 0;
  int rc = nts_ply_count(path, &n);
  if (rc) return rc;
  if (n == 0) return nts_fail(NTS_E_INVALID, "%s holds no splats", path);
  std::vector<float> pos(3 * (size_t)n), rot(4 * (size_t)n), ls(3 * (size_t)n), lg(n), dc(3 * (size_t)n),
      rest(45 * (size_t)n);
  nts_splats_host h;
  h.n = n;
  h.pos = pos.data();
  h.rot = rot.data();
  h.log_scale = ls.data();
  h.logit = lg.data();
  h.sh_dc = dc.data();
  h.sh_rest = rest.data();
  rc = nts_ply_read(path, &h);
  if (rc) return rc;
  return nts_server_add_scene(sv, &h, scene_id);
}

extern "C" int nts_server_submit(nts_server* sv, uint32_t scene_id, const nts_view_desc* view, int format, void* out,
                                 uint64_t* request_id) {
  NTS_CHECK_ARG(sv && view && out && request_id);
  NTS_CHECK_ARG(format == NTS_OUT_RGBA8 || format == NTS_OUT_F32);
  NTS_CHECK_ARG(view->width > 0 && view->height > 0);
  NTS_CHECK_ARG((float)view->width == view->cam.width && (float)view->height == view->cam.height);
  Request* r = new Request();
  r->view = *view;
  r->format = format;
  r->out = out;
  {
    std::lock_guard<std::mutex> g(sv->mu);
    if (scene_id >= sv->scenes.size()) {
      delete r;
      return nts_fail(NTS_E_INVALID, "unknown scene %u", scene_id);
    }
    r->scene = scene_id;
    r->id = sv->next_id++;
    r->t_submit = now_ms();
    sv->reqs[r->id] = r;
    sv->pending.push_back(r);
    sv->stats.submitted++;
    *request_id = r->id;
  }
  sv->cv.notify_one();
  return NTS_OK;
}

static int take_result(nts_server* sv, uint64_t id, Request* r, nts_request_info* info) {
  if (info) {
    info->latency_ms = r->t_done - r->t_submit;
    info->pairs = r->pairs;
    info->batch_views = r->batch_views;
    info->status = r->err;
  }
  const int err = r->err;
  sv->reqs.erase(id);
  delete r;
  return err == NTS_OK ? NTS_OK : nts_fail(err, "request %llu failed (%d)", (unsigned long long)id, err);
}

extern "C" int nts_server_poll(nts_server* sv, uint64_t request_id, int* done, nts_request_info* info) {
  NTS_CHECK_ARG(sv && done);
  std::lock_guard<std::mutex> g(sv->mu);
  auto it = sv->reqs.find(request_id);
  if (it == sv->reqs.end()) return nts_fail(NTS_E_INVALID, "unknown request %llu", (unsigned long long)request_id);
  Request* r = it->second;
  *done = (r->state == REQ_DONE || r->state == REQ_FAILED) ? 1 : 0;
  if (!*done) return NTS_OK;
  return take_result(sv, request_id, r, info);
}

extern "C" int nts_server_wait(nts_server* sv, uint64_t request_id, nts_request_info* info) {
  NTS_CHECK_ARG(sv);
  std::unique_lock<std::mutex> lk(sv->mu);
  auto it = sv->reqs.find(request_id);
  if (it == sv->reqs.end()) return nts_fail(NTS_E_INVALID, "unknown request %llu", (unsigned long long)request_id);
  Request* r = it->second;
  sv->done_cv.wait(lk, [&] { return r->state == REQ_DONE || r->state == REQ_FAILED; });
  return take_result(sv, request_id, r, info);
}

extern "C" int nts_server_stats(nts_server* sv, nts_serve_stats* out) {
  NTS_CHECK_ARG(sv && out);
  std::lock_guard<std::mutex> g(sv->mu);
  *out = sv->stats;
  out->hbm_in_use = sv->used;
  out->scenes = (uint32_t)sv->scenes.size();
  out->pending = (uint32_t)sv->pending.size();
  return NTS_OK;
}

// Free and total device memory (hipMemGetInfo): the serving budget's source.
extern "C" int nts_device_memory(int device, uint64_t* free_bytes, uint64_t* total_bytes) {
  NTS_CHECK_ARG(free_bytes && total_bytes);
  int prev = 0;
  NTS_HIP_TRY(hipGetDevice(&prev));
  NTS_HIP_TRY(hipSetDevice(device));
  size_t f = 0, t = 0;
  const hipError_t e = hipMemGetInfo(&f, &t);
  (void)hipSetDevice(prev);
  if (e != hipSuccess) return nts_fail(NTS_E_HIP, "hipMemGetInfo: %s", hipGetErrorString(e));
  *free_bytes = f;
  *total_bytes = t;
  return NTS_OK;
}
