// Tile binning: per-splat tile counts, exclusive scan, key emission, and the
// binning sort with its per-tile ranges (TileBinner::bin,
// tile_binner.rs:181-382).
#include "nts_bin.h"
#include "nts_raster.h"

// tile_count.wgsl:28-55 (the total comes from the scan instead of an atomic).
__global__ void __launch_bounds__(256) k_tile_count(const Projected* __restrict__ proj, uint32_t n, ViewTiles vt,
                                                    uint32_t* __restrict__ counts) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 a = proj[i].a;
  const float vis = proj[i].b.w;
  uint32_t c = 0;
  int32_t x0, y0, x1, y1;
  if (vis != 0.0f &&
      tile_bbox_dev(a.x, a.y, a.w, (float)vt.tile_size, (int32_t)vt.tiles_x, (int32_t)vt.tiles_y, x0, y0, x1, y1))
    c = (uint32_t)((x1 - x0 + 1) * (y1 - y0 + 1));
  counts[i] = c;
}

// ---------------------------------------------------------------------------
extern "C" int nts_bins_destroy(nts_ctx* ctx, nts_bins* bins) {
  if (!ctx || !bins) return NTS_OK;
  (void)hipStreamSynchronize(ctx->stream);
  nts_dev_free(ctx, bins->base, bins->bytes);
  delete bins;
  return NTS_OK;
}

extern "C" int nts_bins_info(const nts_bins* bins, uint64_t* total_pairs, uint32_t* num_tiles) {
  NTS_CHECK_ARG(bins);
  if (total_pairs) *total_pairs = bins->total_pairs;
  if (num_tiles) *num_tiles = bins->num_tiles;
  return NTS_OK;
}

static inline uint32_t bits_for(uint32_t n) {  // ceil(log2(n)), at least 1
  uint32_t b = 1;
  while (b < 32 && (1ull << b) < n) ++b;
  return b;
}

// Offsets of the bins' single allocation.
struct BinLayout {
  uint64_t keys, keys2, vals, vals2, ranges, order, meta;
};

// Emit, sort, ranges and tile order into an allocated bins object.
static int bin_fill(nts_ctx* ctx, nts_bins* b, const nts_buffer* projected, uint32_t n, const ViewTiles& vt,
                    const uint32_t* counts, const BinLayout& lay, uint32_t key_bits) {
  const uint64_t total = b->total_pairs, num_tiles = b->num_tiles;
  const uint32_t kb = b->key_bytes;
  int rc = NTS_OK;
  char* base = (char*)b->base;
  void* keys_a = base + lay.keys;
  void* keys_b = base + lay.keys2;
  uint32_t* vals_a = (uint32_t*)(base + lay.vals);
  uint32_t* vals_b = (uint32_t*)(base + lay.vals2);
  b->ranges = (uint2*)(base + lay.ranges);
  uint64_t* seg_begin = (uint64_t*)(base + lay.meta);
  uint32_t* tile_base = (uint32_t*)(base + lay.meta + 64);
  uint32_t* order = (uint32_t*)(base + lay.order);
  b->keys = keys_a;
  b->payloads = vals_a;

  rc = nts_fill_u32(ctx, b->ranges, 0, num_tiles * 2);
  if (rc) return rc;
  {
    const uint64_t hs[2] = {0, total};
    const uint32_t ht[2] = {0, (uint32_t)num_tiles};
    NTS_HIP_TRY(hipMemcpyAsync(seg_begin, hs, 16, hipMemcpyHostToDevice, ctx->stream));
    NTS_HIP_TRY(hipMemcpyAsync(tile_base, ht, 8, hipMemcpyHostToDevice, ctx->stream));
    NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  }
  if (total == 0) return NTS_OK;  // empty fast path: ranges all (0, 0)

  bool alt = false;
  TileTable tt;
  memset(&tt, 0, sizeof tt);
  tt.n = 1;
  tt.t[0] = vt;
  RecTable rt;
  memset(&rt, 0, sizeof rt);
  rt.n = 1;
  rt.base[1] = n;
  if (kb == 4) {
    NTS_LAUNCH(ctx, "tile_emit", k_emit_balanced<uint32_t>, dim3(nts_div_up(n, 256)), dim3(256), 0,
               (const Projected*)projected->ptr, rt, tt, (const uint32_t*)counts, (const uint64_t*)seg_begin + 1,
               total, (uint32_t*)keys_a, vals_a);
    rc = nts_bin_sort32_dev(ctx, (uint32_t*)keys_a, vals_a, (uint32_t*)keys_b, vals_b, seg_begin, 1, total,
                            key_bits - 16, tile_base, (uint32_t)num_tiles, b->ranges, order, &alt);
  } else {
    NTS_LAUNCH(ctx, "tile_emit", k_emit_balanced<uint64_t>, dim3(nts_div_up(n, 256)), dim3(256), 0,
               (const Projected*)projected->ptr, rt, tt, (const uint32_t*)counts, (const uint64_t*)seg_begin + 1,
               total, (uint64_t*)keys_a, vals_a);
    rc = nts_bin_sort64_dev(ctx, (uint64_t*)keys_a, vals_a, (uint64_t*)keys_b, vals_b, seg_begin, 1, total,
                            key_bits - 16, tile_base, (uint32_t)num_tiles, b->ranges, order, &alt);
  }
  if (rc) return rc;
  if (alt) {
    b->keys = keys_b;
    b->payloads = vals_b;
  }
  // the blend order (written by the sort) is published only on success
  b->order = order;
  return NTS_OK;
}

extern "C" int nts_bin(nts_ctx* ctx, const nts_buffer* projected, uint32_t n, const nts_tiling* tiling,
                       nts_bins** out) {
  NTS_CHECK_ARG(ctx && projected && tiling && out);
  NTS_CHECK_ARG(tiling->tile_size > 0 && tiling->width > 0 && tiling->height > 0);
  NTS_CHECK_ARG(projected->bytes >= (uint64_t)n * sizeof(Projected));
  *out = nullptr;
  ViewTiles vt;
  vt.tile_size = tiling->tile_size;
  vt.tiles_x = (tiling->width + tiling->tile_size - 1) / tiling->tile_size;
  vt.tiles_y = (tiling->height + tiling->tile_size - 1) / tiling->tile_size;
  vt.depth_max = tiling->depth_max;
  const uint64_t num_tiles = (uint64_t)vt.tiles_x * vt.tiles_y;
  if (num_tiles >= (1ull << 31)) return nts_fail(NTS_E_CAPACITY, "too many tiles");
  const uint32_t tile_bits = bits_for((uint32_t)num_tiles);
  const uint32_t key_bits = 16 + tile_bits;
  const uint32_t kb = key_bits <= 32 ? 4 : 8;

  // counts / offsets and the 64-bit total
  const uint64_t n1 = n > 0 ? n : 1;
  void* tmp = nullptr;
  const uint64_t tmp_bytes = ((n1 * 4 + 255) & ~255ull) + 256;
  int rc = nts_dev_alloc(ctx, &tmp, tmp_bytes);
  if (rc) return rc;
  uint32_t* counts = (uint32_t*)tmp;
  uint64_t* d_total = (uint64_t*)((char*)tmp + ((n1 * 4 + 255) & ~255ull));
  struct Guard {
    nts_ctx* c; void* p; uint64_t b;
    ~Guard() { (void)hipStreamSynchronize(c->stream); nts_dev_free(c, p, b); }
  } guard{ctx, tmp, tmp_bytes};

  if (n > 0)
    NTS_LAUNCH(ctx, "tile_count", k_tile_count, dim3(nts_div_up(n, 256)), dim3(256), 0,
               (const Projected*)projected->ptr, n, vt, counts);
  rc = nts_scan_u32_dev(ctx, counts, counts, n, d_total);
  if (rc) return rc;
  uint64_t total = 0;
  NTS_HIP_TRY(hipMemcpyAsync(&total, d_total, 8, hipMemcpyDeviceToHost, ctx->stream));
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));  // the reference's blocking total readback
  if (total >= (1ull << 30)) return nts_fail(NTS_E_CAPACITY, "%llu pairs exceed 2^30", (unsigned long long)total);

  nts_bins* b = new nts_bins();
  b->total_pairs = total;
  b->num_tiles = (uint32_t)num_tiles;
  b->tiles_x = vt.tiles_x;
  b->tiles_y = vt.tiles_y;
  b->tile_size = vt.tile_size;
  b->key_bytes = kb;
  const uint64_t k1 = total > 0 ? total : 1;
  auto al = [](uint64_t x) { return (x + 255) & ~255ull; };
  const uint64_t off_keys = 0;
  const uint64_t off_keys2 = al(k1 * kb);
  const uint64_t off_vals = off_keys2 + al(k1 * kb);
  const uint64_t off_vals2 = off_vals + al(k1 * 4);
  const uint64_t off_ranges = off_vals2 + al(k1 * 4);
  const uint64_t off_order = off_ranges + al(num_tiles * 8);
  const uint64_t off_meta = off_order + al(num_tiles * 4);
  b->bytes = off_meta + 256;
  rc = nts_dev_alloc(ctx, &b->base, b->bytes);
  if (rc) {
    delete b;
    return rc;
  }
  const BinLayout lay{off_keys, off_keys2, off_vals, off_vals2, off_ranges, off_order, off_meta};
  rc = bin_fill(ctx, b, projected, n, vt, counts, lay, key_bits);
  if (rc) {  // never hand out a half-built bins object
    nts_bins_destroy(ctx, b);
    return rc;
  }
  *out = b;
  return NTS_OK;
}

extern "C" int nts_bins_download_keys_u32(nts_ctx* ctx, const nts_bins* bins, uint32_t* host) {
  NTS_CHECK_ARG(ctx && bins && host);
  if (bins->total_pairs == 0) return NTS_OK;
  if (bins->key_bytes == 4) {
    NTS_HIP_TRY(hipMemcpyAsync(host, bins->keys, bins->total_pairs * 4, hipMemcpyDeviceToHost, ctx->stream));
    NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return NTS_OK;
  }
  std::vector<uint64_t> k(bins->total_pairs);
  NTS_HIP_TRY(hipMemcpyAsync(k.data(), bins->keys, bins->total_pairs * 8, hipMemcpyDeviceToHost, ctx->stream));
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (uint64_t i = 0; i < bins->total_pairs; ++i) {
    if (k[i] >> 32) return nts_fail(NTS_E_CAPACITY, "tile id does not fit the 16-bit reference key");
    host[i] = (uint32_t)k[i];
  }
  return NTS_OK;
}

extern "C" int nts_bins_download_keys_u64(nts_ctx* ctx, const nts_bins* bins, uint64_t* host) {
  NTS_CHECK_ARG(ctx && bins && host);
  if (bins->total_pairs == 0) return NTS_OK;
  if (bins->key_bytes == 8) {
    NTS_HIP_TRY(hipMemcpyAsync(host, bins->keys, bins->total_pairs * 8, hipMemcpyDeviceToHost, ctx->stream));
    NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return NTS_OK;
  }
  std::vector<uint32_t> k(bins->total_pairs);
  NTS_HIP_TRY(hipMemcpyAsync(k.data(), bins->keys, bins->total_pairs * 4, hipMemcpyDeviceToHost, ctx->stream));
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (uint64_t i = 0; i < bins->total_pairs; ++i) host[i] = k[i];
  return NTS_OK;
}

extern "C" int nts_bins_download_payloads(nts_ctx* ctx, const nts_bins* bins, uint32_t* host) {
  NTS_CHECK_ARG(ctx && bins && host);
  if (bins->total_pairs == 0) return NTS_OK;
  NTS_HIP_TRY(hipMemcpyAsync(host, bins->payloads, bins->total_pairs * 4, hipMemcpyDeviceToHost, ctx->stream));
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return NTS_OK;
}

extern "C" int nts_bins_download_ranges(nts_ctx* ctx, const nts_bins* bins, uint32_t* host) {
  NTS_CHECK_ARG(ctx && bins && host);
  NTS_HIP_TRY(hipMemcpyAsync(host, bins->ranges, (uint64_t)bins->num_tiles * 8, hipMemcpyDeviceToHost, ctx->stream));
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return NTS_OK;
}
