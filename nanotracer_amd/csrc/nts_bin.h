// Binning kernels shared by the single-view and multi-view paths.
#pragma once
#include "nts_geom.h"
#include "nts_sort.h"

// Most views one launch bins (the batched step's limit).
#define NTS_MAX_VIEW_TILES 32

// Tile geometry of one view.
struct ViewTiles {
  uint32_t tiles_x, tiles_y;
  uint32_t tile_size;
  float depth_max;
};

// Tile geometry of every view of a launch.
struct TileTable {
  uint32_t n;
  uint32_t pad[3];
  ViewTiles t[NTS_MAX_VIEW_TILES];
};

// Records of a launch: view v owns records [base[v], base[v+1]) (one per
// splat of the view's scene, in splat order).
struct RecTable {
  uint32_t n;
  uint32_t pad[3];
  uint32_t base[NTS_MAX_VIEW_TILES + 1];
};

__device__ __forceinline__ uint32_t rec_view(const RecTable& rt, uint32_t f) {
  uint32_t lo = 0, hi = rt.n;  // base[lo] <= f < base[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (rt.base[mid] <= f) lo = mid; else hi = mid;
  }
  return lo;
}

// tile_emit.wgsl:91-121 with the work spread over pairs instead of splats:
// key = tile << 16 | depth16, payload = splat index, a record's pairs
// written from offsets[record] in ty-major, tx-minor order.
//
// Records are (view, splat): record f = base[v] + i.  A block takes 256
// consecutive records, stages each record's first pair, first tile, bbox
// width and depth bucket in LDS, then its threads walk the block's pair
// range [offsets[f0], offsets[f0 + 256]) together: pair j belongs to the
// last record whose first pair is <= j (binary search in LDS), so every
// store is coalesced whatever the footprints.  offsets is the exclusive
// scan of the per-record tile counts; *total the pair total.
// CULL: only the tiles the record's alpha >= 1/255 box can reach
// (record_tiles), in the same order; the counts must come from the same rule.
// A pair of such a tile that the alpha >= 1/255 region itself cannot reach
// (nts_tile_cull_hits) is written with the all-ones key: it sorts after the
// view's live pairs and lies beyond every tile, so no tile range holds it.
template <typename K, bool CULL = false>
__global__ void __launch_bounds__(256) k_emit_balanced(const Projected* __restrict__ proj, RecTable rt,
                                                       TileTable tt, const uint32_t* __restrict__ offsets,
                                                       const uint64_t* __restrict__ total, uint64_t capacity,
                                                       K* __restrict__ keys, uint32_t* __restrict__ payloads) {
  __shared__ uint32_t s_off[257];
  __shared__ uint32_t s_tile0[256], s_bw[256], s_tx[256], s_d[256], s_pay[256];
  __shared__ NtsTileCull s_cull[CULL ? 256 : 1];
  __shared__ int32_t s_x0[CULL ? 256 : 1], s_y0[CULL ? 256 : 1];
  __shared__ float s_ts[CULL ? 256 : 1];
  const uint32_t t = threadIdx.x;
  const uint32_t nrec = rt.base[rt.n];
  const uint32_t f0 = blockIdx.x * 256u, f = f0 + t;
  uint32_t off = 0xFFFFFFFFu;  // past every pair: records beyond nrec own none
  if (f < nrec) {
    off = offsets[f];
    const uint32_t v = rec_view(rt, f), i = f - rt.base[v];
    const ViewTiles vt = tt.t[v];
    const Projected p = proj[f];
    int32_t x0 = 0, y0 = 0, x1 = -1, y1 = -1;
    record_tiles<CULL>(p, (float)vt.tile_size, (int32_t)vt.tiles_x, (int32_t)vt.tiles_y, x0, y0, x1, y1);
    s_tile0[t] = (uint32_t)y0 * vt.tiles_x + (uint32_t)x0;
    s_bw[t] = x1 >= x0 ? (uint32_t)(x1 - x0 + 1) : 1u;
    s_tx[t] = vt.tiles_x;
    s_d[t] = depth16(p.a.z, vt.depth_max);
    s_pay[t] = i;
    if (CULL) {
      s_cull[t] = nts_tile_cull_setup(p.a.x, p.a.y, p.b.x, p.b.y, p.b.z, p.c.w);
      s_x0[t] = x0;
      s_y0[t] = y0;
      s_ts[t] = (float)vt.tile_size;
    }
  }
  s_off[t] = off;
  if (t == 0) s_off[256] = f0 + 256u < nrec ? offsets[f0 + 256u] : (uint32_t)*total;
  __syncthreads();
  const uint32_t j0 = s_off[0], j1 = s_off[256];
  for (uint32_t j = j0 + t; j < j1; j += 256u) {
    uint32_t lo = 0, hi = 256;  // s_off[lo] <= j < s_off[hi]
#pragma unroll
    for (int it = 0; it < 8; ++it) {
      const uint32_t mid = (lo + hi) >> 1;
      if (s_off[mid] <= j) lo = mid; else hi = mid;
    }
    const uint32_t r = j - s_off[lo], bw = s_bw[lo];
    const uint32_t row = r / bw, col = r - row * bw;
    const uint32_t tile = s_tile0[lo] + row * s_tx[lo] + col;
    bool live = true;
    if (CULL) {
      float x0, x1, y0, y1;
      nts_tile_centres(s_x0[lo] + (int32_t)col, s_y0[lo] + (int32_t)row, s_ts[lo], &x0, &x1, &y0, &y1);
      live = nts_tile_cull_hits(s_cull[lo], x0, x1, y0, y1);
    }
    if ((uint64_t)j < capacity) {
      keys[j] = live ? (((K)tile << 16) | (K)s_d[lo]) : ~(K)0;
      payloads[j] = s_pay[lo];
    }
  }
}

// tile_ranges.wgsl:22-41 over segment s = [seg_begin[s], seg_begin[s+1]):
// ranges[tile_base[s] + tile] = [begin, end); tiles >= tiles_in[s] ignored.
// The segment tables sit in LDS; a thread's indices only grow, so its
// segment only moves forward.
template <typename K>
__global__ void __launch_bounds__(256) k_tile_ranges(const K* __restrict__ keys, const uint64_t* __restrict__ seg_begin,
                                                     uint32_t num_segs, const uint32_t* __restrict__ tile_base,
                                                     uint2* __restrict__ ranges) {
  // Each thread takes E consecutive keys (two 16-byte loads) and the key
  // before them: a tile starts where its key differs from the previous one,
  // which also ends the previous tile; a segment's last key ends its tile.
  constexpr uint32_t E = 32 / sizeof(K);
  __shared__ uint64_t s_sb[NTS_MAX_VIEW_TILES + 1];
  __shared__ uint32_t s_tb[NTS_MAX_VIEW_TILES + 1];
  for (uint32_t k = threadIdx.x; k <= num_segs; k += blockDim.x) {
    s_sb[k] = seg_begin[k];
    s_tb[k] = tile_base[k];
  }
  __syncthreads();
  const uint64_t total = s_sb[num_segs];
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x * E;
  uint32_t s = 0;
  for (uint64_t i0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * E; i0 < total; i0 += stride) {
    K k[E + 1];  // k[e + 1] = keys[i0 + e], k[0] = keys[i0 - 1]
    if (i0 + E <= total) {
      const uint4 v0 = *reinterpret_cast<const uint4*>(keys + i0);
      const uint4 v1 = *reinterpret_cast<const uint4*>(keys + i0 + E / 2);
      const K* a = reinterpret_cast<const K*>(&v0);
      const K* b = reinterpret_cast<const K*>(&v1);
#pragma unroll
      for (uint32_t e = 0; e < E / 2; ++e) {
        k[e + 1] = a[e];
        k[e + 1 + E / 2] = b[e];
      }
    } else {
#pragma unroll
      for (uint32_t e = 0; e < E; ++e) k[e + 1] = i0 + e < total ? keys[i0 + e] : (K)0;
    }
    k[0] = i0 > 0 ? keys[i0 - 1] : (K)0;
#pragma unroll
    for (uint32_t e = 0; e < E; ++e) {
      const uint64_t i = i0 + e;
      if (i >= total) break;
      while (s + 1 < num_segs && i >= s_sb[s + 1]) ++s;
      const uint64_t sb = s_sb[s], se = s_sb[s + 1];
      const uint32_t tb = s_tb[s], tn = s_tb[s + 1] - tb;
      const K tile = k[e + 1] >> 16, prev = k[e] >> 16;
      const bool starts = i == sb || prev != tile;
      if (starts && i > sb && prev < (K)tn) ranges[tb + (uint32_t)prev].y = (uint32_t)i;
      if (tile >= (K)tn) continue;
      if (starts) ranges[tb + (uint32_t)tile].x = (uint32_t)i;
      if (i + 1 == se) ranges[tb + (uint32_t)tile].y = (uint32_t)(i + 1);
    }
  }
}
