// Binning kernels shared by the single-view and multi-view paths.
#pragma once
#include "nts_geom.h"
#include "nts_sort.h"

// Tile geometry of one view.
struct ViewTiles {
  uint32_t tiles_x, tiles_y;
  uint32_t tile_size;
  float depth_max;
};

// tile_emit.wgsl:91-121: key = tile << 16 | depth16, payload = splat index,
// written from offsets[i] in ty-major, tx-minor order.
template <typename K>
__global__ void __launch_bounds__(256) k_tile_emit(const Projected* __restrict__ proj, uint32_t n, ViewTiles vt,
                                                   const uint32_t* __restrict__ offsets, uint64_t capacity,
                                                   K* __restrict__ keys, uint32_t* __restrict__ payloads) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Projected p = proj[i];
  if (p.b.w == 0.0f) return;
  int32_t x0, y0, x1, y1;
  if (!tile_bbox_dev(p.a.x, p.a.y, p.a.w, (float)vt.tile_size, (int32_t)vt.tiles_x, (int32_t)vt.tiles_y, x0, y0, x1,
                     y1))
    return;
  const K d16 = (K)depth16(p.a.z, vt.depth_max);
  uint32_t w = offsets[i];
  for (int32_t ty = y0; ty <= y1; ++ty)
    for (int32_t tx = x0; tx <= x1; ++tx) {
      const K tile = (K)((uint32_t)ty * vt.tiles_x + (uint32_t)tx);
      if ((uint64_t)w < capacity) {
        keys[w] = (tile << 16) | d16;
        payloads[w] = i;
      }
      ++w;
    }
}

// tile_ranges.wgsl:22-41 over segment s = [seg_begin[s], seg_begin[s+1]):
// ranges[tile_base[s] + tile] = [begin, end); tiles >= tiles_in[s] ignored.
template <typename K>
__global__ void __launch_bounds__(256) k_tile_ranges(const K* __restrict__ keys, const uint64_t* __restrict__ seg_begin,
                                                     uint32_t num_segs, const uint32_t* __restrict__ tile_base,
                                                     uint2* __restrict__ ranges) {
  const uint64_t total = seg_begin[num_segs];
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    const uint32_t s = seg_of<uint64_t>(seg_begin, num_segs, i);
    const uint64_t sb = seg_begin[s], se = seg_begin[s + 1];
    const K tile = keys[i] >> 16;
    const bool starts = (i == sb) || ((keys[i - 1] >> 16) != tile);
    const bool ends = (i + 1 == se) || ((keys[i + 1] >> 16) != tile);
    const uint32_t tb = tile_base[s], tn = tile_base[s + 1] - tb;
    if (tile >= (K)tn) continue;
    if (starts) ranges[tb + (uint32_t)tile].x = (uint32_t)i;
    if (ends) ranges[tb + (uint32_t)tile].y = (uint32_t)(i + 1);
  }
}

