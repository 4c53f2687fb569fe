// Device-wide primitives: block scans, the exclusive u32 scan and the
// segmented onesweep LSD radix sort (Adinets & Merrill, "Onesweep: A Faster
// Least Significant Digit Radix Sort for GPUs", 2022): one upfront pass
// builds the 8-bit digit histograms of every pass in LDS, then each pass is a
// single kernel in which a block ranks its 4096 keys per digit with wave64
// ballots, publishes its per-digit counts and looks back over the preceding
// blocks of its segment (decoupled look-back) for its global digit offsets,
// reorders its keys in LDS and writes them out digit-contiguously.  In sorts
// of RS_TABLE_MIN_PAIRS pairs or more, the passes in RS_TABLE_MASK take their
// block offsets from per-partition digit counts instead (a counting pass over
// the pass's input; pass 0's written by the histogram pass when one runs),
// scanned over 16-partition chunks and then over each segment's chunks: no
// block waits on another.  When every pass takes counts, the chunk scans also
// give the digit histograms and no histogram pass runs.
#pragma once
#include "nts_internal.h"

#define RS_BLOCK 256
#ifndef RS_ITEMS
#define RS_ITEMS 16
#endif
#define RS_TILE (RS_BLOCK * RS_ITEMS)
#ifndef RS_WAVES_PER_EU
#define RS_WAVES_PER_EU 3  // onesweep occupancy floor per SIMD (bounds its VGPRs)
#endif
#ifndef RS_LB_SLEEP
#define RS_LB_SLEEP 1  // s_sleep between look-back polls (0: none)
#endif
#ifndef RS_LB_WIN
#define RS_LB_WIN 4  // predecessors whose look-back words are read per round trip
#endif
#define RS_RADIX 256
#define RS_MAX_SEGS 64
#define RS_GRID 2048  // blocks of the partition-walking sort kernels
#ifndef RS_CHUNK
#define RS_CHUNK 16  // partitions per chunk of the digit-count tables
#endif
#ifndef RS_TABLE_MASK
#define RS_TABLE_MASK 3u  // passes (bit ps) whose block offsets come from count tables, not the look-back
#endif
#ifndef RS_PC_COPIES
#define RS_PC_COPIES 4  // private histogram copies of the partition count (a power of two)
#endif
#ifndef RS_HIST_DIRECT0
#define RS_HIST_DIRECT0 1  // the histogram pass adds the first digit key by key, not by runs
#endif
#ifndef RS_HIST_DIRECT1
#define RS_HIST_DIRECT1 1  // ... and the second key by key into private copies
#endif
#ifndef RS_TABLE_MIN_PAIRS
#define RS_TABLE_MIN_PAIRS (16ull << 20)  // sorts sized for fewer pairs keep the look-back in every pass
#endif

// Workspace bytes of the segmented sort for max_pairs pairs in num_segs
// segments and keys of key_bytes bytes (so up to key_bytes passes): segment
// and chunk starts, digit histograms and offsets, counters, per pass one word
// per (partition, digit) (look-back status or partition counts), and one
// word per (chunk, digit) of chunk bases.
constexpr uint64_t rs_ws_bytes(uint64_t num_segs, uint64_t max_pairs, uint64_t key_bytes) {
  return 4 * (256 + 2 * num_segs * key_bytes * RS_RADIX + 64 +
              key_bytes * ((max_pairs + RS_TILE - 1) / RS_TILE + num_segs) * RS_RADIX +
              ((max_pairs + RS_TILE - 1) / RS_TILE / RS_CHUNK + 2 * num_segs + 1) * RS_RADIX);
}

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

// Inclusive scan of v across the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t t = __shfl_up(v, off, 64);
    if (lane >= (uint32_t)off) v += t;
  }
  return v;
}

// Exclusive scan across a block of NT threads (NT a multiple of 64).  s_wave
// must hold NT/64 words.  Contains two barriers.
template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_wave, uint32_t& total) {
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  const uint32_t incl = wave_incl_scan(v);
  if (lane == 63) s_wave[wave] = incl;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) {
    const uint32_t x = s_wave[w];
    if ((uint32_t)w < wave) pre += x;
    tot += x;
  }
  __syncthreads();
  total = tot;
  return pre + incl - v;
}

// Binary search: largest s in [0, n) with a[s] <= x (a ascending, a[0] <= x).
template <typename T>
__device__ __forceinline__ uint32_t seg_of(const T* a, uint32_t n, T x) {
  uint32_t lo = 0, hi = n;  // invariant a[lo] <= x, answer < hi
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] <= x) lo = mid; else hi = mid;
  }
  return lo;
}
