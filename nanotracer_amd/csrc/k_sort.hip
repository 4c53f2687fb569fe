// Exclusive u32 scan and the segmented onesweep radix sort.
#include <mutex>

#include "nts_sort.h"

// ---------------------------------------------------------------------------
// Exclusive scan (reduce, scan of block sums, downsweep), wrapping u32.
// ---------------------------------------------------------------------------
#define SC_ITEMS 16
#define SC_TILE (256 * SC_ITEMS)

__global__ void __launch_bounds__(256) k_scan_reduce(const uint32_t* __restrict__ in, uint64_t n,
                                                     uint32_t* __restrict__ block_sums,
                                                     unsigned long long* __restrict__ block_tot64) {
  __shared__ unsigned long long s_part[4];
  const uint64_t base = (uint64_t)blockIdx.x * SC_TILE;
  unsigned long long sum = 0;
#pragma unroll
  for (int k = 0; k < SC_ITEMS; ++k) {
    const uint64_t i = base + (uint64_t)k * 256 + threadIdx.x;
    if (i < n) sum += in[i];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
  if ((threadIdx.x & 63u) == 0) s_part[threadIdx.x >> 6] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long t = s_part[0] + s_part[1] + s_part[2] + s_part[3];
    block_sums[blockIdx.x] = (uint32_t)t;  // wrapping, as the reference's u32 scan
    block_tot64[blockIdx.x] = t;
  }
}

// One block of 1024 threads scans the block sums in chunks (exclusive) and
// writes the 64-bit grand total.
__global__ void __launch_bounds__(1024) k_scan_top(uint32_t* __restrict__ sums, uint32_t nblocks,
                                                   const unsigned long long* __restrict__ tot64,
                                                   uint64_t* __restrict__ d_total) {
  __shared__ uint32_t s_wave[16];
  __shared__ unsigned long long s_tot[16];
  uint32_t carry = 0;
  unsigned long long t64 = 0;  // the 64-bit grand total, summed across the block
  for (uint32_t c = 0; c < nblocks; c += 1024) {
    const uint32_t i = c + threadIdx.x;
    const uint32_t v = i < nblocks ? sums[i] : 0u;
    if (i < nblocks) t64 += tot64[i];
    uint32_t tot;
    const uint32_t ex = block_excl_scan<1024>(v, s_wave, tot);
    if (i < nblocks) sums[i] = carry + ex;
    carry += tot;
  }
  if (d_total) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) t64 += __shfl_xor(t64, off, 64);
    if ((threadIdx.x & 63u) == 0) s_tot[threadIdx.x >> 6] = t64;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long t = 0;
#pragma unroll
      for (int w = 0; w < 16; ++w) t += s_tot[w];
      *d_total = t;
    }
  }
}

__global__ void __launch_bounds__(256) k_scan_down(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                   uint64_t n, const uint32_t* __restrict__ block_sums) {
  __shared__ uint32_t s_wave[4];
  const uint64_t base = (uint64_t)blockIdx.x * SC_TILE + (uint64_t)threadIdx.x * SC_ITEMS;
  uint32_t v[SC_ITEMS];
  if (base + SC_ITEMS <= n) {
    const uint4* p = reinterpret_cast<const uint4*>(in + base);
#pragma unroll
    for (int k = 0; k < SC_ITEMS / 4; ++k) {
      const uint4 q = p[k];
      v[4 * k + 0] = q.x; v[4 * k + 1] = q.y; v[4 * k + 2] = q.z; v[4 * k + 3] = q.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < SC_ITEMS; ++k) v[k] = (base + k < n) ? in[base + k] : 0u;
  }
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < SC_ITEMS; ++k) sum += v[k];
  uint32_t total;
  uint32_t run = block_excl_scan<256>(sum, s_wave, total) + block_sums[blockIdx.x];
  uint32_t o[SC_ITEMS];
#pragma unroll
  for (int k = 0; k < SC_ITEMS; ++k) {
    o[k] = run;
    run += v[k];
  }
  if (base + SC_ITEMS <= n) {
    uint4* p = reinterpret_cast<uint4*>(out + base);
#pragma unroll
    for (int k = 0; k < SC_ITEMS / 4; ++k) p[k] = make_uint4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
  } else {
#pragma unroll
    for (int k = 0; k < SC_ITEMS; ++k)
      if (base + k < n) out[base + k] = o[k];
  }
}

uint64_t nts_scan_ws_bytes(uint64_t n) {
  const uint64_t nblocks = (n + SC_TILE - 1) / SC_TILE;
  return ((nblocks * 4 + 255) & ~255ull) + nblocks * 8 + 256;
}

int nts_scan_u32_ws(nts_ctx* ctx, const uint32_t* in, uint32_t* out, uint64_t n, uint64_t* d_total, void* ws) {
  if (n == 0) {
    if (d_total) return nts_fill_u32(ctx, d_total, 0, 2);
    return NTS_OK;
  }
  const uint32_t nblocks = nts_div_up(n, SC_TILE);
  const uint64_t sums_bytes = ((uint64_t)nblocks * 4 + 255) & ~255ull;
  uint32_t* sums = (uint32_t*)ws;
  unsigned long long* tot64 = (unsigned long long*)((char*)ws + sums_bytes);
  NTS_LAUNCH(ctx, "scan_reduce", k_scan_reduce, dim3(nblocks), dim3(256), 0, in, n, sums, tot64);
  NTS_LAUNCH(ctx, "scan_top", k_scan_top, dim3(1), dim3(1024), 0, sums, nblocks,
             (const unsigned long long*)tot64, d_total);
  NTS_LAUNCH(ctx, "scan_down", k_scan_down, dim3(nblocks), dim3(256), 0, in, out, n, (const uint32_t*)sums);
  return NTS_OK;
}

int nts_scan_u32_dev(nts_ctx* ctx, const uint32_t* in, uint32_t* out, uint64_t n, uint64_t* d_total) {
  void* scr;
  int rc = nts_scratch(ctx, nts_scan_ws_bytes(n), &scr);
  if (rc) return rc;
  return nts_scan_u32_ws(ctx, in, out, n, d_total, scr);
}

extern "C" int nts_scan_u32(nts_ctx* ctx, nts_buffer* data, uint64_t n) {
  NTS_CHECK_ARG(ctx && data && n * 4 <= data->bytes);
  return nts_scan_u32_dev(ctx, (const uint32_t*)data->ptr, (uint32_t*)data->ptr, n, nullptr);
}

// ---------------------------------------------------------------------------
// Segmented onesweep radix sort.
//
// Segment s holds pairs [seg_begin[s], seg_begin[s+1]); segments are sorted
// independently and stay in place.  Partitions of RS_TILE pairs never span a
// segment; part_begin[s] is the first global partition of segment s.
// ---------------------------------------------------------------------------
#define FLAG_AGG (1u << 30)
#define FLAG_INC (2u << 30)
#define FLAG_MASK (3u << 30)
#define VALUE_MASK ((1u << 30) - 1u)

__global__ void k_sort_parts(const uint64_t* __restrict__ seg_begin, uint32_t num_segs,
                             uint32_t* __restrict__ part_begin) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t p = 0;
  for (uint32_t s = 0; s < num_segs; ++s) {
    part_begin[s] = p;
    const uint64_t len = seg_begin[s + 1] - seg_begin[s];
    p += (uint32_t)((len + RS_TILE - 1) / RS_TILE);
  }
  part_begin[num_segs] = p;
}

// Digit histograms of every pass for every segment.  A fixed grid takes
// contiguous runs of partitions (the live count is only known on the device,
// so a launch sized for the capacity costs nothing for empty partitions) and
// keeps its counts in LDS until the segment changes: the global atomics are
// one set per (block, segment) rather than one per partition.
template <typename K>
__global__ void __launch_bounds__(RS_BLOCK) k_sort_hist(const K* __restrict__ keys, const uint64_t* __restrict__ seg_begin,
                                                        const uint32_t* __restrict__ part_begin, uint32_t num_segs,
                                                        uint32_t passes, uint32_t* __restrict__ hist) {
  __shared__ uint32_t s_hist[8 * RS_RADIX];
  const uint32_t total_parts = part_begin[num_segs];
  const uint32_t per = (total_parts + gridDim.x - 1) / gridDim.x;
  const uint32_t p0 = blockIdx.x * per;
  const uint32_t p1 = p0 + per < total_parts ? p0 + per : total_parts;
  if (p0 >= p1) return;
  const uint32_t nbins = passes * RS_RADIX;
  const auto flush = [&](uint32_t seg) {
    __syncthreads();
    uint32_t* h = hist + (size_t)seg * nbins;
    for (uint32_t i = threadIdx.x; i < nbins; i += RS_BLOCK) {
      const uint32_t c = s_hist[i];
      if (c) atomicAdd(&h[i], c);
      s_hist[i] = 0;
    }
    __syncthreads();
  };
  for (uint32_t i = threadIdx.x; i < nbins; i += RS_BLOCK) s_hist[i] = 0;
  __syncthreads();
  uint32_t s = seg_of<uint32_t>(part_begin, num_segs, p0);
  for (uint32_t p = p0; p < p1; ++p) {
    while (p >= part_begin[s + 1]) {  // next non-empty segment
      flush(s);
      ++s;
    }
    const uint64_t base = seg_begin[s] + (uint64_t)(p - part_begin[s]) * RS_TILE;
    const uint64_t end = seg_begin[s + 1];
    // Neighbouring pairs mostly come from one record (same depth, neighbouring
    // tiles), so a wave's digits come in runs: each run's first lane adds the
    // run's length, one LDS atomic per run instead of one per key.  The valid
    // lanes of a wave are a prefix (consecutive indices).
    const uint32_t lane = threadIdx.x & 63u;
#pragma unroll 4
    for (int k = 0; k < RS_ITEMS; ++k) {
      const uint64_t i = base + (uint64_t)k * RS_BLOCK + threadIdx.x;
      const bool valid = i < end;
      const K key = valid ? keys[i] : (K)0;
      const uint64_t vmask = __ballot(valid);
      if (vmask == 0) continue;
      const uint32_t n_valid = (uint32_t)__popcll(vmask);
      for (uint32_t ps = 0; ps < passes; ++ps) {
        const uint32_t d = (uint32_t)((key >> (8 * ps)) & 255u);
        const uint32_t prev = (uint32_t)__shfl_up((int)d, 1, 64);
        const bool head = valid && (lane == 0 || prev != d);
        const uint64_t heads = __ballot(head);
        if (head) {
          const uint64_t above = lane < 63u ? heads & (~0ull << (lane + 1u)) : 0ull;
          const uint32_t next = above ? (uint32_t)__builtin_ctzll(above) : n_valid;
          atomicAdd(&s_hist[ps * RS_RADIX + d], next - lane);
        }
      }
    }
  }
  flush(s);
}

// Clears the look-back status words of the live partitions of every pass
// (pass ps owns words [ps * max_parts * 256, ...)); the live count is read
// on the device, so a launch sized for the capacity clears only what is used.
__global__ void __launch_bounds__(256) k_clear_status(uint32_t* __restrict__ status,
                                                      const uint32_t* __restrict__ part_begin, uint32_t num_segs,
                                                      uint32_t passes, uint32_t max_parts) {
  const uint64_t words = (uint64_t)part_begin[num_segs] * RS_RADIX;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint32_t ps = 0; ps < passes; ++ps) {
    uint4* st = reinterpret_cast<uint4*>(status + (uint64_t)ps * max_parts * RS_RADIX);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words / 4; i += stride)
      st[i] = make_uint4(0u, 0u, 0u, 0u);
  }
}

// offsets[s][pass][d] = seg_begin[s] + sum of hist[s][pass][d' < d].
__global__ void __launch_bounds__(RS_RADIX) k_sort_offsets(const uint32_t* __restrict__ hist,
                                                           const uint64_t* __restrict__ seg_begin, uint32_t passes,
                                                           uint32_t* __restrict__ offsets) {
  __shared__ uint32_t s_wave[4];
  const uint32_t s = blockIdx.x / passes;
  const uint32_t idx = blockIdx.x * RS_RADIX + threadIdx.x;
  uint32_t total;
  const uint32_t ex = block_excl_scan<RS_RADIX>(hist[idx], s_wave, total);
  offsets[idx] = (uint32_t)seg_begin[s] + ex;
}

template <typename K>
__global__ void __launch_bounds__(RS_BLOCK) k_onesweep(const K* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
                                                       K* __restrict__ keys_out, uint32_t* __restrict__ vals_out,
                                                       const uint64_t* __restrict__ seg_begin,
                                                       const uint32_t* __restrict__ part_begin, uint32_t num_segs,
                                                       uint32_t shift, uint32_t passes, uint32_t pass,
                                                       const uint32_t* __restrict__ offsets, uint32_t* status,
                                                       uint32_t* part_counter, uint32_t* err) {
  __shared__ K s_keys[RS_TILE];
  __shared__ uint32_t s_vals[RS_TILE];
  __shared__ uint32_t s_wave_hist[4][RS_RADIX];
  __shared__ uint32_t s_excl[RS_RADIX];
  __shared__ uint32_t s_gbase[RS_RADIX];
  __shared__ uint32_t s_scan[4];
  __shared__ uint32_t s_part;

  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint32_t total_parts = part_begin[num_segs];
  // A fixed grid takes partitions in the order of an atomic counter, so the
  // partition a block waits on in the look-back was taken earlier by a
  // resident block, which publishes its aggregate before waiting itself.
  for (;;) {
    if (tid == 0) s_part = atomicAdd(part_counter, 1u);
#pragma unroll
    for (int w = 0; w < 4; ++w) s_wave_hist[w][tid] = 0;
    __syncthreads();
    const uint32_t p = s_part;
    if (p >= total_parts) return;
    const uint32_t s = seg_of<uint32_t>(part_begin, num_segs, p);
    const uint32_t first_part = part_begin[s];
    const uint64_t base = seg_begin[s] + (uint64_t)(p - first_part) * RS_TILE;
    const uint64_t end = seg_begin[s + 1];

    // ---- load and rank (order within the block = (wave, round, lane)) ----
    K key[RS_ITEMS];
    uint32_t val[RS_ITEMS];
    uint32_t rank[RS_ITEMS];
    const uint64_t lt_mask = (1ull << lane) - 1ull;
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
      const uint64_t i = base + (uint64_t)wave * (RS_TILE / 4) + (uint64_t)r * 64 + lane;
      const bool valid = i < end;
      key[r] = valid ? keys_in[i] : (K)~(K)0;
      val[r] = valid ? vals_in[i] : 0u;
    }
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
      const uint32_t d = (uint32_t)(key[r] >> shift) & 255u;
      // lanes with this lane's digit: the AND over the digit's bits of the
      // bit's ballot, complemented where this lane's bit is clear, kept as
      // two 32-bit halves (neg is the bit sign-extended: 0 or all ones)
      uint32_t mlo = ~0u, mhi = ~0u;
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const uint32_t neg = (uint32_t)((int32_t)(d << (31 - b)) >> 31);
        const uint64_t bal = __ballot(neg != 0u);
        mlo &= ~((uint32_t)bal ^ neg);
        mhi &= ~((uint32_t)(bal >> 32) ^ neg);
      }
      const uint64_t match = ((uint64_t)mhi << 32) | mlo;
      const uint32_t leader = (uint32_t)__ffsll((unsigned long long)match) - 1u;
      uint32_t old = 0;
      if (lane == leader) {
        old = s_wave_hist[wave][d];
        s_wave_hist[wave][d] = old + (uint32_t)__popcll(match);
      }
      old = __shfl(old, (int)leader, 64);
      rank[r] = old + (uint32_t)__popcll(match & lt_mask);
    }
    __syncthreads();

    // ---- per-digit block totals, wave prefixes, block-local digit offsets ----
    const uint32_t d = tid;
    const uint32_t c0 = s_wave_hist[0][d], c1 = s_wave_hist[1][d], c2 = s_wave_hist[2][d], c3 = s_wave_hist[3][d];
    const uint32_t tot = c0 + c1 + c2 + c3;
    s_wave_hist[0][d] = 0;
    s_wave_hist[1][d] = c0;
    s_wave_hist[2][d] = c0 + c1;
    s_wave_hist[3][d] = c0 + c1 + c2;
    uint32_t block_total;
    const uint32_t ex = block_excl_scan<RS_BLOCK>(tot, s_scan, block_total);
    s_excl[d] = ex;

    // ---- decoupled look-back over the preceding partitions of this segment ----
    uint32_t* st = status;  // this pass's status words (offset by the host)
    uint32_t prefix = 0;
    if (p == first_part) {
      __hip_atomic_store(&st[(size_t)p * RS_RADIX + d], FLAG_INC | (tot & VALUE_MASK), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    } else {
      __hip_atomic_store(&st[(size_t)p * RS_RADIX + d], FLAG_AGG | (tot & VALUE_MASK), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      // The words of up to LB_WIN predecessors are read together (one round
      // trip to the device-coherent level per window, not per partition)
      // and consumed in order until an inclusive one or the segment start.
      constexpr uint32_t LB_WIN = 4;
      uint32_t j = p - 1;  // next predecessor to consume, >= first_part
      uint32_t spins = 0;
      while (true) {
        uint32_t w[LB_WIN];
#pragma unroll
        for (uint32_t q = 0; q < LB_WIN; ++q)
          w[q] = (q <= j - first_part)
                     ? __hip_atomic_load(&st[(size_t)(j - q) * RS_RADIX + d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                     : 0u;
        bool done = false, stalled = false;
#pragma unroll
        for (uint32_t q = 0; q < LB_WIN; ++q) {
          if (done || stalled) continue;
          const uint32_t flag = w[q] & FLAG_MASK;
          if (flag == 0u) {  // not published yet: wait for this one
            stalled = true;
            continue;
          }
          prefix += w[q] & VALUE_MASK;
          if (flag == FLAG_INC || j == first_part) {
            done = true;
            continue;
          }
          --j;
        }
        if (done) break;
        if (stalled) {
          if (++spins > (1u << 22)) {
            atomicOr(err, 1u);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      __hip_atomic_store(&st[(size_t)p * RS_RADIX + d], FLAG_INC | ((prefix + tot) & VALUE_MASK), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
    s_gbase[d] = offsets[((size_t)s * passes + pass) * RS_RADIX + d] + prefix - ex;
    __syncthreads();

    // ---- reorder in LDS by digit, then write digit runs contiguously ----
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
      const uint32_t dd = (uint32_t)(key[r] >> shift) & 255u;
      const uint32_t lp = s_excl[dd] + s_wave_hist[wave][dd] + rank[r];
      s_keys[lp] = key[r];
      s_vals[lp] = val[r];
    }
    __syncthreads();
    const uint64_t valid_n = end - base;
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
      const uint32_t j = (uint32_t)r * RS_BLOCK + tid;
      if (j < valid_n) {
        const K k = s_keys[j];
        const uint32_t dd = (uint32_t)(k >> shift) & 255u;
        const uint32_t pos = s_gbase[dd] + j;
        if ((uint64_t)pos < end) {  // always true unless a look-back timed out
          keys_out[pos] = k;
          vals_out[pos] = s_vals[j];
        }
      }
    }
    __syncthreads();  // LDS and s_part are reused by the next partition
  }
}

// Workspace of the segmented sort, sized for max_pairs pairs in num_segs
// segments and keys of key_bytes bytes (so for up to key_bytes passes).
uint64_t nts_sort_ws_bytes(uint32_t num_segs, uint64_t max_pairs, uint32_t key_bytes) {
  const uint64_t max_parts = (max_pairs + RS_TILE - 1) / RS_TILE + num_segs;
  const uint64_t n_hist = (uint64_t)num_segs * key_bytes * RS_RADIX;
  return 4 * (256 + 2 * n_hist + 64 + (uint64_t)key_bytes * max_parts * RS_RADIX);
}

// Launch the sort of pairs of segments (result in keys/vals, or in the alt
// buffers when *result_in_alt is set) with no host synchronisation.
// seg_begin (device, num_segs + 1 entries) bounds the segments; max_pairs
// bounds seg_begin[num_segs].  Only the low ceil(key_bits / 8) * 8 bits are
// sorted.  A look-back that never completes sets *err (checked by callers).
template <typename K>
static int sort_segmented_ws(nts_ctx* ctx, K* keys, uint32_t* vals, K* keys_alt, uint32_t* vals_alt,
                             const uint64_t* d_seg_begin, uint32_t num_segs, uint64_t max_pairs, uint32_t key_bits,
                             void* ws, bool* result_in_alt, uint32_t** err_out) {
  *result_in_alt = false;
  if (num_segs == 0 || num_segs > RS_MAX_SEGS) return nts_fail(NTS_E_INVALID, "segments %u", num_segs);
  if (max_pairs >= (1ull << 30)) return nts_fail(NTS_E_CAPACITY, "sort supports < 2^30 pairs (%llu)",
                                                 (unsigned long long)max_pairs);
  uint32_t passes = (key_bits + 7) / 8;
  if (passes > sizeof(K)) passes = sizeof(K);
  const uint32_t max_parts = nts_div_up(max_pairs, RS_TILE) + num_segs;
  const uint64_t n_hist = (uint64_t)num_segs * sizeof(K) * RS_RADIX;
  uint32_t* part_begin = (uint32_t*)ws;
  uint32_t* hist = part_begin + 256;
  uint32_t* offsets = hist + n_hist;
  uint32_t* counters = offsets + n_hist;
  uint32_t* err = counters + 32;
  uint32_t* status = counters + 64;
  if (err_out) *err_out = err;
  int rc = nts_fill_u32(ctx, hist, 0, n_hist);
  if (rc) return rc;
  rc = nts_fill_u32(ctx, counters, 0, 64);
  if (rc) return rc;
  if (passes == 0 || max_pairs == 0) return NTS_OK;
  NTS_LAUNCH(ctx, "sort_parts", k_sort_parts, dim3(1), dim3(64), 0, d_seg_begin, num_segs, part_begin);
  NTS_LAUNCH(ctx, "sort_clear", k_clear_status, dim3(1024), dim3(256), 0, status, (const uint32_t*)part_begin,
             num_segs, passes, max_parts);
  const uint32_t grid = max_parts < RS_GRID ? max_parts : RS_GRID;
  NTS_LAUNCH(ctx, "sort_hist", k_sort_hist<K>, dim3(grid), dim3(RS_BLOCK), 0, (const K*)keys, d_seg_begin,
             (const uint32_t*)part_begin, num_segs, passes, hist);
  NTS_LAUNCH(ctx, "sort_offsets", k_sort_offsets, dim3(num_segs * passes), dim3(RS_RADIX), 0, (const uint32_t*)hist,
             d_seg_begin, passes, offsets);
  K* src_k = keys;
  uint32_t* src_v = vals;
  K* dst_k = keys_alt;
  uint32_t* dst_v = vals_alt;
  for (uint32_t ps = 0; ps < passes; ++ps) {
    NTS_LAUNCH(ctx, "sort_onesweep", k_onesweep<K>, dim3(grid), dim3(RS_BLOCK), 0, (const K*)src_k,
               (const uint32_t*)src_v, dst_k, dst_v, d_seg_begin, (const uint32_t*)part_begin, num_segs, 8 * ps,
               passes, ps, (const uint32_t*)offsets, status + (size_t)ps * max_parts * RS_RADIX, counters + ps, err);
    K* tk = src_k; src_k = dst_k; dst_k = tk;
    uint32_t* tv = src_v; src_v = dst_v; dst_v = tv;
  }
  *result_in_alt = (passes & 1u) != 0;
  return NTS_OK;
}

template <typename K>
static int sort_segmented(nts_ctx* ctx, K* keys, uint32_t* vals, K* keys_alt, uint32_t* vals_alt,
                          const uint64_t* d_seg_begin, uint32_t num_segs, uint64_t max_pairs, uint32_t key_bits,
                          bool* result_in_alt) {
  *result_in_alt = false;
  if (max_pairs == 0) return NTS_OK;
  void* scr;
  int rc = nts_scratch(ctx, nts_sort_ws_bytes(num_segs, max_pairs, sizeof(K)), &scr);
  if (rc) return rc;
  uint32_t* err = nullptr;
  rc = sort_segmented_ws<K>(ctx, keys, vals, keys_alt, vals_alt, d_seg_begin, num_segs, max_pairs, key_bits, scr,
                            result_in_alt, &err);
  if (rc) return rc;
  uint32_t* hw = ctx->host_word;
  NTS_HIP_TRY(hipMemcpyAsync(hw, err, 4, hipMemcpyDeviceToHost, ctx->stream));
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (*hw) return nts_fail(NTS_E_HIP, "radix sort look-back timed out");
  return NTS_OK;
}

int nts_sort_pairs_dev(nts_ctx* ctx, uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt,
                       const uint64_t* d_seg_begin, uint32_t num_segs, uint64_t max_pairs, uint32_t key_bits,
                       bool* result_in_alt) {
  return sort_segmented<uint64_t>(ctx, keys, vals, keys_alt, vals_alt, d_seg_begin, num_segs, max_pairs, key_bits,
                                  result_in_alt);
}

int nts_sort_pairs32_dev(nts_ctx* ctx, uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                         const uint64_t* d_seg_begin, uint32_t num_segs, uint64_t max_pairs, uint32_t key_bits,
                         bool* result_in_alt) {
  return sort_segmented<uint32_t>(ctx, keys, vals, keys_alt, vals_alt, d_seg_begin, num_segs, max_pairs, key_bits,
                                  result_in_alt);
}

int nts_sort_ws_launch32(nts_ctx* ctx, uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                         const uint64_t* d_seg_begin, uint32_t num_segs, uint64_t max_pairs, uint32_t key_bits,
                         void* ws, bool* result_in_alt, uint32_t** err) {
  return sort_segmented_ws<uint32_t>(ctx, keys, vals, keys_alt, vals_alt, d_seg_begin, num_segs, max_pairs, key_bits,
                                     ws, result_in_alt, err);
}

int nts_sort_ws_launch64(nts_ctx* ctx, uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt,
                         const uint64_t* d_seg_begin, uint32_t num_segs, uint64_t max_pairs, uint32_t key_bits,
                         void* ws, bool* result_in_alt, uint32_t** err) {
  return sort_segmented_ws<uint64_t>(ctx, keys, vals, keys_alt, vals_alt, d_seg_begin, num_segs, max_pairs, key_bits,
                                     ws, result_in_alt, err);
}

template <typename K>
static int sort_api(nts_ctx* ctx, nts_buffer* keys, nts_buffer* vals, uint64_t n, uint32_t key_bits) {
  NTS_CHECK_ARG(ctx && keys && vals);
  NTS_CHECK_ARG(n * sizeof(K) <= keys->bytes && n * 4 <= vals->bytes);
  NTS_CHECK_ARG(key_bits <= 8 * sizeof(K));
  if (n <= 1) return NTS_OK;
  // ping-pong buffers + the one segment bound live in their own allocation
  // (the sort's own scratch uses the context scratch area)
  void* tmp = nullptr;
  const uint64_t bytes = n * (sizeof(K) + 4) + 256;
  int rc = nts_dev_alloc(ctx, &tmp, bytes);
  if (rc) return rc;
  uint64_t* segb = (uint64_t*)tmp;
  K* k_alt = (K*)((char*)tmp + 64);
  uint32_t* v_alt = (uint32_t*)((char*)tmp + 64 + ((n * sizeof(K) + 63) & ~63ull));
  uint64_t hb[2] = {0, n};
  (void)hipMemcpyAsync(segb, hb, 16, hipMemcpyHostToDevice, ctx->stream);
  bool alt = false;
  rc = sort_segmented<K>(ctx, (K*)keys->ptr, (uint32_t*)vals->ptr, k_alt, v_alt, segb, 1, n, key_bits, &alt);
  if (rc == NTS_OK && alt) {
    (void)hipMemcpyAsync(keys->ptr, k_alt, n * sizeof(K), hipMemcpyDeviceToDevice, ctx->stream);
    (void)hipMemcpyAsync(vals->ptr, v_alt, n * 4, hipMemcpyDeviceToDevice, ctx->stream);
  }
  (void)hipStreamSynchronize(ctx->stream);
  nts_dev_free(ctx, tmp, bytes);
  return rc;
}

extern "C" int nts_sort_pairs_u32(nts_ctx* ctx, nts_buffer* keys, nts_buffer* vals, uint64_t n, uint32_t key_bits) {
  return sort_api<uint32_t>(ctx, keys, vals, n, key_bits);
}

extern "C" int nts_sort_pairs_u64(nts_ctx* ctx, nts_buffer* keys, nts_buffer* vals, uint64_t n, uint32_t key_bits) {
  return sort_api<uint64_t>(ctx, keys, vals, n, key_bits);
}

// ---------------------------------------------------------------------------
// nts_sort_pairs_u32_noalloc: nts_sort_pairs_u32 with every buffer beyond
// the caller's keys and values in one file-scope device array, for callers
// that cannot allocate device memory.  The array holds the segment bounds,
// the ping-pong keys and values and the sort workspace (histograms, digit
// offsets, look-back status words, partition counters, the error word) for
// up to NTS_SORT_NOALLOC_MAX_PAIRS pairs: 18.9 MB, held on each GPU the
// library runs on until the process ends.  Every call on a GPU uses that
// GPU's one copy, so calls that overlap on one GPU race; whatever must start
// at zero is cleared by each call (fill kernels), nothing is allocated.
// ---------------------------------------------------------------------------
namespace {
constexpr uint64_t na_al(uint64_t x) { return (x + 255) & ~255ull; }
constexpr uint64_t kNaMax = NTS_SORT_NOALLOC_MAX_PAIRS;
constexpr uint64_t kNaParts = (kNaMax + RS_TILE - 1) / RS_TILE + 1;
constexpr uint64_t kNaWs = 4 * (256 + 2 * (4 * RS_RADIX) + 64 + 4 * kNaParts * RS_RADIX);  // nts_sort_ws_bytes(1, max, 4)
constexpr uint64_t kNaOffKeys = 256;
constexpr uint64_t kNaOffVals = kNaOffKeys + na_al(4 * kNaMax);
constexpr uint64_t kNaOffWs = kNaOffVals + na_al(4 * kNaMax);
constexpr uint64_t kNaBytes = kNaOffWs + na_al(kNaWs);
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNaBytes];

namespace {
std::mutex g_na_mu;
void* g_na_ptr[64] = {};  // the device address of `scratch` on each device

// The address of this device's copy of `scratch`, looked up once per device
// with that device current.
int noalloc_scratch(nts_ctx* ctx, char** out) {
  if (ctx->device < 0 || ctx->device >= 64) return nts_fail(NTS_E_INVALID, "device %d", ctx->device);
  std::lock_guard<std::mutex> g(g_na_mu);
  if (!g_na_ptr[ctx->device]) {
    int prev = 0;
    NTS_HIP_TRY(hipGetDevice(&prev));
    NTS_HIP_TRY(hipSetDevice(ctx->device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    (void)hipSetDevice(prev);
    if (e != hipSuccess || !p)
      return nts_fail(NTS_E_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_na_ptr[ctx->device] = p;
  }
  *out = (char*)g_na_ptr[ctx->device];
  return NTS_OK;
}
}  // namespace

__global__ void k_set_one_segment(uint64_t* __restrict__ seg, uint64_t n) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    seg[0] = 0;
    seg[1] = n;
  }
}

__global__ void __launch_bounds__(256) k_copy_pairs_u32(const uint32_t* __restrict__ ks, const uint32_t* __restrict__ vs,
                                                        uint32_t* __restrict__ kd, uint32_t* __restrict__ vd,
                                                        uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    kd[i] = ks[i];
    vd[i] = vs[i];
  }
}

extern "C" int nts_sort_pairs_u32_noalloc(nts_ctx* ctx, nts_buffer* keys, nts_buffer* vals, uint64_t n,
                                          uint32_t key_bits) {
  NTS_CHECK_ARG(ctx && keys && vals);
  NTS_CHECK_ARG(n * 4 <= keys->bytes && n * 4 <= vals->bytes);
  NTS_CHECK_ARG(key_bits <= 32);
  if (n > kNaMax)
    return nts_fail(NTS_E_CAPACITY, "nts_sort_pairs_u32_noalloc sorts at most %llu pairs (got %llu)",
                    (unsigned long long)kNaMax, (unsigned long long)n);
  if (n <= 1) return NTS_OK;
  char* base = nullptr;
  int rc = noalloc_scratch(ctx, &base);
  if (rc) return rc;
  uint64_t* seg = (uint64_t*)base;
  uint32_t* k_alt = (uint32_t*)(base + kNaOffKeys);
  uint32_t* v_alt = (uint32_t*)(base + kNaOffVals);
  void* ws = base + kNaOffWs;
  NTS_LAUNCH(ctx, "sort_noalloc_segment", k_set_one_segment, dim3(1), dim3(64), 0, seg, n);
  bool alt = false;
  uint32_t* err = nullptr;
  rc = sort_segmented_ws<uint32_t>(ctx, (uint32_t*)keys->ptr, (uint32_t*)vals->ptr, k_alt, v_alt, seg, 1, n,
                                   key_bits, ws, &alt, &err);
  if (rc) return rc;
  if (alt) {
    uint32_t grid = nts_div_up(n, 256);
    if (grid > 4096) grid = 4096;
    NTS_LAUNCH(ctx, "sort_noalloc_copy", k_copy_pairs_u32, dim3(grid), dim3(256), 0, (const uint32_t*)k_alt,
               (const uint32_t*)v_alt, (uint32_t*)keys->ptr, (uint32_t*)vals->ptr, n);
  }
  // the look-back error word, read like nts_sort_pairs_u32 reads it (the
  // context's pinned word: no allocation)
  uint32_t* hw = ctx->host_word;
  NTS_HIP_TRY(hipMemcpyAsync(hw, err, 4, hipMemcpyDeviceToHost, ctx->stream));
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (*hw) return nts_fail(NTS_E_HIP, "radix sort look-back timed out");
  return NTS_OK;
}
