// Exclusive u32 scan and the segmented onesweep radix sort.
#include <mutex>

#include "nts_bin.h"
#include "nts_raster.h"
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

// part_begin[s]: the first partition of segment s; chunk_begin[s]: its first
// chunk (RS_CHUNK consecutive partitions of one segment).
__global__ void k_sort_parts(const uint64_t* __restrict__ seg_begin, uint32_t num_segs,
                             uint32_t* __restrict__ part_begin, uint32_t* __restrict__ chunk_begin) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t p = 0, c = 0;
  for (uint32_t s = 0; s < num_segs; ++s) {
    part_begin[s] = p;
    chunk_begin[s] = c;
    const uint64_t len = seg_begin[s + 1] - seg_begin[s];
    const uint32_t np = (uint32_t)((len + RS_TILE - 1) / RS_TILE);
    p += np;
    c += (np + RS_CHUNK - 1) / RS_CHUNK;
  }
  part_begin[num_segs] = p;
  chunk_begin[num_segs] = c;
}

// Chunks the count tables hold for max_pairs pairs in num_segs segments (the
// stride of the chunk-base table, which is digit-major).
static inline uint32_t rs_max_chunks(uint32_t num_segs, uint64_t max_pairs) {
  return (uint32_t)((max_pairs + RS_TILE - 1) / RS_TILE / RS_CHUNK + 2ull * num_segs + 1);
}

// Digit histograms of every pass for every segment (pass 0 sorts bits
// [lo_bit, lo_bit + first_bits), pass ps > 0 the 8 bits from
// lo_bit + first_bits + 8 (ps - 1)), and, where cnt is given, pass 0's
// count table: cnt[p][d] = the keys of digit d in partition p.  A fixed grid
// takes contiguous runs of partitions (the live count is only known on the
// device, so a launch sized for the capacity costs nothing for empty
// partitions) and keeps the histograms in LDS until the segment changes: the
// global atomics are one set per (block, segment) rather than one per
// partition.  The partition counts go through two LDS rows used in turn, so
// one barrier per partition orders both the row's readers and its reuse.
template <typename K>
__device__ __forceinline__ void hist_load(const K* __restrict__ keys, uint64_t base, uint64_t end, uint32_t tid,
                                          K (&key)[RS_ITEMS]) {
#pragma unroll
  for (int k = 0; k < RS_ITEMS; ++k) {
    const uint64_t i = base + (uint64_t)k * RS_BLOCK + tid;
    key[k] = i < end ? keys[i] : (K)0;
  }
}

template <typename K>
__global__ void __launch_bounds__(RS_BLOCK) k_sort_hist(const K* __restrict__ keys, const uint64_t* __restrict__ seg_begin,
                                                        const uint32_t* __restrict__ part_begin, uint32_t num_segs,
                                                        uint32_t lo_bit, uint32_t first_bits, uint32_t passes,
                                                        uint32_t* __restrict__ hist, uint32_t* __restrict__ cnt) {
  __shared__ uint32_t s_hist[8 * RS_RADIX];
  __shared__ uint32_t s_part[2][RS_RADIX];
  // the second digit, key by key into private copies (lane l into copy
  // l % 4, rows of 257 words: one digit's copies in different banks)
  constexpr uint32_t C1 = RS_HIST_DIRECT1 ? 4 : 1, ROW1 = RS_RADIX + 1;
  __shared__ uint32_t s_h1[C1 * ROW1];
  const uint32_t total_parts = part_begin[num_segs];
  const uint32_t per = (total_parts + gridDim.x - 1) / gridDim.x;
  const uint32_t p0 = blockIdx.x * per;
  const uint32_t p1 = p0 + per < total_parts ? p0 + per : total_parts;
  if (p0 >= p1) return;
  const uint32_t tid = threadIdx.x, lane = tid & 63u;
  const uint32_t nbins = passes * RS_RADIX;
  const uint32_t mask0 = (1u << first_bits) - 1u;
  const bool direct1 = RS_HIST_DIRECT1 && passes > 1;
  uint32_t* mine1 = s_h1 + (tid & (C1 - 1u)) * ROW1;
  const auto flush = [&](uint32_t seg) {
    __syncthreads();
    uint32_t* h = hist + (size_t)seg * nbins;
    for (uint32_t i = tid; i < nbins; i += RS_BLOCK) {
      uint32_t c = s_hist[i];
      if (direct1 && i >= RS_RADIX && i < 2 * RS_RADIX) {
#pragma unroll
        for (uint32_t q = 0; q < C1; ++q) {
          c += s_h1[q * ROW1 + i - RS_RADIX];
          s_h1[q * ROW1 + i - RS_RADIX] = 0;
        }
      }
      if (c) atomicAdd(&h[i], c);
      s_hist[i] = 0;
    }
    __syncthreads();
  };
  for (uint32_t i = tid; i < nbins; i += RS_BLOCK) s_hist[i] = 0;
  for (uint32_t i = tid; i < C1 * ROW1; i += RS_BLOCK) s_h1[i] = 0;
  s_part[0][tid] = 0;
  s_part[1][tid] = 0;
  __syncthreads();
  uint32_t s = seg_of<uint32_t>(part_begin, num_segs, p0);
  for (uint32_t p = p0; p < p1; ++p) {
    while (p >= part_begin[s + 1]) {  // next non-empty segment
      flush(s);
      ++s;
    }
    const uint64_t base = seg_begin[s] + (uint64_t)(p - part_begin[s]) * RS_TILE;
    const uint64_t end = seg_begin[s + 1];
    uint32_t* row = s_part[p & 1u];
    K key[RS_ITEMS];
    hist_load<K>(keys, base, end, tid, key);
    // The first digit (neighbouring tiles differ in it) and the second add
    // key by key, the second into private copies; the digits of later
    // passes come in runs in emit order (a record's pairs share them), so
    // each run's first lane adds the run's length, one LDS atomic per run.
    // The valid lanes of a wave are a prefix (consecutive indices).
#pragma unroll
    for (int k = 0; k < RS_ITEMS; ++k) {
      const uint64_t i = base + (uint64_t)k * RS_BLOCK + tid;
      const bool valid = i < end;
      const uint64_t vmask = __ballot(valid);
      if (vmask == 0) continue;
      const uint32_t n_valid = (uint32_t)__popcll(vmask);
      // runs of equal digits: the run's first lane adds its length
      const auto add_runs = [&](uint32_t d, uint32_t* h, uint32_t* h2) {
        const uint32_t prev = (uint32_t)__shfl_up((int)d, 1, 64);
        const bool head = valid && (lane == 0 || prev != d);
        const uint64_t heads = __ballot(head);
        if (head) {
          const uint64_t above = lane < 63u ? heads & (~0ull << (lane + 1u)) : 0ull;
          const uint32_t next = above ? (uint32_t)__builtin_ctzll(above) : n_valid;
          atomicAdd(&h[d], next - lane);
          if (h2) atomicAdd(&h2[d], next - lane);
        }
      };
      if (RS_HIST_DIRECT0) {  // the first digit's runs are short (neighbouring tiles): one add per key
        if (valid) {
          const uint32_t d0 = (uint32_t)(key[k] >> lo_bit) & mask0;
          atomicAdd(&s_hist[d0], 1u);
          if (cnt) atomicAdd(&row[d0], 1u);
        }
      } else {
        add_runs((uint32_t)(key[k] >> lo_bit) & mask0, s_hist, cnt ? row : nullptr);
      }
      for (uint32_t ps = 1; ps < passes; ++ps) {
        const uint32_t d = (uint32_t)(key[k] >> (lo_bit + first_bits + 8 * (ps - 1))) & 255u;
        if (ps == 1 && direct1) {
          if (valid) atomicAdd(&mine1[d], 1u);
        } else {
          add_runs(d, s_hist + ps * RS_RADIX, nullptr);
        }
      }
    }
    if (cnt) {
      __syncthreads();
      cnt[(size_t)p * RS_RADIX + tid] = row[tid];
      row[tid] = 0;
    }
  }
  flush(s);
}

// The count table of the digit (key >> shift) & dmask over keys in
// partition order: cnt[p][d] = the keys of digit d in partition p (the
// digits of a sorted pass's output repeat without forming runs, so each key
// adds one into a private copy of the histogram).
template <typename K>
__global__ void __launch_bounds__(RS_BLOCK) k_part_count(const K* __restrict__ keys, const uint64_t* __restrict__ seg_begin,
                                                         const uint32_t* __restrict__ part_begin, uint32_t num_segs,
                                                         uint32_t shift, uint32_t dmask, uint32_t* __restrict__ cnt) {
  // RS_PC_COPIES private copies of the histogram, lane l adding into copy
  // l % RS_PC_COPIES; rows padded to 257 words so that the copies of one
  // digit sit in different banks
  constexpr uint32_t C = RS_PC_COPIES, ROW = RS_RADIX + 1;
  __shared__ uint32_t s_h[C * ROW];
  const uint32_t total_parts = part_begin[num_segs];
  const uint32_t tid = threadIdx.x;
  uint32_t* mine = s_h + (tid & (C - 1u)) * ROW;
  for (uint32_t p = blockIdx.x; p < total_parts; p += gridDim.x) {
    for (uint32_t w = tid; w < C * ROW; w += RS_BLOCK) s_h[w] = 0;
    const uint32_t s = seg_of<uint32_t>(part_begin, num_segs, p);
    const uint64_t base = seg_begin[s] + (uint64_t)(p - part_begin[s]) * RS_TILE;
    const uint64_t end = seg_begin[s + 1];
    K key[RS_ITEMS];
    hist_load<K>(keys, base, end, tid, key);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < RS_ITEMS; ++k) {
      const uint64_t i = base + (uint64_t)k * RS_BLOCK + tid;
      if (i < end) atomicAdd(&mine[(uint32_t)(key[k] >> shift) & dmask], 1u);
    }
    __syncthreads();
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t c = 0; c < C; ++c) sum += s_h[c * ROW + tid];
    cnt[(size_t)p * RS_RADIX + tid] = sum;
    __syncthreads();  // the copies are cleared for the next partition
  }
}

// Chunk prefixes of a count table: within each chunk (RS_CHUNK consecutive
// partitions of one segment) cnt[p][d] becomes the keys of digit d in the
// chunk's partitions before p, and ctot[d][c] receives the chunk's total
// (digit-major, stride max_chunks).  One block per chunk, one thread per digit.
__global__ void __launch_bounds__(RS_RADIX) k_chunk_prefix(uint32_t* __restrict__ cnt,
                                                           const uint32_t* __restrict__ part_begin,
                                                           const uint32_t* __restrict__ chunk_begin, uint32_t num_segs,
                                                           uint32_t* __restrict__ ctot, uint32_t max_chunks) {
  const uint32_t total_chunks = chunk_begin[num_segs];
  const uint32_t d = threadIdx.x;
  for (uint32_t c = blockIdx.x; c < total_chunks; c += gridDim.x) {
    const uint32_t s = seg_of<uint32_t>(chunk_begin, num_segs, c);
    const uint32_t pa = part_begin[s] + (c - chunk_begin[s]) * RS_CHUNK;
    const uint32_t pb = pa + RS_CHUNK < part_begin[s + 1] ? pa + RS_CHUNK : part_begin[s + 1];
    uint32_t v[RS_CHUNK];
#pragma unroll
    for (int k = 0; k < RS_CHUNK; ++k) v[k] = pa + k < pb ? cnt[(size_t)(pa + k) * RS_RADIX + d] : 0u;
    uint32_t run = 0;
#pragma unroll
    for (int k = 0; k < RS_CHUNK; ++k) {
      if (pa + k < pb) cnt[(size_t)(pa + k) * RS_RADIX + d] = run;
      run += v[k];
    }
    ctot[(size_t)d * max_chunks + c] = run;
  }
}

// Chunk bases of one pass: ctot[d][c] (the keys of digit d in chunk c) becomes
// the keys of digit d in the chunks of segment s before c, and, where hist is
// given, hist[s][pass][d] receives the segment's keys of digit d (the digit
// histogram the offsets come from).  One wave per (segment, digit), 64
// chunks per step.
__global__ void __launch_bounds__(256) k_chunk_scan(uint32_t* __restrict__ ctot, const uint32_t* __restrict__ chunk_begin,
                                                    uint32_t num_segs, uint32_t* __restrict__ hist,
                                                    uint32_t passes, uint32_t pass, uint32_t max_chunks) {
  const uint32_t wave = blockIdx.x * 4u + (threadIdx.x >> 6), lane = threadIdx.x & 63u;
  const uint32_t s = wave >> 8, d = wave & 255u;
  if (s >= num_segs) return;
  const uint32_t cb = chunk_begin[s], ce = chunk_begin[s + 1];
  uint32_t carry = 0;
  uint32_t* col = ctot + (size_t)d * max_chunks;
  for (uint32_t c = cb; c < ce; c += 64u) {
    const uint32_t i = c + lane;
    const uint32_t v = i < ce ? col[i] : 0u;
    const uint32_t incl = wave_incl_scan(v);
    if (i < ce) col[i] = carry + incl - v;
    carry += __shfl(incl, 63, 64);
  }
  if (hist && lane == 0) hist[((size_t)s * passes + pass) * RS_RADIX + d] = carry;
}

// Clears the look-back status words of the live partitions of every pass
// (pass ps owns words [ps * max_parts * 256, ...)); the live count is read
// on the device, so a launch sized for the capacity clears only what is used.
__global__ void __launch_bounds__(256) k_clear_status(uint32_t* __restrict__ status,
                                                      const uint32_t* __restrict__ part_begin, uint32_t num_segs,
                                                      uint32_t passes, uint32_t max_parts, uint32_t skip_mask) {
  const uint64_t words = (uint64_t)part_begin[num_segs] * RS_RADIX;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint32_t ps = 0; ps < passes; ++ps) {
    if ((skip_mask >> ps) & 1u) continue;  // a pass that takes its offsets from a count table
    uint4* st = reinterpret_cast<uint4*>(status + (uint64_t)ps * max_parts * RS_RADIX);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words / 4; i += stride)
      st[i] = make_uint4(0u, 0u, 0u, 0u);
  }
}

// offsets[s][pass][d] = seg_begin[s] + sum of hist[s][pass][d' < d].
// One block per (segment, pass) of passes [first_pass, first_pass + npass).
__global__ void __launch_bounds__(RS_RADIX) k_sort_offsets(const uint32_t* __restrict__ hist,
                                                           const uint64_t* __restrict__ seg_begin, uint32_t passes,
                                                           uint32_t first_pass, uint32_t npass,
                                                           uint32_t* __restrict__ offsets) {
  __shared__ uint32_t s_wave[4];
  const uint32_t s = blockIdx.x / npass, ps = first_pass + blockIdx.x % npass;
  const uint32_t idx = ((s * passes) + ps) * RS_RADIX + threadIdx.x;
  uint32_t total;
  const uint32_t ex = block_excl_scan<RS_RADIX>(hist[idx], s_wave, total);
  offsets[idx] = (uint32_t)seg_begin[s] + ex;
}

//
// TABLE: the block's digit bases come from count tables written before the
// pass (tcnt[p][d]: digit d's keys in the partitions of p's chunk before p;
// tctot[d][c]: where chunk c's keys of digit d begin) instead of the
// look-back, so no block waits on another.
template <typename K, bool TABLE>
__global__ void __launch_bounds__(RS_BLOCK) __attribute__((amdgpu_waves_per_eu(RS_WAVES_PER_EU))) k_onesweep(const K* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
                                                       K* __restrict__ keys_out, uint32_t* __restrict__ vals_out,
                                                       const uint64_t* __restrict__ seg_begin,
                                                       const uint32_t* __restrict__ part_begin, uint32_t num_segs,
                                                       uint32_t shift, uint32_t dmask, uint32_t passes, uint32_t pass,
                                                       const uint32_t* __restrict__ offsets, uint32_t* status,
                                                       uint32_t* part_counter, uint32_t* err,
                                                       const uint32_t* __restrict__ tcnt, const uint32_t* __restrict__ tctot,
                                                       const uint32_t* __restrict__ chunk_begin, uint32_t max_chunks) {
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
    uint32_t tbase = 0;  // TABLE: digit tid's first output position for this block
    if (TABLE)
      tbase = offsets[((size_t)s * passes + pass) * RS_RADIX + tid] +
              tctot[(size_t)tid * max_chunks + chunk_begin[s] + (p - first_part) / RS_CHUNK] +
              tcnt[(size_t)p * RS_RADIX + tid];

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
      const uint32_t d = (uint32_t)(key[r] >> shift) & dmask;
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

    if (TABLE) {
      s_gbase[d] = tbase - ex;
    } else {
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
      constexpr uint32_t LB_WIN = RS_LB_WIN;
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
          if (RS_LB_SLEEP) __builtin_amdgcn_s_sleep(RS_LB_SLEEP);
        }
      }
      __hip_atomic_store(&st[(size_t)p * RS_RADIX + d], FLAG_INC | ((prefix + tot) & VALUE_MASK), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
    s_gbase[d] = offsets[((size_t)s * passes + pass) * RS_RADIX + d] + prefix - ex;
    }
    __syncthreads();

    // ---- reorder in LDS by digit, then write digit runs contiguously ----
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
      const uint32_t dd = (uint32_t)(key[r] >> shift) & dmask;
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
        const uint32_t dd = (uint32_t)(k >> shift) & dmask;
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
  return rs_ws_bytes(num_segs, max_pairs, key_bytes);
}

// Launch the onesweep passes over bits [lo_bit, lo_bit + first_bits +
// 8 (passes - 1)) of pairs of segments (pass 0 sorts the first first_bits
// bits, each later pass 8; result in keys/vals, or in the alt buffers when
// *result_in_alt is set) with no host synchronisation.  seg_begin (device,
// num_segs + 1 entries) bounds the segments; max_pairs bounds
// seg_begin[num_segs].  A look-back that never completes sets *err (checked
// by callers).  *offsets_out receives the digit starts: offsets[(s * passes +
// ps) * 256 + d] is where digit d of pass ps begins in segment s.
template <typename K>
static int onesweep_ws(nts_ctx* ctx, K* keys, uint32_t* vals, K* keys_alt, uint32_t* vals_alt,
                       const uint64_t* d_seg_begin, uint32_t num_segs, uint64_t max_pairs, uint32_t lo_bit,
                       uint32_t passes, void* ws, bool* result_in_alt, uint32_t** err_out,
                       const uint32_t** offsets_out = nullptr, uint32_t** counters_out = nullptr,
                       uint32_t first_bits = 8) {
  *result_in_alt = false;
  if (num_segs == 0 || num_segs > RS_MAX_SEGS) return nts_fail(NTS_E_INVALID, "segments %u", num_segs);
  if (max_pairs >= (1ull << 30)) return nts_fail(NTS_E_CAPACITY, "sort supports < 2^30 pairs (%llu)",
                                                 (unsigned long long)max_pairs);
  if (first_bits == 0 || first_bits > 8) return nts_fail(NTS_E_INVALID, "first digit of %u bits", first_bits);
  if (passes && lo_bit + first_bits + 8 * (passes - 1) > 8 * sizeof(K))
    return nts_fail(NTS_E_INVALID, "sort bits beyond the key");
  const uint32_t max_parts = nts_div_up(max_pairs, RS_TILE) + num_segs;
  const uint32_t max_chunks = rs_max_chunks(num_segs, max_pairs);
  const uint64_t n_hist = (uint64_t)num_segs * sizeof(K) * RS_RADIX;
  uint32_t* part_begin = (uint32_t*)ws;
  uint32_t* chunk_begin = part_begin + 128;
  uint32_t* hist = part_begin + 256;
  uint32_t* offsets = hist + n_hist;
  uint32_t* counters = offsets + n_hist;
  uint32_t* err = counters + 32;
  uint32_t* status = counters + 64;  // per pass: look-back words, or the count table
  uint32_t* ctot = status + (size_t)sizeof(K) * max_parts * RS_RADIX;
  // Count tables only for large sorts: below RS_TABLE_MIN_PAIRS their extra
  // launches cost more than the look-back they replace.
  const uint32_t table = max_pairs >= RS_TABLE_MIN_PAIRS ? RS_TABLE_MASK & ((1u << passes) - 1u) : 0u;
  if (err_out) *err_out = err;
  if (offsets_out) *offsets_out = offsets;
  if (counters_out) *counters_out = counters;  // words 40-63 are free for the caller
  int rc = nts_fill_u32(ctx, hist, 0, n_hist);
  if (rc) return rc;
  rc = nts_fill_u32(ctx, counters, 0, 64);
  if (rc) return rc;
  if (passes == 0 || max_pairs == 0) return NTS_OK;
  NTS_LAUNCH(ctx, "sort_parts", k_sort_parts, dim3(1), dim3(64), 0, d_seg_begin, num_segs, part_begin, chunk_begin);
  if (table != (1u << passes) - 1u)
    NTS_LAUNCH(ctx, "sort_clear", k_clear_status, dim3(1024), dim3(256), 0, status, (const uint32_t*)part_begin,
               num_segs, passes, max_parts, table);
  const uint32_t grid = max_parts < RS_GRID ? max_parts : RS_GRID;
  const uint32_t cgrid = max_chunks < RS_GRID ? max_chunks : RS_GRID;
  const uint32_t sgrid = nts_div_up(num_segs * RS_RADIX, 4);  // one wave per (segment, digit)
  // With every pass on count tables, each pass's digit histogram is its
  // table's total (k_chunk_scan writes it) and no histogram pass runs;
  // otherwise the histogram pass counts every pass's digits, and pass 0's
  // count table if it has one.
  const bool all_tables = table == (1u << passes) - 1u;
  if (!all_tables) {
    NTS_LAUNCH(ctx, "sort_hist", k_sort_hist<K>, dim3(grid), dim3(RS_BLOCK), 0, (const K*)keys, d_seg_begin,
               (const uint32_t*)part_begin, num_segs, lo_bit, first_bits, passes, hist,
               (table & 1u) ? status : (uint32_t*)nullptr);
    NTS_LAUNCH(ctx, "sort_offsets", k_sort_offsets, dim3(num_segs * passes), dim3(RS_RADIX), 0,
               (const uint32_t*)hist, d_seg_begin, passes, 0u, passes, offsets);
  }
  K* src_k = keys;
  uint32_t* src_v = vals;
  K* dst_k = keys_alt;
  uint32_t* dst_v = vals_alt;
  for (uint32_t ps = 0; ps < passes; ++ps) {
    uint32_t* st = status + (size_t)ps * max_parts * RS_RADIX;
    const uint32_t shift = ps ? lo_bit + first_bits + 8 * (ps - 1) : lo_bit;
    const uint32_t dmask = ps ? 255u : (1u << first_bits) - 1u;
    if ((table >> ps) & 1u) {
      if (ps > 0 || all_tables)  // the count table of this pass's digit over this pass's input
        NTS_LAUNCH(ctx, "sort_part_count", k_part_count<K>, dim3(max_parts), dim3(RS_BLOCK), 0, (const K*)src_k,
                   d_seg_begin, (const uint32_t*)part_begin, num_segs, shift, dmask, st);
      NTS_LAUNCH(ctx, "sort_chunk_prefix", k_chunk_prefix, dim3(cgrid), dim3(RS_RADIX), 0, st,
                 (const uint32_t*)part_begin, (const uint32_t*)chunk_begin, num_segs, ctot, max_chunks);
      NTS_LAUNCH(ctx, "sort_chunk_scan", k_chunk_scan, dim3(sgrid), dim3(256), 0, ctot,
                 (const uint32_t*)chunk_begin, num_segs, all_tables ? hist : (uint32_t*)nullptr, passes, ps,
                 max_chunks);
      if (all_tables)
        NTS_LAUNCH(ctx, "sort_offsets", k_sort_offsets, dim3(num_segs), dim3(RS_RADIX), 0, (const uint32_t*)hist,
                   d_seg_begin, passes, ps, 1u, offsets);
      NTS_LAUNCH(ctx, "sort_onesweep", (k_onesweep<K, true>), dim3(grid), dim3(RS_BLOCK), 0, (const K*)src_k,
                 (const uint32_t*)src_v, dst_k, dst_v, d_seg_begin, (const uint32_t*)part_begin, num_segs,
                 shift, dmask, passes, ps, (const uint32_t*)offsets, st, counters + ps, err, (const uint32_t*)st,
                 (const uint32_t*)ctot, (const uint32_t*)chunk_begin, max_chunks);
    } else {
      NTS_LAUNCH(ctx, "sort_onesweep", (k_onesweep<K, false>), dim3(grid), dim3(RS_BLOCK), 0, (const K*)src_k,
                 (const uint32_t*)src_v, dst_k, dst_v, d_seg_begin, (const uint32_t*)part_begin, num_segs,
                 shift, dmask, passes, ps, (const uint32_t*)offsets, st, counters + ps, err,
                 (const uint32_t*)nullptr, (const uint32_t*)nullptr, (const uint32_t*)chunk_begin, max_chunks);
    }
    K* tk = src_k; src_k = dst_k; dst_k = tk;
    uint32_t* tv = src_v; src_v = dst_v; dst_v = tv;
  }
  *result_in_alt = (passes & 1u) != 0;
  return NTS_OK;
}

// The full sort: the low ceil(key_bits / 8) * 8 bits.
template <typename K>
static int sort_segmented_ws(nts_ctx* ctx, K* keys, uint32_t* vals, K* keys_alt, uint32_t* vals_alt,
                             const uint64_t* d_seg_begin, uint32_t num_segs, uint64_t max_pairs, uint32_t key_bits,
                             void* ws, bool* result_in_alt, uint32_t** err_out) {
  uint32_t passes = (key_bits + 7) / 8;
  if (passes > sizeof(K)) passes = sizeof(K);
  return onesweep_ws<K>(ctx, keys, vals, keys_alt, vals_alt, d_seg_begin, num_segs, max_pairs, 0, passes, ws,
                        result_in_alt, err_out);
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

// ---------------------------------------------------------------------------
// Binning sort of (tile << 16 | depth16) keys, segment by segment, into the
// order of the full-key stable sort (ascending key, ties in emit order).
//
// The onesweep passes sort only the tile bits from bit 16 + g up.  Each group
// of 2^g consecutive tiles is then contiguous and, inside it, still in emit
// order, so sorting every group stably by its low 16 + g bits completes the
// order.  One workgroup sorts a group in LDS (stable LSD, 8-bit digits ranked
// with wave64 ballots as the onesweep blocks rank theirs) and writes the
// group's tile ranges from the sorted keys.  A group beyond the LDS capacity
// goes in chunks through the other ping-pong buffer: one stable pass by its g
// tile bits, then each tile by depth.  A group is 1/256 of its view's pairs,
// which outgrows LDS for views of a few million pairs, so groups span at most
// 2^BS_GMAX tiles; views with more tiles take the onesweep over every tile
// bit, the tile ranges from the keys, and one workgroup per tile for the
// depth (k_tile_sort: 8192 pairs in LDS).
//
// The bits sorted are [0, 16 + tile_bits): keys must be below
// 2^(16 + tile_bits) or all ones (a culled pair), and with culled pairs the
// tile ids stay below 2^tile_bits - 1, so the all-ones key sorts last.
// ---------------------------------------------------------------------------
#define BS_GMAX 0         // tile bits a group takes in LDS (0: a group is one tile)
#define GS_GROUP_NT 1024  // threads of a group workgroup
#define GS_TILE_NT 256    // threads of a tile workgroup

// Ranks lane digits within a wave: returns whist[d] (the wave's running count
// of digit d) plus the lanes below this one with the same digit, and adds the
// wave's lanes of each digit to whist[d].  Digits have w <= 8 bits; lanes
// with valid false take no part (their result is meaningless).
__device__ __forceinline__ uint32_t gs_wave_rank(uint32_t d, uint32_t w, bool valid, uint32_t* whist) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t vb = __ballot(valid);
  uint32_t mlo = (uint32_t)vb, mhi = (uint32_t)(vb >> 32);
#pragma unroll
  for (uint32_t b = 0; b < 8; ++b) {
    if (b < w) {
      const uint32_t neg = (uint32_t)((int32_t)(d << (31 - b)) >> 31);
      const uint64_t bal = __ballot(neg != 0u);
      mlo &= ~((uint32_t)bal ^ neg);
      mhi &= ~((uint32_t)(bal >> 32) ^ neg);
    }
  }
  const uint64_t match = valid ? (((uint64_t)mhi << 32) | mlo) : 0ull;
  const uint32_t leader = match ? (uint32_t)__builtin_ctzll(match) : 64u;
  uint32_t old = 0;
  if (lane == leader) {
    old = whist[d];
    whist[d] = old + (uint32_t)__popcll(match);
  }
  old = __shfl(old, (int)(leader & 63u), 64);
  return old + (uint32_t)__popcll(match & ((1ull << lane) - 1ull));
}

// After ranking: the per-wave rows of s_hist become exclusive prefixes over
// the waves, and the thread of digit t (t < 256) gets the digit's total.
template <int NT>
__device__ __forceinline__ uint32_t gs_wave_prefix(uint32_t* s_hist) {
  uint32_t tot = 0;
  if (threadIdx.x < 256) {
#pragma unroll 4
    for (int w = 0; w < NT / 64; ++w) {
      const uint32_t c = s_hist[w * 256 + threadIdx.x];
      s_hist[w * 256 + threadIdx.x] = tot;
      tot += c;
    }
  }
  return tot;
}

// Sorts n <= NT * ITEMS pairs of (in_k, in_v) by bits [0, bits) into (out_k,
// out_v) through LDS; in and out may be the same arrays (every pair is read
// before any is written).  Wave w ranks the R * 64 consecutive pairs from
// w * R * 64 (R = ceil(n / NT) rounds); the slots past n hold all-ones keys,
// which stay behind every pair.  With ranges set, the ranges of the sorted
// tiles below tn go to ranges[tbase + tile], positions offset by gpos.
template <typename K, int NT, int ITEMS>
__device__ void gs_lds_sort(const K* in_k, const uint32_t* in_v, K* out_k, uint32_t* out_v, uint32_t n,
                            uint32_t bits, K* s_keys, uint32_t* s_vals, uint32_t* s_hist, uint32_t* s_excl,
                            uint32_t* s_scan, uint2* ranges, uint32_t tbase, uint32_t tn, uint64_t gpos) {
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  const uint32_t R = (n + NT - 1) / NT;
  const uint32_t wbase = wave * R * 64u;
  K key[ITEMS];
  uint32_t val[ITEMS], rank[ITEMS];
#pragma unroll
  for (int r = 0; r < ITEMS; ++r) {
    if ((uint32_t)r < R) {
      const uint32_t i = wbase + (uint32_t)r * 64u + lane;
      const bool valid = i < n;
      key[r] = valid ? in_k[i] : ~(K)0;
      val[r] = valid ? in_v[i] : 0u;
    }
  }
  for (uint32_t shift = 0; shift < bits; shift += 8) {
    const uint32_t w = bits - shift < 8u ? bits - shift : 8u;
    const uint32_t mask = (1u << w) - 1u;
    for (uint32_t i = t; i < (NT / 64) * 256; i += NT) s_hist[i] = 0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ITEMS; ++r)
      if ((uint32_t)r < R) rank[r] = gs_wave_rank((uint32_t)(key[r] >> shift) & mask, w, true, s_hist + wave * 256);
    __syncthreads();
    const uint32_t tot = gs_wave_prefix<NT>(s_hist);
    uint32_t all;
    const uint32_t ex = block_excl_scan<NT>(tot, s_scan, all);
    if (t < 256) s_excl[t] = ex;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
      if ((uint32_t)r < R) {
        const uint32_t d = (uint32_t)(key[r] >> shift) & mask;
        const uint32_t lp = s_excl[d] + s_hist[wave * 256 + d] + rank[r];
        s_keys[lp] = key[r];
        s_vals[lp] = val[r];
      }
    }
    __syncthreads();
    if (shift + 8 < bits) {
#pragma unroll
      for (int r = 0; r < ITEMS; ++r) {
        if ((uint32_t)r < R) {
          const uint32_t i = wbase + (uint32_t)r * 64u + lane;
          key[r] = s_keys[i];
          val[r] = s_vals[i];
        }
      }
    }
  }
  for (uint32_t j = t; j < n; j += NT) {
    const K k = s_keys[j];
    out_k[j] = k;
    out_v[j] = s_vals[j];
    if (ranges) {
      const K tile = k >> 16;
      if (tile < (K)tn) {
        if (j == 0 || (s_keys[j - 1] >> 16) != tile) ranges[tbase + (uint32_t)tile].x = (uint32_t)(gpos + j);
        if (j + 1 == n || (s_keys[j + 1] >> 16) != tile) ranges[tbase + (uint32_t)tile].y = (uint32_t)(gpos + j + 1);
      }
    }
  }
  __syncthreads();  // the caller reuses the LDS
}

// One stable counting pass by digit (key >> shift) & (2^w - 1) of n pairs of
// (in_k, in_v) into (out_k, out_v) (other arrays), NT * ITEMS pairs at a
// time.  With s_start set, digit d's run is left at [s_start[d], s_start[d] +
// s_count[d]).
template <typename K, int NT, int ITEMS>
__device__ void gs_chunk_pass(const K* __restrict__ in_k, const uint32_t* __restrict__ in_v, K* __restrict__ out_k,
                              uint32_t* __restrict__ out_v, uint32_t n, uint32_t shift, uint32_t w, uint32_t* s_hist,
                              uint32_t* s_run, uint32_t* s_scan, uint32_t* s_start, uint32_t* s_count) {
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  const uint32_t mask = (1u << w) - 1u;
  uint32_t* wh = s_hist + wave * 256;
  for (uint32_t i = t; i < (NT / 64) * 256; i += NT) s_hist[i] = 0;
  __syncthreads();
  // digit counts, a chunk at a time (its loads issued together)
  for (uint32_t c0 = 0; c0 < n; c0 += NT * ITEMS) {
    const uint32_t wbase = c0 + wave * ITEMS * 64u;
    uint32_t dg[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
      const uint32_t i = wbase + (uint32_t)r * 64u + lane;
      dg[r] = i < n ? (uint32_t)(in_k[i] >> shift) & mask : 0u;
    }
#pragma unroll
    for (int r = 0; r < ITEMS; ++r)
      if (wbase + (uint32_t)r * 64u < n) (void)gs_wave_rank(dg[r], w, wbase + (uint32_t)r * 64u + lane < n, wh);
  }
  __syncthreads();
  const uint32_t tot = gs_wave_prefix<NT>(s_hist);
  uint32_t all;
  const uint32_t ex = block_excl_scan<NT>(tot, s_scan, all);
  if (t < 256) {
    s_run[t] = ex;
    if (s_start) {
      s_start[t] = ex;
      s_count[t] = tot;
    }
  }
  __syncthreads();
  K key[ITEMS];
  uint32_t val[ITEMS], rank[ITEMS];
  for (uint32_t c0 = 0; c0 < n; c0 += NT * ITEMS) {
    for (uint32_t i = t; i < (NT / 64) * 256; i += NT) s_hist[i] = 0;
    __syncthreads();
    const uint32_t wbase = c0 + wave * ITEMS * 64u;
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
      rank[r] = 0;
      if (wbase + (uint32_t)r * 64u < n) {  // the wave has pairs in this round
        const uint32_t i = wbase + (uint32_t)r * 64u + lane;
        const bool valid = i < n;
        key[r] = valid ? in_k[i] : (K)0;
        val[r] = valid ? in_v[i] : 0u;
        rank[r] = gs_wave_rank((uint32_t)(key[r] >> shift) & mask, w, valid, wh);
      }
    }
    __syncthreads();
    const uint32_t ctot = gs_wave_prefix<NT>(s_hist);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
      const uint32_t i = wbase + (uint32_t)r * 64u + lane;
      if (i < n) {
        const uint32_t d = (uint32_t)(key[r] >> shift) & mask;
        const uint32_t pos = s_run[d] + s_hist[wave * 256 + d] + rank[r];
        out_k[pos] = key[r];
        out_v[pos] = val[r];
      }
    }
    __syncthreads();
    if (t < 256) s_run[t] += ctot;
  }
  __syncthreads();
}

template <typename K, int NT>
__device__ void gs_copy(const K* __restrict__ in_k, const uint32_t* __restrict__ in_v, K* __restrict__ out_k,
                        uint32_t* __restrict__ out_v, uint32_t n) {
  for (uint32_t j = threadIdx.x; j < n; j += NT) {
    out_k[j] = in_k[j];
    out_v[j] = in_v[j];
  }
  __syncthreads();
}

// One workgroup per group, sorting it in place in (keys, vals) with (tmp_k,
// tmp_v) as scratch over the same positions.
// BY_DIGIT: group blockIdx.x = (segment s, digit d) of the onesweep pass,
// [digit_start[s * 256 + d], the next start or the segment end); the sort
// takes bits [0, 16 + g) and writes the group's tile ranges (tile ids below
// the segment's tile count, ranges[tile_base[s] + tile]).
// Otherwise: group blockIdx.x = groups[blockIdx.x] (one tile), depth bits only.
template <typename K, bool BY_DIGIT, int NT, int ITEMS>
__global__ void __launch_bounds__(NT) k_group_sort(K* keys, uint32_t* vals, K* tmp_k, uint32_t* tmp_v,
                                                   const uint32_t* __restrict__ digit_start,
                                                   const uint64_t* __restrict__ seg_begin,
                                                   const uint2* __restrict__ groups, uint32_t g,
                                                   const uint32_t* __restrict__ tile_base, uint2* ranges) {
  constexpr uint32_t CAP = NT * ITEMS;
  __shared__ K s_keys[CAP];
  __shared__ uint32_t s_vals[CAP];
  __shared__ uint32_t s_hist[(NT / 64) * 256];
  __shared__ uint32_t s_excl[256], s_run[256], s_start[256], s_count[256];
  __shared__ uint32_t s_scan[NT / 64];
  uint64_t gb, ge;
  uint32_t d = 0, tbase = 0, tn = 0;
  if (BY_DIGIT) {
    const uint32_t s = blockIdx.x >> 8;
    d = blockIdx.x & 255u;
    gb = digit_start[blockIdx.x];
    ge = d < 255u ? (uint64_t)digit_start[blockIdx.x + 1] : seg_begin[s + 1];
    tbase = tile_base[s];
    tn = tile_base[s + 1] - tbase;
  } else {
    const uint2 r = groups[blockIdx.x];
    gb = r.x;
    ge = r.y;
  }
  if (ge <= gb) return;
  const uint32_t n = (uint32_t)(ge - gb);
  K* rk = keys + gb;
  uint32_t* rv = vals + gb;
  K* sk = tmp_k + gb;
  uint32_t* sv = tmp_v + gb;
  if (n <= CAP) {
    gs_lds_sort<K, NT, ITEMS>(rk, rv, rk, rv, n, 16 + g, s_keys, s_vals, s_hist, s_excl, s_scan,
                              BY_DIGIT ? ranges : nullptr, tbase, tn, gb);
    return;
  }
  if (BY_DIGIT && g > 0) {
    // by tile within the group, then each tile by depth
    gs_chunk_pass<K, NT, ITEMS>(rk, rv, sk, sv, n, 16, g, s_hist, s_run, s_scan, s_start, s_count);
    for (uint32_t b = 0; b < (1u << g); ++b) {
      const uint32_t bs = s_start[b], bc = s_count[b];
      if (bc == 0) continue;
      if (bc <= CAP) {
        gs_lds_sort<K, NT, ITEMS>(sk + bs, sv + bs, rk + bs, rv + bs, bc, 16, s_keys, s_vals, s_hist, s_excl, s_scan,
                                  nullptr, 0, 0, 0);
      } else {
        gs_chunk_pass<K, NT, ITEMS>(sk + bs, sv + bs, rk + bs, rv + bs, bc, 0, 8, s_hist, s_run, s_scan, nullptr,
                                    nullptr);
        gs_chunk_pass<K, NT, ITEMS>(rk + bs, rv + bs, sk + bs, sv + bs, bc, 8, 8, s_hist, s_run, s_scan, nullptr,
                                    nullptr);
        gs_copy<K, NT>(sk + bs, sv + bs, rk + bs, rv + bs, bc);
      }
      const uint32_t tile = (d << g) | b;
      if (tile < tn && threadIdx.x == 0)
        ranges[tbase + tile] = make_uint2((uint32_t)(gb + bs), (uint32_t)(gb + bs + bc));
    }
  } else {
    // one tile: its two depth digits, ending in place
    gs_chunk_pass<K, NT, ITEMS>(rk, rv, sk, sv, n, 0, 8, s_hist, s_run, s_scan, nullptr, nullptr);
    gs_chunk_pass<K, NT, ITEMS>(sk, sv, rk, rv, n, 8, 8, s_hist, s_run, s_scan, nullptr, nullptr);
    if (BY_DIGIT && d < tn && threadIdx.x == 0) ranges[tbase + d] = make_uint2((uint32_t)gb, (uint32_t)ge);
  }
}

// Compiler barrier between LDS accesses of one wave that different lanes
// make: the hardware runs a wave's LDS instructions in order, so only the
// compiler could reorder them.
__device__ __forceinline__ void wave_lds_order() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// One wave sorts the n <= 64 * WITEMS pairs of a tile in place by depth, as
// the workgroup path of k_tile_sort does but with no barrier: the packed
// words go through the wave's own LDS (s_pack: 64 * WITEMS words, wh: 256
// digit counters), and one wave scan turns the counts into digit offsets.
template <typename K, int WITEMS>
__device__ void gs_wave_sort(K* rk, uint32_t* rv, uint32_t n, uint32_t* s_pack, uint32_t* wh) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t R = (n + 63u) / 64u;
  const K hi = rk[0] >> 16;
  uint32_t p[WITEMS], rank[WITEMS];
#pragma unroll
  for (int r = 0; r < WITEMS; ++r) {
    if ((uint32_t)r < R) {
      const uint32_t i = (uint32_t)r * 64u + lane;
      p[r] = i < n ? ((uint32_t)(rk[i] & (K)0xFFFFu) << 16) | i : 0xFFFFFFFFu;
    }
  }
#pragma unroll
  for (uint32_t shift = 16; shift < 32; shift += 8) {
    reinterpret_cast<uint4*>(wh)[lane] = make_uint4(0u, 0u, 0u, 0u);
    wave_lds_order();
#pragma unroll
    for (int r = 0; r < WITEMS; ++r)
      if ((uint32_t)r < R) rank[r] = gs_wave_rank((p[r] >> shift) & 255u, 8, true, wh);
    wave_lds_order();
    // digit offsets: lane l owns digits 4 l .. 4 l + 3
    const uint4 c = reinterpret_cast<const uint4*>(wh)[lane];
    const uint32_t sum = c.x + c.y + c.z + c.w;
    const uint32_t ex = wave_incl_scan(sum) - sum;
    reinterpret_cast<uint4*>(wh)[lane] = make_uint4(ex, ex + c.x, ex + c.x + c.y, ex + c.x + c.y + c.z);
    wave_lds_order();
#pragma unroll
    for (int r = 0; r < WITEMS; ++r)
      if ((uint32_t)r < R) s_pack[wh[(p[r] >> shift) & 255u] + rank[r]] = p[r];
    wave_lds_order();
    if (shift == 16) {
#pragma unroll
      for (int r = 0; r < WITEMS; ++r)
        if ((uint32_t)r < R) p[r] = s_pack[(uint32_t)r * 64u + lane];
      wave_lds_order();
    }
  }
#pragma unroll
  for (int r = 0; r < WITEMS; ++r) {
    const uint32_t j = (uint32_t)r * 64u + lane;
    if ((uint32_t)r < R && j < n) {
      p[r] = s_pack[j];
      rank[r] = rv[p[r] & 0xFFFFu];
    }
  }
  // every payload gathered (vmcnt(0)) before any is overwritten
  __builtin_amdgcn_s_waitcnt(0x0F70);
#pragma unroll
  for (int r = 0; r < WITEMS; ++r) {
    const uint32_t j = (uint32_t)r * 64u + lane;
    if ((uint32_t)r < R && j < n) {
      rk[j] = (hi << 16) | (K)(p[r] >> 16);
      rv[j] = rank[r];
    }
  }
}

// Sorts every tile's pairs in place by depth (the low 16 key bits; the pairs
// of a tile share the bits above), heaviest tiles first (order: a large tile
// started late would run on alone at the end).  The first *n_heavy tiles of
// the order, those of NTS_TILE_HEAVY pairs or more, take one workgroup each:
// up to NT * ITEMS pairs go through LDS as one u32 each, depth << 16 | the
// pair's index in the tile, the stable LSD over the two depth digits leaves
// them in (depth, index) order, and the payloads are then gathered by index
// (from L2: the tile was just read); 4 B of LDS per pair keeps several tiles
// in flight per CU.  Larger tiles take two chunked passes through (tmp_k,
// tmp_v), CITEMS pairs per thread per chunk.  The lighter tiles take one wave
// each (gs_wave_sort), NT / 64 per workgroup, with no barrier.  The grid has
// one workgroup per tile, which covers both.
template <typename K, int NT, int ITEMS, int CITEMS>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) k_tile_sort(K* keys, uint32_t* vals, K* tmp_k, uint32_t* tmp_v,
                                                  const uint2* __restrict__ ranges, const uint32_t* __restrict__ order,
                                                  const uint32_t* __restrict__ n_heavy, uint32_t n_tiles) {
  constexpr uint32_t CAP = NT * ITEMS;
  static_assert(CAP <= 65536, "the pair index takes 16 bits");
  static_assert(64u * ITEMS == NTS_TILE_HEAVY, "a wave sorts every tile below the heavy class");
  __shared__ uint32_t s_pack[CAP];
  __shared__ __align__(16) uint32_t s_hist[(NT / 64) * 256];  // the wave path reads it as uint4
  __shared__ uint32_t s_excl[256], s_run[256];
  __shared__ uint32_t s_scan[NT / 64];
  const uint32_t heavy = *n_heavy;
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  if (blockIdx.x >= heavy) {
    const uint32_t k = heavy + (blockIdx.x - heavy) * (NT / 64) + wave;
    if (k >= n_tiles) return;
    const uint2 rg = ranges[order[k]];
    if (rg.y <= rg.x + 1u) return;  // zero or one pair: in order
    gs_wave_sort<K, ITEMS>(keys + rg.x, vals + rg.x, rg.y - rg.x, s_pack + wave * 64u * ITEMS, s_hist + wave * 256u);
    return;
  }
  const uint2 rg = ranges[order[blockIdx.x]];
  const uint32_t n = rg.y - rg.x;
  K* rk = keys + rg.x;
  uint32_t* rv = vals + rg.x;
  if (n > CAP) {
    gs_chunk_pass<K, NT, CITEMS>(rk, rv, tmp_k + rg.x, tmp_v + rg.x, n, 0, 8, s_hist, s_run, s_scan, nullptr,
                                 nullptr);
    gs_chunk_pass<K, NT, CITEMS>(tmp_k + rg.x, tmp_v + rg.x, rk, rv, n, 8, 8, s_hist, s_run, s_scan, nullptr,
                                 nullptr);
    return;
  }
  const uint32_t R = (n + NT - 1) / NT;
  const uint32_t wbase = wave * R * 64u;
  const K hi = rk[0] >> 16;  // the tile's key bits above the depth
  uint32_t p[ITEMS], rank[ITEMS];
#pragma unroll
  for (int r = 0; r < ITEMS; ++r) {
    if ((uint32_t)r < R) {
      const uint32_t i = wbase + (uint32_t)r * 64u + lane;
      p[r] = i < n ? ((uint32_t)(rk[i] & (K)0xFFFFu) << 16) | i : 0xFFFFFFFFu;
    }
  }
#pragma unroll
  for (uint32_t shift = 16; shift < 32; shift += 8) {
    for (uint32_t i = t; i < (NT / 64) * 256; i += NT) s_hist[i] = 0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ITEMS; ++r)
      if ((uint32_t)r < R) rank[r] = gs_wave_rank((p[r] >> shift) & 255u, 8, true, s_hist + wave * 256);
    __syncthreads();
    const uint32_t tot = gs_wave_prefix<NT>(s_hist);
    uint32_t all;
    const uint32_t ex = block_excl_scan<NT>(tot, s_scan, all);
    if (t < 256) s_excl[t] = ex;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
      if ((uint32_t)r < R) {
        const uint32_t d = (p[r] >> shift) & 255u;
        s_pack[s_excl[d] + s_hist[wave * 256 + d] + rank[r]] = p[r];
      }
    }
    __syncthreads();
    if (shift == 16) {
#pragma unroll
      for (int r = 0; r < ITEMS; ++r)
        if ((uint32_t)r < R) p[r] = s_pack[wbase + (uint32_t)r * 64u + lane];
    }
  }
  // gather every payload before any is overwritten
#pragma unroll
  for (int r = 0; r < ITEMS; ++r) {
    const uint32_t j = (uint32_t)r * NT + t;
    if ((uint32_t)r < R && j < n) {
      p[r] = s_pack[j];
      rank[r] = rv[p[r] & 0xFFFFu];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < ITEMS; ++r) {
    const uint32_t j = (uint32_t)r * NT + t;
    if ((uint32_t)r < R && j < n) {
      rk[j] = (hi << 16) | (K)(p[r] >> 16);
      rv[j] = rank[r];
    }
  }
}

// Launch the binning sort (result in keys/vals, or in the alt buffers when
// *result_in_alt is set) and the tile ranges, with no host synchronisation.
// ranges (total_tiles entries, zero-filled by the caller) receives [begin,
// end) of every tile; d_tile_base (num_segs + 1 entries) holds each
// segment's first tile.  ws as for the full sort.
template <typename K>
static int bin_sort_ws(nts_ctx* ctx, K* keys, uint32_t* vals, K* keys_alt, uint32_t* vals_alt,
                       const uint64_t* d_seg_begin, uint32_t num_segs, uint64_t max_pairs, uint32_t tile_bits,
                       const uint32_t* d_tile_base, uint32_t total_tiles, uint2* ranges, uint32_t* order,
                       void* ws, bool* result_in_alt, uint32_t** err_out) {
  if (tile_bits == 0 || 16 + tile_bits > 8 * sizeof(K))
    return nts_fail(NTS_E_INVALID, "%u tile bits do not fit %u-byte keys", tile_bits, (unsigned)sizeof(K));
  const bool by_digit = tile_bits <= 8 + BS_GMAX;
  const uint32_t g = by_digit && tile_bits > 8 ? tile_bits - 8 : 0;
  // Exactly the tile bits: the first pass takes the remainder of a whole
  // number of bytes (fewer digits, so longer runs of each digit per block
  // and fuller cache lines in its scatter), the later passes 8 bits each.
  const uint32_t passes = by_digit ? 1 : (tile_bits + 7) / 8;
  const uint32_t first_bits = by_digit ? 8 : tile_bits - 8 * (passes - 1);
  const uint32_t* offsets = nullptr;
  bool alt = false;
  uint32_t* counters = nullptr;
  int rc = onesweep_ws<K>(ctx, keys, vals, keys_alt, vals_alt, d_seg_begin, num_segs, max_pairs, 16 + g, passes, ws,
                          &alt, err_out, &offsets, &counters, first_bits);
  if (rc) return rc;
  *result_in_alt = alt;
  if (max_pairs == 0) return nts_tile_order(ctx, ranges, total_tiles, order);
  K* rk = alt ? keys_alt : keys;
  uint32_t* rv = alt ? vals_alt : vals;
  K* sk = alt ? keys : keys_alt;
  uint32_t* sv = alt ? vals : vals_alt;
  if (by_digit) {
    constexpr int IA = sizeof(K) == 4 ? 16 : 8;
    NTS_LAUNCH(ctx, "bin_group_sort", (k_group_sort<K, true, GS_GROUP_NT, IA>), dim3(num_segs * 256),
               dim3(GS_GROUP_NT), 0, rk, rv, sk, sv, offsets, d_seg_begin, (const uint2*)nullptr, g, d_tile_base,
               ranges);
    return nts_tile_order(ctx, ranges, total_tiles, order);
  } else {
    uint32_t grid = nts_div_up(max_pairs, 256);
    if (grid > 8192) grid = 8192;
    NTS_LAUNCH(ctx, "tile_ranges", k_tile_ranges<K>, dim3(grid), dim3(256), 0, (const K*)rk, d_seg_begin, num_segs,
               d_tile_base, ranges);
    uint32_t* n_heavy = counters + 40;
    rc = nts_tile_order(ctx, ranges, total_tiles, order, n_heavy);
    if (rc) return rc;
    if (total_tiles > 0)
      NTS_LAUNCH(ctx, "bin_tile_sort", (k_tile_sort<K, GS_TILE_NT, 32, 16>), dim3(total_tiles), dim3(GS_TILE_NT), 0,
                 rk, rv, sk, sv, (const uint2*)ranges, (const uint32_t*)order, (const uint32_t*)n_heavy,
                 total_tiles);
  }
  return NTS_OK;
}

int nts_bin_sort_ws_launch32(nts_ctx* ctx, uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                             const uint64_t* d_seg_begin, uint32_t num_segs, uint64_t max_pairs, uint32_t tile_bits,
                             const uint32_t* d_tile_base, uint32_t total_tiles, uint2* ranges, uint32_t* order,
                             void* ws, bool* result_in_alt, uint32_t** err) {
  return bin_sort_ws<uint32_t>(ctx, keys, vals, keys_alt, vals_alt, d_seg_begin, num_segs, max_pairs, tile_bits,
                               d_tile_base, total_tiles, ranges, order, ws, result_in_alt, err);
}

int nts_bin_sort_ws_launch64(nts_ctx* ctx, uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt,
                             const uint64_t* d_seg_begin, uint32_t num_segs, uint64_t max_pairs, uint32_t tile_bits,
                             const uint32_t* d_tile_base, uint32_t total_tiles, uint2* ranges, uint32_t* order,
                             void* ws, bool* result_in_alt, uint32_t** err) {
  return bin_sort_ws<uint64_t>(ctx, keys, vals, keys_alt, vals_alt, d_seg_begin, num_segs, max_pairs, tile_bits,
                               d_tile_base, total_tiles, ranges, order, ws, result_in_alt, err);
}

// The binning sort with its workspace from the context scratch, checked:
// returns once the sort has finished.
template <typename K>
static int bin_sort_sync(nts_ctx* ctx, K* keys, uint32_t* vals, K* keys_alt, uint32_t* vals_alt,
                         const uint64_t* d_seg_begin, uint32_t num_segs, uint64_t max_pairs, uint32_t tile_bits,
                         const uint32_t* d_tile_base, uint32_t total_tiles, uint2* ranges, uint32_t* order,
                         bool* result_in_alt) {
  *result_in_alt = false;
  if (max_pairs == 0) return nts_tile_order(ctx, ranges, total_tiles, order);
  void* scr;
  int rc = nts_scratch(ctx, nts_sort_ws_bytes(num_segs, max_pairs, sizeof(K)), &scr);
  if (rc) return rc;
  uint32_t* err = nullptr;
  rc = bin_sort_ws<K>(ctx, keys, vals, keys_alt, vals_alt, d_seg_begin, num_segs, max_pairs, tile_bits, d_tile_base,
                      total_tiles, ranges, order, scr, result_in_alt, &err);
  if (rc) return rc;
  uint32_t* hw = ctx->host_word;
  NTS_HIP_TRY(hipMemcpyAsync(hw, err, 4, hipMemcpyDeviceToHost, ctx->stream));
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (*hw) return nts_fail(NTS_E_HIP, "radix sort look-back timed out");
  return NTS_OK;
}

int nts_bin_sort32_dev(nts_ctx* ctx, uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                       const uint64_t* d_seg_begin, uint32_t num_segs, uint64_t max_pairs, uint32_t tile_bits,
                       const uint32_t* d_tile_base, uint32_t total_tiles, uint2* ranges, uint32_t* order,
                       bool* result_in_alt) {
  return bin_sort_sync<uint32_t>(ctx, keys, vals, keys_alt, vals_alt, d_seg_begin, num_segs, max_pairs, tile_bits,
                                 d_tile_base, total_tiles, ranges, order, result_in_alt);
}

int nts_bin_sort64_dev(nts_ctx* ctx, uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt,
                       const uint64_t* d_seg_begin, uint32_t num_segs, uint64_t max_pairs, uint32_t tile_bits,
                       const uint32_t* d_tile_base, uint32_t total_tiles, uint2* ranges, uint32_t* order,
                       bool* result_in_alt) {
  return bin_sort_sync<uint64_t>(ctx, keys, vals, keys_alt, vals_alt, d_seg_begin, num_segs, max_pairs, tile_bits,
                                 d_tile_base, total_tiles, ranges, order, result_in_alt);
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

template <typename K>
static int tile_sort_api(nts_ctx* ctx, nts_buffer* keys, nts_buffer* vals, uint64_t n, uint32_t tile_bits,
                         nts_buffer* ranges, uint32_t num_tiles) {
  NTS_CHECK_ARG(n * sizeof(K) <= keys->bytes && n * 4 <= vals->bytes && (uint64_t)num_tiles * 8 <= ranges->bytes);
  NTS_CHECK_ARG(tile_bits >= 1 && 16 + tile_bits <= 8 * sizeof(K) && (uint64_t)num_tiles <= (1ull << tile_bits));
  if (n >= (1ull << 30)) return nts_fail(NTS_E_CAPACITY, "sort supports < 2^30 pairs (%llu)", (unsigned long long)n);
  int rc = nts_fill_u32(ctx, ranges->ptr, 0, 2ull * num_tiles);
  if (rc) return rc;
  if (n == 0) {
    NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return NTS_OK;
  }
  // ping-pong buffers, the one segment bound and the tile base
  void* tmp = nullptr;
  const uint64_t kbytes = (n * sizeof(K) + 255) & ~255ull;
  const uint64_t vbytes = (n * 4 + 255) & ~255ull;
  const uint64_t bytes = 256 + kbytes + vbytes + (uint64_t)num_tiles * 4 + 4;
  rc = nts_dev_alloc(ctx, &tmp, bytes);
  if (rc) return rc;
  uint64_t* segb = (uint64_t*)tmp;
  uint32_t* tbase = (uint32_t*)((char*)tmp + 64);
  K* k_alt = (K*)((char*)tmp + 256);
  uint32_t* v_alt = (uint32_t*)((char*)tmp + 256 + kbytes);
  uint32_t* order = (uint32_t*)((char*)tmp + 256 + kbytes + vbytes);
  uint64_t hb[2] = {0, n};
  uint32_t ht[2] = {0, num_tiles};
  bool alt = false;
  if (hipMemcpyAsync(segb, hb, 16, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
      hipMemcpyAsync(tbase, ht, 8, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
    rc = nts_fail(NTS_E_HIP, "tile sort: bounds upload failed");
  if (rc == NTS_OK)
    rc = bin_sort_sync<K>(ctx, (K*)keys->ptr, (uint32_t*)vals->ptr, k_alt, v_alt, segb, 1, n, tile_bits, tbase,
                          num_tiles, (uint2*)ranges->ptr, order, &alt);
  if (rc == NTS_OK && alt) {
    (void)hipMemcpyAsync(keys->ptr, k_alt, n * sizeof(K), hipMemcpyDeviceToDevice, ctx->stream);
    (void)hipMemcpyAsync(vals->ptr, v_alt, n * 4, hipMemcpyDeviceToDevice, ctx->stream);
  }
  (void)hipStreamSynchronize(ctx->stream);
  nts_dev_free(ctx, tmp, bytes);
  return rc;
}

extern "C" int nts_sort_tile_pairs(nts_ctx* ctx, nts_buffer* keys, nts_buffer* vals, uint64_t n, uint32_t key_bytes,
                                   uint32_t tile_bits, nts_buffer* ranges, uint32_t num_tiles) {
  NTS_CHECK_ARG(ctx && keys && vals && ranges && (key_bytes == 4 || key_bytes == 8));
  if (key_bytes == 4) return tile_sort_api<uint32_t>(ctx, keys, vals, n, tile_bits, ranges, num_tiles);
  return tile_sort_api<uint64_t>(ctx, keys, vals, n, tile_bits, ranges, num_tiles);
}

// ---------------------------------------------------------------------------
// nts_sort_pairs_u32_noalloc: nts_sort_pairs_u32 with every buffer beyond
// the caller's keys and values in one file-scope device array, for callers
// that cannot allocate device memory.  The array holds the segment bounds,
// the ping-pong keys and values and the sort workspace (histograms, digit
// offsets, look-back status words, partition counters, the error word) for
// up to NTS_SORT_NOALLOC_MAX_PAIRS pairs: 19.0 MB, held on each GPU the
// library runs on until the process ends.  Every call on a GPU uses that
// GPU's one copy, so calls that overlap on one GPU race; whatever must start
// at zero is cleared by each call (fill kernels), nothing is allocated.
// ---------------------------------------------------------------------------
namespace {
constexpr uint64_t na_al(uint64_t x) { return (x + 255) & ~255ull; }
constexpr uint64_t kNaMax = NTS_SORT_NOALLOC_MAX_PAIRS;
constexpr uint64_t kNaWs = rs_ws_bytes(1, kNaMax, 4);  // nts_sort_ws_bytes(1, max, 4)
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
