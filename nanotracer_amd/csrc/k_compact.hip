// Prune and densify on the device (train.rs:414-602): stream compaction of
// the scene planes that reproduces the reference's CPU bookkeeping exactly.
//
// Removal.  The reference removes a set R (|R| = m) of an n-splat scene by
// swap_remove in descending index order: step j (j = m-1 .. 0) moves the
// element at position n-m+j into r_j, the j-th smallest index of R.  So the
// final array keeps every position below n-m that is not in R, and hole r_j
// receives what position n-m+j held at step j: the original element there
// unless that position is itself r_j' (j' > j), whose content came from
// n-m+j' at step j', and so on.  Following that chain from tail slot j is
// "next(j) = rank(n-m+j) if n-m+j is in R", which strictly increases j and
// ends at a slot that is not removed.  On the device: flags -> exclusive
// scan (ranks) -> next pointers over the m tail slots -> pointer jumping
// (ceil(log2 m) rounds) -> every hole copies its chain end.  Sources lie in
// the tail [n-m, n) and destinations below n-m, so the copy is in place.
//
// Densify.  Candidates (grad_acc / count > threshold) are taken in index
// order up to the budget max_splats - n; each is a split (max log-scale above
// the split threshold) or a clone.  Clones are appended in index order, then
// two children per split, whose positions draw from the LCG stream of the
// step (12 draws per split: each split parent jumps the stream to 12 x its
// rank); then the split parents are removed as above.
#include "nts_internal.h"

namespace {

__global__ void __launch_bounds__(256) k_prune_mark(const float* __restrict__ logit, uint32_t n, float thr,
                                                    uint32_t* __restrict__ flags) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flags[i] = logit[i] < thr ? 1u : 0u;
}

__global__ void __launch_bounds__(256) k_dens_mark(const float* __restrict__ P, const float* __restrict__ grad_acc,
                                                   size_t cap, uint32_t n, float count_f, float thr, float split_thr,
                                                   uint32_t* __restrict__ cand, uint32_t* __restrict__ split) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float avg = grad_acc[i] / count_f;
  const bool c = avg > thr;  // avg <= threshold skips (train.rs:437-440)
  float mx = -__builtin_inff();
#pragma unroll
  for (int k = 0; k < 3; ++k) mx = nts_maxf(mx, P[(size_t)(P_SCALE + k) * cap + i]);
  cand[i] = c ? 1u : 0u;
  split[i] = (c && mx > split_thr) ? 1u : 0u;
}

__global__ void __launch_bounds__(256) k_dens_accept(const uint32_t* __restrict__ cand,
                                                     const uint32_t* __restrict__ split,
                                                     const uint32_t* __restrict__ crank, uint32_t n, uint32_t budget,
                                                     uint32_t* __restrict__ acc_clone,
                                                     uint32_t* __restrict__ acc_split) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool a = cand[i] && crank[i] < budget;
  acc_clone[i] = (a && !split[i]) ? 1u : 0u;
  acc_split[i] = (a && split[i]) ? 1u : 0u;
}

__device__ __forceinline__ void zero_state(float* M, float* V, float* grad_acc, size_t cap, uint32_t dst) {
  for (int k = 0; k < P_COUNT; ++k) {
    M[(size_t)k * cap + dst] = 0.0f;
    V[(size_t)k * cap + dst] = 0.0f;
  }
  grad_acc[dst] = 0.0f;
}

__global__ void __launch_bounds__(256) k_dens_emit(float* __restrict__ P, float* __restrict__ M,
                                                   float* __restrict__ V, float* __restrict__ grad_acc, size_t cap,
                                                   uint32_t n, const uint32_t* __restrict__ acc_clone,
                                                   const uint32_t* __restrict__ clone_rank,
                                                   const uint32_t* __restrict__ acc_split,
                                                   const uint32_t* __restrict__ split_rank, uint32_t n_clone,
                                                   uint32_t seed) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (acc_clone[i]) {
    // train.rs:466-481: the copy and the parent both get logit + ln(1/2)
    const uint32_t dst = n + clone_rank[i];
    const float op_new = P[(size_t)P_LOGIT * cap + i] + NTS_LN_HALF;
    for (int k = 0; k < P_COUNT; ++k) P[(size_t)k * cap + dst] = P[(size_t)k * cap + i];
    P[(size_t)P_LOGIT * cap + dst] = op_new;
    P[(size_t)P_LOGIT * cap + i] = op_new;
    zero_state(M, V, grad_acc, cap, dst);
  } else if (acc_split[i]) {
    // train.rs:486-525: two children at parent + q * (N(0,1) * sigma)
    const uint32_t r = split_rank[i];
    uint32_t st = nts_lcg_jump(seed, 12ull * r);
    float pos[3], q[4], sc[3], sg[3], nsc[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      pos[k] = P[(size_t)(P_POS + k) * cap + i];
      sc[k] = P[(size_t)(P_SCALE + k) * cap + i];
      nsc[k] = sc[k] + NTS_NEG_LN_1_6;
      sg[k] = nts_expf(sc[k]);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) q[k] = P[(size_t)(P_ROT + k) * cap + i];
    const float qw = q[0], qx = q[1], qy = q[2], qz = q[3];
    for (int child = 0; child < 2; ++child) {
      float off[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) off[k] = nts_box_muller(&st) * sg[k];
      // unit quaternion times vector: v + 2w (q x v) + 2 q x (q x v)
      const float tx = 2.0f * (qy * off[2] - qz * off[1]);
      const float ty = 2.0f * (qz * off[0] - qx * off[2]);
      const float tz = 2.0f * (qx * off[1] - qy * off[0]);
      const float rx = off[0] + qw * tx + (qy * tz - qz * ty);
      const float ry = off[1] + qw * ty + (qz * tx - qx * tz);
      const float rz = off[2] + qw * tz + (qx * ty - qy * tx);
      const uint32_t dst = n + n_clone + 2u * r + (uint32_t)child;
      for (int k = 0; k < P_COUNT; ++k) P[(size_t)k * cap + dst] = P[(size_t)k * cap + i];
      P[(size_t)(P_POS + 0) * cap + dst] = pos[0] + rx;
      P[(size_t)(P_POS + 1) * cap + dst] = pos[1] + ry;
      P[(size_t)(P_POS + 2) * cap + dst] = pos[2] + rz;
#pragma unroll
      for (int k = 0; k < 3; ++k) P[(size_t)(P_SCALE + k) * cap + dst] = nsc[k];
      zero_state(M, V, grad_acc, cap, dst);
    }
  }
}

// next pointer of tail slot j (position N - m + j), tail-relative
__global__ void __launch_bounds__(256) k_tail_next(const uint32_t* __restrict__ flags,
                                                   const uint32_t* __restrict__ rank, uint32_t N, uint32_t m,
                                                   uint32_t* __restrict__ nxt) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const uint32_t t = N - m + j;
  nxt[j] = flags[t] ? rank[t] : j;
}

__global__ void __launch_bounds__(256) k_jump(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                              uint32_t m) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m) out[j] = in[in[j]];
}

// Hole p (< N - m, flagged) takes the chain end of tail slot rank[p]: one
// thread per (hole candidate, plane group).
__global__ void __launch_bounds__(256) k_fill_holes(float* __restrict__ P, float* __restrict__ M,
                                                    float* __restrict__ V, float* __restrict__ grad_acc, size_t cap,
                                                    const uint32_t* __restrict__ flags,
                                                    const uint32_t* __restrict__ rank,
                                                    const uint32_t* __restrict__ chain_end, uint32_t N, uint32_t m) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t plane_set = blockIdx.y;  // 0: params, 1: m, 2: v (+ grad_acc)
  if (p >= N - m || !flags[p]) return;
  const uint32_t src = N - m + chain_end[rank[p]];
  float* base = plane_set == 0 ? P : (plane_set == 1 ? M : V);
  for (int k = 0; k < P_COUNT; ++k) base[(size_t)k * cap + p] = base[(size_t)k * cap + src];
  if (plane_set == 2) grad_acc[p] = grad_acc[src];
}

struct Work {
  void* base = nullptr;
  uint64_t bytes = 0;
  uint32_t *flags, *rank, *a, *b, *c, *d, *e, *nxt0, *nxt1;
  uint64_t* totals;
  void* scan_ws;
};

int work_alloc(nts_ctx* ctx, uint32_t N, Work& w) {
  const uint64_t words = ((uint64_t)N + 63) & ~63ull;
  const uint64_t scan_bytes = (nts_scan_ws_bytes(N) + 255) & ~255ull;
  w.bytes = words * 4 * 9 + 256 + scan_bytes;
  int rc = nts_dev_alloc(ctx, &w.base, w.bytes);
  if (rc) return rc;
  uint32_t* q = (uint32_t*)w.base;
  uint32_t** arrs[9] = {&w.flags, &w.rank, &w.a, &w.b, &w.c, &w.d, &w.e, &w.nxt0, &w.nxt1};
  for (int k = 0; k < 9; ++k) *arrs[k] = q + (uint64_t)k * words;
  w.totals = (uint64_t*)(q + 9 * words);
  w.scan_ws = (char*)w.base + words * 4 * 9 + 256;
  return NTS_OK;
}

int read_totals(nts_ctx* ctx, const uint64_t* d, uint64_t* h, int count) {
  NTS_HIP_TRY(hipMemcpyAsync(h, d, 8 * (size_t)count, hipMemcpyDeviceToHost, ctx->stream));
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return NTS_OK;
}

// Remove the flagged splats of [0, N) (flags, exclusive-scan ranks, m
// flagged) in the order of descending swap_remove; the scene then holds
// N - m splats.
int remove_flagged(nts_ctx* ctx, nts_scene* s, Work& w, uint32_t N, uint32_t m) {
  if (m == 0) return NTS_OK;
  const size_t cap = s->cap;
  NTS_LAUNCH(ctx, "compact_next", k_tail_next, dim3(nts_div_up(m, 256)), dim3(256), 0, (const uint32_t*)w.flags,
             (const uint32_t*)w.rank, N, m, w.nxt0);
  uint32_t* in = w.nxt0;
  uint32_t* out = w.nxt1;
  for (uint32_t span = 1; span < m; span <<= 1) {  // chains are shorter than m
    NTS_LAUNCH(ctx, "compact_jump", k_jump, dim3(nts_div_up(m, 256)), dim3(256), 0, (const uint32_t*)in, out, m);
    uint32_t* t = in;
    in = out;
    out = t;
  }
  if (N > m)
    NTS_LAUNCH(ctx, "compact_fill", k_fill_holes, dim3(nts_div_up(N - m, 256), 3), dim3(256), 0, s->params, s->m,
               s->v, s->grad_acc, cap, (const uint32_t*)w.flags, (const uint32_t*)w.rank, (const uint32_t*)in, N, m);
  s->n = N - m;
  return NTS_OK;
}

}  // namespace

extern "C" int nts_scene_zero_grad_acc(nts_ctx* ctx, nts_scene* s) {
  NTS_CHECK_ARG(ctx && s);
  NTS_NEED_TRAINABLE(s);
  return nts_fill_u32(ctx, s->grad_acc, 0, s->cap);
}

extern "C" int nts_prune(nts_ctx* ctx, nts_scene* s, float logit_threshold, uint32_t* removed) {
  NTS_CHECK_ARG(ctx && s);
  NTS_NEED_TRAINABLE(s);
  if (removed) *removed = 0;
  const uint32_t n = s->n;
  if (n == 0) return NTS_OK;
  Work w;
  int rc = work_alloc(ctx, n, w);
  if (rc) return rc;
  struct Guard {
    nts_ctx* c; Work* w;
    ~Guard() { (void)hipStreamSynchronize(c->stream); nts_dev_free(c, w->base, w->bytes); }
  } guard{ctx, &w};
  NTS_LAUNCH(ctx, "prune_mark", k_prune_mark, dim3(nts_div_up(n, 256)), dim3(256), 0,
             (const float*)(s->params + (size_t)P_LOGIT * s->cap), n, logit_threshold, w.flags);
  rc = nts_scan_u32_ws(ctx, w.flags, w.rank, n, w.totals, w.scan_ws);
  if (rc) return rc;
  uint64_t m = 0;
  rc = read_totals(ctx, w.totals, &m, 1);
  if (rc) return rc;
  rc = remove_flagged(ctx, s, w, n, (uint32_t)m);
  if (rc) return rc;
  if (removed) *removed = (uint32_t)m;
  return NTS_OK;
}

extern "C" int nts_densify(nts_ctx* ctx, nts_scene* s, uint32_t grad_count, float grad_threshold,
                           float split_log_scale, uint32_t max_splats, uint32_t iter, uint32_t counts[3]) {
  NTS_CHECK_ARG(ctx && s);
  NTS_NEED_TRAINABLE(s);
  if (counts) counts[0] = counts[1] = counts[2] = 0;
  const uint32_t n = s->n;
  if (n == 0) return NTS_OK;
  const uint32_t budget = max_splats > n ? max_splats - n : 0u;
  const float count_f = (float)(grad_count > 1u ? grad_count : 1u);
  // the count can grow to n + budget + (splits <= budget) before parents go
  const uint64_t worst = (uint64_t)n + 2ull * budget;
  if (worst >= (1ull << 32)) return nts_fail(NTS_E_CAPACITY, "densify would exceed 2^32 splats");
  Work w;
  int rc = work_alloc(ctx, (uint32_t)(worst > n ? worst : n), w);
  if (rc) return rc;
  struct Guard {
    nts_ctx* c; Work* w;
    ~Guard() { (void)hipStreamSynchronize(c->stream); nts_dev_free(c, w->base, w->bytes); }
  } guard{ctx, &w};
  const dim3 g1(nts_div_up(n, 256));
  uint32_t *cand = w.a, *split = w.b, *crank = w.c, *acc_clone = w.d, *acc_split = w.e;
  NTS_LAUNCH(ctx, "densify_mark", k_dens_mark, g1, dim3(256), 0, (const float*)s->params,
             (const float*)s->grad_acc, (size_t)s->cap, n, count_f, grad_threshold, split_log_scale, cand, split);
  rc = nts_scan_u32_ws(ctx, cand, crank, n, w.totals, w.scan_ws);
  if (rc) return rc;
  NTS_LAUNCH(ctx, "densify_accept", k_dens_accept, g1, dim3(256), 0, (const uint32_t*)cand,
             (const uint32_t*)split, (const uint32_t*)crank, n, budget, acc_clone, acc_split);
  // ranks among the accepted clones / splits (cand and crank are free now)
  uint32_t* clone_rank = cand;
  uint32_t* split_rank = w.rank;
  rc = nts_scan_u32_ws(ctx, acc_clone, clone_rank, n, w.totals, w.scan_ws);
  if (rc) return rc;
  rc = nts_scan_u32_ws(ctx, acc_split, split_rank, n, w.totals + 1, w.scan_ws);
  if (rc) return rc;
  uint64_t tot[2] = {0, 0};
  rc = read_totals(ctx, w.totals, tot, 2);
  if (rc) return rc;
  const uint32_t n_clone = (uint32_t)tot[0], n_split = (uint32_t)tot[1];
  if (n_clone + n_split == 0) return NTS_OK;
  const uint64_t grown = (uint64_t)n + n_clone + 2ull * n_split;
  if (grown > s->cap)
    return nts_fail(NTS_E_CAPACITY, "densify needs %llu splats of room, the scene holds %u",
                    (unsigned long long)grown, s->cap);
  const uint32_t seed = iter * 1664525u + 1013904223u;  // train.rs:460
  NTS_LAUNCH(ctx, "densify_emit", k_dens_emit, g1, dim3(256), 0, s->params, s->m, s->v, s->grad_acc, (size_t)s->cap,
             n, (const uint32_t*)acc_clone, (const uint32_t*)clone_rank, (const uint32_t*)acc_split,
             (const uint32_t*)split_rank, n_clone, seed);
  // remove the split parents: flags = acc_split on [0, n), 0 on the appended
  const uint32_t N = (uint32_t)grown;
  NTS_HIP_TRY(hipMemcpyAsync(w.flags, acc_split, (size_t)n * 4, hipMemcpyDeviceToDevice, ctx->stream));
  rc = nts_fill_u32(ctx, w.flags + n, 0, N - n);
  if (rc) return rc;
  rc = nts_scan_u32_ws(ctx, w.flags, w.rank, N, w.totals, w.scan_ws);
  if (rc) return rc;
  s->n = N;
  rc = remove_flagged(ctx, s, w, N, n_split);
  if (rc) return rc;
  if (counts) {
    counts[0] = n_clone;
    counts[1] = 2u * n_split;
    counts[2] = n_split;
  }
  return NTS_OK;
}
