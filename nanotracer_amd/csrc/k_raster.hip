// Per-tile alpha blend, forward and backward, plus loss and tonemap.
//
// One workgroup per (view, tile).  Splat records of the tile's depth-sorted
// slice are staged through LDS in batches; every thread blends its own
// pixel(s) front to back with exactly the arithmetic of the host
// restatement (bit-identical images).  A wave skips a splat whose
// alpha >= 1/255 region misses the wave's pixels, a pixel skips the
// exponential below the splat's power threshold, and a block-wide vote ends
// the tile once every pixel is saturated.
#include <map>
#include <mutex>

#include "nts_raster.h"

// Thread layout of a (view, tile) workgroup with at most MAXT threads.
// Every wave covers a square sub-tile: 8x8 pixels one per lane, or 16x16
// as a 2x2 quad per lane when the tile has more pixels than threads (each
// row of 16 lanes then an 8x8 square of it); the waves tile the tile
// row-major.  Square rectangles are what the per-wave culling rejects best.  The forward runs one pixel per thread
// (up to 1024 threads); the backward caps the workgroup at 256 threads,
// which keeps its per-batch barriers and the LDS partials of its waves
// small.
template <int TS, int MAXT>
struct Blend {
  static constexpr int NPIX = TS * TS;
  static constexpr int NT = NPIX < MAXT ? NPIX : MAXT;
  static constexpr int PPT = NPIX / NT;  // 1 or 4
  static constexpr int SUB = PPT == 1 ? 8 : 16;
  static_assert(PPT == 1 || PPT == 4, "one pixel or a 2x2 quad per thread");
};
#define NTS_FWD_THREADS 1024
#define NTS_BWD_THREADS 256
#ifndef NTS_FWD_LIVE_CHECK
#define NTS_FWD_LIVE_CHECK 1  // the forward leaves a batch once every pixel of the wave is finished
#endif
#ifndef NTS_BWD_QUARTERS
// The quad backward's rows of 16 lanes walk their own splat streams: each
// row takes the splats that reach its 8x8-pixel square, independently of the
// other rows of the wave.
#define NTS_BWD_QUARTERS 1
#endif
#define NTS_NO_SLOT 0xFFFFFFFFu
#ifndef NTS_BWD_WAVES
#define NTS_BWD_WAVES 8  // occupancy floor of the one-pixel backward per SIMD (bounds its VGPRs)
#endif
#ifndef NTS_BWD_QUAD_WAVES
#define NTS_BWD_QUAD_WAVES 7  // the same for the quad backward (72 VGPRs: at 64 it spills in its loop)
#endif
// Occupancy floor of a 256-thread backward workgroup with PPT pixels per lane.
#define NTS_BWD_OCC(PPT) ((PPT) == 4 && NTS_BWD_QUARTERS ? NTS_BWD_QUAD_WAVES : NTS_BWD_WAVES)

// Tile-local pixel j of thread t.
template <int TS, int MAXT>
__device__ __forceinline__ void tile_pixel(uint32_t t, int j, uint32_t& lx, uint32_t& ly) {
  using B = Blend<TS, MAXT>;
  constexpr uint32_t WX = TS / B::SUB;  // waves per tile row
  const uint32_t w = t >> 6, l = t & 63u;
  const uint32_t bx = B::SUB * (w % WX), by = B::SUB * (w / WX);
  if (B::PPT == 1) {
    lx = bx + (l & 7u);
    ly = by + (l >> 3);
  } else if (NTS_BWD_QUARTERS) {
    // each row of 16 lanes covers an 8x8-pixel square of the wave (4x4
    // quads): row r = l >> 4 at (8 (r & 1), 8 (r >> 1)), its lane q = l & 15
    // at quad (q & 3, q >> 2)
    const uint32_t r = l >> 4, q = l & 15u;
    lx = bx + 8u * (r & 1u) + 2u * (q & 3u) + ((uint32_t)j & 1u);
    ly = by + 8u * (r >> 1) + 2u * (q >> 2) + ((uint32_t)j >> 1);
  } else {
    lx = bx + 2u * (l & 7u) + ((uint32_t)j & 1u);
    ly = by + 2u * (l >> 3) + ((uint32_t)j >> 1);
  }
}

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

// Pixel-centre rectangle of the wave: bounds, centre and half extents.
struct WaveRect {
  float x0, x1, y0, y1;
  float cx, cy, hw, hh;
};

template <int TS, int MAXT>
__device__ __forceinline__ WaveRect wave_rect(uint32_t tx, uint32_t ty) {
  float x0 = 3.4e38f, x1 = -3.4e38f, y0 = 3.4e38f, y1 = -3.4e38f;
#pragma unroll
  for (int j = 0; j < Blend<TS, MAXT>::PPT; ++j) {
    uint32_t lx, ly;
    tile_pixel<TS, MAXT>(threadIdx.x, j, lx, ly);
    const float fx = (float)(tx * TS + lx) + 0.5f, fy = (float)(ty * TS + ly) + 0.5f;
    x0 = fminf(x0, fx);
    x1 = fmaxf(x1, fx);
    y0 = fminf(y0, fy);
    y1 = fmaxf(y1, fy);
  }
  x0 = wave_min(x0);
  x1 = wave_max(x1);
  y0 = wave_min(y0);
  y1 = wave_max(y1);
  return WaveRect{x0, x1, y0, y1, 0.5f * (x0 + x1), 0.5f * (y0 + y1), 0.5f * (x1 - x0), 0.5f * (y1 - y0)};
}

// Staged splat (mean, conic a, 2b, c, opacity, colour) plus a conservative
// bound of the region where alpha >= 1/255:  -0.5 x^T Q x >= -ln(255 o) is
// the ellipse x^T Q x <= R = 2 ln(255 o), whose half extents are
// sqrt(R Sigma_xx) and sqrt(R Sigma_yy) with Sigma = Q^-1.  Extents get a
// 1% + 0.05 px margin and the power threshold a 1e-3 margin, so a culled
// evaluation is one that would have failed the alpha >= 1/255 test: images
// and gradients are unchanged.  The margin covers the float evaluation of
// the power for kappa < NTS_CULL_KAPPA_MAX (nts_cull_conic_ok); other
// conics, and a NaN or infinite opacity, are not culled.  Records:
// a = (mean, -a/2, -b), b = (-c/2, opacity, red, green),
// c = (blue, power threshold, box half extents x / y); the fourth record
// holds what the per-wave test needs that depends on the splat alone
// (computed once here, not once per wave): the threshold of that test with
// its float-error margin (-inf where nothing is culled), the edge
// minimisers' slopes -b/c and -b/a, and -c/2; the test's quadratic form and
// threshold carry the power's factor -1/2 too.
__device__ __forceinline__ void stage_splat(float4 pa, float4 pb, float4 pc, float4& a, float4& b, float4& c,
                                            float4& d) {
  // The conic is staged as the power's own coefficients: power =
  // fma(-c/2 dy, dy, fma(-a/2 dx, dx, -b dx dy)) rounds at every step
  // exactly as nts_blend_power's -0.5 fma(c dy, dy, fma(a dx, dx, 2b dx dy))
  // does (scaling by a power of two commutes with rounding; where an
  // intermediate is subnormal the power is within 2^-126 of zero and exp
  // gives 1 either way), one multiply fewer.
  a = make_float4(pa.x, pa.y, -0.5f * pb.x, -pb.y);
  b = make_float4(-0.5f * pb.z, pc.w, pc.x, pc.y);
  const float o = pc.w, qa = pb.x, qb = pb.y, qc = pb.z;
  const float det = qa * qc - qb * qb;
  float ex = 3.4e38f, ey = 3.4e38f, pmin, R = __builtin_inff(), rbc = 0.0f, rba = 0.0f;
  if (o * 255.0f < 0.999f) {  // alpha = o G < 1/255 everywhere
    ex = ey = -3.4e38f;
    pmin = 3.4e38f;
  } else if (!(det > 0.0f)) {  // degenerate conic: no culling
    pmin = -3.4e38f;
  } else {
    const float l = logf(o * 255.0f);
    pmin = -l - 1e-3f;
    if (l <= 3.4e38f && nts_cull_conic_ok(qa, qb, qc, det)) {  // not for a NaN or infinite opacity
      const float r2 = fmaxf(2.0f * l, 0.0f);
      ex = sqrtf(r2 * (qc / det)) * 1.01f + 0.05f;
      ey = sqrtf(r2 * (qa / det)) * 1.01f + 0.05f;
      const float s = sqrtf(qa * qc) + fabsf(qb);
      const float kap = s * s / det;
      R = -2.0f * pmin * (1.0f + 1e-5f + 4e-6f * kap);
      rbc = -qb / qc;
      rba = -qb / qa;
    }
  }
  c = make_float4(pc.z, pmin, ex, ey);
  d = make_float4(-0.5f * R, rbc, rba, -0.5f * qc);  // the test's threshold and c, scaled as the power
}
__device__ __forceinline__ void stage_splat(const Projected& p, float4& a, float4& b, float4& c, float4& d) {
  stage_splat(p.a, p.b, p.c, a, b, c, d);
}


// Whether any pixel centre of the wave's rectangle can pass the blend's
// power test  q(d) = a dx^2 + 2b dx dy + c dy^2 <= -2 pmin  for this
// splat.  First the bounding-box test of the alpha >= 1/255 ellipse, then
// the exact minimum of the convex q over the rectangle: it is 0 when the
// centre lies inside, otherwise it lies on an edge that faces the centre,
// where it is a clamped 1-D minimum (at dy = -b U / c on the edge dx = U,
// dx = -b V / a on the edge dy = V).  Every pixel's float offset lies in the
// float rectangle [x0 - mx, x1 - mx] x [y0 - my, y1 - my], and the float
// evaluation of q at any point is within 6u * kappa * q of the exact value,
// kappa = (sqrt(ac) + |b|)^2 / det (u = 2^-24); the threshold carries a
// larger margin, 4e-6 kappa + 1e-5, and only splats with kappa below
// NTS_CULL_KAPPA_MAX are culled at all.  The
// rounding of the minimiser moves q by a second-order amount.  A rejected
// splat therefore fails every pixel's power test: results are unchanged.
// Both edges are evaluated without branches (a wave's lanes hold different
// splats).
__device__ __forceinline__ bool wave_hits(const WaveRect& r, float4 A, float4 Cc, float4 D) {
  if (fabsf(r.cx - A.x) > r.hw + Cc.z || fabsf(r.cy - A.y) > r.hh + Cc.w) return false;
  // in the staged scale: -q/2 >= -R/2 (the same decision, bit for bit)
  const float u0 = r.x0 - A.x, u1 = r.x1 - A.x, v0 = r.y0 - A.y, v1 = r.y1 - A.y;
  const bool out_x = u0 > 0.0f || u1 < 0.0f, out_y = v0 > 0.0f || v1 < 0.0f;
  const float U = u0 > 0.0f ? u0 : u1;
  const float dy = __builtin_amdgcn_fmed3f(U * D.y, v0, v1);
  // (fused multiply-adds: fewer roundings than the bound above allows for)
  const float qx = __builtin_fmaf(A.z * U, U, __builtin_fmaf(A.w * U, dy, D.w * dy * dy));
  const float V = v0 > 0.0f ? v0 : v1;
  const float dx = __builtin_amdgcn_fmed3f(V * D.z, u0, u1);
  const float qy = __builtin_fmaf(A.z * dx, dx, __builtin_fmaf(A.w * dx, V, D.w * V * V));
  const float q = fmaxf(out_x ? qx : -3.4e38f, out_y ? qy : -3.4e38f);
  return (!out_x && !out_y) || q >= D.x;
}

// Per-wave hit masks of the staged batch: bit b of mk[i] is set when slot
// 64 i + b (< klim) can reach one of the wave's pixels.  Each lane tests
// NM slots; the masks are wave-uniform.
template <int NM>
__device__ __forceinline__ void wave_masks(const WaveRect& wr, const float4* s_a, const float4* s_c,
                                           const float4* s_d, uint32_t klim, uint64_t (&mk)[NM]) {
  const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
  for (int i = 0; i < NM; ++i) {
    const uint32_t k = (uint32_t)i * 64u + lane;
    bool h = false;
    if (k < klim) h = wave_hits(wr, s_a[k], s_c[k], s_d[k]);
    mk[i] = __ballot(h);
  }
}

// The 8x8-pixel square of row q (16 lanes) inside the wave's 16x16
// rectangle (the quad layout with NTS_BWD_QUARTERS).  Pixel centres are
// integers + 0.5, so the offsets are exact.
__device__ __forceinline__ WaveRect quarter_rect(const WaveRect& w, int q) {
  const float x0 = w.x0 + ((q & 1) ? 8.0f : 0.0f), y0 = w.y0 + ((q & 2) ? 8.0f : 0.0f);
  const float x1 = x0 + 7.0f, y1 = y0 + 7.0f;
  return WaveRect{x0, x1, y0, y1, 0.5f * (x0 + x1), 0.5f * (y0 + y1), 3.5f, 3.5f};
}

// Hit masks of the staged batch per row of 16 lanes: bit b of mq[r][i] is
// set when slot 64 i + 63 - b (< klq[r]) can reach a pixel of row r's square
// (reversed, so that the lowest bit is the last slot).  Each lane tests its
// slots against the four squares directly: a test of the wave's rectangle
// first measured slower (nearly every batch has a slot that passes it).
template <int NM>
__device__ __forceinline__ void quarter_masks(const WaveRect& wr, const float4* s_a, const float4* s_c,
                                              const float4* s_d, const uint32_t (&klq)[4], uint64_t (&mq)[4][NM]) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t klim = max(max(klq[0], klq[1]), max(klq[2], klq[3]));
#pragma unroll
  for (int i = 0; i < NM; ++i) {
    const uint32_t k = (uint32_t)i * 64u + 63u - lane;
    bool h0 = false, h1 = false, h2 = false, h3 = false;
    if (k < klim) {
      const float4 A = s_a[k], Cc = s_c[k], D = s_d[k];
      h0 = k < klq[0] && wave_hits(quarter_rect(wr, 0), A, Cc, D);
      h1 = k < klq[1] && wave_hits(quarter_rect(wr, 1), A, Cc, D);
      h2 = k < klq[2] && wave_hits(quarter_rect(wr, 2), A, Cc, D);
      h3 = k < klq[3] && wave_hits(quarter_rect(wr, 3), A, Cc, D);
    }
    mq[0][i] = __ballot(h0);
    mq[1][i] = __ballot(h1);
    mq[2][i] = __ballot(h2);
    mq[3][i] = __ballot(h3);
  }
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const uint32_t o = (uint32_t)__shfl_xor((int)v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}

// ---------------------------------------------------------------------------
// forward (rasterize.wgsl:48-118)
// ---------------------------------------------------------------------------
// dL/dC of one pixel (L1 with sign(0) = 0, or MSE as loss.rs:163-179) and
// its loss term, with exactly the arithmetic of the standalone loss kernel.
__device__ __forceinline__ float4 pixel_loss_grad(float4 p, float4 t, float inv_n, int kind, double& acc) {
  const float d0 = p.x - t.x, d1 = p.y - t.y, d2 = p.z - t.z;
  if (kind == NTS_LOSS_MSE) {
    acc += (double)(d0 * d0 + d1 * d1 + d2 * d2);
    return make_float4(d0 * inv_n, d1 * inv_n, d2 * inv_n, 0.0f);
  }
  acc += (double)fabsf(d0) + (double)fabsf(d1) + (double)fabsf(d2);
  return make_float4(d0 > 0.f ? inv_n : (d0 < 0.f ? -inv_n : 0.f), d1 > 0.f ? inv_n : (d1 < 0.f ? -inv_n : 0.f),
                     d2 > 0.f ? inv_n : (d2 < 0.f ? -inv_n : 0.f), 0.0f);
}

// Counting form of the blend kernels: adds a wave's evaluation count (the
// same in every lane) and the sum of its lanes' pixel counts into
// counts[0] and counts[1].
__device__ __forceinline__ void flush_counts(unsigned long long* counts, uint32_t wave_count, uint32_t lane_count) {
  uint32_t sum = lane_count;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
  if ((threadIdx.x & 63u) == 0) {
    atomicAdd(counts, (unsigned long long)wave_count);
    atomicAdd(counts + 1, (unsigned long long)sum);
  }
}

// A 1024-thread workgroup is 4 waves per SIMD: two of them share a CU only
// within 64 VGPRs (8 waves per SIMD), which the attribute holds the compiler
// to; one per CU leaves half the SIMD slots idle at every barrier.
template <int TS, int BATCH, bool LOSS, int MODE>
__global__ void __launch_bounds__((Blend<TS, NTS_FWD_THREADS>::NT))
__attribute__((amdgpu_waves_per_eu(Blend<TS, NTS_FWD_THREADS>::NT == 1024 ? 8 : 1))) k_blend_fwd(const Projected* __restrict__ proj,
                                                             const uint32_t* __restrict__ payloads,
                                                             const uint2* __restrict__ ranges, ViewTable vt,
                                                             float4* __restrict__ out, float* __restrict__ final_t,
                                                             uint32_t* __restrict__ last_idx, LossArgs la) {
  using BL = Blend<TS, NTS_FWD_THREADS>;
  constexpr int NT = BL::NT, PPT = BL::PPT, NM = BATCH / 64;
  constexpr bool PERSIST = MODE == NTS_BLEND_PERSIST, COUNT = MODE == NTS_BLEND_COUNT;
  uint32_t n_hits = 0, n_blend = 0;  // counting form: wave evaluations, lane blends
  __shared__ float4 s_a[BATCH];  // mean.x, mean.y, conic.a, conic.b
  __shared__ float4 s_b[BATCH];  // conic.c, opacity, r, g
  __shared__ float4 s_c[BATCH];  // b, power threshold, cull half-extents x / y
  __shared__ float4 s_d[BATCH];  // per-wave test: threshold, -b/c, -b/a, c
  __shared__ double s_loss[NT / 64];
  __shared__ uint32_t s_ticket;
  // One pass per tile; `slot` is the tile's place in the blend order.  The
  // persistent form takes tiles from the forward counter until it passes the
  // last, so every workgroup leaves once the tiles are handed out (the
  // barrier first lets the previous tile's LDS reads finish).
  for (uint32_t slot = blockIdx.x;;) {
  if (PERSIST) {
    __syncthreads();  // the previous tile's LDS reads are done
    if (threadIdx.x == 0) s_ticket = atomicAdd(vt.tickets, 1u);
    // The ticket's LDS write completes before the barrier (lgkmcnt(0)), so
    // no wave can read the previous tile's ticket after it.
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __syncthreads();
    slot = __builtin_amdgcn_readfirstlane(s_ticket);  // uniform: the tile's values stay scalar
    if (slot >= vt.total_tiles) break;
  }
  const uint32_t gt = vt.order ? vt.order[slot] : slot;
  const uint32_t vi = view_of_tile(vt, gt);
  const ViewDesc vd = vt.v[vi];
  const uint32_t lt = gt - vd.tile_base;
  const uint32_t tx = lt % vd.tiles_x, ty = lt / vd.tiles_x;
  const uint2 rg = ranges[gt];
  const uint32_t begin = rg.x, end = rg.y;
  const WaveRect wr = wave_rect<TS, NTS_FWD_THREADS>(tx, ty);

  // A pixel that is finished (saturated, or outside the image) has its
  // centre set to NaN: every later power test then fails, which costs the
  // evaluation loop no separate check.
  float pxf[PPT], pyf[PPT], T[PPT], C0[PPT], C1[PPT], C2[PPT];
  uint32_t last[PPT], pix_x[PPT], pix_y[PPT];
  bool inb[PPT];
#pragma unroll
  for (int j = 0; j < PPT; ++j) {
    uint32_t lx, ly;
    tile_pixel<TS, NTS_FWD_THREADS>(threadIdx.x, j, lx, ly);
    pix_x[j] = tx * TS + lx;
    pix_y[j] = ty * TS + ly;
    inb[j] = pix_x[j] < vd.width && pix_y[j] < vd.height;
    pxf[j] = inb[j] ? (float)pix_x[j] + 0.5f : __builtin_nanf("");
    pyf[j] = (float)pix_y[j] + 0.5f;
    T[j] = 1.0f;
    C0[j] = C1[j] = C2[j] = 0.0f;
    last[j] = 0;
  }
  for (uint32_t base = begin; base < end; base += BATCH) {
    int live = 0;
#pragma unroll
    for (int j = 0; j < PPT; ++j) live |= pxf[j] == pxf[j];
    if (!__syncthreads_or(live)) break;
    const uint32_t nb = (end - base) < (uint32_t)BATCH ? (end - base) : (uint32_t)BATCH;
    for (uint32_t k = threadIdx.x; k < nb; k += NT)
      stage_splat(proj[vd.proj_base + payloads[base + k]], s_a[k], s_b[k], s_c[k], s_d[k]);
    __syncthreads();
    uint64_t mk[NM];
    wave_masks<NM>(wr, s_a, s_c, s_d, __ballot(live) ? nb : 0u, mk);
#pragma unroll
    for (int i = 0; i < NM; ++i) {
      uint64_t m = mk[i];
      while (m) {
        if (NTS_FWD_LIVE_CHECK) {
          int lv = 0;
#pragma unroll
          for (int j = 0; j < PPT; ++j) lv |= pxf[j] == pxf[j];
          if (!__ballot(lv)) break;  // every pixel of the wave is finished
        }
        const uint32_t k = (uint32_t)i * 64u + (uint32_t)__builtin_ctzll(m);
        m &= m - 1;
        if (COUNT) ++n_hits;
        const float4 A = s_a[k], B = s_b[k], Cc = s_c[k];
        // All of the record is read here, from one address register: left to
        // itself the compiler sinks each read into the branch that uses it
        // and rebuilds the address there (a move per read).
        __asm__ volatile("" ::"v"(B.y), "v"(B.z), "v"(B.w), "v"(Cc.x));
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
          const float dx = pxf[j] - A.x, dy = pyf[j] - A.y;
          // nts_blend_power with the staged coefficients (-a/2, -b, -c/2)
          const float power = __builtin_fmaf(B.x * dy, dy, __builtin_fmaf(A.z * dx, dx, A.w * dx * dy));
          // power > 0, a finished pixel (NaN), or below the alpha >= 1/255 bound
          if (!(power <= 0.0f) || power < Cc.y) continue;
          const float g = nts_exp_blend(power);
          const float ar = B.y * g;
          const float alpha = fminf(ar, NTS_ALPHA_MAX);  // min(alpha_raw, 0.99): a NaN gives 0.99
          if (alpha < NTS_ALPHA_MIN) continue;
          if (COUNT) ++n_blend;
          const float ta = T[j] * alpha;
          // color + (T alpha) c, contracted (one rounding), as the host
          // restatement does
          C0[j] = __builtin_fmaf(ta, B.z, C0[j]);
          C1[j] = __builtin_fmaf(ta, B.w, C1[j]);
          C2[j] = __builtin_fmaf(ta, Cc.x, C2[j]);
          T[j] = T[j] * (1.0f - alpha);
          last[j] = base + k - begin + 1;
          pxf[j] = T[j] < NTS_T_MIN ? __builtin_nanf("") : pxf[j];
        }
      }
    }
  }
  double lacc = 0.0;
#pragma unroll
  for (int j = 0; j < PPT; ++j) {
    if (!inb[j]) continue;
    const uint64_t pix = vd.pix_base + (uint64_t)pix_y[j] * vd.width + pix_x[j];
    const float4 px = make_float4(C0[j], C1[j], C2[j], 1.0f - T[j]);
    if (out) out[pix] = px;
    if (final_t) final_t[pix] = T[j];
    if (last_idx) last_idx[pix] = last[j];
    if (LOSS) la.dl_dc[pix] = pixel_loss_grad(px, la.target[pix], vd.inv_n, la.kind, lacc);
  }
  if (LOSS) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) lacc += __shfl_xor(lacc, off, 64);
    if ((threadIdx.x & 63u) == 0) s_loss[threadIdx.x >> 6] = lacc;
    __syncthreads();
    if (threadIdx.x == 0) {
      double sum = 0.0;
#pragma unroll
      for (int w = 0; w < NT / 64; ++w) sum += s_loss[w];
      unsafeAtomicAdd(la.losses + vi, sum);
    }
  }
  if (!PERSIST) break;
  }
  if (COUNT) flush_counts(vt.counts, n_hits, n_blend);
  if (la.clear) {
    const uint64_t per = (la.clear_n4 + gridDim.x - 1) / gridDim.x;
    const uint64_t c0 = (uint64_t)blockIdx.x * per;
    const uint64_t c1 = c0 + per < la.clear_n4 ? c0 + per : la.clear_n4;
    for (uint64_t i = c0 + threadIdx.x; i < c1; i += NT) la.clear[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
}

// ---------------------------------------------------------------------------
// backward
//   reference-parity mode (rasterize_backward.wgsl:73-210): T = 1 - out.w,
//     walk from the slice end, T /= max(1 - a, 1e-4), bail once T > 1.0001;
//   recompute mode: T = saved final T, walk from the pixel's last
//     contributing splat, no bail (the exact chain rule of the forward);
//     the three divisions by 1 - a become one hardware reciprocal (1 ulp).
// Both modes take every power / alpha decision exactly as the forward does;
// the gradient arithmetic uses fused multiply-adds, and the recompute mode
// the hardware exponential (within 1e-6 relative of the exact forms).
// Each splat's nine per-pixel terms are reduce-scattered across the wave,
// the waves' sums are added in LDS (float atomics into one slot table, which
// keeps the workgroup's LDS small), and one set of global float atomics per
// (tile, splat) commits them.
// ---------------------------------------------------------------------------

// Lane exchanges within a row of 16 lanes (DPP): the mirror partners differ
// from the lane in bit 3 (row mirror) or bit 2 (half-row mirror); quad
// permutes give the xor-2 and xor-1 partners.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
#define NTS_DPP_ROW_MIRROR 0x140
#define NTS_DPP_ROW_HALF_MIRROR 0x141
#define NTS_DPP_QUAD_XOR2 0x4E  // quad_perm [2, 3, 0, 1]
#define NTS_DPP_QUAD_XOR1 0xB1  // quad_perm [1, 0, 3, 2]

// v_permlane32_swap(a, b) leaves (a[0:32], b[0:32]) and (a[32:64],
// b[32:64]) in its two registers; their sum holds, in lanes 0-31, a summed
// over each lane and lane + 32, and in lanes 32-63 the same for b: two
// values reduce-scattered across bit 5 in one swap and one add.
// v_permlane16_swap does the same across bit 4 (rows of 16 lanes).
__device__ __forceinline__ float swap_sum32(float a, float b) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ float swap_sum16(float a, float b) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

// One DPP halving stage: lanes whose bit is clear keep terms [0, H) of the
// N, the others terms [H, N), each adding the partner's copy of the terms it
// keeps.
template <int N, int H, int CTRL>
__device__ __forceinline__ void rs_stage(const float (&v)[N], float (&w)[H], bool hi_half) {
#pragma unroll
  for (int j = 0; j < H; ++j) {
    const float hi = (j + H < N) ? v[j + H] : 0.0f;
    const float keep = hi_half ? hi : v[j];
    const float send = hi_half ? v[j] : hi;
    w[j] = keep + dpp_f<CTRL>(send);
  }
}

// Reduce-scatter of the 9 per-lane gradient terms across the wave: swaps
// across bits 5 and 4 (9 -> 5 -> 3 terms), DPP halvings across bits 3 and
// 2 (3 -> 2 -> 1), then full sums across bits 1 and 0.  Afterwards every
// lane of quad q = lane >> 2 holds the wave sum of term grad_term_of_quad(q)
// (none when that is >= 9).
__device__ __forceinline__ float wave_reduce9(const float (&v)[9]) {
  const uint32_t lane = threadIdx.x & 63u;
  float w[5], x[3], y[2], z[1];
#pragma unroll
  for (int j = 0; j < 4; ++j) w[j] = swap_sum32(v[j], v[j + 5]);
  w[4] = swap_sum32(v[4], 0.0f);
  x[0] = swap_sum16(w[0], w[3]);
  x[1] = swap_sum16(w[1], w[4]);
  x[2] = swap_sum16(w[2], 0.0f);
  rs_stage<3, 2, NTS_DPP_ROW_MIRROR>(x, y, lane & 8u);
  rs_stage<2, 1, NTS_DPP_ROW_HALF_MIRROR>(y, z, lane & 4u);
  float r = z[0] + dpp_f<NTS_DPP_QUAD_XOR2>(z[0]);
  r += dpp_f<NTS_DPP_QUAD_XOR1>(r);
  return r;
}

// Reduce-scatter of the 9 per-lane terms within each row of 16 lanes (DPP
// halvings across bits 3, 2, 1 and 0: 9 -> 5 -> 3 -> 2 -> 1).  Afterwards
// lane l holds its row's sum of term grad_term_of_quad(l & 15) (none when
// that is >= 9).
__device__ __forceinline__ float row_reduce9(const float (&v)[9]) {
  const uint32_t lane = threadIdx.x & 63u;
  float w[5], x[3], y[2], z[1];
  rs_stage<9, 5, NTS_DPP_ROW_MIRROR>(v, w, lane & 8u);
  rs_stage<5, 3, NTS_DPP_ROW_HALF_MIRROR>(w, x, lane & 4u);
  rs_stage<3, 2, NTS_DPP_QUAD_XOR2>(x, y, lane & 2u);
  rs_stage<2, 1, NTS_DPP_QUAD_XOR1>(y, z, lane & 1u);
  return z[0];
}

// Term index held by quad q after wave_reduce9 (9 = none): bit 5 of the
// lane picks terms 5-8, bit 4 the upper of 5, bit 3 the upper of 3, bit 2
// the upper of 2.
__device__ __forceinline__ uint32_t grad_term_of_quad(uint32_t q) {
  const uint32_t xs = ((q & 2u) ? 2u : 0u) + (q & 1u);
  if (xs >= 3u) return 9u;
  const uint32_t ws = ((q & 4u) ? 3u : 0u) + xs;
  if (ws >= 5u) return 9u;
  const uint32_t vi = ((q & 8u) ? 5u : 0u) + ws;
  return vi < 9u ? vi : 9u;
}

// The backward of one staged splat over one lane's PPT pixels: updates the
// pixels' T and U, leaves the splat's nine per-lane gradient moments in acc
// and returns whether any of the lane's pixels took part.  relk is the
// splat's slice-relative index.  With COUNT, n_ok counts the lane's pixels
// that took part.  (Shared by the backward kernel and its instruction-mix
// probe, so that both run the same code.)
template <int PPT, bool RECOMP, bool COUNT>
__device__ __forceinline__ bool bwd_hit(const float4 A, const float4 B, const float4 Cc, uint32_t relk,
                                        const float (&pxf)[PPT], const float (&pyf)[PPT], float (&T)[PPT],
                                        float (&U)[PPT], const float (&G0)[PPT], const float (&G1)[PPT],
                                        const float (&G2)[PPT], uint32_t (&lastj)[PPT], float (&acc)[9],
                                        uint32_t& n_ok) {
  // Per-pixel moments of the splat's dL/dpower, mapped through the
  // conic once per splat at the commit:
  //   0, 1   sum gpow dx, sum gpow dy        -> dL/dmean = Q (X, Y)
  //   2      sum dL/dalpha G                  -> dL/dopacity
  //   3-5    sum gpow dx^2, dx dy, dy^2       -> dL/dconic (-1/2, -1, -1/2)
  //   6-8    sum dL/dC_i T alpha              -> dL/dcolour
  // (recompute mode: acc is set by the quad's first pixel)
  if (!RECOMP) {
#pragma unroll
    for (int q = 0; q < 9; ++q) acc[q] = 0.0f;
  }
  bool contrib = false;
  const float sigma = B.y, c0 = B.z, c1 = B.w, c2 = Cc.x;
  // A lane's 2x2 quad (PPT 4): pixel j lies at x + (j & 1), y + (j >> 1).
  // Its five positional moments (terms 0, 1, 3-5) follow from the four
  // pixels' dL/dpower at the end, Sum g (dx0 + ox)^a (dy0 + oy)^b with
  // ox, oy in {0, 1}, instead of two products and five sums per pixel.
  constexpr bool QUAD = PPT == 4;
  float gp[PPT];
#pragma unroll
  for (int j = 0; j < PPT; ++j) {
    const float dx = pxf[j] - A.x, dy = pyf[j] - A.y;
    // the forward's power and thresholds, bit for bit
    const float power = __builtin_fmaf(B.x * dy, dy, __builtin_fmaf(A.z * dx, dx, A.w * dx * dy));
    // alpha for every lane, then one branch on all the forward's tests
    // recompute mode: hardware 2^x (error well under 1e-6 relative),
    // and where alpha lies that close to 1/255 or 0.99 the forward's
    // exact exponential, so every alpha test decides as in the
    // forward.  Reference-parity mode keeps the exact exponential: its
    // 1.0001 bail on the divided-back T is sensitive to the last bit.
    // alpha = min(alpha_raw, 0.99), a NaN giving 0.99, taken next to each
    // product (a product is canonical: the min needs no quieting first)
    // The alpha >= 1/255 decision of the hardware value is taken as
    // final only outside the band [1 - 4e-6, 1 + 4e-6] 1/255; inside
    // it, and from 0.99 - 4e-6 up (the clamp), the exact exponential
    // decides.  Three compares: alpha above the band, alpha in or above
    // it, ar at or above the clamp's band.  A NaN ar gives alpha 0.99:
    // kept, no exact pass, as in the forward.
    float g, ar, alpha;
    bool aok;
    if (RECOMP) {
      g = __builtin_amdgcn_exp2f(power * 1.44269504088896341f);
      ar = sigma * g;
      alpha = fminf(ar, NTS_ALPHA_MAX);
      aok = alpha > NTS_ALPHA_MIN * (1.0f + 4e-6f);
      if ((!aok && alpha >= NTS_ALPHA_MIN * (1.0f - 4e-6f)) || ar >= NTS_ALPHA_MAX - 4e-6f) {
        g = nts_exp_blend(power);
        ar = sigma * g;
        alpha = fminf(ar, NTS_ALPHA_MAX);
        aok = !(alpha < NTS_ALPHA_MIN);
      }
    } else {
      g = nts_exp_blend(power);
      ar = sigma * g;
      alpha = fminf(ar, NTS_ALPHA_MAX);
      aok = !(alpha < NTS_ALPHA_MIN);
    }
    // the exact complement of the forward's skip tests (a NaN alpha must
    // be kept here as it is there).  The forward's power threshold needs no
    // test of its own: it lies 1e-3 below -ln(255 opacity), so a power under
    // it gives alpha below 0.999 / 255, which aok rejects (outside the band
    // of the exact exponential); a NaN threshold skips nothing on either side.
    const bool ok = relk < lastj[j] && power <= 0.0f && aok;
    if (RECOMP && j == 0) {
      // The first pixel of the quad without a branch: a lane that fails
      // the tests sets exact zeros and keeps T (acc needs no zeroing
      // on a path around it, which costs nine moves per join, and
      // starts from this pixel's terms rather than from zero).
      const float r = __builtin_amdgcn_rcpf(1.0f - alpha);
      const float Tn = T[j] * r, Un = U[j] * r;
      const float gc = __builtin_fmaf(G2[j], c2, __builtin_fmaf(G1[j], c1, G0[j] * c0));
      const float gal = __builtin_fmaf(Tn, gc, -Un);
      const float gar = (ar < NTS_ALPHA_MAX) ? gal : 0.0f;
      const float gpow = ok ? gar * ar : 0.0f;
      const float og = ok ? gar * g : 0.0f;
      const float ta = ok ? Tn * alpha : 0.0f;
      if (QUAD) {
        gp[0] = gpow;
      } else {
        const float gx = gpow * dx, gy = gpow * dy;
        acc[0] = gx;
        acc[1] = gy;
        acc[3] = gx * dx;
        acc[4] = gx * dy;
        acc[5] = gy * dy;
      }
      acc[2] = og;
      acc[6] = G0[j] * ta;
      acc[7] = G1[j] * ta;
      acc[8] = G2[j] * ta;
      U[j] = __builtin_fmaf(ta, gc, U[j]);
      T[j] = ok ? Tn : T[j];
      contrib = contrib || ok;
      if (COUNT) n_ok += ok ? 1u : 0u;
      continue;
    }
    if (QUAD) gp[j] = 0.0f;
    if (!ok) continue;
    float Tn, Un;
    if (RECOMP) {
      const float r = __builtin_amdgcn_rcpf(1.0f - alpha);
      Tn = T[j] * r;
      Un = U[j] * r;
    } else {
      const float oma = nts_maxf(1.0f - alpha, NTS_ONE_MINUS_ALPHA_EPS);
      Tn = T[j] / oma;
      T[j] = Tn;
      if (Tn > NTS_T_BAIL) {
        lastj[j] = 0u;
        continue;
      }
      Un = U[j] / oma;
    }
    T[j] = Tn;
    // dL/dalpha = T (dL/dC . c) - dL/dC . S_behind (fused multiply-adds;
    // no bit parity is needed in the gradient)
    const float gc = __builtin_fmaf(G2[j], c2, __builtin_fmaf(G1[j], c1, G0[j] * c0));
    const float gal = __builtin_fmaf(Tn, gc, -Un);
    const float gar = (ar < NTS_ALPHA_MAX) ? gal : 0.0f;  // alpha clamp gate
    const float gpow = gar * ar;                           // dL/dpower = dL/dalpha sigma G
    const float ta = Tn * alpha;
    if (QUAD) {
      gp[j] = gpow;
    } else {
      const float gx = gpow * dx, gy = gpow * dy;
      acc[0] += gx;
      acc[1] += gy;
      acc[3] = __builtin_fmaf(gx, dx, acc[3]);
      acc[4] = __builtin_fmaf(gx, dy, acc[4]);
      acc[5] = __builtin_fmaf(gy, dy, acc[5]);
    }
    acc[2] = __builtin_fmaf(gar, g, acc[2]);
    acc[6] = __builtin_fmaf(G0[j], ta, acc[6]);
    acc[7] = __builtin_fmaf(G1[j], ta, acc[7]);
    acc[8] = __builtin_fmaf(G2[j], ta, acc[8]);
    U[j] = __builtin_fmaf(ta, gc, U[j]);  // S += T alpha c (S itself is not rescaled)
    contrib = true;
    if (COUNT) ++n_ok;
  }
  if (QUAD) {
    const float dx0 = pxf[0] - A.x, dy0 = pyf[0] - A.y;
    const float S = (gp[0] + gp[1]) + (gp[2] + gp[3]);
    const float Sx = gp[1] + gp[3], Sy = gp[2] + gp[3];
    acc[0] = __builtin_fmaf(S, dx0, Sx);
    acc[1] = __builtin_fmaf(S, dy0, Sy);
    acc[3] = __builtin_fmaf(dx0, __builtin_fmaf(S, dx0, Sx + Sx), Sx);
    acc[4] = __builtin_fmaf(dx0, __builtin_fmaf(S, dy0, Sy), __builtin_fmaf(dy0, Sx, gp[3]));
    acc[5] = __builtin_fmaf(dy0, __builtin_fmaf(S, dy0, Sy + Sy), Sy);
  }
  return contrib;
}

// A 256-thread workgroup is one wave per SIMD.  The kernel issues its own
// arithmetic about as fast as a replay of it with no memory traffic or
// barriers does (k_probe_bwd_mix), so more waves to cover that arithmetic's
// latencies is what shortens it, up to the point where the loop spills: the
// one-pixel form is held to 8 waves per SIMD (64 VGPRs), the quad form, whose
// rows walk their own splat streams, to 7 (72 VGPRs; 4.31 ms on 1M splats x
// 16 views against 5.0 ms at 8 waves, where it reloads spilled registers in
// its loop, and 4.45 ms at 6).  The persistent form keeps some per-thread
// setup live across its tile loop: it is held to 5 waves per SIMD (96 VGPRs).
template <int TS, int BATCH, bool RECOMP, int MODE>
__global__ void __launch_bounds__((Blend<TS, NTS_BWD_THREADS>::NT))
__attribute__((amdgpu_waves_per_eu(Blend<TS, NTS_BWD_THREADS>::NT == 256 ? (MODE == NTS_BLEND_PERSIST ? 5 : NTS_BWD_OCC((Blend<TS, NTS_BWD_THREADS>::PPT))) : 1))) k_blend_bwd(
    const Projected* __restrict__ proj, const uint32_t* __restrict__ payloads, const uint2* __restrict__ ranges,
    ViewTable vt, const float4* __restrict__ fwd_out, const float* __restrict__ final_t,
    const uint32_t* __restrict__ last_idx, const float4* __restrict__ dl_dc, float* __restrict__ pgrad) {
  using BL = Blend<TS, NTS_BWD_THREADS>;
  constexpr int NT = BL::NT, PPT = BL::PPT, NM = BATCH / 64;
  constexpr bool PERSIST = MODE == NTS_BLEND_PERSIST, COUNT = MODE == NTS_BLEND_COUNT;
  uint32_t n_hits = 0, n_ok = 0;  // counting form: wave evaluations, contributing lane pixels
  __shared__ float4 s_a[BATCH];
  __shared__ float4 s_b[BATCH];
  __shared__ float4 s_c[BATCH];
  __shared__ float4 s_d[BATCH];
  __shared__ uint32_t s_idx[BATCH];
  __shared__ float s_acc[BATCH][9];  // each slot's terms summed over the waves (LDS float atomics)
  __shared__ uint32_t s_max;
  __shared__ uint32_t s_ticket;
  // One pass per tile, as the forward (with the backward's counter).
  for (uint32_t slot = blockIdx.x;;) {
  if (PERSIST) {
    __syncthreads();  // the previous tile's LDS reads are done
    if (threadIdx.x == 0) s_ticket = atomicAdd(vt.tickets + 1, 1u);
    // The ticket's LDS write completes before the barrier (lgkmcnt(0)), so
    // no wave can read the previous tile's ticket after it.
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __syncthreads();
    slot = __builtin_amdgcn_readfirstlane(s_ticket);  // uniform: the tile's values stay scalar
    if (slot >= vt.total_tiles) break;
  }
  const uint32_t gt = vt.order ? vt.order[slot] : slot;
  const uint32_t vi = view_of_tile(vt, gt);
  const ViewDesc vd = vt.v[vi];
  const uint32_t lt = gt - vd.tile_base;
  const uint32_t tx = lt % vd.tiles_x, ty = lt / vd.tiles_x;
  const uint2 rg = ranges[gt];
  const uint32_t begin = rg.x, end = rg.y;
  const uint32_t lane = threadIdx.x & 63u;
  constexpr bool QUART = PPT == 4 && NTS_BWD_QUARTERS;
  const uint32_t my_term = QUART ? grad_term_of_quad(lane & 15u)
                                 : ((lane & 3u) == 0u ? grad_term_of_quad(lane >> 2) : 9u);
  const bool term_lane = my_term < 9u;
  const uint32_t row = lane >> 4;
  const WaveRect wr = wave_rect<TS, NTS_BWD_THREADS>(tx, ty);

  // Per pixel: T walked back, dL/dC, U = dL/dC . S with S the sum of
  // T alpha c over the splats behind the current one (the gradient needs S
  // only through dL/dC . S / (1 - alpha)), and the slice-relative end of its
  // splats (0 once the pixel needs no more: no contributor, or the
  // reference-parity bail).
  float pxf[PPT], pyf[PPT], T[PPT], G0[PPT], G1[PPT], G2[PPT], U[PPT];
  uint32_t lastj[PPT];
  uint32_t my_max = 0;
  if (threadIdx.x == 0) s_max = 0;
#pragma unroll
  for (int j = 0; j < PPT; ++j) {
    uint32_t lx, ly;
    tile_pixel<TS, NTS_BWD_THREADS>(threadIdx.x, j, lx, ly);
    const uint32_t px = tx * TS + lx, py = ty * TS + ly;
    const bool inb = px < vd.width && py < vd.height;
    pxf[j] = (float)px + 0.5f;
    pyf[j] = (float)py + 0.5f;
    T[j] = 1.0f;
    G0[j] = G1[j] = G2[j] = 0.0f;
    lastj[j] = 0;
    if (inb) {
      const uint64_t pix = vd.pix_base + (uint64_t)py * vd.width + px;
      if (RECOMP) {
        T[j] = final_t[pix];
        lastj[j] = last_idx[pix];
      } else {
        T[j] = 1.0f - fwd_out[pix].w;
        lastj[j] = end > begin ? end - begin : 0;
      }
      const float4 g = dl_dc[pix];
      G0[j] = g.x;
      G1[j] = g.y;
      G2[j] = g.z;
    }
    U[j] = 0.0f;
    my_max = lastj[j] > my_max ? lastj[j] : my_max;
  }
  const uint32_t wave_last = wave_max_u32(my_max);  // slice entries this wave can need
  // and each row of 16 lanes (the quad form's splat streams)
  uint32_t row_last[4] = {0u, 0u, 0u, 0u};
  if constexpr (QUART) {
    uint32_t rl = my_max;
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) {
      const uint32_t o = (uint32_t)__shfl_xor((int)rl, off, 64);
      rl = o > rl ? o : rl;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) row_last[q] = (uint32_t)__builtin_amdgcn_readlane((int)rl, 16 * q);
  }
  __syncthreads();
  if (my_max) atomicMax(&s_max, my_max);
  __syncthreads();
  uint32_t rem = s_max;  // entries [begin, begin + rem) can matter to this tile
  while (rem > 0) {
    int live = 0;
#pragma unroll
    for (int j = 0; j < PPT; ++j) live |= lastj[j] != 0u;
    if (!__syncthreads_or(live)) break;
    const uint32_t take = rem < (uint32_t)BATCH ? rem : (uint32_t)BATCH;
    const uint32_t rel0 = rem - take;  // slice-relative index of LDS slot 0
    for (uint32_t k = threadIdx.x; k < take; k += NT) {
      const uint32_t sidx = payloads[begin + rel0 + k];
      stage_splat(proj[vd.proj_base + sidx], s_a[k], s_b[k], s_c[k], s_d[k]);
      s_idx[k] = sidx;
    }
    for (uint32_t i = threadIdx.x; i < take * 9u; i += NT) (&s_acc[0][0])[i] = 0.0f;
    __syncthreads();
    uint32_t klim = 0;
    if (__ballot(live) && wave_last > rel0) klim = wave_last - rel0 < take ? wave_last - rel0 : take;
    if constexpr (QUART) {
      // Every row walks the slots that reach its own square, last first; a
      // row whose slots are used up evaluates slot 0 with an index past
      // every pixel's last contributor, so none of its pixels takes part.
      uint32_t klq[4];
#pragma unroll
      for (int q = 0; q < 4; ++q)
        klq[q] = (klim && row_last[q] > rel0) ? (row_last[q] - rel0 < take ? row_last[q] - rel0 : take) : 0u;
      uint64_t mq[4][NM];
      quarter_masks<NM>(wr, s_a, s_c, s_d, klq, mq);
#pragma unroll
      for (int i = NM - 1; i >= 0; --i) {
        // this lane's row's mask of word i, walked lowest bit (last slot)
        // first in vector registers: no scalar work per evaluation
        uint64_t m = (row & 2u) ? ((row & 1u) ? mq[3][i] : mq[2][i]) : ((row & 1u) ? mq[1][i] : mq[0][i]);
        while (__ballot(m != 0ull)) {
          if (COUNT) ++n_hits;
          const bool none = m == 0ull;
          const uint32_t k = none ? 0u : (uint32_t)i * 64u + 63u - (uint32_t)__builtin_ctzll(m | (1ull << 63));
          m &= m - 1ull;
          const float4 A = s_a[k], B = s_b[k], Cc = s_c[k];
          float acc[9];
          const bool contrib = bwd_hit<PPT, RECOMP, COUNT>(A, B, Cc, none ? 0xFFFFFFFFu : rel0 + k, pxf, pyf, T,
                                                           U, G0, G1, G2, lastj, acc, n_ok);
          const uint64_t cm = __ballot(contrib);
          if (cm) {
            const float r = row_reduce9(acc);
            // (a row none of whose lanes took part adds exact zeros: cheaper
            // than testing for it)
            if (term_lane) atomicAdd(&s_acc[k][my_term], r);
          }
        }
      }
    } else {
      uint64_t mk[NM];
      wave_masks<NM>(wr, s_a, s_c, s_d, klim, mk);
#pragma unroll
      for (int i = NM - 1; i >= 0; --i) {
        uint64_t m = mk[i];
        while (m) {
          const int bt = 63 - __builtin_clzll(m);
          m &= ~(1ull << bt);
          const uint32_t k = (uint32_t)i * 64u + (uint32_t)bt;
          if (COUNT) ++n_hits;
          const float4 A = s_a[k], B = s_b[k], Cc = s_c[k];
          float acc[9];
          const bool contrib =
              bwd_hit<PPT, RECOMP, COUNT>(A, B, Cc, rel0 + k, pxf, pyf, T, U, G0, G1, G2, lastj, acc, n_ok);
          if (__ballot(contrib)) {
            const float r = wave_reduce9(acc);
            if (term_lane) atomicAdd(&s_acc[k][my_term], r);
          }
        }
      }
    }
    __syncthreads();
    // one set of float atomics per (tile, splat): each thread maps one
    // (slot, term) of the waves' sums, so the nine terms of a splat are
    // consecutive lanes and consecutive words
    for (uint32_t it = threadIdx.x; it < take * 9u; it += NT) {
      const uint32_t k = it / 9u, q = it - k * 9u;
      // dL/dmean needs both first moments
      float sum = s_acc[k][q < 2u ? 0u : q];
      const float sum2 = s_acc[k][1];
      if (q < 2u) {
        const float4 A = s_a[k];
        const float qa = -2.0f * A.z, qb = -A.w, qc = -2.0f * s_b[k].x;  // the conic, from the staged -q/2
        sum = q == 0u ? __builtin_fmaf(qa, sum, qb * sum2) : __builtin_fmaf(qb, sum, qc * sum2);
      } else if (q == 3u || q == 5u) {
        sum = -0.5f * sum;
      } else if (q == 4u) {
        sum = -sum;
      }
      if (sum != 0.0f) {
        float* o = pgrad + (size_t)(vd.proj_base + s_idx[k]) * vt.grad_words;
        unsafeAtomicAdd(o + (vt.grad_words == NTS_GRAD9_WORDS ? q : (q < 3u ? q : (q < 6u ? q + 1u : q + 2u))), sum);
      }
    }
    rem -= take;
  }
  if (!PERSIST) break;
  }
  if (COUNT) flush_counts(vt.counts + 2, n_hits, n_ok);
}

// Instruction-mix probe of the recompute backward: every wave runs `hits`
// evaluations of staged splats through bwd_hit, wave_reduce9 and the slot
// table's LDS float atomic, as the kernel does for each (splat, wave) it
// evaluates, with every pixel of the wave taking part, and with no global
// memory access and no barrier in the loop.  Workgroup size, LDS and
// occupancy bound are those of k_blend_bwd<TS, BATCH, true, grid>, so the
// launch measures how fast the kernel's own instruction mix issues (the
// host pads the LDS so that no more workgroups share a CU than the kernel's
// VGPRs allow).  The staged splats are wide (every pixel of the tile well
// inside their alpha band, alpha near 0.25); the pixel state comes from the
// runtime value s1 (1.0) so that the compiler cannot fold it; T and U
// restart every 64 evaluations so that they stay finite.
template <int TS, int BATCH>
__global__ void __launch_bounds__((Blend<TS, NTS_BWD_THREADS>::NT))
__attribute__((amdgpu_waves_per_eu(Blend<TS, NTS_BWD_THREADS>::NT == 256 ? NTS_BWD_OCC((Blend<TS, NTS_BWD_THREADS>::PPT)) : 1))) k_probe_bwd_mix(
    uint32_t hits, float s1, float* __restrict__ sink) {
  using BL = Blend<TS, NTS_BWD_THREADS>;
  constexpr int NT = BL::NT, PPT = BL::PPT;
  __shared__ float4 s_a[BATCH];
  __shared__ float4 s_b[BATCH];
  __shared__ float4 s_c[BATCH];
  __shared__ float4 s_d[BATCH];
  __shared__ float s_acc[BATCH][9];
  for (uint32_t k = threadIdx.x; k < (uint32_t)BATCH; k += NT) {
    Projected p;
    const float off = 0.01f * (float)k;
    p.a = make_float4(0.5f * TS + off, 0.5f * TS - off, 3.0f, 2.0f * TS);
    p.b = make_float4(0.4f / (TS * TS), 0.0f, 0.4f / (TS * TS), 1.0f);
    p.c = make_float4(0.3f, 0.5f, 0.7f, 0.3f);
    stage_splat(p, s_a[k], s_b[k], s_c[k], s_d[k]);
#pragma unroll
    for (int q = 0; q < 9; ++q) s_acc[k][q] = 0.0f;
  }
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63u;
  constexpr bool QUART = PPT == 4 && NTS_BWD_QUARTERS;
  const uint32_t my_term = QUART ? grad_term_of_quad(lane & 15u)
                                 : ((lane & 3u) == 0u ? grad_term_of_quad(lane >> 2) : 9u);
  const bool term_lane = my_term < 9u;
  const uint32_t row = lane >> 4;
  float pxf[PPT], pyf[PPT], T[PPT], U[PPT], G0[PPT], G1[PPT], G2[PPT];
  uint32_t lastj[PPT];
#pragma unroll
  for (int j = 0; j < PPT; ++j) {
    uint32_t lx, ly;
    tile_pixel<TS, NTS_BWD_THREADS>(threadIdx.x, j, lx, ly);
    pxf[j] = (float)lx + 0.5f;
    pyf[j] = (float)ly + 0.5f;
    T[j] = 0.5f * s1;
    U[j] = 0.0f * s1;
    G0[j] = 0.01f * s1;
    G1[j] = 0.02f * (float)(j + 1) * s1;
    G2[j] = -0.01f * s1;
    lastj[j] = 0xFFFFFF00u | (threadIdx.x & 0xFFu);
  }
  uint32_t n_ok = 0;
  // the waves of a workgroup walk the slots from different places, as their
  // own hit masks make them do in the kernel (in step, their LDS atomics
  // would all meet on one slot)
  const uint32_t k0 = __builtin_amdgcn_readfirstlane((threadIdx.x >> 6) * 37u);
  for (uint32_t it = 0; it < hits; ++it) {
    uint32_t k = (it + k0) & (uint32_t)(BATCH - 1);
    if (QUART) {
      // the rows take different slots, as their own masks make them do
      uint32_t kq[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) kq[q] = __builtin_amdgcn_readfirstlane((k + 5u * (uint32_t)q) & (uint32_t)(BATCH - 1));
      k = (row & 2u) ? ((row & 1u) ? kq[3] : kq[2]) : ((row & 1u) ? kq[1] : kq[0]);
    }
    const float4 A = s_a[k], B = s_b[k], Cc = s_c[k];
    float acc[9];
    const bool contrib = bwd_hit<PPT, true, false>(A, B, Cc, k, pxf, pyf, T, U, G0, G1, G2, lastj, acc, n_ok);
    const uint64_t cm = __ballot(contrib);
    if (cm) {
      if (QUART) {
        const float r = row_reduce9(acc);
        if (term_lane) atomicAdd(&s_acc[k][my_term], r);
      } else {
        const float r = wave_reduce9(acc);
        if (term_lane) atomicAdd(&s_acc[k][my_term], r);
      }
    }
    if ((it & 63u) == 63u) {
#pragma unroll
      for (int j = 0; j < PPT; ++j) {
        T[j] = 0.5f * s1;
        U[j] = 0.0f * s1;
      }
    }
  }
  __syncthreads();
  float keep = s_acc[threadIdx.x % BATCH][threadIdx.x % 9];
#pragma unroll
  for (int j = 0; j < PPT; ++j) keep += T[j] + U[j];
  if (keep == 1234.5f) sink[blockIdx.x] = keep;  // a result no run produces: keeps the loop live
}

// ---------------------------------------------------------------------------
// dispatch
// ---------------------------------------------------------------------------
static int fill_single_view(const nts_bins* bins, const nts_tiling* t, ViewTable& vt) {
  memset(&vt, 0, sizeof vt);
  vt.n = 1;
  vt.total_tiles = bins->num_tiles;
  vt.v[0].width = t->width;
  vt.v[0].height = t->height;
  vt.v[0].tiles_x = bins->tiles_x;
  vt.v[0].tiles_y = bins->tiles_y;
  vt.v[0].tile_base = 0;
  vt.v[0].proj_base = 0;
  vt.v[0].pix_base = 0;
  vt.order = bins->order;
  vt.grad_words = NTS_GRAD_WORDS;
  if (bins->tile_size != t->tile_size || bins->tiles_x != (t->width + t->tile_size - 1) / t->tile_size ||
      bins->tiles_y != (t->height + t->tile_size - 1) / t->tile_size)
    return nts_fail(NTS_E_INVALID, "tiling does not match the bins");
  return NTS_OK;
}

bool nts_blend_config_ok(uint32_t tile_size, uint32_t fwd_batch, uint32_t bwd_batch) {
  const bool ts = tile_size == 8 || tile_size == 16 || tile_size == 32;
  const bool fb = fwd_batch == 64 || fwd_batch == 128 || fwd_batch == 256 || fwd_batch == 512;
  const bool bb = bwd_batch == 64 || bwd_batch == 128 || bwd_batch == 256;
  return ts && fb && bb;
}

// Workgroups of `kernel` (block threads) resident on the device at once:
// the persistent scheduler's grid.  Looked up once per kernel and device.
template <typename KernelT>
static uint32_t resident_workgroups(KernelT kernel, int block) {
  static std::mutex mu;
  static std::map<std::pair<const void*, int>, uint32_t> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    return 256;
  }
  const auto key = std::make_pair(reinterpret_cast<const void*>(kernel), dev);
  std::lock_guard<std::mutex> g(mu);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int cus = 0, per_cu = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess) {
    (void)hipGetLastError();
    return 256;
  }
  const uint32_t r = (uint32_t)((per_cu > 0 ? per_cu : 1) * (cus > 0 ? cus : 1));
  cache[key] = r;
  return r;
}

template <int TS, int B, bool LOSS>
static int launch_fwd_kernel(nts_ctx* ctx, const char* name, const Projected* proj, const uint32_t* payloads,
                             const uint2* ranges, const ViewTable& vt, float4* out, float* final_t,
                             uint32_t* last_idx, const LossArgs& la) {
  constexpr int NT = Blend<TS, NTS_FWD_THREADS>::NT;
  if (vt.tickets) {
    const uint32_t r = resident_workgroups(k_blend_fwd<TS, B, LOSS, NTS_BLEND_PERSIST>, NT);
    const dim3 grid(r < vt.total_tiles ? r : vt.total_tiles);
    NTS_LAUNCH(ctx, name, (k_blend_fwd<TS, B, LOSS, NTS_BLEND_PERSIST>), grid, dim3(NT), 0, proj, payloads, ranges,
               vt, out, final_t, last_idx, la);
  } else if (vt.counts) {
    NTS_LAUNCH(ctx, name, (k_blend_fwd<TS, B, LOSS, NTS_BLEND_COUNT>), dim3(vt.total_tiles), dim3(NT), 0, proj,
               payloads, ranges, vt, out, final_t, last_idx, la);
  } else {
    NTS_LAUNCH(ctx, name, (k_blend_fwd<TS, B, LOSS, NTS_BLEND_GRID>), dim3(vt.total_tiles), dim3(NT), 0, proj,
               payloads, ranges, vt, out, final_t, last_idx, la);
  }
  return NTS_OK;
}

template <int TS, int B>
static int launch_fwd(nts_ctx* ctx, dim3 grid, bool loss, const Projected* proj, const uint32_t* payloads,
                       const uint2* ranges, const ViewTable& vt, float4* out, float* final_t, uint32_t* last_idx,
                       const LossArgs& la) {
  if (loss)
    return launch_fwd_kernel<TS, B, true>(ctx, "blend_fwd_loss", proj, payloads, ranges, vt, out, final_t, last_idx,
                                          la);
  return launch_fwd_kernel<TS, B, false>(ctx, "blend_fwd", proj, payloads, ranges, vt, out, final_t, last_idx, la);
}

template <int TS>
static int fwd_batch_switch(nts_ctx* ctx, uint32_t lds_batch, dim3 grid, bool loss, const Projected* proj,
                            const uint32_t* payloads, const uint2* ranges, const ViewTable& vt, float4* out,
                            float* final_t, uint32_t* last_idx, const LossArgs& la) {
  switch (lds_batch) {
    case 64: return launch_fwd<TS, 64>(ctx, grid, loss, proj, payloads, ranges, vt, out, final_t, last_idx, la);
    case 128: return launch_fwd<TS, 128>(ctx, grid, loss, proj, payloads, ranges, vt, out, final_t, last_idx, la);
    case 256: return launch_fwd<TS, 256>(ctx, grid, loss, proj, payloads, ranges, vt, out, final_t, last_idx, la);
    case 512: return launch_fwd<TS, 512>(ctx, grid, loss, proj, payloads, ranges, vt, out, final_t, last_idx, la);
    default: return nts_fail(NTS_E_INVALID, "forward LDS batch %u not in {64, 128, 256, 512}", lds_batch);
  }
}

int nts_blend_forward(nts_ctx* ctx, uint32_t tile_size, uint32_t lds_batch, const Projected* proj,
                      const uint32_t* payloads, const uint2* ranges, const ViewTable& vt, float4* out, float* final_t,
                      uint32_t* last_idx, const LossArgs* loss) {
  if (vt.total_tiles == 0) return NTS_OK;
  const dim3 grid(vt.total_tiles);
  LossArgs la{};
  if (loss) la = *loss;
  switch (tile_size) {
    case 8: return fwd_batch_switch<8>(ctx, lds_batch, grid, loss, proj, payloads, ranges, vt, out, final_t, last_idx, la);
    case 16: return fwd_batch_switch<16>(ctx, lds_batch, grid, loss, proj, payloads, ranges, vt, out, final_t, last_idx, la);
    case 32: return fwd_batch_switch<32>(ctx, lds_batch, grid, loss, proj, payloads, ranges, vt, out, final_t, last_idx, la);
    default: return nts_fail(NTS_E_INVALID, "tile size %u not in {8, 16, 32}", tile_size);
  }
}

template <int TS, int B, bool RECOMP>
static int launch_bwd_kernel(nts_ctx* ctx, const char* name, const Projected* proj, const uint32_t* payloads,
                             const uint2* ranges, const ViewTable& vt, const float4* fwd_out, const float* final_t,
                             const uint32_t* last_idx, const float4* dl_dc, float* pgrad) {
  constexpr int NT = Blend<TS, NTS_BWD_THREADS>::NT;
  if (vt.tickets) {
    const uint32_t r = resident_workgroups(k_blend_bwd<TS, B, RECOMP, NTS_BLEND_PERSIST>, NT);
    const dim3 grid(r < vt.total_tiles ? r : vt.total_tiles);
    NTS_LAUNCH(ctx, name, (k_blend_bwd<TS, B, RECOMP, NTS_BLEND_PERSIST>), grid, dim3(NT), 0, proj, payloads,
               ranges, vt, fwd_out, final_t, last_idx, dl_dc, pgrad);
  } else if (vt.counts) {
    NTS_LAUNCH(ctx, name, (k_blend_bwd<TS, B, RECOMP, NTS_BLEND_COUNT>), dim3(vt.total_tiles), dim3(NT), 0, proj,
               payloads, ranges, vt, fwd_out, final_t, last_idx, dl_dc, pgrad);
  } else {
    NTS_LAUNCH(ctx, name, (k_blend_bwd<TS, B, RECOMP, NTS_BLEND_GRID>), dim3(vt.total_tiles), dim3(NT), 0, proj,
               payloads, ranges, vt, fwd_out, final_t, last_idx, dl_dc, pgrad);
  }
  return NTS_OK;
}

template <int TS, int B>
static int launch_bwd(nts_ctx* ctx, dim3 grid, int mode, const Projected* proj, const uint32_t* payloads,
                       const uint2* ranges, const ViewTable& vt, const float4* fwd_out, const float* final_t,
                       const uint32_t* last_idx, const float4* dl_dc, float* pgrad) {
  if (mode == NTS_BWD_RECOMPUTE)
    return launch_bwd_kernel<TS, B, true>(ctx, "blend_bwd_recompute", proj, payloads, ranges, vt, fwd_out, final_t,
                                          last_idx, dl_dc, pgrad);
  return launch_bwd_kernel<TS, B, false>(ctx, "blend_bwd", proj, payloads, ranges, vt, fwd_out, final_t, last_idx,
                                         dl_dc, pgrad);
}

template <int TS>
static int bwd_batch_switch(nts_ctx* ctx, uint32_t lds_batch, dim3 grid, int mode, const Projected* proj,
                            const uint32_t* payloads, const uint2* ranges, const ViewTable& vt, const float4* fwd_out,
                            const float* final_t, const uint32_t* last_idx, const float4* dl_dc,
                            float* pgrad) {
  switch (lds_batch) {
    case 64: return launch_bwd<TS, 64>(ctx, grid, mode, proj, payloads, ranges, vt, fwd_out, final_t, last_idx, dl_dc, pgrad);
    case 128: return launch_bwd<TS, 128>(ctx, grid, mode, proj, payloads, ranges, vt, fwd_out, final_t, last_idx, dl_dc, pgrad);
    case 256: return launch_bwd<TS, 256>(ctx, grid, mode, proj, payloads, ranges, vt, fwd_out, final_t, last_idx, dl_dc, pgrad);
    default: return nts_fail(NTS_E_INVALID, "backward LDS batch %u not in {64, 128, 256}", lds_batch);
  }
}

int nts_blend_backward(nts_ctx* ctx, uint32_t tile_size, uint32_t lds_batch, int mode, const Projected* proj,
                       const uint32_t* payloads, const uint2* ranges, const ViewTable& vt, const float4* fwd_out,
                       const float* final_t, const uint32_t* last_idx, const float4* dl_dc, float* pgrad) {
  if (vt.total_tiles == 0) return NTS_OK;
  const dim3 grid(vt.total_tiles);
  switch (tile_size) {
    case 8: return bwd_batch_switch<8>(ctx, lds_batch, grid, mode, proj, payloads, ranges, vt, fwd_out, final_t, last_idx, dl_dc, pgrad);
    case 16: return bwd_batch_switch<16>(ctx, lds_batch, grid, mode, proj, payloads, ranges, vt, fwd_out, final_t, last_idx, dl_dc, pgrad);
    case 32: return bwd_batch_switch<32>(ctx, lds_batch, grid, mode, proj, payloads, ranges, vt, fwd_out, final_t, last_idx, dl_dc, pgrad);
    default: return nts_fail(NTS_E_INVALID, "tile size %u not in {8, 16, 32}", tile_size);
  }
}

template <int TS, int B>
static int probe_bwd_mix(nts_ctx* ctx, double* hits_per_s) {
  constexpr int NT = Blend<TS, NTS_BWD_THREADS>::NT;
  // as many workgroups per CU as the backward kernel holds: LDS padding
  // caps the probe (whose VGPRs would allow more) at the kernel's count
  int dev = 0, cus = 0, per_cu = 0, lds_cu = 0;
  NTS_HIP_TRY(hipGetDevice(&dev));
  NTS_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  NTS_HIP_TRY(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
  NTS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_blend_bwd<TS, B, true, NTS_BLEND_GRID>, NT, 0));
  if (per_cu < 1 || cus < 1 || lds_cu < 1) return nts_fail(NTS_E_HIP, "occupancy query");
  hipFuncAttributes fa;
  NTS_HIP_TRY(hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(k_probe_bwd_mix<TS, B>)));
  const size_t per_wg = (size_t)lds_cu / (size_t)per_cu;
  const size_t pad = per_wg > fa.sharedSizeBytes + 256 ? per_wg - fa.sharedSizeBytes - 256 : 0;
  const uint32_t grid = (uint32_t)(per_cu * cus);
  const uint32_t hits = 4096;
  void* sink = nullptr;
  int rc = nts_dev_alloc(ctx, &sink, (uint64_t)grid * 4 + 256);
  if (rc) return rc;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double best = 1e30;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = nts_fail(NTS_E_HIP, "events");
  for (int r = 0; r < 4 && rc == NTS_OK; ++r) {
    (void)hipEventRecord(e0, ctx->stream);
    hipLaunchKernelGGL((k_probe_bwd_mix<TS, B>), dim3(grid), dim3(NT), pad, ctx->stream, hits, 1.0f,
                       (float*)sink);
    (void)hipEventRecord(e1, ctx->stream);
    if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
      rc = nts_fail(NTS_E_HIP, "blend mix probe failed");
      break;
    }
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (r > 0 && ms < best) best = ms;  // the first launch warms up
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  nts_dev_free(ctx, sink, (uint64_t)grid * 4 + 256);
  if (rc) return rc;
  *hits_per_s = (double)grid * (NT / 64) * hits / (best * 1e-3);
  return NTS_OK;
}

extern "C" int nts_probe_blend_mix(nts_ctx* ctx, uint32_t tile_size, uint32_t bwd_batch, double* bwd_hits_per_s) {
  NTS_CHECK_ARG(ctx && bwd_hits_per_s);
  switch (tile_size * 1000 + bwd_batch) {
    case 8064: return probe_bwd_mix<8, 64>(ctx, bwd_hits_per_s);
    case 8128: return probe_bwd_mix<8, 128>(ctx, bwd_hits_per_s);
    case 8256: return probe_bwd_mix<8, 256>(ctx, bwd_hits_per_s);
    case 16064: return probe_bwd_mix<16, 64>(ctx, bwd_hits_per_s);
    case 16128: return probe_bwd_mix<16, 128>(ctx, bwd_hits_per_s);
    case 16256: return probe_bwd_mix<16, 256>(ctx, bwd_hits_per_s);
    case 32064: return probe_bwd_mix<32, 64>(ctx, bwd_hits_per_s);
    case 32128: return probe_bwd_mix<32, 128>(ctx, bwd_hits_per_s);
    case 32256: return probe_bwd_mix<32, 256>(ctx, bwd_hits_per_s);
    default: return nts_fail(NTS_E_INVALID, "tile %u / backward batch %u", tile_size, bwd_batch);
  }
}

extern "C" int nts_composite_ex(nts_ctx* ctx, const nts_buffer* projected, const nts_bins* bins, nts_buffer* image,
                                const nts_tiling* tiling, nts_buffer* final_t, nts_buffer* last_idx) {
  NTS_CHECK_ARG(ctx && projected && bins && image && tiling);
  const uint64_t np = (uint64_t)tiling->width * tiling->height;
  NTS_CHECK_ARG(image->bytes >= np * 16);
  NTS_CHECK_ARG(!final_t || final_t->bytes >= np * 4);
  NTS_CHECK_ARG(!last_idx || last_idx->bytes >= np * 4);
  ViewTable vt;
  int rc = fill_single_view(bins, tiling, vt);
  if (rc) return rc;
  return nts_blend_forward(ctx, tiling->tile_size, NTS_DEFAULT_FWD_BATCH, (const Projected*)projected->ptr,
                           bins->payloads, bins->ranges, vt,
                           (float4*)image->ptr, final_t ? (float*)final_t->ptr : nullptr,
                           last_idx ? (uint32_t*)last_idx->ptr : nullptr, nullptr);
}

extern "C" int nts_composite(nts_ctx* ctx, const nts_buffer* projected, const nts_bins* bins, nts_buffer* image,
                             const nts_tiling* tiling) {
  return nts_composite_ex(ctx, projected, bins, image, tiling, nullptr, nullptr);
}

extern "C" int nts_composite_backward_ex(nts_ctx* ctx, const nts_buffer* projected, const nts_bins* bins,
                                         const nts_buffer* forward_out, const nts_buffer* final_t,
                                         const nts_buffer* last_idx, const nts_buffer* dl_dc,
                                         nts_buffer* projected_grad, const nts_tiling* tiling, int mode) {
  NTS_CHECK_ARG(ctx && projected && bins && dl_dc && projected_grad && tiling);
  NTS_CHECK_ARG(mode == NTS_BWD_REF_PARITY || mode == NTS_BWD_RECOMPUTE);
  const uint64_t np = (uint64_t)tiling->width * tiling->height;
  NTS_CHECK_ARG(dl_dc->bytes >= np * 16);
  if (mode == NTS_BWD_REF_PARITY) NTS_CHECK_ARG(forward_out && forward_out->bytes >= np * 16);
  if (mode == NTS_BWD_RECOMPUTE)
    NTS_CHECK_ARG(final_t && last_idx && final_t->bytes >= np * 4 && last_idx->bytes >= np * 4);
  NTS_CHECK_ARG(projected_grad->bytes >= projected->bytes);
  ViewTable vt;
  int rc = fill_single_view(bins, tiling, vt);
  if (rc) return rc;
  return nts_blend_backward(ctx, tiling->tile_size, nts_default_bwd_batch(tiling->tile_size), mode, (const Projected*)projected->ptr,
                            bins->payloads,
                            bins->ranges, vt, forward_out ? (const float4*)forward_out->ptr : nullptr,
                            final_t ? (const float*)final_t->ptr : nullptr,
                            last_idx ? (const uint32_t*)last_idx->ptr : nullptr, (const float4*)dl_dc->ptr,
                            (float*)projected_grad->ptr);
}

extern "C" int nts_composite_backward(nts_ctx* ctx, const nts_buffer* projected, const nts_bins* bins,
                                      const nts_buffer* forward_out, const nts_buffer* dl_dc,
                                      nts_buffer* projected_grad, const nts_tiling* tiling) {
  return nts_composite_backward_ex(ctx, projected, bins, forward_out, nullptr, nullptr, dl_dc, projected_grad, tiling,
                                   NTS_BWD_REF_PARITY);
}

// ---------------------------------------------------------------------------
// loss (L1 over 3WH with sign(0) = 0, or MSE as loss.rs:163-179)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_loss(const float4* __restrict__ pred, const float4* __restrict__ target,
                                              float4* __restrict__ dl_dc, uint64_t np, float inv_n, int kind,
                                              double* __restrict__ loss_acc) {
  __shared__ double s_part[4];
  double local = 0.0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < np; i += stride) {
    const float4 p = pred[i], t = target[i];
    const float d0 = p.x - t.x, d1 = p.y - t.y, d2 = p.z - t.z;
    float4 g;
    if (kind == NTS_LOSS_MSE) {
      local += (double)(d0 * d0 + d1 * d1 + d2 * d2);
      g = make_float4(d0 * inv_n, d1 * inv_n, d2 * inv_n, 0.0f);
    } else {
      local += (double)fabsf(d0) + (double)fabsf(d1) + (double)fabsf(d2);
      g = make_float4(d0 > 0.f ? inv_n : (d0 < 0.f ? -inv_n : 0.f), d1 > 0.f ? inv_n : (d1 < 0.f ? -inv_n : 0.f),
                      d2 > 0.f ? inv_n : (d2 < 0.f ? -inv_n : 0.f), 0.0f);
    }
    dl_dc[i] = g;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) local += __shfl_xor(local, off, 64);
  if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = local;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(loss_acc, s_part[0] + s_part[1] + s_part[2] + s_part[3]);
}

extern "C" int nts_loss_grad(nts_ctx* ctx, const nts_buffer* image, const nts_buffer* target, nts_buffer* dl_dc,
                             uint32_t width, uint32_t height, int kind, double* loss) {
  NTS_CHECK_ARG(ctx && image && target && dl_dc);
  NTS_CHECK_ARG(kind == NTS_LOSS_L1 || kind == NTS_LOSS_MSE);
  const uint64_t np = (uint64_t)width * height;
  NTS_CHECK_ARG(image->bytes >= np * 16 && target->bytes >= np * 16 && dl_dc->bytes >= np * 16);
  if (np == 0) return NTS_OK;
  void* scr;
  int rc = nts_scratch(ctx, 256, &scr);
  if (rc) return rc;
  rc = nts_fill_u32(ctx, scr, 0, 2);
  if (rc) return rc;
  const float n_scalars = (float)width * (float)height * 3.0f;
  const float inv_n = kind == NTS_LOSS_MSE ? 2.0f / n_scalars : 1.0f / n_scalars;
  uint32_t grid = nts_div_up(np, 256);
  if (grid > 2048) grid = 2048;
  NTS_LAUNCH(ctx, "loss", k_loss, dim3(grid), dim3(256), 0, (const float4*)image->ptr, (const float4*)target->ptr,
             (float4*)dl_dc->ptr, np, inv_n, kind, (double*)scr);
  double acc = 0.0;
  NTS_HIP_TRY(hipMemcpyAsync(&acc, scr, 8, hipMemcpyDeviceToHost, ctx->stream));
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (loss) *loss = acc / ((double)np * 3.0);
  return NTS_OK;
}

// ---------------------------------------------------------------------------
// tonemap.wgsl:27-41 (Reinhard, sRGB, rgba8; alpha = coverage)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t to_u8(float s) {
  s = nts_minf(nts_maxf(s, 0.0f), 1.0f) * 255.0f;
  return (uint32_t)__builtin_rintf(s);
}

__global__ void __launch_bounds__(256) k_tonemap(const float4* __restrict__ img, uint32_t* __restrict__ out,
                                                 uint64_t np) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const float4 p = img[i];
  const float c[3] = {p.x, p.y, p.z};
  uint32_t o = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float tm = c[k] / (1.0f + nts_maxf(c[k], 0.0f));
    tm = nts_minf(nts_maxf(tm, 0.0f), 1.0f);
    const float s = tm > 0.0031308f ? 1.055f * nts_srgb_pow(tm) - 0.055f : tm * 12.92f;
    o |= to_u8(s) << (8 * k);
  }
  o |= to_u8(p.w) << 24;
  out[i] = o;
}

int nts_tonemap_dev(nts_ctx* ctx, const float4* image, uint32_t* rgba8, uint64_t np) {
  if (np == 0) return NTS_OK;
  NTS_LAUNCH(ctx, "tonemap", k_tonemap, dim3(nts_div_up(np, 256)), dim3(256), 0, image, rgba8, np);
  return NTS_OK;
}

extern "C" int nts_tonemap(nts_ctx* ctx, const nts_buffer* image, nts_buffer* rgba8, uint32_t width, uint32_t height) {
  NTS_CHECK_ARG(ctx && image && rgba8);
  const uint64_t np = (uint64_t)width * height;
  NTS_CHECK_ARG(image->bytes >= np * 16 && rgba8->bytes >= np * 4);
  if (np == 0) return NTS_OK;
  NTS_LAUNCH(ctx, "tonemap", k_tonemap, dim3(nts_div_up(np, 256)), dim3(256), 0, (const float4*)image->ptr,
             (uint32_t*)rgba8->ptr, np);
  return NTS_OK;
}
