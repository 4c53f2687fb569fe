// Combined MSE + SSIM loss on the device (loss.rs:23-415): per channel,
// SSIM over an 11-tap separable Gaussian (sigma 1.5, normalised,
// reflect-101 borders), C1 = 1e-4, C2 = 9e-4, loss 1 - mean(SSIM) averaged
// over the three channels, the gradient through the transpose of the blur.
//
// Two LDS-tiled passes over 64 x 16 pixel tiles of every view of a batch:
//   moments  loads the tile's window (5-pixel reflected halo) of prediction
//            and target once, then per channel blurs x, y, x^2, y^2, xy
//            horizontally into LDS and vertically into registers, forms the
//            SSIM map (summed in float64 per view and channel, with the MSE
//            sum) and its partials P = dS/dmu_x, Q = dS/dsigma_x^2,
//            R = dS/dsigma_xy and the centring products mu_x Q, mu_y R
//            (15 planes, 60 B per pixel);
//   grad     per channel loads those 5 planes over the halo, applies the
//            vertical then horizontal transpose of the blur in LDS, and
//            combines dL/dx with the MSE gradient into dL/dC.
// The transpose is gathered in the scatter form's summation order, and
// every float32 expression keeps the host restatement's operation order, so
// dL/dC matches it bit for bit; the float64 sums differ only in order.
#include "nts_raster.h"

#define SSIM_TAPS 11
#define SSIM_R 5
#define SSIM_TW 64
#define SSIM_TH 16
#define SSIM_WW (SSIM_TW + 2 * SSIM_R)
#define SSIM_WH (SSIM_TH + 2 * SSIM_R)
#define SSIM_NT 256
// window cells per thread, rounded up to even
#define WIN_ITEMS ((((SSIM_WH * SSIM_WW + SSIM_NT - 1) / SSIM_NT) + 1) & ~1)

// reflect-101 over a period of 2(n - 1) (loss.rs:57-67); n >= 6 needs at
// most one reflection for offsets up to 5
__device__ __forceinline__ int refl101(int i, int n) {
  if (n >= 6) {
    if (i < 0) i = -i;
    if (i >= n) i = 2 * (n - 1) - i;
    return i;
  }
  if (n == 1) return 0;
  const int period = 2 * (n - 1);
  int x = i % period;
  if (x < 0) x += period;
  if (x >= n) x = period - x;
  return x;
}

// Transpose of one blur pass along an axis of n samples, for source index
// i: value(j) is the pass's output-side value at index j.  Contributions are
// summed in the order of the scatter form (output index j ascending, then
// tap t ascending), so the result is bit-identical to scattering.  For
// n >= 6 the taps of j that land on i are t = 5 - i - j (through the low
// reflection, i != 0), t = i - j + 5 (direct) and t = 2(n - 1) - i - j + 5
// (through the high reflection, i != n - 1), in that order, and only
// j in [i - 5, i + 5] has any; for 6 <= i <= n - 7 only the direct tap
// exists, with static tap indices.  Below 6 samples every j in [0, n) lies
// within 4 of i.  kl: the taps in LDS, for the data-dependent tap indices
// of the border.
template <typename F>
__device__ __forceinline__ float adj_gather(int i, int n, const SsimKernel& K, const float* kl, F value) {
  float acc = 0.0f;
  if (i >= SSIM_R + 1 && i <= n - 2 - SSIM_R) {
#pragma unroll
    for (int d = -SSIM_R; d <= SSIM_R; ++d) acc = acc + value(i + d) * K.k[SSIM_R - d];
    return acc;
  }
  if (n >= 6) {
    const int lo = i - SSIM_R < 0 ? 0 : i - SSIM_R;
    const int hi = i + SSIM_R > n - 1 ? n - 1 : i + SSIM_R;
    for (int j = lo; j <= hi; ++j) {
      const float v = value(j);
      const int t2 = SSIM_R - i - j;
      if (i != 0 && t2 >= 0) acc = acc + v * kl[t2];
      acc = acc + v * kl[i - j + SSIM_R];
      const int t3 = 2 * (n - 1) - i - j + SSIM_R;
      if (i != n - 1 && t3 < SSIM_TAPS) acc = acc + v * kl[t3];
    }
    return acc;
  }
  for (int j = 0; j < n; ++j) {
    const float v = value(j);
    for (int t = 0; t < SSIM_TAPS; ++t)
      if (refl101(j + t - SSIM_R, n) == i) acc = acc + v * kl[t];
  }
  return acc;
}

struct TileLoc {
  uint32_t view, w, h;
  int x0, y0;
  uint64_t pix_base;
};

__device__ __forceinline__ TileLoc locate(const SsimViews& sv, uint32_t b) {
  uint32_t v = 0;
  while (v + 1 < sv.n && b >= sv.tile_base[v + 1]) ++v;
  const uint32_t t = b - sv.tile_base[v];
  TileLoc L;
  L.view = v;
  L.w = sv.w[v];
  L.h = sv.h[v];
  L.x0 = (int)(t % sv.tiles_x[v]) * SSIM_TW;
  L.y0 = (int)(t / sv.tiles_x[v]) * SSIM_TH;
  L.pix_base = sv.pix_base[v];
  return L;
}

__global__ void __launch_bounds__(SSIM_NT) k_ssim_moments(const float4* __restrict__ pred,
                                                          const float4* __restrict__ target, SsimViews sv,
                                                          SsimKernel K, float* __restrict__ T5, uint64_t plane,
                                                          double* __restrict__ sums) {
  __shared__ float s_in[6][SSIM_WH][SSIM_WW];   // x r g b, y r g b
  __shared__ float s_h[SSIM_WH][SSIM_TW];       // horizontal blur of one moment plane
  __shared__ double s_red[4][SSIM_NT / 64];
  const TileLoc L = locate(sv, blockIdx.x);
  const int w = (int)L.w, h = (int)L.h;
  const int tid = threadIdx.x;
  // the window in two halves of statically unrolled loads, so each thread
  // has up to 8 loads in flight
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    float4 pv[WIN_ITEMS / 2], gv[WIN_ITEMS / 2];
#pragma unroll
    for (int u = 0; u < WIN_ITEMS / 2; ++u) {
      const int k = tid + (half * (WIN_ITEMS / 2) + u) * SSIM_NT;
      const int ly = k / SSIM_WW, lx = k - ly * SSIM_WW;
      const int ry = L.y0 - SSIM_R + ly, rx = L.x0 - SSIM_R + lx;
      // cells beyond 5 past the last row or column feed no output pixel
      const int gy = ry < h + SSIM_R ? refl101(ry, h) : 0;
      const int gx = rx < w + SSIM_R ? refl101(rx, w) : 0;
      const uint64_t idx = L.pix_base + (uint64_t)(k < SSIM_WH * SSIM_WW ? gy : 0) * w + gx;
      pv[u] = pred[idx];
      gv[u] = target[idx];
    }
#pragma unroll
    for (int u = 0; u < WIN_ITEMS / 2; ++u) {
      const int k = tid + (half * (WIN_ITEMS / 2) + u) * SSIM_NT;
      if (k >= SSIM_WH * SSIM_WW) continue;
      const int ly = k / SSIM_WW, lx = k - ly * SSIM_WW;
      s_in[0][ly][lx] = pv[u].x;
      s_in[1][ly][lx] = pv[u].y;
      s_in[2][ly][lx] = pv[u].z;
      s_in[3][ly][lx] = gv[u].x;
      s_in[4][ly][lx] = gv[u].y;
      s_in[5][ly][lx] = gv[u].z;
    }
  }
  __syncthreads();
  const int lx = tid & (SSIM_TW - 1), ly0 = tid / SSIM_TW;
  constexpr int RS = SSIM_NT / SSIM_TW;  // rows per sweep of the block
  constexpr int PER = SSIM_TH / RS;
  double acc_s[3] = {0.0, 0.0, 0.0};
#pragma unroll 1
  for (int c = 0; c < 3; ++c) {
    float m[PER][5];
    // one moment plane at a time: x, y, x^2, y^2, xy blurred horizontally
    // into LDS over the halo rows, then vertically into registers
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      for (int k = tid; k < SSIM_WH * SSIM_TW; k += SSIM_NT) {
        const int ry = k / SSIM_TW, rx = k - ry * SSIM_TW;
        float acc = 0.0f;
#pragma unroll
        for (int t = 0; t < SSIM_TAPS; ++t) {
          const float a = s_in[c][ry][rx + t], b = s_in[3 + c][ry][rx + t];
          const float f = q == 0 ? a : (q == 1 ? b : (q == 2 ? a * a : (q == 3 ? b * b : a * b)));
          acc = acc + f * K.k[t];
        }
        s_h[ry][rx] = acc;
      }
      __syncthreads();
#pragma unroll
      for (int r = 0; r < PER; ++r) {
        const int ly = ly0 + r * RS;
        float acc = 0.0f;
#pragma unroll
        for (int t = 0; t < SSIM_TAPS; ++t) acc = acc + s_h[ly + t][lx] * K.k[t];
        m[r][q] = acc;
      }
      __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < PER; ++r) {
      const int ly = ly0 + r * RS;
      const int ox = L.x0 + lx, oy = L.y0 + ly;
      if (ox >= w || oy >= h) continue;
      const float mux = m[r][0], muy = m[r][1];
      const float mux2 = mux * mux, muy2 = muy * muy;
      const float sxx = m[r][2] - mux2, syy = m[r][3] - muy2, sxy = m[r][4] - mux * muy;
      const float av = 2.0f * mux * muy + 1e-4f;
      const float bv = 2.0f * sxy + 9e-4f;
      const float cv = mux2 + muy2 + 1e-4f;
      const float dv = sxx + syy + 9e-4f;
      acc_s[c] += (double)((av * bv) / (cv * dv));
      const float cd = cv * dv;
      const float P = 2.0f * bv * (muy * cv - mux * av) / (cv * cd);
      const float Q = -av * bv / (cv * dv * dv);
      const float R = 2.0f * av / cd;
      float* base = T5 + (uint64_t)(c * 5) * plane + L.pix_base + (uint64_t)oy * w + ox;
      base[0] = P;
      base[plane] = Q;
      base[2 * plane] = R;
      base[3 * plane] = mux * Q;
      base[4 * plane] = muy * R;
    }
  }
  double acc_m = 0.0;
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int ly = ly0 + r * RS;
    if (L.x0 + lx >= w || L.y0 + ly >= h) continue;
    const float d0 = s_in[0][ly + SSIM_R][lx + SSIM_R] - s_in[3][ly + SSIM_R][lx + SSIM_R];
    const float d1 = s_in[1][ly + SSIM_R][lx + SSIM_R] - s_in[4][ly + SSIM_R][lx + SSIM_R];
    const float d2 = s_in[2][ly + SSIM_R][lx + SSIM_R] - s_in[5][ly + SSIM_R][lx + SSIM_R];
    acc_m += (double)(d0 * d0 + d1 * d1 + d2 * d2);
  }
  double part[4] = {acc_s[0], acc_s[1], acc_s[2], acc_m};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double v = part[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((tid & 63) == 0) s_red[q][tid >> 6] = v;
  }
  __syncthreads();
  if (tid < 4) {
    double v = 0.0;
#pragma unroll
    for (int k = 0; k < SSIM_NT / 64; ++k) v += s_red[tid][k];
    unsafeAtomicAdd(sums + 4 * L.view + tid, v);
  }
}

__global__ void __launch_bounds__(SSIM_NT) k_ssim_grad(const float* __restrict__ T5, uint64_t plane,
                                                       const float4* __restrict__ pred,
                                                       const float4* __restrict__ target, SsimViews sv,
                                                       SsimKernel K, float lam, float4* __restrict__ dl_dc) {
  __shared__ float s_t[SSIM_WH][SSIM_WW];   // one plane over the halo
  __shared__ float s_v[SSIM_TH][SSIM_WW];   // after the vertical transpose
  __shared__ float s_k[SSIM_TAPS];
  const TileLoc L = locate(sv, blockIdx.x);
  const int w = (int)L.w, h = (int)L.h;
  const int tid = threadIdx.x;
  if (tid < SSIM_TAPS) s_k[tid] = K.k[tid];  // published by the first barrier below
  const int lx = tid & (SSIM_TW - 1), ly0 = tid / SSIM_TW;
  constexpr int RS = SSIM_NT / SSIM_TW;
  constexpr int PER = SSIM_TH / RS;
  float out[PER][3];
  const float inv_n = 1.0f / (float)((uint64_t)w * h);
  const float inv_mse = sv.inv_mse[L.view];
  // Window cell u of this thread: k = tid + u * 256.  Cells outside the
  // view are never read by the transposes; they load a valid dummy.
  uint32_t cell_off[WIN_ITEMS];
#pragma unroll
  for (int u = 0; u < WIN_ITEMS; ++u) {
    const int k = tid + u * SSIM_NT;
    const int ly = k / SSIM_WW, lxx = k - ly * SSIM_WW;
    const int gy = L.y0 - SSIM_R + ly, gx = L.x0 - SSIM_R + lxx;
    const bool in = k < SSIM_WH * SSIM_WW && gy >= 0 && gy < h && gx >= 0 && gx < w;
    cell_off[u] = in ? (uint32_t)(gy * w + gx) : 0u;
  }
  float pre[WIN_ITEMS];
  auto load_plane = [&](int p) {
    const float* src = T5 + (uint64_t)p * plane + L.pix_base;
#pragma unroll
    for (int u = 0; u < WIN_ITEMS; ++u) pre[u] = src[cell_off[u]];
  };
  load_plane(0);
#pragma unroll 1
  for (int c = 0; c < 3; ++c) {
    float a[PER][5];
    // one plane at a time: the vertical transpose over the halo columns into
    // LDS, then the horizontal transpose at the block's own pixels; the next
    // plane's loads are in flight meanwhile
#pragma unroll
    for (int q = 0; q < 5; ++q) {
#pragma unroll
      for (int u = 0; u < WIN_ITEMS; ++u) {
        const int k = tid + u * SSIM_NT;
        if (k < SSIM_WH * SSIM_WW) (&s_t[0][0])[k] = pre[u];
      }
      __syncthreads();
      if (c * 5 + q + 1 < 15) load_plane(c * 5 + q + 1);
      for (int k = tid; k < SSIM_TH * SSIM_WW; k += SSIM_NT) {
        const int ly = k / SSIM_WW, lxx = k - ly * SSIM_WW;
        const int gy = L.y0 + ly, gx = L.x0 - SSIM_R + lxx;
        if (gy >= h || gx < 0 || gx >= w) continue;
        s_v[ly][lxx] = adj_gather(gy, h, K, s_k, [&](int j) { return s_t[j - (L.y0 - SSIM_R)][lxx]; });
      }
      __syncthreads();
#pragma unroll
      for (int r = 0; r < PER; ++r) {
        const int ly = ly0 + r * RS;
        const int ox = L.x0 + lx;
        a[r][q] = ox < w && L.y0 + ly < h
                      ? adj_gather(ox, w, K, s_k, [&](int j) { return s_v[ly][j - (L.x0 - SSIM_R)]; })
                      : 0.0f;
      }
      // the next plane's stores to s_t follow this plane's last read of it
      // (before the barrier above); its writes to s_v follow the next
      // barrier, which every thread reaches after these reads of s_v
    }
#pragma unroll
    for (int r = 0; r < PER; ++r) {
      const int ox = L.x0 + lx, oy = L.y0 + ly0 + r * RS;
      if (ox >= w || oy >= h) continue;
      const uint64_t o = L.pix_base + (uint64_t)oy * w + ox;
      const float4 xv = pred[o], yv = target[o];
      const float xs = c == 0 ? xv.x : (c == 1 ? xv.y : xv.z);
      const float ys = c == 0 ? yv.x : (c == 1 ? yv.y : yv.z);
      const float d = a[r][0] + 2.0f * (xs * a[r][1] - a[r][3]) + (ys * a[r][2] - a[r][4]);
      const float gs = (-inv_n * d) / 3.0f;    // ssim_loss_grad's channel average
      const float gm = (xs - ys) * inv_mse;    // 2 (x - y) / (3WH)
      out[r][c] = gm * (1.0f - lam) + gs * lam;
    }
  }
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int ox = L.x0 + lx, oy = L.y0 + ly0 + r * RS;
    if (ox >= w || oy >= h) continue;
    dl_dc[L.pix_base + (uint64_t)oy * w + ox] = make_float4(out[r][0], out[r][1], out[r][2], 0.0f);
  }
}

// Per-view loss (loss.rs:404-409): (1 - lam) MSE + lam (1 - SSIM), or the
// SSIM term alone at lam = 1.
__device__ __host__ __forceinline__ void ssim_totals(const double* s, uint64_t np, float lam, float* total,
                                                     float* mse, float* ssim) {
  const float m = (float)(s[3] / ((double)np * 3.0));
  float l[3];
  for (int c = 0; c < 3; ++c) l[c] = 1.0f - (float)(s[c] / (double)np);
  const float sl = (l[0] + l[1] + l[2]) / 3.0f;
  *mse = lam == 1.0f ? 0.0f : m;
  *ssim = sl;
  *total = lam == 1.0f ? sl : (1.0f - lam) * m + lam * sl;
}

__global__ void k_ssim_finalize(const double* __restrict__ sums, SsimViews sv, float lam,
                                double* __restrict__ losses) {
  const uint32_t v = threadIdx.x;
  if (v >= sv.n) return;
  float t, m, s;
  ssim_totals(sums + 4 * v, (uint64_t)sv.w[v] * sv.h[v], lam, &t, &m, &s);
  losses[v] = (double)t;
}

static SsimKernel ssim_kernel() {
  SsimKernel K;
  const float centre = (float)(SSIM_TAPS - 1) * 0.5f;
  const float two_sigma_sq = 2.0f * 1.5f * 1.5f;
  float sum = 0.0f;
  for (int i = 0; i < SSIM_TAPS; ++i) {
    const float d = (float)i - centre;
    K.k[i] = nts_expf(-d * d / two_sigma_sq);
    sum += K.k[i];
  }
  for (int i = 0; i < SSIM_TAPS; ++i) K.k[i] = K.k[i] / sum;
  return K;
}

int nts_ssim_views(uint32_t n_views, const uint32_t* widths, const uint32_t* heights, const uint64_t* pix_base,
                   SsimViews* out) {
  if (n_views == 0 || n_views > NTS_MAX_VIEWS) return nts_fail(NTS_E_INVALID, "SSIM over %u views", n_views);
  memset(out, 0, sizeof *out);
  out->n = n_views;
  uint64_t tiles = 0;
  for (uint32_t v = 0; v < n_views; ++v) {
    const uint32_t w = widths[v], h = heights[v];
    if (w == 0 || h == 0) return nts_fail(NTS_E_INVALID, "SSIM of an empty image (view %u)", v);
    out->w[v] = w;
    out->h[v] = h;
    out->tiles_x[v] = (w + SSIM_TW - 1) / SSIM_TW;
    out->tile_base[v] = (uint32_t)tiles;
    out->pix_base[v] = pix_base[v];
    out->inv_mse[v] = 2.0f / ((float)w * (float)h * 3.0f);
    tiles += (uint64_t)out->tiles_x[v] * ((h + SSIM_TH - 1) / SSIM_TH);
  }
  if (tiles >= (1ull << 31)) return nts_fail(NTS_E_CAPACITY, "SSIM over %llu tiles", (unsigned long long)tiles);
  out->tile_base[n_views] = (uint32_t)tiles;
  return NTS_OK;
}

int nts_ssim_launch(nts_ctx* ctx, const SsimViews& sv, const float4* pred, const float4* target, float* planes,
                    uint64_t plane, double* sums, float4* dl_dc, float lam, double* losses) {
  int rc = nts_fill_u32(ctx, sums, 0, 8 * (uint64_t)sv.n);
  if (rc) return rc;
  const SsimKernel K = ssim_kernel();
  const dim3 grid(sv.tile_base[sv.n]);
  NTS_LAUNCH(ctx, "ssim_moments", k_ssim_moments, grid, dim3(SSIM_NT), 0, pred, target, sv, K, planes, plane, sums);
  NTS_LAUNCH(ctx, "ssim_grad", k_ssim_grad, grid, dim3(SSIM_NT), 0, (const float*)planes, plane, pred, target, sv, K,
             lam, dl_dc);
  if (losses)
    NTS_LAUNCH(ctx, "ssim_finalize", k_ssim_finalize, dim3(1), dim3(64), 0, (const double*)sums, sv, lam, losses);
  return NTS_OK;
}

extern "C" int nts_loss_grad_ssim(nts_ctx* ctx, const nts_buffer* image, const nts_buffer* target,
                                  nts_buffer* dl_dc, uint32_t width, uint32_t height, float lambda, double* total,
                                  double* mse, double* ssim) {
  NTS_CHECK_ARG(ctx && image && target && dl_dc);
  const uint64_t n = (uint64_t)width * height;
  NTS_CHECK_ARG(image->bytes >= n * 16 && target->bytes >= n * 16 && dl_dc->bytes >= n * 16);
  const float lam = lambda < 0.0f ? 0.0f : (lambda > 1.0f ? 1.0f : lambda);
  if (lam == 0.0f) {
    // loss.rs:398-402: the plain MSE gradient, SSIM skipped
    double m = 0.0;
    const int rc0 = nts_loss_grad(ctx, image, target, dl_dc, width, height, NTS_LOSS_MSE, &m);
    if (rc0) return rc0;
    if (total) *total = (double)(float)m;
    if (mse) *mse = (double)(float)m;
    if (ssim) *ssim = 0.0;
    return NTS_OK;
  }
  SsimViews sv;
  const uint64_t zero = 0;
  int rc = nts_ssim_views(1, &width, &height, &zero, &sv);
  if (rc) return rc;
  void* scr;
  rc = nts_scratch(ctx, 256 + 15 * n * 4, &scr);
  if (rc) return rc;
  double* sums = (double*)scr;
  rc = nts_ssim_launch(ctx, sv, (const float4*)image->ptr, (const float4*)target->ptr, (float*)((char*)scr + 256),
                       n, sums, (float4*)dl_dc->ptr, lam, nullptr);
  if (rc) return rc;
  double h_sums[4];
  NTS_HIP_TRY(hipMemcpyAsync(h_sums, sums, 32, hipMemcpyDeviceToHost, ctx->stream));
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  float t, m, s;
  ssim_totals(h_sums, n, lam, &t, &m, &s);
  if (total) *total = t;
  if (mse) *mse = m;
  if (ssim) *ssim = s;
  return NTS_OK;
}
