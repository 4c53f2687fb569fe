// Combined MSE + SSIM loss on the device (loss.rs:23-415): per channel,
// SSIM over an 11-tap separable Gaussian (sigma 1.5, normalised, reflect-101
// borders), C1 = 1e-4, C2 = 9e-4, loss 1 - mean(SSIM) averaged over the
// three channels, the gradient through the transpose of the blur.
//
// Four passes over planar intermediates (3 channels x 5 planes):
//   h      horizontal blur of x, y, x^2, y^2, xy;
//   v      vertical blur, the SSIM map (summed in float64 per channel) and
//          its partials P = dS/dmu_x, Q = dS/dsigma_x^2, R = dS/dsigma_xy with
//          the centring products mu_x Q and mu_y R;
//   vadj   the vertical transpose;
//   hadj   the horizontal transpose, then dL/dx and the weighted sum with
//          the MSE gradient into dL/dC.
// The transpose is written as a gather in the scatter's summation order, so
// every value matches the scatter form bit for bit; the per-channel SSIM sums
// are float64 atomics (order-dependent only in the last bits of a double).
#include "nts_raster.h"

#define SSIM_TAPS 11
#define SSIM_R 5

struct SsimKernel {
  float k[SSIM_TAPS];
};

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

__global__ void __launch_bounds__(256) k_ssim_h(const float4* __restrict__ pred, const float4* __restrict__ target,
                                                uint32_t w, uint32_t h, SsimKernel K, float* __restrict__ H5) {
  const uint64_t n = (uint64_t)w * h;
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int x = (int)(i % w);
  const uint64_t row = i - (uint64_t)x;
  float s[3][5];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int q = 0; q < 5; ++q) s[c][q] = 0.0f;
#pragma unroll
  for (int t = 0; t < SSIM_TAPS; ++t) {
    const uint64_t j = row + (uint64_t)refl101(x + t - SSIM_R, (int)w);
    const float4 p = pred[j], g = target[j];
    const float pc[3] = {p.x, p.y, p.z}, gc[3] = {g.x, g.y, g.z};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float a = pc[c], b = gc[c];
      s[c][0] = s[c][0] + a * K.k[t];
      s[c][1] = s[c][1] + b * K.k[t];
      s[c][2] = s[c][2] + (a * a) * K.k[t];
      s[c][3] = s[c][3] + (b * b) * K.k[t];
      s[c][4] = s[c][4] + (a * b) * K.k[t];
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int q = 0; q < 5; ++q) H5[(uint64_t)(c * 5 + q) * n + i] = s[c][q];
}

__global__ void __launch_bounds__(256) k_ssim_v(const float* __restrict__ H5, uint32_t w, uint32_t h, SsimKernel K,
                                                float* __restrict__ T5, double* __restrict__ sums) {
  const uint64_t n = (uint64_t)w * h;
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double local[3] = {0.0, 0.0, 0.0};
  if (i < n) {
    const int y = (int)(i / w), x = (int)(i - (uint64_t)y * w);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      float m[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int t = 0; t < SSIM_TAPS; ++t) {
        const uint64_t j = (uint64_t)refl101(y + t - SSIM_R, (int)h) * w + x;
#pragma unroll
        for (int q = 0; q < 5; ++q) m[q] = m[q] + H5[(uint64_t)(c * 5 + q) * n + j] * K.k[t];
      }
      const float mux = m[0], muy = m[1];
      const float mux2 = mux * mux, muy2 = muy * muy;
      const float sxx = m[2] - mux2, syy = m[3] - muy2, sxy = m[4] - mux * muy;
      const float av = 2.0f * mux * muy + 1e-4f;
      const float bv = 2.0f * sxy + 9e-4f;
      const float cv = mux2 + muy2 + 1e-4f;
      const float dv = sxx + syy + 9e-4f;
      local[c] = (double)((av * bv) / (cv * dv));
      const float cd = cv * dv;
      const float P = 2.0f * bv * (muy * cv - mux * av) / (cv * cd);
      const float Q = -av * bv / (cv * dv * dv);
      const float R = 2.0f * av / cd;
      T5[(uint64_t)(c * 5 + 0) * n + i] = P;
      T5[(uint64_t)(c * 5 + 1) * n + i] = Q;
      T5[(uint64_t)(c * 5 + 2) * n + i] = R;
      T5[(uint64_t)(c * 5 + 3) * n + i] = mux * Q;
      T5[(uint64_t)(c * 5 + 4) * n + i] = muy * R;
    }
  }
  __shared__ double s_sum[3][4];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double v = local[c];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63u) == 0) s_sum[c][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const double v = s_sum[threadIdx.x][0] + s_sum[threadIdx.x][1] + s_sum[threadIdx.x][2] + s_sum[threadIdx.x][3];
    unsafeAtomicAdd(sums + threadIdx.x, v);
  }
}

// Transpose of one blur pass along an axis of n samples, for source index
// i: value(j) is the pass's output-side value at index j.  Contributions are
// summed in the order of the scatter form (output index j ascending, then
// tap t ascending), so the result is bit-identical to scattering.  For
// n >= 6 the taps of j that land on i are t = 5 - i - j (through the low
// reflection, i != 0), t = i - j + 5 (direct) and t = 2(n - 1) - i - j + 5
// (through the high reflection, i != n - 1), in that order, and only
// j in [i - 5, i + 5] has any.
template <typename F>
__device__ __forceinline__ float adj_gather(int i, int n, const SsimKernel& K, F value) {
  float acc = 0.0f;
  if (n >= 6) {
    const int lo = i - SSIM_R < 0 ? 0 : i - SSIM_R;
    const int hi = i + SSIM_R > n - 1 ? n - 1 : i + SSIM_R;
    for (int j = lo; j <= hi; ++j) {
      const float v = value(j);
      const int t2 = SSIM_R - i - j;
      if (i != 0 && t2 >= 0) acc = acc + v * K.k[t2];
      acc = acc + v * K.k[i - j + SSIM_R];
      const int t3 = 2 * (n - 1) - i - j + SSIM_R;
      if (i != n - 1 && t3 < SSIM_TAPS) acc = acc + v * K.k[t3];
    }
    return acc;
  }
  for (int j = 0; j < n; ++j) {
    const float v = value(j);
    for (int t = 0; t < SSIM_TAPS; ++t)
      if (refl101(j + t - SSIM_R, n) == i) acc = acc + v * K.k[t];
  }
  return acc;
}

__global__ void __launch_bounds__(256) k_ssim_vadj(const float* __restrict__ T5, uint32_t w, uint32_t h, SsimKernel K,
                                                   float* __restrict__ A5) {
  const uint64_t n = (uint64_t)w * h;
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int y = (int)(i / w), x = (int)(i - (uint64_t)y * w);
  for (int p = 0; p < 15; ++p) {
    const float* src = T5 + (uint64_t)p * n;
    A5[(uint64_t)p * n + i] = adj_gather(y, (int)h, K, [&](int j) { return src[(uint64_t)j * w + x]; });
  }
}

__global__ void __launch_bounds__(256) k_ssim_hadj(const float* __restrict__ A5, const float4* __restrict__ pred,
                                                   const float4* __restrict__ target, uint32_t w, uint32_t h,
                                                   SsimKernel K, float lam, float inv_mse, float4* __restrict__ dl_dc) {
  const uint64_t n = (uint64_t)w * h;
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int x = (int)(i % w);
  const uint64_t row = i - (uint64_t)x;
  float adj[15];
#pragma unroll
  for (int p = 0; p < 15; ++p) {
    const float* src = A5 + (uint64_t)p * n + row;
    adj[p] = adj_gather(x, (int)w, K, [&](int j) { return src[j]; });
  }
  const float4 pv = pred[i], tv = target[i];
  const float xs[3] = {pv.x, pv.y, pv.z}, ys[3] = {tv.x, tv.y, tv.z};
  const float inv_n = 1.0f / (float)n;
  float out[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float* a = adj + 5 * c;
    const float d = a[0] + 2.0f * (xs[c] * a[1] - a[3]) + (ys[c] * a[2] - a[4]);
    const float gs = (-inv_n * d) / 3.0f;          // ssim_loss_grad's channel average
    const float gm = (xs[c] - ys[c]) * inv_mse;    // 2 (x - y) / (3WH)
    out[c] = gm * (1.0f - lam) + gs * lam;
  }
  dl_dc[i] = make_float4(out[0], out[1], out[2], 0.0f);
}

__global__ void __launch_bounds__(256) k_mse_sum(const float4* __restrict__ pred, const float4* __restrict__ target,
                                                 uint64_t n, double* __restrict__ acc) {
  double local = 0.0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const float4 p = pred[i], t = target[i];
    const float d0 = p.x - t.x, d1 = p.y - t.y, d2 = p.z - t.z;
    local += (double)(d0 * d0 + d1 * d1 + d2 * d2);
  }
  __shared__ double s_part[4];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) local += __shfl_xor(local, off, 64);
  if ((threadIdx.x & 63u) == 0) s_part[threadIdx.x >> 6] = local;
  __syncthreads();
  if (threadIdx.x == 0) unsafeAtomicAdd(acc, s_part[0] + s_part[1] + s_part[2] + s_part[3]);
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

extern "C" int nts_loss_grad_ssim(nts_ctx* ctx, const nts_buffer* image, const nts_buffer* target,
                                  nts_buffer* dl_dc, uint32_t width, uint32_t height, float lambda, double* total,
                                  double* mse, double* ssim) {
  NTS_CHECK_ARG(ctx && image && target && dl_dc);
  const uint64_t n = (uint64_t)width * height;
  NTS_CHECK_ARG(image->bytes >= n * 16 && target->bytes >= n * 16 && dl_dc->bytes >= n * 16);
  if (n == 0) return nts_fail(NTS_E_INVALID, "SSIM of an empty image");
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
  void* scr;
  const uint64_t plane_bytes = ((n * 4 + 255) & ~255ull);
  int rc = nts_scratch(ctx, 256 + 30 * plane_bytes, &scr);
  if (rc) return rc;
  double* sums = (double*)scr;  // [0..2] SSIM per channel, [3] MSE
  float* H5 = (float*)((char*)scr + 256);
  float* T5 = H5 + 15 * (plane_bytes / 4);
  rc = nts_fill_u32(ctx, sums, 0, 8);
  if (rc) return rc;
  const float4* P = (const float4*)image->ptr;
  const float4* G = (const float4*)target->ptr;
  const dim3 grid(nts_div_up(n, 256));
  uint32_t mgrid = nts_div_up(n, 256);
  if (mgrid > 2048) mgrid = 2048;
  NTS_LAUNCH(ctx, "mse_sum", k_mse_sum, dim3(mgrid), dim3(256), 0, P, G, n, sums + 3);
  const SsimKernel K = ssim_kernel();
  const float inv_mse = 2.0f / ((float)width * (float)height * 3.0f);
  {
    NTS_LAUNCH(ctx, "ssim_h", k_ssim_h, grid, dim3(256), 0, P, G, width, height, K, H5);
    NTS_LAUNCH(ctx, "ssim_v", k_ssim_v, grid, dim3(256), 0, (const float*)H5, width, height, K, T5, sums);
    NTS_LAUNCH(ctx, "ssim_vadj", k_ssim_vadj, grid, dim3(256), 0, (const float*)T5, width, height, K, H5);
    NTS_LAUNCH(ctx, "ssim_hadj", k_ssim_hadj, grid, dim3(256), 0, (const float*)H5, P, G, width, height, K, lam,
               inv_mse, (float4*)dl_dc->ptr);
  }
  double h_sums[4];
  NTS_HIP_TRY(hipMemcpyAsync(h_sums, sums, 32, hipMemcpyDeviceToHost, ctx->stream));
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  const float m = (float)(h_sums[3] / ((double)n * 3.0));
  float l[3];
  for (int c = 0; c < 3; ++c) l[c] = 1.0f - (float)(h_sums[c] / (double)n);
  const float sl = (l[0] + l[1] + l[2]) / 3.0f;
  // loss.rs:404-409
  if (mse) *mse = lam == 1.0f ? 0.0 : (double)m;
  if (ssim) *ssim = (double)sl;
  if (total) *total = lam == 1.0f ? (double)sl : (double)((1.0f - lam) * m + lam * sl);
  return NTS_OK;
}
