// Per-tile alpha blend, forward and backward, plus loss and tonemap.
//
// One workgroup per (view, tile).  Splat records of the tile's depth-sorted
// slice are staged through LDS in batches; every thread blends its own
// pixel(s) front to back with exactly the arithmetic of the host
// restatement (bit-identical images).  A block-wide vote ends the tile as
// soon as every pixel is saturated.  The backward walks the slice in
// reverse, sums each splat's per-pixel gradient across the wave with
// shuffles and commits one set of float atomics per wave.
#include "nts_raster.h"

template <int TS>
struct Blend {
  static constexpr int NPIX = TS * TS;
  static constexpr int NT = NPIX >= 256 ? 256 : NPIX;
  static constexpr int PPT = NPIX / NT;
};

// ---------------------------------------------------------------------------
// forward (rasterize.wgsl:48-118)
// ---------------------------------------------------------------------------
template <int TS, int BATCH>
__global__ void __launch_bounds__(Blend<TS>::NT) k_blend_fwd(const Projected* __restrict__ proj,
                                                             const uint32_t* __restrict__ payloads,
                                                             const uint2* __restrict__ ranges, ViewTable vt,
                                                             float4* __restrict__ out, float* __restrict__ final_t,
                                                             uint32_t* __restrict__ last_idx) {
  constexpr int NT = Blend<TS>::NT, PPT = Blend<TS>::PPT;
  __shared__ float4 s_a[BATCH];  // mean.x, mean.y, conic.a, conic.b
  __shared__ float4 s_b[BATCH];  // conic.c, opacity, r, g
  __shared__ float s_c[BATCH];   // b
  const uint32_t gt = blockIdx.x;
  const uint32_t vi = view_of_tile(vt, gt);
  const ViewDesc vd = vt.v[vi];
  const uint32_t lt = gt - vd.tile_base;
  const uint32_t tx = lt % vd.tiles_x, ty = lt / vd.tiles_x;
  const uint2 rg = ranges[gt];
  const uint32_t begin = rg.x, end = rg.y;

  float pxf[PPT], pyf[PPT], T[PPT], C0[PPT], C1[PPT], C2[PPT];
  uint32_t last[PPT], pix_x[PPT], pix_y[PPT];
  bool done[PPT], inb[PPT];
#pragma unroll
  for (int j = 0; j < PPT; ++j) {
    const uint32_t li = threadIdx.x + j * NT;
    pix_x[j] = tx * TS + li % TS;
    pix_y[j] = ty * TS + li / TS;
    inb[j] = pix_x[j] < vd.width && pix_y[j] < vd.height;
    pxf[j] = (float)pix_x[j] + 0.5f;
    pyf[j] = (float)pix_y[j] + 0.5f;
    T[j] = 1.0f;
    C0[j] = C1[j] = C2[j] = 0.0f;
    last[j] = 0;
    done[j] = !inb[j];
  }
  for (uint32_t base = begin; base < end; base += BATCH) {
    int live = 0;
#pragma unroll
    for (int j = 0; j < PPT; ++j) live |= !done[j];
    if (!__syncthreads_or(live)) break;
    const uint32_t nb = (end - base) < (uint32_t)BATCH ? (end - base) : (uint32_t)BATCH;
    for (uint32_t k = threadIdx.x; k < nb; k += NT) {
      const Projected p = proj[vd.proj_base + payloads[base + k]];
      s_a[k] = make_float4(p.a.x, p.a.y, p.b.x, p.b.y);
      s_b[k] = make_float4(p.b.z, p.c.w, p.c.x, p.c.y);
      s_c[k] = p.c.z;
    }
    __syncthreads();
    if (PPT == 1) {
      if (!done[0]) {
        for (uint32_t k = 0; k < nb; ++k) {
          const float4 A = s_a[k], B = s_b[k];
          const float dx = pxf[0] - A.x, dy = pyf[0] - A.y;
          const float power = -0.5f * (A.z * dx * dx + 2.0f * A.w * dx * dy + B.x * dy * dy);
          if (power > 0.0f) continue;
          const float g = nts_expf_nonpos(power);
          const float ar = B.y * g;
          const float alpha = nts_minf(ar, NTS_ALPHA_MAX);
          if (alpha < NTS_ALPHA_MIN) continue;
          const float ta = T[0] * alpha;
          C0[0] = C0[0] + ta * B.z;
          C1[0] = C1[0] + ta * B.w;
          C2[0] = C2[0] + ta * s_c[k];
          T[0] = T[0] * (1.0f - alpha);
          last[0] = base + k - begin + 1;
          if (T[0] < NTS_T_MIN) {
            done[0] = true;
            break;
          }
        }
      }
    } else {
      for (uint32_t k = 0; k < nb; ++k) {
        const float4 A = s_a[k], B = s_b[k];
        const float cb = s_c[k];
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
          if (done[j]) continue;
          const float dx = pxf[j] - A.x, dy = pyf[j] - A.y;
          const float power = -0.5f * (A.z * dx * dx + 2.0f * A.w * dx * dy + B.x * dy * dy);
          if (power > 0.0f) continue;
          const float g = nts_expf_nonpos(power);
          const float ar = B.y * g;
          const float alpha = nts_minf(ar, NTS_ALPHA_MAX);
          if (alpha < NTS_ALPHA_MIN) continue;
          const float ta = T[j] * alpha;
          C0[j] = C0[j] + ta * B.z;
          C1[j] = C1[j] + ta * B.w;
          C2[j] = C2[j] + ta * cb;
          T[j] = T[j] * (1.0f - alpha);
          last[j] = base + k - begin + 1;
          if (T[j] < NTS_T_MIN) done[j] = true;
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < PPT; ++j) {
    if (!inb[j]) continue;
    const uint64_t pix = vd.pix_base + (uint64_t)pix_y[j] * vd.width + pix_x[j];
    out[pix] = make_float4(C0[j], C1[j], C2[j], 1.0f - T[j]);
    if (final_t) final_t[pix] = T[j];
    if (last_idx) last_idx[pix] = last[j];
  }
}

// ---------------------------------------------------------------------------
// backward, reference-parity mode (rasterize_backward.wgsl:73-210)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int TS, int BATCH>
__global__ void __launch_bounds__(Blend<TS>::NT) k_blend_bwd_ref(
    const Projected* __restrict__ proj, const uint32_t* __restrict__ payloads, const uint2* __restrict__ ranges,
    ViewTable vt, const float4* __restrict__ fwd_out, const float4* __restrict__ dl_dc,
    ProjectedGrad* __restrict__ pgrad) {
  constexpr int NT = Blend<TS>::NT, PPT = Blend<TS>::PPT;
  __shared__ float4 s_a[BATCH];
  __shared__ float4 s_b[BATCH];
  __shared__ float s_c[BATCH];
  __shared__ uint32_t s_idx[BATCH];
  const uint32_t gt = blockIdx.x;
  const uint32_t vi = view_of_tile(vt, gt);
  const ViewDesc vd = vt.v[vi];
  const uint32_t lt = gt - vd.tile_base;
  const uint32_t tx = lt % vd.tiles_x, ty = lt / vd.tiles_x;
  const uint2 rg = ranges[gt];
  const uint32_t begin = rg.x, end = rg.y;
  const uint32_t lane = threadIdx.x & 63u;

  float pxf[PPT], pyf[PPT], T[PPT], G0[PPT], G1[PPT], G2[PPT], S0[PPT], S1[PPT], S2[PPT];
  bool done[PPT];
#pragma unroll
  for (int j = 0; j < PPT; ++j) {
    const uint32_t li = threadIdx.x + j * NT;
    const uint32_t px = tx * TS + li % TS, py = ty * TS + li / TS;
    const bool inb = px < vd.width && py < vd.height;
    pxf[j] = (float)px + 0.5f;
    pyf[j] = (float)py + 0.5f;
    T[j] = 1.0f;
    G0[j] = G1[j] = G2[j] = 0.0f;
    if (inb) {
      const uint64_t pix = vd.pix_base + (uint64_t)py * vd.width + px;
      T[j] = 1.0f - fwd_out[pix].w;
      const float4 g = dl_dc[pix];
      G0[j] = g.x;
      G1[j] = g.y;
      G2[j] = g.z;
    }
    S0[j] = S1[j] = S2[j] = 0.0f;
    done[j] = !inb || end <= begin;
  }
  uint32_t rem = end > begin ? end - begin : 0;
  while (rem > 0) {
    int live = 0;
#pragma unroll
    for (int j = 0; j < PPT; ++j) live |= !done[j];
    if (!__syncthreads_or(live)) break;
    const uint32_t take = rem < (uint32_t)BATCH ? rem : (uint32_t)BATCH;
    const uint32_t cs = begin + rem - take;
    for (uint32_t k = threadIdx.x; k < take; k += NT) {
      const uint32_t sidx = payloads[cs + k];
      const Projected p = proj[vd.proj_base + sidx];
      s_a[k] = make_float4(p.a.x, p.a.y, p.b.x, p.b.y);
      s_b[k] = make_float4(p.b.z, p.c.w, p.c.x, p.c.y);
      s_c[k] = p.c.z;
      s_idx[k] = sidx;
    }
    __syncthreads();
    for (int k = (int)take - 1; k >= 0; --k) {
      const float4 A = s_a[k], B = s_b[k];
      const float cb = s_c[k];
      float gmx = 0.f, gmy = 0.f, gsig = 0.f, ga_ = 0.f, gb_ = 0.f, gc_ = 0.f, gr = 0.f, gg = 0.f, gbl = 0.f;
      bool contrib = false;
#pragma unroll
      for (int j = 0; j < PPT; ++j) {
        if (done[j]) continue;
        const float dx = pxf[j] - A.x, dy = pyf[j] - A.y;
        const float power = -0.5f * (A.z * dx * dx + 2.0f * A.w * dx * dy + B.x * dy * dy);
        if (power > 0.0f) continue;
        const float g = nts_expf_nonpos(power);
        const float sigma = B.y;
        const float ar = sigma * g;
        const float alpha = nts_minf(ar, NTS_ALPHA_MAX);
        if (alpha < NTS_ALPHA_MIN) continue;
        const float oma = nts_maxf(1.0f - alpha, NTS_ONE_MINUS_ALPHA_EPS);
        T[j] = T[j] / oma;
        if (T[j] > NTS_T_BAIL) {
          done[j] = true;
          continue;
        }
        const float ta = T[j] * alpha;
        const float c0 = B.z, c1 = B.w, c2 = cb;
        const float ct0 = ta * c0, ct1 = ta * c1, ct2 = ta * c2;
        const float d0 = T[j] * c0 - S0[j] / oma;
        const float d1 = T[j] * c1 - S1[j] / oma;
        const float d2 = T[j] * c2 - S2[j] / oma;
        const float gal = G0[j] * d0 + G1[j] * d1 + G2[j] * d2;
        const float sat = (ar < NTS_ALPHA_MAX) ? 1.0f : 0.0f;
        const float gar = gal * sat;
        const float gpow = gar * sigma * g;
        gr += G0[j] * ta;
        gg += G1[j] * ta;
        gbl += G2[j] * ta;
        gsig += gar * g;
        ga_ += gpow * -0.5f * (dx * dx);
        gb_ += gpow * -0.5f * (2.0f * dx * dy);
        gc_ += gpow * -0.5f * (dy * dy);
        gmx += gpow * (A.z * dx + A.w * dy);
        gmy += gpow * (A.w * dx + B.x * dy);
        S0[j] = S0[j] + ct0;
        S1[j] = S1[j] + ct1;
        S2[j] = S2[j] + ct2;
        contrib = true;
      }
      if (__ballot(contrib)) {
        gmx = wave_sum(gmx);
        gmy = wave_sum(gmy);
        gsig = wave_sum(gsig);
        ga_ = wave_sum(ga_);
        gb_ = wave_sum(gb_);
        gc_ = wave_sum(gc_);
        gr = wave_sum(gr);
        gg = wave_sum(gg);
        gbl = wave_sum(gbl);
        if (lane == 0) {
          float* o = reinterpret_cast<float*>(pgrad + vd.proj_base + s_idx[k]);
          unsafeAtomicAdd(o + 0, gmx);
          unsafeAtomicAdd(o + 1, gmy);
          unsafeAtomicAdd(o + 2, gsig);
          unsafeAtomicAdd(o + 4, ga_);
          unsafeAtomicAdd(o + 5, gb_);
          unsafeAtomicAdd(o + 6, gc_);
          unsafeAtomicAdd(o + 8, gr);
          unsafeAtomicAdd(o + 9, gg);
          unsafeAtomicAdd(o + 10, gbl);
        }
      }
    }
    rem -= take;
  }
}

// ---------------------------------------------------------------------------
// dispatch
// ---------------------------------------------------------------------------
#define NTS_BATCH 256

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
  if (bins->tile_size != t->tile_size || bins->tiles_x != (t->width + t->tile_size - 1) / t->tile_size ||
      bins->tiles_y != (t->height + t->tile_size - 1) / t->tile_size)
    return nts_fail(NTS_E_INVALID, "tiling does not match the bins");
  return NTS_OK;
}

int nts_blend_forward(nts_ctx* ctx, uint32_t tile_size, const Projected* proj, const uint32_t* payloads,
                      const uint2* ranges, const ViewTable& vt, float4* out, float* final_t, uint32_t* last_idx) {
  if (vt.total_tiles == 0) return NTS_OK;
  const dim3 grid(vt.total_tiles);
  switch (tile_size) {
    case 8:
      NTS_LAUNCH(ctx, "blend_fwd", (k_blend_fwd<8, NTS_BATCH>), grid, dim3(Blend<8>::NT), 0, proj, payloads, ranges,
                 vt, out, final_t, last_idx);
      break;
    case 16:
      NTS_LAUNCH(ctx, "blend_fwd", (k_blend_fwd<16, NTS_BATCH>), grid, dim3(Blend<16>::NT), 0, proj, payloads,
                 ranges, vt, out, final_t, last_idx);
      break;
    case 32:
      NTS_LAUNCH(ctx, "blend_fwd", (k_blend_fwd<32, NTS_BATCH>), grid, dim3(Blend<32>::NT), 0, proj, payloads,
                 ranges, vt, out, final_t, last_idx);
      break;
    default:
      return nts_fail(NTS_E_INVALID, "tile size %u not in {8, 16, 32}", tile_size);
  }
  return NTS_OK;
}

int nts_blend_backward_ref(nts_ctx* ctx, uint32_t tile_size, const Projected* proj, const uint32_t* payloads,
                           const uint2* ranges, const ViewTable& vt, const float4* fwd_out, const float4* dl_dc,
                           ProjectedGrad* pgrad) {
  if (vt.total_tiles == 0) return NTS_OK;
  const dim3 grid(vt.total_tiles);
  switch (tile_size) {
    case 8:
      NTS_LAUNCH(ctx, "blend_bwd", (k_blend_bwd_ref<8, NTS_BATCH>), grid, dim3(Blend<8>::NT), 0, proj, payloads,
                 ranges, vt, fwd_out, dl_dc, pgrad);
      break;
    case 16:
      NTS_LAUNCH(ctx, "blend_bwd", (k_blend_bwd_ref<16, NTS_BATCH>), grid, dim3(Blend<16>::NT), 0, proj, payloads,
                 ranges, vt, fwd_out, dl_dc, pgrad);
      break;
    case 32:
      NTS_LAUNCH(ctx, "blend_bwd", (k_blend_bwd_ref<32, NTS_BATCH>), grid, dim3(Blend<32>::NT), 0, proj, payloads,
                 ranges, vt, fwd_out, dl_dc, pgrad);
      break;
    default:
      return nts_fail(NTS_E_INVALID, "tile size %u not in {8, 16, 32}", tile_size);
  }
  return NTS_OK;
}

extern "C" int nts_composite(nts_ctx* ctx, const nts_buffer* projected, const nts_bins* bins, nts_buffer* image,
                             const nts_tiling* tiling) {
  NTS_CHECK_ARG(ctx && projected && bins && image && tiling);
  NTS_CHECK_ARG(image->bytes >= (uint64_t)tiling->width * tiling->height * 16);
  ViewTable vt;
  int rc = fill_single_view(bins, tiling, vt);
  if (rc) return rc;
  return nts_blend_forward(ctx, tiling->tile_size, (const Projected*)projected->ptr, bins->payloads, bins->ranges, vt,
                           (float4*)image->ptr, nullptr, nullptr);
}

extern "C" int nts_composite_backward(nts_ctx* ctx, const nts_buffer* projected, const nts_bins* bins,
                                      const nts_buffer* forward_out, const nts_buffer* dl_dc,
                                      nts_buffer* projected_grad, const nts_tiling* tiling) {
  NTS_CHECK_ARG(ctx && projected && bins && forward_out && dl_dc && projected_grad && tiling);
  const uint64_t np = (uint64_t)tiling->width * tiling->height;
  NTS_CHECK_ARG(forward_out->bytes >= np * 16 && dl_dc->bytes >= np * 16);
  NTS_CHECK_ARG(projected_grad->bytes * sizeof(Projected) >= projected->bytes * sizeof(ProjectedGrad));
  ViewTable vt;
  int rc = fill_single_view(bins, tiling, vt);
  if (rc) return rc;
  return nts_blend_backward_ref(ctx, tiling->tile_size, (const Projected*)projected->ptr, bins->payloads,
                                bins->ranges, vt, (const float4*)forward_out->ptr, (const float4*)dl_dc->ptr,
                                (ProjectedGrad*)projected_grad->ptr);
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
    const float s = tm > 0.0031308f ? 1.055f * powf(tm, 1.0f / 2.4f) - 0.055f : tm * 12.92f;
    o |= to_u8(s) << (8 * k);
  }
  o |= to_u8(p.w) << 24;
  out[i] = o;
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
