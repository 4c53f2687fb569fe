// Scalar maths shared by the gfx950 kernels and the host-side restatement.
//
// Every function here is written with IEEE-754 operations only (+ - * /,
// sqrt, fma, floor/ceil/rint, bit casts), so a translation unit compiled
// with -ffp-contract=off produces the same bits on the host (x86-64, SSE/FMA)
// and on gfx950 (hipcc rounds fp32 division and sqrt correctly and keeps
// fp32 denormals).  That is what makes tile keys, tile ranges and rendered
// images bit-identical between the device path and the host checker.
//
// nts_expf: Cody & Waite (1980) argument reduction x = n*ln2 + r with a
// two-word ln2 (the high word has 9 significant bits, so n*ln2_hi is exact
// for |n| <= 150), then e^r on |r| <= ln2/2 as 1 + r + r^2 Q(r), Q the
// degree-4 minimax fit of (e^r - 1 - r) / r^2 in relative error (3.8e-9 with
// the float coefficients below, under 0.07 ulp), in Horner form with fused
// multiply-adds, then scaling by 2^n built from the exponent field.
// Accuracy: at most 0.90 ulp over every float argument with a normal result
// (checked exhaustively against a float64 exp; the test suite samples it).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define NTS_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#define NTS_HD static inline
#endif

NTS_HD float nts_bits_to_f32(uint32_t u) { return __builtin_bit_cast(float, u); }
NTS_HD uint32_t nts_f32_to_bits(float f) { return __builtin_bit_cast(uint32_t, f); }

NTS_HD float nts_fmaf(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// The blend's power -1/2 (a dx^2 + 2 b dx dy + c dy^2) of conic (a, b, c)
// at offset (dx, dy) (rasterize.wgsl:91-95), in its contracted form: the
// sum of the three products taken as two fused multiply-adds,
// -1/2 fma(c dy, dy, fma(a dx, dx, (2 b dx) dy)).  The device evaluates the
// same with the factor -1/2 folded into the coefficients (exact: a power of
// two).
NTS_HD float nts_blend_power(float a, float b, float c, float dx, float dy) {
  return -0.5f * nts_fmaf(c * dy, dy, nts_fmaf(a * dx, dx, 2.0f * b * dx * dy));
}

// 2^k for k in [-126, 127] as an exact float.
NTS_HD float nts_pow2i(int k) { return nts_bits_to_f32((uint32_t)(k + 127) << 23); }

// e^r for the reduced argument |r| <= ln2/2: six fused multiply-adds.
NTS_HD float nts_exp_poly(float r) {
  float p = 0.0013814572f;
  p = nts_fmaf(p, r, 0.008368721f);
  p = nts_fmaf(p, r, 0.04166839f);
  p = nts_fmaf(p, r, 0.16666521f);
  p = nts_fmaf(p, r, 0.49999994f);
  p = nts_fmaf(p, r, 1.0f);
  return nts_fmaf(p, r, 1.0f);
}

NTS_HD float nts_expf(float x) {
  if (!(x > -104.0f)) {            // deep underflow (and NaN)
    return (x != x) ? x : 0.0f;
  }
  if (x > 88.72284f) return nts_bits_to_f32(0x7f800000u);  // +inf
  const float n = __builtin_rintf(x * 1.44269504088896341f);
  float r = nts_fmaf(n, -0.693359375f, x);          // ln2_hi = 355/512
  r = nts_fmaf(n, 2.12194440e-4f, r);               // ln2 - ln2_hi
  float p = nts_exp_poly(r);
  int k = (int)n;
  if (k < -125) {                                   // result near/below FLT_MIN
    p = p * nts_pow2i(-64);
    k += 64;
  } else if (k > 127) {
    p = p * 2.0f;
    k -= 1;
  }
  return p * nts_pow2i(k);
}

NTS_HD float nts_sigmoidf(float x) { return 1.0f / (1.0f + nts_expf(-x)); }

// Saturating float -> int32 conversion of an integral-valued float (the
// result of floor).  Out-of-range values clamp to the int32 range.
NTS_HD int32_t nts_f2i_sat(float x) {
  if (x != x) return 0;
  if (x <= -2147483648.0f) return (int32_t)0x80000000u;
  if (x >= 2147483520.0f) return 2147483520;
  return (int32_t)x;
}

// x / ts for a (positive, normal) tile size: a power of two divides exactly
// as a multiplication by its reciprocal, which is what it costs then.
NTS_HD float nts_div_tile(float x, float ts) {
  const uint32_t b = nts_f32_to_bits(ts);
  return (b & 0x007FFFFFu) == 0u ? x * nts_bits_to_f32(0x7F000000u - (b & 0x7F800000u)) : x / ts;
}

NTS_HD float nts_minf(float a, float b) { return (b < a) ? b : a; }
NTS_HD float nts_maxf(float a, float b) { return (a < b) ? b : a; }
NTS_HD int32_t nts_mini(int32_t a, int32_t b) { return (b < a) ? b : a; }
NTS_HD int32_t nts_maxi(int32_t a, int32_t b) { return (a < b) ? b : a; }

// Real spherical-harmonic constants of the rasteriser (the band-1 constant is
// 0.488602, as in crates/nano-optimize/src/shaders/project_gaussians.wgsl:51-64).
#define NTS_SH_C0 0.2820948f
#define NTS_SH_C1 0.488602f
#define NTS_SH_C2_0 1.0925485f
#define NTS_SH_C2_1 (-1.0925485f)
#define NTS_SH_C2_2 0.31539157f
#define NTS_SH_C2_3 (-1.0925485f)
#define NTS_SH_C2_4 0.54627424f
#define NTS_SH_C3_0 (-0.5900436f)
#define NTS_SH_C3_1 2.8906114f
#define NTS_SH_C3_2 (-0.4570458f)
#define NTS_SH_C3_3 0.37317634f
#define NTS_SH_C3_4 (-0.4570458f)
#define NTS_SH_C3_5 1.4453057f
#define NTS_SH_C3_6 (-0.5900436f)

// Blend thresholds (crates/nano-optimize/src/shaders/rasterize.wgsl:87-107,
// rasterize_backward.wgsl:48-51).
#define NTS_ALPHA_MIN (1.0f / 255.0f)
#define NTS_ALPHA_MAX 0.99f
#define NTS_T_MIN 1e-4f
#define NTS_ONE_MINUS_ALPHA_EPS 1e-4f
#define NTS_T_BAIL 1.0001f
#define NTS_COV_DILATION 0.3f

// ---- densify random numbers (train.rs:552-564) ------------------------------
// nts_logf: x = 2^e m with m in [sqrt(1/2), sqrt(2)), ln m = 2 atanh(s),
// s = (m - 1)/(m + 1), |s| <= 0.1716, series to s^11 (truncation below
// 1e-10), e ln2 with the two-word ln2 of nts_expf.  Positive normal x only
// (the densify draws are in [1e-8, 1)).
NTS_HD float nts_logf(float x) {
  uint32_t b = nts_f32_to_bits(x);
  int e = (int)((b >> 23) & 0xffu) - 127;
  float m = nts_bits_to_f32((b & 0x007fffffu) | 0x3f800000u);  // [1, 2)
  if (m > 1.41421356f) {
    m = m * 0.5f;
    e += 1;
  }
  const float s = (m - 1.0f) / (m + 1.0f);
  const float s2 = s * s;
  float p = 1.0f / 11.0f;
  p = nts_fmaf(p, s2, 1.0f / 9.0f);
  p = nts_fmaf(p, s2, 1.0f / 7.0f);
  p = nts_fmaf(p, s2, 1.0f / 5.0f);
  p = nts_fmaf(p, s2, 1.0f / 3.0f);
  const float lm = nts_fmaf(2.0f * s * s2, p, 2.0f * s);
  const float fe = (float)e;
  return nts_fmaf(fe, 0.693359375f, nts_fmaf(fe, -2.12194440e-4f, lm));
}

// t^(1/2.4) of the sRGB transfer (tonemap.wgsl:27-41) for the t above its
// cutoff (0.0031308 < t <= 1: positive normal): exp(ln(t) / 2.4) with
// nts_logf and nts_expf, so that the device and the host restatement give
// the same bits (a libm powf and the device's do not always).
NTS_HD float nts_srgb_pow(float t) { return nts_expf(nts_logf(t) * (1.0f / 2.4f)); }

// cos(2 pi u): v = u - rint(u) in [-1/2, 1/2] (exact), a = |v|, then the
// octant reduction a -> 1/4 - a or 1/2 - a (exact for the octants that use
// them) to an argument x = 2 pi w with |x| <= pi/4 and the Taylor series of
// cos (to x^10) or sin (to x^11).
NTS_HD float nts_cos2pi(float u) {
  const float a = __builtin_fabsf(u - __builtin_rintf(u));
  float w, sign = 1.0f;
  bool use_sin = false;
  if (a <= 0.125f) {
    w = a;
  } else if (a <= 0.375f) {
    w = 0.25f - a;
    use_sin = true;
  } else {
    w = 0.5f - a;
    sign = -1.0f;
  }
  const float x = w * 6.28318548f;
  const float x2 = x * x;
  float r;
  if (use_sin) {
    float p = 2.50521084e-8f;            //  1/11!
    p = nts_fmaf(p, -x2, 2.75573192e-6f);  //  1/9!
    p = nts_fmaf(p, -x2, 1.98412698e-4f);  //  1/7!
    p = nts_fmaf(p, -x2, 8.33333333e-3f);  //  1/5!
    p = nts_fmaf(p, -x2, 1.66666667e-1f);  //  1/3!
    r = nts_fmaf(-x * x2, p, x);
  } else {
    float p = 2.75573192e-7f;            //  1/10!
    p = nts_fmaf(p, -x2, 2.48015873e-5f);  //  1/8!
    p = nts_fmaf(p, -x2, 1.38888889e-3f);  //  1/6!
    p = nts_fmaf(p, -x2, 4.16666667e-2f);  //  1/4!
    p = nts_fmaf(p, -x2, 0.5f);            //  1/2!
    r = nts_fmaf(-x2, p, 1.0f);
  }
  return sign * r;
}

// The densify LCG (train.rs:552-556): state' = 1664525 state + 1013904223
// (mod 2^32), uniform = state' / 2^32.
NTS_HD float nts_lcg01(uint32_t* st) {
  *st = *st * 1664525u + 1013904223u;
  return (float)(*st) / 4294967296.0f;
}

// State after k steps, by squaring the affine map s -> a s + c.
NTS_HD uint32_t nts_lcg_jump(uint32_t st, uint64_t k) {
  uint32_t a = 1664525u, c = 1013904223u;  // the map applied 2^bit times
  uint32_t A = 1u, C = 0u;                 // accumulated map
  while (k) {
    if (k & 1u) {
      A = a * A;
      C = a * C + c;
    }
    c = a * c + c;
    a = a * a;
    k >>= 1;
  }
  return A * st + C;
}

// train.rs:558-564: sqrt(-2 ln max(u1, 1e-8)) cos(2 pi u2).
NTS_HD float nts_box_muller(uint32_t* st) {
  const float u1 = nts_maxf(nts_lcg01(st), 1e-8f);
  const float u2 = nts_lcg01(st);
  return __builtin_sqrtf(-2.0f * nts_logf(u1)) * nts_cos2pi(u2);
}

// ln(1/2) and -ln(1.6) rounded to float: the clone opacity step and the
// split scale step (train.rs:455, 489).
#define NTS_LN_HALF (-0.693147182f)
#define NTS_NEG_LN_1_6 (-0.470003635f)

// The tiles of a splat that can hold a pixel with alpha >= 1/255: the
// reference's tile rectangle [x0, x1] x [y0, y1] (in: non-empty) cut to the
// tiles overlapping the axis-aligned box of the ellipse x^T Q x <= 2 ln(255 o)
// (half extents sqrt(2 ln(255 o) Sigma_xx) and sqrt(2 ln(255 o) Sigma_yy),
// Sigma = Q^-1, with a 1% + 0.05 px margin).  A pixel outside that box has
// o exp(-x^T Q x / 2) < 1/255, so it fails the blend's alpha test: dropping
// the other tiles' (splat, tile) pairs leaves every image unchanged.  Conic
// Q = (qa, qb; qb, qc), o = opacity.  False when no tile remains (opacity
// below 1/255 included); a conic that is not positive definite, or too
// ill-conditioned for the margin (nts_cull_conic_ok), and a NaN or infinite
// opacity (whose alpha is not below 1/255 anywhere) keep the reference
// rectangle.
//
// The margin: the float evaluation of q(d) = qa dx^2 + 2 qb dx dy + qc dy^2
// (the blend's power) is within 6 u kappa q of the exact value, with
// kappa = (sqrt(qa qc) + |qb|)^2 / det the largest ratio of
// qa dx^2 + 2 |qb dx dy| + qc dy^2 to q (u = 2^-24).  A point outside a
// half-width grown by 1% has q >= 1.0201 R; with kappa < 2e4 the evaluation
// error (7.2e-3 R) and the rounding of det and of the half-widths (below
// 8e-3 R together) cannot bring its float power back over the threshold.
#define NTS_CULL_KAPPA_MAX 2.0e4f
NTS_HD bool nts_cull_conic_ok(float qa, float qb, float qc, float det) {
  if (!(det > 0.0f) || !(qa > 0.0f) || !(qc > 0.0f)) return false;
  const float s = __builtin_sqrtf(qa * qc) + __builtin_fabsf(qb);
  return s * s / det < NTS_CULL_KAPPA_MAX;
}
NTS_HD bool nts_alpha_tile_rect(float mx, float my, float qa, float qb, float qc, float o, float ts, int32_t* x0,
                                int32_t* y0, int32_t* x1, int32_t* y1) {
  if (o * 255.0f < 0.999f) return false;
  const float det = qa * qc - qb * qb;
  if (!nts_cull_conic_ok(qa, qb, qc, det) || !(o * 255.0f <= 3.4e38f)) return true;  // NaN opacity: kept
  const float l = nts_logf(o * 255.0f);
  const float r2 = nts_maxf(2.0f * l, 0.0f);
  const float ex = __builtin_sqrtf(r2 * (qc / det)) * 1.01f + 0.05f;
  const float ey = __builtin_sqrtf(r2 * (qa / det)) * 1.01f + 0.05f;
  *x0 = nts_maxi(*x0, nts_f2i_sat(__builtin_floorf(nts_div_tile(mx - ex, ts))));
  *y0 = nts_maxi(*y0, nts_f2i_sat(__builtin_floorf(nts_div_tile(my - ey, ts))));
  *x1 = nts_mini(*x1, nts_f2i_sat(__builtin_floorf(nts_div_tile(mx + ex, ts))));
  *y1 = nts_mini(*y1, nts_f2i_sat(__builtin_floorf(nts_div_tile(my + ey, ts))));
  return !(*x1 < *x0 || *y1 < *y0);
}

// Culled binning, per (splat, tile): whether the splat's alpha >= 1/255
// region can reach a pixel centre of the tile's rectangle
// [x0, x1] x [y0, y1].  The test is the exact minimum of
// q(d) = qa dx^2 + 2 qb dx dy + qc dy^2 over the rectangle. It is 0 when
// the mean lies inside; otherwise the minimum is a clamped 1-D minimum on
// an edge facing the mean, at dy = -qb U / qc on dx = U and dx = -qb V / qa
// on dy = V. The result is compared with R = 2 ln(255 o) + 2e-3, which
// carries a float-error margin of 4e-6 kappa + 1e-5 (the float evaluation
// of q is within 6 u kappa q, u = 2^-24).  A failing pair
// fails every pixel's alpha test.  Set up once per record; a splat the
// alpha box keeps whole (nts_alpha_tile_rect: degenerate or ill-conditioned
// conic, NaN or infinite opacity) gets R = +inf and keeps every tile.
struct NtsTileCull {
  float mx, my, qa, qb2, qc, R, rbc, rba;
};
NTS_HD NtsTileCull nts_tile_cull_setup(float mx, float my, float qa, float qb, float qc, float o) {
  NtsTileCull c{mx, my, qa, 2.0f * qb, qc, __builtin_inff(), 0.0f, 0.0f};
  const float det = qa * qc - qb * qb;
  if (!(o * 255.0f >= 0.999f) || !(o * 255.0f <= 3.4e38f) || !nts_cull_conic_ok(qa, qb, qc, det)) return c;
  const float pmin = -nts_logf(o * 255.0f) - 1e-3f;
  const float s = __builtin_sqrtf(qa * qc) + __builtin_fabsf(qb);
  const float kap = s * s / det;
  c.R = -2.0f * pmin * (1.0f + 1e-5f + 4e-6f * kap);
  c.rbc = -qb / qc;
  c.rba = -qb / qa;
  return c;
}
NTS_HD bool nts_tile_cull_hits(const NtsTileCull& c, float x0, float x1, float y0, float y1) {
  const float u0 = x0 - c.mx, u1 = x1 - c.mx, v0 = y0 - c.my, v1 = y1 - c.my;
  const bool out_x = u0 > 0.0f || u1 < 0.0f, out_y = v0 > 0.0f || v1 < 0.0f;
  if (!out_x && !out_y) return true;
  const float U = u0 > 0.0f ? u0 : u1;
  const float dy = nts_minf(nts_maxf(U * c.rbc, v0), v1);
  const float qx = c.qa * U * U + c.qb2 * U * dy + c.qc * dy * dy;
  const float V = v0 > 0.0f ? v0 : v1;
  const float dx = nts_minf(nts_maxf(V * c.rba, u0), u1);
  const float qy = c.qa * dx * dx + c.qb2 * dx * V + c.qc * V * V;
  const float q = nts_minf(out_x ? qx : 3.4e38f, out_y ? qy : 3.4e38f);
  return !(q > c.R);
}

// The pixel-centre rectangle of tile (tx, ty) of size ts (exact in float).
NTS_HD void nts_tile_centres(int32_t tx, int32_t ty, float ts, float* x0, float* x1, float* y0, float* y1) {
  *x0 = (float)tx * ts + 0.5f;
  *x1 = *x0 + (ts - 1.0f);
  *y0 = (float)ty * ts + 0.5f;
  *y1 = *y0 + (ts - 1.0f);
}
