// Host restatement of nano-optimize's splat rasteriser.  TEST INFRASTRUCTURE
// ONLY: the checker for the gfx950 path (see nts_oracle.h).
//
// Operation order follows the WGSL text: `a * b * c` is (a*b)*c, sums are
// left to right, a WGSL matrix product (A*B)[c][r] = A[0][r]*B[c][0] +
// A[1][r]*B[c][1] + A[2][r]*B[c][2].  Built with -ffp-contract=off.
#include "nts_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "nts_math.h"

namespace {

template <typename T> struct M;
template <> struct M<float> {
  static float exp(float x) { return nts_expf(x); }
  static float sqrt(float x) { return __builtin_sqrtf(x); }
  static float ceil(float x) { return __builtin_ceilf(x); }
  static float floor(float x) { return __builtin_floorf(x); }
  static float fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
};
template <> struct M<double> {
  static double exp(double x) { return ::exp(x); }
  static double sqrt(double x) { return ::sqrt(x); }
  static double ceil(double x) { return ::ceil(x); }
  static double floor(double x) { return ::floor(x); }
  static double fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
};

template <typename T> struct V3 { T x, y, z; };

// Column-major 3x3: c[col][row].
template <typename T> struct M3 { T c[3][3]; };

template <typename T> M3<T> mul(const M3<T>& a, const M3<T>& b) {
  M3<T> o;
  for (int col = 0; col < 3; ++col)
    for (int row = 0; row < 3; ++row)
      o.c[col][row] = a.c[0][row] * b.c[col][0] + a.c[1][row] * b.c[col][1] +
                      a.c[2][row] * b.c[col][2];
  return o;
}
template <typename T> M3<T> transpose(const M3<T>& a) {
  M3<T> o;
  for (int col = 0; col < 3; ++col)
    for (int row = 0; row < 3; ++row) o.c[col][row] = a.c[row][col];
  return o;
}
template <typename T> V3<T> mulv(const M3<T>& a, const V3<T>& v) {
  // WGSL mat * vec = col0*v.x + col1*v.y + col2*v.z
  V3<T> o;
  o.x = a.c[0][0] * v.x + a.c[1][0] * v.y + a.c[2][0] * v.z;
  o.y = a.c[0][1] * v.x + a.c[1][1] * v.y + a.c[2][1] * v.z;
  o.z = a.c[0][2] * v.x + a.c[1][2] * v.y + a.c[2][2] * v.z;
  return o;
}
template <typename T> T dot3(const V3<T>& a, const V3<T>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// project_gaussians.wgsl:68-78 (rotation stored w, x, y, z).
template <typename T> M3<T> quat_to_mat3(const T* q) {
  const T w = q[0], x = q[1], y = q[2], z = q[3];
  M3<T> r;
  r.c[0][0] = T(1) - T(2) * (y * y + z * z);
  r.c[0][1] = T(2) * (x * y + w * z);
  r.c[0][2] = T(2) * (x * z - w * y);
  r.c[1][0] = T(2) * (x * y - w * z);
  r.c[1][1] = T(1) - T(2) * (x * x + z * z);
  r.c[1][2] = T(2) * (y * z + w * x);
  r.c[2][0] = T(2) * (x * z + w * y);
  r.c[2][1] = T(2) * (y * z - w * x);
  r.c[2][2] = T(1) - T(2) * (x * x + y * y);
  return r;
}

// SH basis of bands 1..3 (project_gaussians.wgsl:102-128,
// project_backward.wgsl:166-182).
template <typename T> void sh_basis(T x, T y, T z, T b[15]) {
  const T xx = x * x, yy = y * y, zz = z * z;
  b[0] = -T(NTS_SH_C1) * y;
  b[1] = T(NTS_SH_C1) * z;
  b[2] = -T(NTS_SH_C1) * x;
  b[3] = T(NTS_SH_C2_0) * x * y;
  b[4] = T(NTS_SH_C2_1) * y * z;
  b[5] = T(NTS_SH_C2_2) * (T(2) * zz - xx - yy);
  b[6] = T(NTS_SH_C2_3) * x * z;
  b[7] = T(NTS_SH_C2_4) * (xx - yy);
  b[8] = T(NTS_SH_C3_0) * y * (T(3) * xx - yy);
  b[9] = T(NTS_SH_C3_1) * x * y * z;
  b[10] = T(NTS_SH_C3_2) * y * (T(4) * zz - xx - yy);
  b[11] = T(NTS_SH_C3_3) * z * (T(2) * zz - T(3) * xx - T(3) * yy);
  b[12] = T(NTS_SH_C3_4) * x * (T(4) * zz - xx - yy);
  b[13] = T(NTS_SH_C3_5) * z * (xx - yy);
  b[14] = T(NTS_SH_C3_6) * x * (xx - T(3) * yy);
}

// d b_k / d (x, y, z) of the 15 basis functions above.
static void sh_basis_grad(float x, float y, float z, float g[15][3]) {
  const float xx = x * x, yy = y * y, zz = z * z;
  const float C1 = NTS_SH_C1;
  g[0][0] = 0.0f; g[0][1] = -C1; g[0][2] = 0.0f;
  g[1][0] = 0.0f; g[1][1] = 0.0f; g[1][2] = C1;
  g[2][0] = -C1; g[2][1] = 0.0f; g[2][2] = 0.0f;
  g[3][0] = NTS_SH_C2_0 * y; g[3][1] = NTS_SH_C2_0 * x; g[3][2] = 0.0f;
  g[4][0] = 0.0f; g[4][1] = NTS_SH_C2_1 * z; g[4][2] = NTS_SH_C2_1 * y;
  g[5][0] = -2.0f * NTS_SH_C2_2 * x; g[5][1] = -2.0f * NTS_SH_C2_2 * y; g[5][2] = 4.0f * NTS_SH_C2_2 * z;
  g[6][0] = NTS_SH_C2_3 * z; g[6][1] = 0.0f; g[6][2] = NTS_SH_C2_3 * x;
  g[7][0] = 2.0f * NTS_SH_C2_4 * x; g[7][1] = -2.0f * NTS_SH_C2_4 * y; g[7][2] = 0.0f;
  g[8][0] = 6.0f * NTS_SH_C3_0 * x * y; g[8][1] = NTS_SH_C3_0 * (3.0f * xx - 3.0f * yy); g[8][2] = 0.0f;
  g[9][0] = NTS_SH_C3_1 * y * z; g[9][1] = NTS_SH_C3_1 * x * z; g[9][2] = NTS_SH_C3_1 * x * y;
  g[10][0] = -2.0f * NTS_SH_C3_2 * x * y; g[10][1] = NTS_SH_C3_2 * (4.0f * zz - xx - 3.0f * yy);
  g[10][2] = 8.0f * NTS_SH_C3_2 * y * z;
  g[11][0] = -6.0f * NTS_SH_C3_3 * x * z; g[11][1] = -6.0f * NTS_SH_C3_3 * y * z;
  g[11][2] = NTS_SH_C3_3 * (6.0f * zz - 3.0f * xx - 3.0f * yy);
  g[12][0] = NTS_SH_C3_4 * (4.0f * zz - 3.0f * xx - yy); g[12][1] = -2.0f * NTS_SH_C3_4 * x * y;
  g[12][2] = 8.0f * NTS_SH_C3_4 * x * z;
  g[13][0] = 2.0f * NTS_SH_C3_5 * x * z; g[13][1] = -2.0f * NTS_SH_C3_5 * y * z; g[13][2] = NTS_SH_C3_5 * (xx - yy);
  g[14][0] = NTS_SH_C3_6 * (3.0f * xx - 3.0f * yy); g[14][1] = -6.0f * NTS_SH_C3_6 * x * y; g[14][2] = 0.0f;
}

template <typename T> struct Splats {
  const T *pos, *rot, *log_scale, *logit, *sh_dc, *sh_rest;
};

template <typename T> M3<T> view_rot(const orc_camera* cam) {
  M3<T> w;
  for (int col = 0; col < 3; ++col)
    for (int row = 0; row < 3; ++row) w.c[col][row] = T(cam->view[col * 4 + row]);
  return w;
}

// view * vec4(p, 1) then divide by w (project_gaussians.wgsl:170-172).
template <typename T> V3<T> to_camera(const orc_camera* cam, const T* p) {
  const float* v = cam->view;
  T o[4];
  for (int row = 0; row < 4; ++row)
    o[row] = T(v[0 * 4 + row]) * p[0] + T(v[1 * 4 + row]) * p[1] + T(v[2 * 4 + row]) * p[2] +
             T(v[3 * 4 + row]) * T(1);
  return V3<T>{o[0] / o[3], o[1] / o[3], o[2] / o[3]};
}

// project_gaussians.wgsl:156-259.
template <typename T> void project_one(const Splats<T>& s, uint32_t idx, const orc_camera* cam, T* out) {
  for (int k = 0; k < 12; ++k) out[k] = T(0);
  const T* pw = s.pos + 3 * (size_t)idx;
  const V3<T> cx = to_camera<T>(cam, pw);
  const T depth = -cx.z;
  if (depth < T(cam->near_)) return;
  const V3<T> t{cx.x, cx.y, depth};
  const T W = T(cam->width), H = T(cam->height), f = T(cam->focal);
  const T cxp = W * T(0.5), cyp = H * T(0.5);
  const T mx = cxp + (f * t.x) / t.z;
  const T my = cyp - (f * t.y) / t.z;

  const M3<T> R = quat_to_mat3<T>(s.rot + 4 * (size_t)idx);
  const T* ls = s.log_scale + 3 * (size_t)idx;
  const T sg[3] = {M<T>::exp(ls[0]), M<T>::exp(ls[1]), M<T>::exp(ls[2])};
  const T s2[3] = {sg[0] * sg[0], sg[1] * sg[1], sg[2] * sg[2]};
  M3<T> Ms;
  for (int col = 0; col < 3; ++col)
    for (int row = 0; row < 3; ++row) Ms.c[col][row] = R.c[col][row] * s2[col];
  const M3<T> cov_w = mul(Ms, transpose(R));
  const M3<T> Wm = view_rot<T>(cam);
  const M3<T> cov_c = mul(mul(Wm, cov_w), transpose(Wm));

  const T inv_z = T(1) / t.z;
  const T inv_z2 = inv_z * inv_z;
  const V3<T> j0{f * inv_z, T(0), -f * t.x * inv_z2};
  const V3<T> j1{T(0), -f * inv_z, f * t.y * inv_z2};
  V3<T> r0, r1;
  r0.x = cov_c.c[0][0] * j0.x + cov_c.c[1][0] * j0.y + cov_c.c[2][0] * j0.z;
  r0.y = cov_c.c[0][1] * j0.x + cov_c.c[1][1] * j0.y + cov_c.c[2][1] * j0.z;
  r0.z = cov_c.c[0][2] * j0.x + cov_c.c[1][2] * j0.y + cov_c.c[2][2] * j0.z;
  r1.x = cov_c.c[0][0] * j1.x + cov_c.c[1][0] * j1.y + cov_c.c[2][0] * j1.z;
  r1.y = cov_c.c[0][1] * j1.x + cov_c.c[1][1] * j1.y + cov_c.c[2][1] * j1.z;
  r1.z = cov_c.c[0][2] * j1.x + cov_c.c[1][2] * j1.y + cov_c.c[2][2] * j1.z;
  T s00 = dot3(r0, j0);
  const T s01 = dot3(r0, j1);
  T s11 = dot3(r1, j1);
  s00 = s00 + T(NTS_COV_DILATION);
  s11 = s11 + T(NTS_COV_DILATION);
  const T det = s00 * s11 - s01 * s01;
  if (det <= T(0)) return;
  const T inv_det = T(1) / det;
  const T ca = s11 * inv_det, cb = -s01 * inv_det, cc = s00 * inv_det;
  const T mid = T(0.5) * (s00 + s11);
  const T disc = mid * mid - det;
  const T lambda = mid + M<T>::sqrt(disc > T(0) ? disc : T(0));
  const T radius = M<T>::ceil(T(3) * M<T>::sqrt(lambda));
  const T maxpix = (W > H ? W : H) + radius;
  if (mx < -radius || mx >= W + radius || my < -radius || my >= H + radius || radius >= maxpix) return;

  // colour (project_gaussians.wgsl:88-154, 250-252)
  V3<T> d{T(cam->cam_pos[0]) - pw[0], T(cam->cam_pos[1]) - pw[1], T(cam->cam_pos[2]) - pw[2]};
  const T inv_len = T(1) / M<T>::sqrt(dot3(d, d));
  d.x = d.x * inv_len;
  d.y = d.y * inv_len;
  d.z = d.z * inv_len;
  T b[15];
  sh_basis<T>(d.x, d.y, d.z, b);
  const T* dc = s.sh_dc + 3 * (size_t)idx;
  const T* rest = s.sh_rest + 45 * (size_t)idx;
  T rgb[3] = {T(NTS_SH_C0) * dc[0], T(NTS_SH_C0) * dc[1], T(NTS_SH_C0) * dc[2]};
  T acc[3] = {T(0), T(0), T(0)};
  for (int i = 0; i < 3; ++i)
    for (int ch = 0; ch < 3; ++ch) acc[ch] = acc[ch] + rest[15 * ch + i] * b[i];
  for (int i = 0; i < 5; ++i)
    for (int ch = 0; ch < 3; ++ch) acc[ch] = acc[ch] + rest[15 * ch + 3 + i] * b[3 + i];
  for (int i = 0; i < 7; ++i)
    for (int ch = 0; ch < 3; ++ch) acc[ch] = acc[ch] + rest[15 * ch + 8 + i] * b[8 + i];
  for (int ch = 0; ch < 3; ++ch) {
    rgb[ch] = rgb[ch] + acc[ch];
    const T c = rgb[ch] + T(0.5);
    rgb[ch] = c > T(0) ? c : T(0);
  }
  const T lg = s.logit[idx];
  const T alpha = T(1) / (T(1) + M<T>::exp(-lg));

  out[0] = mx; out[1] = my; out[2] = depth; out[3] = radius;
  out[4] = ca; out[5] = cb; out[6] = cc; out[7] = T(1);
  out[8] = rgb[0]; out[9] = rgb[1]; out[10] = rgb[2]; out[11] = alpha;
}

template <typename T> void project_all(const Splats<T>& s, uint32_t n, const orc_camera* cam, T* projected) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) project_one<T>(s, (uint32_t)i, cam, projected + 12 * i);
}

// Tile bbox of a visible record (tile_count.wgsl:39-50); returns false when empty.
template <typename T>
bool tile_bbox(const T* p, const orc_tiling* t, int32_t* x0, int32_t* y0, int32_t* x1, int32_t* y1) {
  if (p[7] == T(0)) return false;
  const uint32_t tiles_x = (t->width + t->tile_size - 1) / t->tile_size;
  const uint32_t tiles_y = (t->height + t->tile_size - 1) / t->tile_size;
  const float mx = (float)p[0], my = (float)p[1], r = (float)p[3];
  const float ts = (float)t->tile_size;
  *x0 = nts_maxi(nts_f2i_sat(__builtin_floorf((mx - r) / ts)), 0);
  *y0 = nts_maxi(nts_f2i_sat(__builtin_floorf((my - r) / ts)), 0);
  *x1 = nts_mini(nts_f2i_sat(__builtin_floorf((mx + r) / ts)), (int32_t)tiles_x - 1);
  *y1 = nts_mini(nts_f2i_sat(__builtin_floorf((my + r) / ts)), (int32_t)tiles_y - 1);
  return !(*x1 < *x0 || *y1 < *y0);
}

template <typename T>
void composite_t(const T* projected, const uint32_t* payloads, const uint32_t* ranges,
                 const orc_tiling* t, T* out, float* final_t, uint32_t* last_idx, uint32_t* n_used = nullptr) {
  const uint32_t ts = t->tile_size;
  const uint32_t tiles_x = (t->width + ts - 1) / ts;
  const uint32_t tiles_y = (t->height + ts - 1) / ts;
  const int64_t ntiles = (int64_t)tiles_x * tiles_y;
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t tile = 0; tile < ntiles; ++tile) {
    const uint32_t tx = (uint32_t)(tile % tiles_x), ty = (uint32_t)(tile / tiles_x);
    const uint32_t begin = ranges[2 * tile], end = ranges[2 * tile + 1];
    for (uint32_t ly = 0; ly < ts; ++ly)
      for (uint32_t lx = 0; lx < ts; ++lx) {
        const uint32_t px = tx * ts + lx, py = ty * ts + ly;
        if (px >= t->width || py >= t->height) continue;
        const T pxf = T((float)px + 0.5f), pyf = T((float)py + 0.5f);
        T C[3] = {T(0), T(0), T(0)};
        T Tr = T(1);
        uint32_t last = 0, used = 0;
        for (uint32_t i = begin; i < end; ++i) {
          const T* p = projected + 12 * (size_t)payloads[i];
          const T dx = pxf - p[0], dy = pyf - p[1];
          const T power = T(-0.5) * M<T>::fma(p[6] * dy, dy, M<T>::fma(p[4] * dx, dx, T(2) * p[5] * dx * dy));
          if (!(power <= T(0))) continue;  // also a NaN power
          const T g = M<T>::exp(power);
          const T ar = p[11] * g;
          const T alpha = ar < T(NTS_ALPHA_MAX) ? ar : T(NTS_ALPHA_MAX);
          if (alpha < T(NTS_ALPHA_MIN)) continue;
          // color = color + T alpha c (rasterize.wgsl:103), with the
          // product-add contracted into one fused multiply-add (a form the
          // WGSL's compilation may take; one rounding)
          const T ta = Tr * alpha;
          C[0] = M<T>::fma(ta, p[8], C[0]);
          C[1] = M<T>::fma(ta, p[9], C[1]);
          C[2] = M<T>::fma(ta, p[10], C[2]);
          Tr = Tr * (T(1) - alpha);
          last = i - begin + 1;
          ++used;
          if (Tr < T(NTS_T_MIN)) break;
        }
        const size_t pix = (size_t)py * t->width + px;
        out[4 * pix + 0] = C[0];
        out[4 * pix + 1] = C[1];
        out[4 * pix + 2] = C[2];
        out[4 * pix + 3] = T(1) - Tr;
        if (final_t) final_t[pix] = (float)Tr;
        if (last_idx) last_idx[pix] = last;
        if (n_used) n_used[pix] = used;
      }
  }
}

// compiler-rt __powisf2 (what Rust's f32::powi lowers to).
float powi_f32(float a, int32_t b) {
  const int recip = b < 0;
  float r = 1.0f;
  while (1) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return recip ? 1.0f / r : r;
}

}  // namespace

extern "C" {

int orc_version(void) { return 1; }

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_expf_array(const float* x, float* y, uint64_t n) {
  for (uint64_t i = 0; i < n; ++i) y[i] = nts_expf(x[i]);
}

void orc_logf_array(const float* x, float* y, uint64_t n) {
  for (uint64_t i = 0; i < n; ++i) y[i] = nts_logf(x[i]);
}

void orc_cos2pi_array(const float* x, float* y, uint64_t n) {
  for (uint64_t i = 0; i < n; ++i) y[i] = nts_cos2pi(x[i]);
}

uint32_t orc_lcg_jump(uint32_t state, uint64_t k) { return nts_lcg_jump(state, k); }

// n draws of the densify Box-Muller stream from `state`, plus the uniforms
// each consumed (2 per draw) for checking the transform.
void orc_box_muller_stream(uint32_t state, uint64_t n, float* out, float* uniforms) {
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t probe = state;
    uniforms[2 * i] = nts_lcg01(&probe);
    uniforms[2 * i + 1] = nts_lcg01(&probe);
    out[i] = nts_box_muller(&state);
  }
}

// raster.rs:79-99 CameraUniform::from_pose with glam's right-handed look-at
// (view = [s | u | -f] rows, translation -eye.s, -eye.u, eye.f).
void orc_camera_from_pose(const float eye[3], const float target[3], const float up[3], float fov_y,
                          uint32_t width, uint32_t height, uint32_t n_splats, orc_camera* out) {
  float f[3] = {target[0] - eye[0], target[1] - eye[1], target[2] - eye[2]};
  float il = 1.0f / sqrtf(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  for (int i = 0; i < 3; ++i) f[i] = f[i] * il;
  float s[3] = {f[1] * up[2] - up[1] * f[2], f[2] * up[0] - up[2] * f[0], f[0] * up[1] - up[0] * f[1]};
  il = 1.0f / sqrtf(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
  for (int i = 0; i < 3; ++i) s[i] = s[i] * il;
  const float u[3] = {s[1] * f[2] - f[1] * s[2], s[2] * f[0] - f[2] * s[0], s[0] * f[1] - f[0] * s[1]};
  float* v = out->view;
  for (int c = 0; c < 3; ++c) {
    v[c * 4 + 0] = s[c];
    v[c * 4 + 1] = u[c];
    v[c * 4 + 2] = -f[c];
    v[c * 4 + 3] = 0.0f;
  }
  v[12] = -(eye[0] * s[0] + eye[1] * s[1] + eye[2] * s[2]);
  v[13] = -(eye[0] * u[0] + eye[1] * u[1] + eye[2] * u[2]);
  v[14] = eye[0] * f[0] + eye[1] * f[1] + eye[2] * f[2];
  v[15] = 1.0f;
  out->cam_pos[0] = eye[0];
  out->cam_pos[1] = eye[1];
  out->cam_pos[2] = eye[2];
  out->near_ = 0.01f;
  out->focal = (float)height / (2.0f * tanf(fov_y * 0.5f));
  out->width = (float)width;
  out->height = (float)height;
  out->n_splats = n_splats;
}

void orc_project_f32(const orc_splats_f32* s, uint32_t n, const orc_camera* cam, float* projected) {
  Splats<float> sp{s->pos, s->rot, s->log_scale, s->logit, s->sh_dc, s->sh_rest};
  project_all<float>(sp, n, cam, projected);
}

void orc_project_f64(const orc_splats_f64* s, uint32_t n, const orc_camera* cam, double* projected) {
  Splats<double> sp{s->pos, s->rot, s->log_scale, s->logit, s->sh_dc, s->sh_rest};
  project_all<double>(sp, n, cam, projected);
}

// The tile rectangle of record p: the reference's, or with cull the tiles
// its alpha >= 1/255 ellipse can reach (nts_alpha_tile_rect; the device's
// batched and serving paths bin this way).
static bool record_tiles(const float* p, const orc_tiling* t, int cull, int32_t* x0, int32_t* y0, int32_t* x1,
                         int32_t* y1) {
  if (!tile_bbox<float>(p, t, x0, y0, x1, y1)) return false;
  return !cull || nts_alpha_tile_rect(p[0], p[1], p[4], p[5], p[6], p[11], (float)t->tile_size, x0, y0, x1, y1);
}

static uint64_t tile_count_ex(const float* projected, uint32_t n, const orc_tiling* t, int cull, uint32_t* counts) {
  uint64_t total = 0;
#pragma omp parallel for reduction(+ : total) schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    int32_t x0, y0, x1, y1;
    uint32_t c = 0;
    if (record_tiles(projected + 12 * i, t, cull, &x0, &y0, &x1, &y1))
      c = (uint32_t)((x1 - x0 + 1) * (y1 - y0 + 1));
    counts[i] = c;
    total += c;
  }
  return total;
}

uint64_t orc_tile_count(const float* projected, uint32_t n, const orc_tiling* t, uint32_t* counts) {
  return tile_count_ex(projected, n, t, 0, counts);
}

// scan_block.wgsl + scan_add_offsets.wgsl: exclusive, wrapping u32.
void orc_exclusive_scan_u32(const uint32_t* in, uint64_t n, uint32_t* out) {
  uint32_t acc = 0;
  for (uint64_t i = 0; i < n; ++i) {
    const uint32_t v = in[i];
    out[i] = acc;
    acc += v;
  }
}

// tile_emit.wgsl:91-121.
static void tile_emit_ex(const float* projected, uint32_t n, const uint32_t* offsets, const orc_tiling* t, int cull,
                         uint32_t* keys, uint32_t* payloads) {
  const uint32_t tiles_x = (t->width + t->tile_size - 1) / t->tile_size;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const float* p = projected + 12 * i;
    int32_t x0, y0, x1, y1;
    if (!record_tiles(p, t, cull, &x0, &y0, &x1, &y1)) continue;
    float dn = p[2] / t->depth_max;
    dn = dn < 0.0f ? 0.0f : (dn > 1.0f ? 1.0f : dn);
    const uint32_t d16 = (uint32_t)(dn * 65535.0f);
    // with cull, a tile the alpha >= 1/255 region cannot reach gets the
    // all-ones key (after every live pair of the view, in no tile's range)
    const NtsTileCull tc = nts_tile_cull_setup(p[0], p[1], p[4], p[5], p[6], p[11]);
    uint32_t w = offsets[i];
    for (int32_t ty = y0; ty <= y1; ++ty)
      for (int32_t tx = x0; tx <= x1; ++tx) {
        const uint32_t tile = (uint32_t)ty * tiles_x + (uint32_t)tx;
        bool live = true;
        if (cull) {
          float cx0, cx1, cy0, cy1;
          nts_tile_centres(tx, ty, (float)t->tile_size, &cx0, &cx1, &cy0, &cy1);
          live = nts_tile_cull_hits(tc, cx0, cx1, cy0, cy1);
        }
        keys[w] = live ? ((tile << 16) | d16) : 0xFFFFFFFFu;
        payloads[w] = (uint32_t)i;
        ++w;
      }
  }
}

void orc_tile_emit(const float* projected, uint32_t n, const uint32_t* offsets, const orc_tiling* t,
                   uint32_t* keys, uint32_t* payloads) {
  tile_emit_ex(projected, n, offsets, t, 0, keys, payloads);
}

// radix_sort.rs:1-26: 32 stable 1-bit LSD passes == a stable sort by key.
void orc_sort_pairs_u32(uint32_t* keys, uint32_t* vals, uint64_t n) {
  std::vector<uint64_t> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) { return keys[a] < keys[b]; });
  std::vector<uint32_t> k2(n), v2(n);
  for (uint64_t i = 0; i < n; ++i) {
    k2[i] = keys[idx[i]];
    v2[i] = vals[idx[i]];
  }
  memcpy(keys, k2.data(), n * 4);
  memcpy(vals, v2.data(), n * 4);
}

void orc_sort_pairs_u64(uint64_t* keys, uint32_t* vals, uint64_t n) {
  std::vector<uint64_t> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) { return keys[a] < keys[b]; });
  std::vector<uint64_t> k2(n);
  std::vector<uint32_t> v2(n);
  for (uint64_t i = 0; i < n; ++i) {
    k2[i] = keys[idx[i]];
    v2[i] = vals[idx[i]];
  }
  memcpy(keys, k2.data(), n * 8);
  memcpy(vals, v2.data(), n * 4);
}

// tile_ranges.wgsl:22-41 (ranges zero-initialised: empty tiles stay (0,0)).
void orc_tile_ranges(const uint32_t* sorted_keys, uint64_t k, uint32_t num_tiles, uint32_t* ranges) {
  memset(ranges, 0, sizeof(uint32_t) * 2 * (size_t)num_tiles);
  for (uint64_t i = 0; i < k; ++i) {
    const uint32_t tile = sorted_keys[i] >> 16;
    const bool starts = (i == 0) || ((sorted_keys[i - 1] >> 16) != tile);
    const bool ends = (i + 1 == k) || ((sorted_keys[i + 1] >> 16) != tile);
    if (tile >= num_tiles) continue;
    if (starts) ranges[2 * tile] = (uint32_t)i;
    if (ends) ranges[2 * tile + 1] = (uint32_t)(i + 1);
  }
}

uint64_t orc_bin_ex(const float* projected, uint32_t n, const orc_tiling* t, int cull, uint32_t* keys,
                    uint32_t* payloads, uint32_t* ranges) {
  std::vector<uint32_t> counts(n > 0 ? n : 1), offsets(n > 0 ? n : 1);
  const uint64_t total = tile_count_ex(projected, n, t, cull, counts.data());
  if (!keys) return total;
  const uint32_t tiles_x = (t->width + t->tile_size - 1) / t->tile_size;
  const uint32_t tiles_y = (t->height + t->tile_size - 1) / t->tile_size;
  const uint32_t num_tiles = tiles_x * tiles_y;
  if (total == 0) {
    memset(ranges, 0, sizeof(uint32_t) * 2 * (size_t)num_tiles);
    return 0;
  }
  orc_exclusive_scan_u32(counts.data(), n, offsets.data());
  tile_emit_ex(projected, n, offsets.data(), t, cull, keys, payloads);
  orc_sort_pairs_u32(keys, payloads, total);
  orc_tile_ranges(keys, total, num_tiles, ranges);
  return total;
}

uint64_t orc_bin(const float* projected, uint32_t n, const orc_tiling* t, uint32_t* keys,
                 uint32_t* payloads, uint32_t* ranges) {
  return orc_bin_ex(projected, n, t, 0, keys, payloads, ranges);
}

void orc_composite(const float* projected, const uint32_t* payloads, const uint32_t* ranges,
                   const orc_tiling* t, float* out, float* final_t, uint32_t* last_idx) {
  composite_t<float>(projected, payloads, ranges, t, out, final_t, last_idx);
}

void orc_composite_used(const float* projected, const uint32_t* payloads, const uint32_t* ranges,
                        const orc_tiling* t, float* out, uint32_t* n_used) {
  composite_t<float>(projected, payloads, ranges, t, out, nullptr, nullptr, n_used);
}

void orc_composite_f64(const double* projected, const uint32_t* payloads, const uint32_t* ranges,
                       const orc_tiling* t, double* out, uint32_t* last_idx, uint32_t* n_used) {
  composite_t<double>(projected, payloads, ranges, t, out, nullptr, last_idx, n_used);
}

// rasterize_backward.wgsl:73-210.  Per-pixel terms are computed exactly as
// the kernel does (fp32); the per-splat sums are accumulated in float64.
void orc_composite_backward_ref(const float* projected, const uint32_t* payloads, const uint32_t* ranges,
                                const orc_tiling* t, uint32_t n, const float* forward_out,
                                const float* dl_dc, float* projected_grad) {
  const uint32_t ts = t->tile_size;
  const uint32_t tiles_x = (t->width + ts - 1) / ts;
  const uint32_t tiles_y = (t->height + ts - 1) / ts;
  const int64_t ntiles = (int64_t)tiles_x * tiles_y;
  std::vector<double> acc((size_t)n * 9, 0.0);
  int nthreads = orc_num_threads();
  std::vector<std::vector<double>> partial(nthreads);
#pragma omp parallel
  {
#ifdef _OPENMP
    const int tid = omp_get_thread_num();
#else
    const int tid = 0;
#endif
    std::vector<double>& mine = partial[tid];
    mine.assign((size_t)n * 9, 0.0);
#pragma omp for schedule(dynamic, 4)
    for (int64_t tile = 0; tile < ntiles; ++tile) {
      const uint32_t tx = (uint32_t)(tile % tiles_x), ty = (uint32_t)(tile / tiles_x);
      const uint32_t begin = ranges[2 * tile], end = ranges[2 * tile + 1];
      if (end <= begin) continue;
      for (uint32_t ly = 0; ly < ts; ++ly)
        for (uint32_t lx = 0; lx < ts; ++lx) {
          const uint32_t px = tx * ts + lx, py = ty * ts + ly;
          if (px >= t->width || py >= t->height) continue;
          const size_t pix = (size_t)py * t->width + px;
          const float pxf = (float)px + 0.5f, pyf = (float)py + 0.5f;
          float T = 1.0f - forward_out[4 * pix + 3];
          const float gC[3] = {dl_dc[4 * pix + 0], dl_dc[4 * pix + 1], dl_dc[4 * pix + 2]};
          float S[3] = {0.0f, 0.0f, 0.0f};
          for (int64_t i = (int64_t)end - 1; i >= (int64_t)begin; --i) {
            const uint32_t sidx = payloads[i];
            const float* p = projected + 12 * (size_t)sidx;
            const float dx = pxf - p[0], dy = pyf - p[1];
            const float ca = p[4], cb = p[5], cc = p[6];
            const float power = nts_blend_power(ca, cb, cc, dx, dy);
            if (!(power <= 0.0f)) continue;  // also a NaN power
            const float g = nts_expf(power);
            const float sigma = p[11];
            const float ar = sigma * g;
            const float alpha = ar < NTS_ALPHA_MAX ? ar : NTS_ALPHA_MAX;  // a NaN clamps, as in the forward
            if (alpha < NTS_ALPHA_MIN) continue;
            const float oma = nts_maxf(1.0f - alpha, NTS_ONE_MINUS_ALPHA_EPS);
            T = T / oma;
            if (T > NTS_T_BAIL) break;
            const float ta = T * alpha;
            const float col[3] = {p[8], p[9], p[10]};
            const float contrib[3] = {ta * col[0], ta * col[1], ta * col[2]};
            const float gcol[3] = {gC[0] * ta, gC[1] * ta, gC[2] * ta};
            const float dCa[3] = {T * col[0] - S[0] / oma, T * col[1] - S[1] / oma, T * col[2] - S[2] / oma};
            const float ga = gC[0] * dCa[0] + gC[1] * dCa[1] + gC[2] * dCa[2];
            const float sat = (ar < NTS_ALPHA_MAX) ? 1.0f : 0.0f;
            const float gar = ga * sat;
            const float gsig = gar * g;
            const float gpow = gar * sigma * g;
            const float g_a = gpow * -0.5f * (dx * dx);
            const float g_b = gpow * -0.5f * (2.0f * dx * dy);
            const float g_c = gpow * -0.5f * (dy * dy);
            const float gmx = gpow * (ca * dx + cb * dy);
            const float gmy = gpow * (cb * dx + cc * dy);
            double* a = &mine[(size_t)sidx * 9];
            a[0] += gmx; a[1] += gmy; a[2] += gsig;
            a[3] += g_a; a[4] += g_b; a[5] += g_c;
            a[6] += gcol[0]; a[7] += gcol[1]; a[8] += gcol[2];
            S[0] = S[0] + contrib[0];
            S[1] = S[1] + contrib[1];
            S[2] = S[2] + contrib[2];
          }
        }
    }
  }
  for (int th = 0; th < nthreads; ++th) {
    const std::vector<double>& pv = partial[th];
    if (pv.empty()) continue;
    for (size_t j = 0; j < acc.size(); ++j) acc[j] += pv[j];
  }
  for (uint32_t i = 0; i < n; ++i) {
    float* o = projected_grad + 12 * (size_t)i;
    const double* a = &acc[(size_t)i * 9];
    o[0] = (float)a[0]; o[1] = (float)a[1]; o[2] = (float)a[2]; o[3] = 0.0f;
    o[4] = (float)a[3]; o[5] = (float)a[4]; o[6] = (float)a[5]; o[7] = 0.0f;
    o[8] = (float)a[6]; o[9] = (float)a[7]; o[10] = (float)a[8]; o[11] = 0.0f;
  }
}

// Recompute-mode backward (the Differentiation form): each pixel starts
// at its last contributing splat with the final transmittance saved by the
// forward, walks back to the slice start recovering T_i = T_{i+1}/(1 - a_i),
// and has no 1.0001 bail.  Per-pixel terms otherwise as
// rasterize_backward.wgsl:130-200; sums accumulated in float64.
void orc_composite_backward_recompute(const float* projected, const uint32_t* payloads, const uint32_t* ranges,
                                      const orc_tiling* t, uint32_t n, const float* final_t,
                                      const uint32_t* last_idx, const float* dl_dc, float* projected_grad) {
  const uint32_t ts = t->tile_size;
  const uint32_t tiles_x = (t->width + ts - 1) / ts;
  const uint32_t tiles_y = (t->height + ts - 1) / ts;
  const int64_t ntiles = (int64_t)tiles_x * tiles_y;
  int nthreads = orc_num_threads();
  std::vector<std::vector<double>> partial(nthreads);
#pragma omp parallel
  {
#ifdef _OPENMP
    const int tid = omp_get_thread_num();
#else
    const int tid = 0;
#endif
    std::vector<double>& mine = partial[tid];
    mine.assign((size_t)n * 9, 0.0);
#pragma omp for schedule(dynamic, 4)
    for (int64_t tile = 0; tile < ntiles; ++tile) {
      const uint32_t tx = (uint32_t)(tile % tiles_x), ty = (uint32_t)(tile / tiles_x);
      const uint32_t begin = ranges[2 * tile];
      for (uint32_t ly = 0; ly < ts; ++ly)
        for (uint32_t lx = 0; lx < ts; ++lx) {
          const uint32_t px = tx * ts + lx, py = ty * ts + ly;
          if (px >= t->width || py >= t->height) continue;
          const size_t pix = (size_t)py * t->width + px;
          const float pxf = (float)px + 0.5f, pyf = (float)py + 0.5f;
          float T = final_t[pix];
          const float gC[3] = {dl_dc[4 * pix + 0], dl_dc[4 * pix + 1], dl_dc[4 * pix + 2]};
          float S[3] = {0.0f, 0.0f, 0.0f};
          for (int64_t i = (int64_t)begin + (int64_t)last_idx[pix] - 1; i >= (int64_t)begin; --i) {
            const uint32_t sidx = payloads[i];
            const float* p = projected + 12 * (size_t)sidx;
            const float dx = pxf - p[0], dy = pyf - p[1];
            const float ca = p[4], cb = p[5], cc = p[6];
            const float power = nts_blend_power(ca, cb, cc, dx, dy);
            if (!(power <= 0.0f)) continue;  // also a NaN power
            const float g = nts_expf(power);
            const float sigma = p[11];
            const float ar = sigma * g;
            const float alpha = ar < NTS_ALPHA_MAX ? ar : NTS_ALPHA_MAX;  // a NaN clamps, as in the forward
            if (alpha < NTS_ALPHA_MIN) continue;
            const float oma = 1.0f - alpha;
            T = T / oma;
            const float ta = T * alpha;
            const float col[3] = {p[8], p[9], p[10]};
            const float contrib[3] = {ta * col[0], ta * col[1], ta * col[2]};
            const float dCa[3] = {T * col[0] - S[0] / oma, T * col[1] - S[1] / oma, T * col[2] - S[2] / oma};
            const float ga = gC[0] * dCa[0] + gC[1] * dCa[1] + gC[2] * dCa[2];
            const float sat = (ar < NTS_ALPHA_MAX) ? 1.0f : 0.0f;
            const float gar = ga * sat;
            const float gpow = gar * sigma * g;
            double* a = &mine[(size_t)sidx * 9];
            a[0] += gpow * (ca * dx + cb * dy);
            a[1] += gpow * (cb * dx + cc * dy);
            a[2] += gar * g;
            a[3] += gpow * -0.5f * (dx * dx);
            a[4] += gpow * -0.5f * (2.0f * dx * dy);
            a[5] += gpow * -0.5f * (dy * dy);
            a[6] += gC[0] * ta;
            a[7] += gC[1] * ta;
            a[8] += gC[2] * ta;
            S[0] = S[0] + contrib[0];
            S[1] = S[1] + contrib[1];
            S[2] = S[2] + contrib[2];
          }
        }
    }
  }
  std::vector<double> acc((size_t)n * 9, 0.0);
  for (int th = 0; th < nthreads; ++th)
    if (!partial[th].empty())
      for (size_t j = 0; j < acc.size(); ++j) acc[j] += partial[th][j];
  for (uint32_t i = 0; i < n; ++i) {
    float* o = projected_grad + 12 * (size_t)i;
    const double* a = &acc[(size_t)i * 9];
    o[0] = (float)a[0]; o[1] = (float)a[1]; o[2] = (float)a[2]; o[3] = 0.0f;
    o[4] = (float)a[3]; o[5] = (float)a[4]; o[6] = (float)a[5]; o[7] = 0.0f;
    o[8] = (float)a[6]; o[9] = (float)a[7]; o[10] = (float)a[8]; o[11] = 0.0f;
  }
}

// project_backward.wgsl:79-304.  With `full` set, the exact chain rule of
// the forward: the colour clamp gate, the SH direction's dependence on the
// position (d dir / d p) and the Jacobian's dependence on the camera-space
// mean (d J / d t), which the reference drops (project_backward.wgsl:17-24).
static void project_backward_impl(const orc_splats_f32* sp, uint32_t n, const orc_camera* cam,
                                  const float* projected_grad, orc_grads_f32* g, int full_chain) {
#pragma omp parallel for schedule(static)
  for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
    const uint32_t idx = (uint32_t)ii;
    float* gp = g->pos + 3 * (size_t)idx;
    float* gr = g->rot + 4 * (size_t)idx;
    float* gs = g->log_scale + 3 * (size_t)idx;
    float* go = g->logit + idx;
    float* gd = g->sh_dc + 3 * (size_t)idx;
    float* grest = g->sh_rest + 45 * (size_t)idx;
    const float* pw = sp->pos + 3 * (size_t)idx;
    const V3<float> cx = to_camera<float>(cam, pw);
    const float depth = -cx.z;
    auto zero_all = [&]() {
      for (int k = 0; k < 3; ++k) gp[k] = gs[k] = gd[k] = 0.0f;
      for (int k = 0; k < 4; ++k) gr[k] = 0.0f;
      *go = 0.0f;
      for (int k = 0; k < 45; ++k) grest[k] = 0.0f;
    };
    if (depth < cam->near_) { zero_all(); continue; }
    const V3<float> t{cx.x, cx.y, depth};
    const float inv_z = 1.0f / t.z;
    const float inv_z2 = inv_z * inv_z;
    const float f = cam->focal;
    const M3<float> R = quat_to_mat3<float>(sp->rot + 4 * (size_t)idx);
    const float* ls = sp->log_scale + 3 * (size_t)idx;
    const float sg[3] = {nts_expf(ls[0]), nts_expf(ls[1]), nts_expf(ls[2])};
    const float s2[3] = {sg[0] * sg[0], sg[1] * sg[1], sg[2] * sg[2]};
    M3<float> Ms;
    for (int col = 0; col < 3; ++col)
      for (int row = 0; row < 3; ++row) Ms.c[col][row] = R.c[col][row] * s2[col];
    const M3<float> cov_world = mul(Ms, transpose(R));
    const M3<float> Wm = view_rot<float>(cam);
    const M3<float> cov_cam = mul(mul(Wm, cov_world), transpose(Wm));
    const V3<float> j0{f * inv_z, 0.0f, -f * t.x * inv_z2};
    const V3<float> j1{0.0f, -f * inv_z, f * t.y * inv_z2};
    V3<float> r0, r1;
    r0.x = cov_cam.c[0][0] * j0.x + cov_cam.c[1][0] * j0.y + cov_cam.c[2][0] * j0.z;
    r0.y = cov_cam.c[0][1] * j0.x + cov_cam.c[1][1] * j0.y + cov_cam.c[2][1] * j0.z;
    r0.z = cov_cam.c[0][2] * j0.x + cov_cam.c[1][2] * j0.y + cov_cam.c[2][2] * j0.z;
    r1.x = cov_cam.c[0][0] * j1.x + cov_cam.c[1][0] * j1.y + cov_cam.c[2][0] * j1.z;
    r1.y = cov_cam.c[0][1] * j1.x + cov_cam.c[1][1] * j1.y + cov_cam.c[2][1] * j1.z;
    r1.z = cov_cam.c[0][2] * j1.x + cov_cam.c[1][2] * j1.y + cov_cam.c[2][2] * j1.z;
    const float s00 = dot3(r0, j0) + NTS_COV_DILATION;
    const float s01 = dot3(r0, j1);
    const float s11 = dot3(r1, j1) + NTS_COV_DILATION;
    const float det = s00 * s11 - s01 * s01;
    if (det <= 0.0f) { zero_all(); continue; }
    const float inv_det = 1.0f / det;
    const float ca = s11 * inv_det, cb = -s01 * inv_det, cc = s00 * inv_det;

    const float* pg = projected_grad + 12 * (size_t)idx;
    const float dmx = pg[0], dmy = pg[1], dop = pg[2];
    const float dca = pg[4], dcb = pg[5], dcc = pg[6];
    const float dcol[3] = {pg[8], pg[9], pg[10]};

    const float os = 1.0f / (1.0f + nts_expf(-sp->logit[idx]));
    *go = dop * os * (1.0f - os);

    V3<float> d{cam->cam_pos[0] - pw[0], cam->cam_pos[1] - pw[1], cam->cam_pos[2] - pw[2]};
    const float inv_len = 1.0f / __builtin_sqrtf(dot3(d, d));
    d.x = d.x * inv_len; d.y = d.y * inv_len; d.z = d.z * inv_len;
    float b[15];
    sh_basis<float>(d.x, d.y, d.z, b);
    const float* dcp = sp->sh_dc + 3 * (size_t)idx;
    const float* rest = sp->sh_rest + 45 * (size_t)idx;
    float dcu[3] = {dcol[0], dcol[1], dcol[2]};
    if (full_chain) {  // max(SH + 0.5, 0) passes no gradient where it clamps
      for (int ch = 0; ch < 3; ++ch) {
        float acc = 0.0f;
        for (int k = 0; k < 15; ++k) acc = acc + rest[15 * ch + k] * b[k];
        const float raw = (NTS_SH_C0 * dcp[ch] + acc) + 0.5f;
        if (!(raw > 0.0f)) dcu[ch] = 0.0f;
      }
    }
    for (int ch = 0; ch < 3; ++ch) gd[ch] = dcu[ch] * NTS_SH_C0;
    for (int k = 0; k < 15; ++k) {
      grest[k] = dcu[0] * b[k];
      grest[15 + k] = dcu[1] * b[k];
      grest[30 + k] = dcu[2] * b[k];
    }

    // dSigma2d = -(cm * dconic_mat * cm), symmetric halving (203-217).
    const float cm[2][2] = {{ca, cb}, {cb, cc}};                 // [col][row]
    const float dm[2][2] = {{dca, dcb * 0.5f}, {dcb * 0.5f, dcc}};
    float t1[2][2], full[2][2];
    for (int col = 0; col < 2; ++col)
      for (int row = 0; row < 2; ++row) t1[col][row] = cm[0][row] * dm[col][0] + cm[1][row] * dm[col][1];
    for (int col = 0; col < 2; ++col)
      for (int row = 0; row < 2; ++row)
        full[col][row] = -1.0f * (t1[0][row] * cm[col][0] + t1[1][row] * cm[col][1]);
    const float ds00 = full[0][0], ds11 = full[1][1];
    const float ds01 = full[0][1] + full[1][0];
    const float dsig[2][2] = {{ds00, ds01 * 0.5f}, {ds01 * 0.5f, ds11}};
    float d00 = 0.0f, d01 = 0.0f, d02 = 0.0f, d11 = 0.0f, d12 = 0.0f, d22 = 0.0f;
    for (int a = 0; a < 2; ++a)
      for (int bb = 0; bb < 2; ++bb) {
        const float coef = dsig[bb][a];
        const V3<float> ja = (a == 0) ? j0 : j1;
        const V3<float> jb = (bb == 0) ? j0 : j1;
        d00 = d00 + ja.x * coef * jb.x;
        d01 = d01 + ja.x * coef * jb.y;
        d02 = d02 + ja.x * coef * jb.z;
        d11 = d11 + ja.y * coef * jb.y;
        d12 = d12 + ja.y * coef * jb.z;
        d22 = d22 + ja.z * coef * jb.z;
      }
    M3<float> dcov_cam;
    dcov_cam.c[0][0] = d00; dcov_cam.c[0][1] = d01; dcov_cam.c[0][2] = d02;
    dcov_cam.c[1][0] = d01; dcov_cam.c[1][1] = d11; dcov_cam.c[1][2] = d12;
    dcov_cam.c[2][0] = d02; dcov_cam.c[2][1] = d12; dcov_cam.c[2][2] = d22;
    const M3<float> dcw = mul(mul(transpose(Wm), dcov_cam), Wm);
    M3<float> RS2;
    for (int col = 0; col < 3; ++col)
      for (int row = 0; row < 3; ++row) RS2.c[col][row] = R.c[col][row] * s2[col];
    M3<float> sym;
    const M3<float> dcwT = transpose(dcw);
    for (int col = 0; col < 3; ++col)
      for (int row = 0; row < 3; ++row) sym.c[col][row] = dcw.c[col][row] + dcwT.c[col][row];
    const M3<float> dR = mul(sym, RS2);
    const float G00 = dR.c[0][0], G10 = dR.c[0][1], G20 = dR.c[0][2];
    const float G01 = dR.c[1][0], G11 = dR.c[1][1], G21 = dR.c[1][2];
    const float G02 = dR.c[2][0], G12 = dR.c[2][1], G22 = dR.c[2][2];
    for (int k = 0; k < 3; ++k) {
      const V3<float> rk{R.c[k][0], R.c[k][1], R.c[k][2]};
      const float ds2 = dot3(rk, mulv(dcw, rk));
      gs[k] = ds2 * 2.0f * s2[k];
    }
    const float* q = sp->rot + 4 * (size_t)idx;
    const float qw = q[0], qx = q[1], qy = q[2], qz = q[3];
    gr[0] = 2.0f * (qz * (G10 - G01) + qy * (G02 - G20) + qx * (G21 - G12));
    gr[1] = 2.0f * (qy * (G01 + G10) + qz * (G02 + G20) + qw * (G21 - G12)) - 4.0f * qx * (G11 + G22);
    gr[2] = -4.0f * qy * (G00 + G22) + 2.0f * (qx * (G01 + G10) + qw * (G02 - G20) + qz * (G12 + G21));
    gr[3] = -4.0f * qz * (G00 + G11) + 2.0f * (qw * (G10 - G01) + qx * (G02 + G20) + qy * (G12 + G21));

    const V3<float> dtx{f * inv_z, 0.0f, -f * t.x * inv_z2};
    const V3<float> dty{0.0f, -f * inv_z, f * t.y * inv_z2};
    V3<float> dLdt{dtx.x * dmx + dty.x * dmy, dtx.y * dmx + dty.y * dmy, dtx.z * dmx + dty.z * dmy};
    if (full_chain) {
      // dL/dJ = 2 dSigma2d J Sigma_cam, with (J Sigma_cam) rows r0, r1
      const float h01 = ds01 * 0.5f;
      const V3<float> dJ0{2.0f * (ds00 * r0.x + h01 * r1.x), 2.0f * (ds00 * r0.y + h01 * r1.y),
                          2.0f * (ds00 * r0.z + h01 * r1.z)};
      const V3<float> dJ1{2.0f * (h01 * r0.x + ds11 * r1.x), 2.0f * (h01 * r0.y + ds11 * r1.y),
                          2.0f * (h01 * r0.z + ds11 * r1.z)};
      const float fz2 = f * inv_z2;
      const float fz3 = fz2 * inv_z;
      const float jtx = dJ0.z * -fz2;
      const float jty = dJ1.z * fz2;
      const float jtz = dJ0.x * -fz2 + dJ0.z * (2.0f * t.x * fz3) + dJ1.y * fz2 + dJ1.z * (-2.0f * t.y * fz3);
      dLdt.x = dLdt.x + jtx;
      dLdt.y = dLdt.y + jty;
      dLdt.z = dLdt.z + jtz;
    }
    const V3<float> dLdcam{dLdt.x, dLdt.y, -dLdt.z};
    V3<float> dp = mulv(transpose(Wm), dLdcam);
    if (full_chain) {
      // colour through the view direction: dir = (cam - p) / |cam - p|
      float db[15][3];
      sh_basis_grad(d.x, d.y, d.z, db);
      V3<float> v{0.0f, 0.0f, 0.0f};
      for (int k = 0; k < 15; ++k) {
        const float wk = dcu[0] * rest[k] + dcu[1] * rest[15 + k] + dcu[2] * rest[30 + k];
        v.x = v.x + wk * db[k][0];
        v.y = v.y + wk * db[k][1];
        v.z = v.z + wk * db[k][2];
      }
      const float pr = dot3(d, v);
      dp.x = dp.x - (v.x - d.x * pr) * inv_len;
      dp.y = dp.y - (v.y - d.y * pr) * inv_len;
      dp.z = dp.z - (v.z - d.z * pr) * inv_len;
    }
    gp[0] = dp.x; gp[1] = dp.y; gp[2] = dp.z;
  }
}

void orc_project_backward(const orc_splats_f32* sp, uint32_t n, const orc_camera* cam,
                          const float* projected_grad, orc_grads_f32* g) {
  project_backward_impl(sp, n, cam, projected_grad, g, 0);
}

void orc_project_backward_full(const orc_splats_f32* sp, uint32_t n, const orc_camera* cam,
                               const float* projected_grad, orc_grads_f32* g) {
  project_backward_impl(sp, n, cam, projected_grad, g, 1);
}

float orc_powi_f32(float a, int32_t b) { return powi_f32(a, b); }

// adam.rs:70-90; bias corrections computed once per step (adam_gpu.rs:275-276).
void orc_adam_step(float* params, const float* grads, float* m, float* v, uint64_t n, float lr,
                   float beta1, float beta2, float eps, uint64_t t) {
  const float bc1 = 1.0f - powi_f32(beta1, (int32_t)t);
  const float bc2 = 1.0f - powi_f32(beta2, (int32_t)t);
  for (uint64_t i = 0; i < n; ++i) {
    const float g = grads[i];
    const float mn = beta1 * m[i] + (1.0f - beta1) * g;
    const float vn = beta2 * v[i] + (1.0f - beta2) * g * g;
    m[i] = mn;
    v[i] = vn;
    const float mh = mn / bc1;
    const float vh = vn / bc2;
    params[i] = params[i] - lr * mh / (__builtin_sqrtf(vh) + eps);
  }
}

// apply_constraints.wgsl:60-77.
void orc_apply_constraints(float* rot, float* log_scale, float* logit, uint32_t n, float smin, float smax,
                           float omax) {
  for (uint32_t i = 0; i < n; ++i) {
    float* q = rot + 4 * (size_t)i;
    const float n2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    const float nrm = nts_maxf(__builtin_sqrtf(n2), 1e-8f);
    for (int k = 0; k < 4; ++k) q[k] = q[k] / nrm;
    for (int k = 0; k < 3; ++k) {
      float s = log_scale[3 * (size_t)i + k];
      s = nts_minf(nts_maxf(s, smin), smax);
      log_scale[3 * (size_t)i + k] = s;
    }
    logit[i] = nts_minf(nts_maxf(logit[i], -omax), omax);
  }
}

// accumulate_grad_pos.wgsl:17-23.
void orc_accumulate_grad(const float* dpos, float* acc, uint32_t n) {
  for (uint32_t i = 0; i < n; ++i)
    acc[i] = acc[i] + fabsf(dpos[3 * (size_t)i]) + fabsf(dpos[3 * (size_t)i + 1]) + fabsf(dpos[3 * (size_t)i + 2]);
}

double orc_loss_grad(const float* pred, const float* target, uint32_t w, uint32_t h, int kind, float* dl_dc) {
  const float n_scalars = (float)w * (float)h * 3.0f;
  const uint64_t np = (uint64_t)w * h;
  double sum = 0.0;
  if (kind == 1) {  // MSE, loss.rs:163-179
    const float inv_n = 2.0f / n_scalars;
    for (uint64_t i = 0; i < np; ++i) {
      float d[3];
      for (int c = 0; c < 3; ++c) d[c] = pred[4 * i + c] - target[4 * i + c];
      sum += (double)(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      for (int c = 0; c < 3; ++c) dl_dc[4 * i + c] = d[c] * inv_n;
      dl_dc[4 * i + 3] = 0.0f;
    }
  } else {  // L1: mean |d| over 3WH, d/dC = sign(d)/(3WH)
    const float inv_n = 1.0f / n_scalars;
    for (uint64_t i = 0; i < np; ++i) {
      for (int c = 0; c < 3; ++c) {
        const float d = pred[4 * i + c] - target[4 * i + c];
        sum += (double)fabsf(d);
        dl_dc[4 * i + c] = d > 0.0f ? inv_n : (d < 0.0f ? -inv_n : 0.0f);
      }
      dl_dc[4 * i + 3] = 0.0f;
    }
  }
  return sum / ((double)np * 3.0);
}

// -- prune / densify (train.rs:414-602) ------------------------------------

namespace {
const int kStride[6] = {3, 4, 3, 1, 3, 45};

float* attr_ptr(orc_splats_mut* s, int a) {
  switch (a) {
    case 0: return s->pos;
    case 1: return s->rot;
    case 2: return s->log_scale;
    case 3: return s->logit;
    case 4: return s->sh_dc;
    default: return s->sh_rest;
  }
}

// SplatBuffer::swap_remove + AdamState::swap_remove_range + grad_acc swap/pop.
void swap_remove(orc_splats_mut* s, uint32_t n, uint32_t i, float** moments, float* grad_acc) {
  const uint32_t last = n - 1;
  for (int a = 0; a < 6; ++a) {
    const int st = kStride[a];
    float* p = attr_ptr(s, a);
    for (int k = 0; k < st; ++k) p[(size_t)i * st + k] = p[(size_t)last * st + k];
    if (moments)
      for (int mm = 0; mm < 2; ++mm) {
        float* q = moments[2 * a + mm];
        for (int k = 0; k < st; ++k) q[(size_t)i * st + k] = q[(size_t)last * st + k];
      }
  }
  if (grad_acc) grad_acc[i] = grad_acc[last];
}

void push_splat(orc_splats_mut* s, uint32_t n, const float* pos, const float* rot, const float* scale,
                float logit, const float* dc, const float* rest, float** moments, float* grad_acc) {
  const float* src[6] = {pos, rot, scale, &logit, dc, rest};
  for (int a = 0; a < 6; ++a) {
    const int st = kStride[a];
    float* p = attr_ptr(s, a);
    for (int k = 0; k < st; ++k) p[(size_t)n * st + k] = src[a][k];
    if (moments)
      for (int mm = 0; mm < 2; ++mm) {
        float* q = moments[2 * a + mm];
        for (int k = 0; k < st; ++k) q[(size_t)n * st + k] = 0.0f;
      }
  }
  if (grad_acc) grad_acc[n] = 0.0f;
}

// train.rs:552-564 with the shared scalar functions (nts_box_muller): the
// logarithm and cosine bits of the reference's libm are not reproducible,
// so both sides use the same restated functions.
float box_muller(uint32_t* st) { return nts_box_muller(st); }
}  // namespace

uint32_t orc_prune(orc_splats_mut* s, uint32_t n, float threshold, float** moments, float* grad_acc) {
  // Indices collected ascending, applied high-to-low (train.rs:586-601).
  std::vector<uint32_t> rm;
  for (uint32_t i = 0; i < n; ++i)
    if (s->logit[i] < threshold) rm.push_back(i);
  for (size_t j = rm.size(); j-- > 0;) {
    swap_remove(s, n, rm[j], moments, grad_acc);
    --n;
  }
  return n;
}

uint32_t orc_densify(orc_splats_mut* s, uint32_t n, uint32_t grad_count, float grad_threshold,
                     float split_thr, uint32_t max_splats, uint32_t iter, float** moments, float* grad_acc,
                     uint32_t counts_out[3]) {
  const float count_f = (float)(grad_count > 1 ? grad_count : 1);
  std::vector<uint32_t> to_clone, to_split;
  uint32_t projected = n;
  for (uint32_t i = 0; i < n; ++i) {
    const float avg = grad_acc[i] / count_f;
    if (avg <= grad_threshold) continue;
    const float* ls = s->log_scale + 3 * (size_t)i;
    float mx = -INFINITY;
    for (int k = 0; k < 3; ++k) mx = nts_maxf(mx, ls[k]);
    if (mx > split_thr) {
      if (projected < max_splats) { to_split.push_back(i); ++projected; }
    } else if (projected < max_splats) {
      to_clone.push_back(i);
      ++projected;
    }
  }
  uint32_t rng = iter * 1664525u + 1013904223u;
  const float half = NTS_LN_HALF;          // -LN_2 as f32 (train.rs:462)
  const float log_split = NTS_NEG_LN_1_6;  // -(1.6f32).ln() (train.rs:463)
  for (uint32_t i : to_clone) {
    const float op_new = s->logit[i] + half;
    float pos[3], rot[4], sc[3], dc[3], rest[45];
    memcpy(pos, s->pos + 3 * (size_t)i, 12);
    memcpy(rot, s->rot + 4 * (size_t)i, 16);
    memcpy(sc, s->log_scale + 3 * (size_t)i, 12);
    memcpy(dc, s->sh_dc + 3 * (size_t)i, 12);
    memcpy(rest, s->sh_rest + 45 * (size_t)i, 180);
    push_splat(s, n, pos, rot, sc, op_new, dc, rest, moments, grad_acc);
    ++n;
    s->logit[i] = op_new;
  }
  std::vector<uint32_t> parents;
  for (uint32_t i : to_split) {
    float pos[3], rot[4], sc[3], dc[3], rest[45];
    memcpy(pos, s->pos + 3 * (size_t)i, 12);
    memcpy(rot, s->rot + 4 * (size_t)i, 16);
    memcpy(sc, s->log_scale + 3 * (size_t)i, 12);
    memcpy(dc, s->sh_dc + 3 * (size_t)i, 12);
    memcpy(rest, s->sh_rest + 45 * (size_t)i, 180);
    const float op = s->logit[i];
    const float nsc[3] = {sc[0] + log_split, sc[1] + log_split, sc[2] + log_split};
    const float sg[3] = {nts_expf(sc[0]), nts_expf(sc[1]), nts_expf(sc[2])};
    // glam Quat (x, y, z, w) * Vec3: v + 2w(q x v) + 2 q x (q x v)
    const float qw = rot[0], qx = rot[1], qy = rot[2], qz = rot[3];
    for (int child = 0; child < 2; ++child) {
      float off[3];
      for (int k = 0; k < 3; ++k) off[k] = box_muller(&rng) * sg[k];
      const float tx = 2.0f * (qy * off[2] - qz * off[1]);
      const float ty = 2.0f * (qz * off[0] - qx * off[2]);
      const float tz = 2.0f * (qx * off[1] - qy * off[0]);
      const float rx = off[0] + qw * tx + (qy * tz - qz * ty);
      const float ry = off[1] + qw * ty + (qz * tx - qx * tz);
      const float rz = off[2] + qw * tz + (qx * ty - qy * tx);
      const float np[3] = {pos[0] + rx, pos[1] + ry, pos[2] + rz};
      push_splat(s, n, np, rot, nsc, op, dc, rest, moments, grad_acc);
      ++n;
    }
    parents.push_back(i);
  }
  std::sort(parents.begin(), parents.end(), [](uint32_t a, uint32_t b) { return a > b; });
  for (uint32_t i : parents) {
    swap_remove(s, n, i, moments, grad_acc);
    --n;
  }
  if (counts_out) {
    counts_out[0] = (uint32_t)to_clone.size();
    counts_out[1] = (uint32_t)to_split.size() * 2;
    counts_out[2] = (uint32_t)parents.size();
  }
  return n;
}

// ---- local central differences for the large-scene gradient check ----------
// The mean-square loss L = sum (C - target)^2 / (3WH) depends on one splat's
// parameters only through the tiles its record is binned to, so a central
// difference needs only those tiles re-blended (in float64, with the splat's
// record projected at theta +- h and every other splat at its float64
// record), over the fixed float32 binning that the device gradient uses.
// A parameter whose +-h evaluations change the splat's tile set, depth
// bucket or visibility, or any pixel's sequence of contributing splats,
// blend termination or alpha clamp, is reported as discontinuous (1); one
// whose float64 pattern differs from the float32 forward's (the device's) on
// any of those pixels is reported as 2.
namespace {

uint32_t depth16_of(double depth, float depth_max) {
  float dn = (float)depth / depth_max;
  dn = dn < 0.0f ? 0.0f : (dn > 1.0f ? 1.0f : dn);
  return (uint32_t)(dn * 65535.0f);
}

// blend of one pixel over [begin, end) with splat `idx` drawn from `rec`;
// returns sum_c (C_c - target_c)^2 and a signature of the contributions
double pixel_sq_err(const double* proj64, const uint32_t* payloads, uint32_t begin, uint32_t end, uint32_t idx,
                    const double* rec, double pxf, double pyf, const float* tgt, const float* w, uint64_t* sig,
                    bool* touched) {
  *touched = false;
  double C[3] = {0.0, 0.0, 0.0}, Tr = 1.0;
  uint64_t h = 1469598103934665603ull;
  for (uint32_t i = begin; i < end; ++i) {
    const uint32_t sidx = payloads[i];
    const double* p = sidx == idx ? rec : proj64 + 12 * (size_t)sidx;
    const double dx = pxf - p[0], dy = pyf - p[1];
    const double power = -0.5 * (p[4] * dx * dx + 2.0 * p[5] * dx * dy + p[6] * dy * dy);
    if (!(power <= 0.0)) continue;  // also a NaN power
    const double ar = p[11] * exp(power);
    const bool clamped = !(ar < (double)NTS_ALPHA_MAX);  // a NaN clamps, as in the forward
    const double alpha = clamped ? (double)NTS_ALPHA_MAX : ar;
    if (alpha < (double)NTS_ALPHA_MIN) continue;
    h = (h ^ ((uint64_t)i * 2 + (clamped ? 1 : 0))) * 1099511628211ull;
    if (sidx == idx) *touched = true;
    const double ta = Tr * alpha;
    C[0] += ta * p[8];
    C[1] += ta * p[9];
    C[2] += ta * p[10];
    Tr = Tr * (1.0 - alpha);
    if (Tr < (double)NTS_T_MIN) {
      h = (h ^ 0xABCDull) * 1099511628211ull;
      break;
    }
  }
  *sig = h;
  double e = 0.0;
  for (int c = 0; c < 3; ++c) {
    if (w) {
      e += (double)w[c] * C[c];  // the linear functional <w, C>
    } else {
      const double d = C[c] - (double)tgt[c];
      e += d * d;
    }
  }
  return e;
}

// Contribution signature of one pixel under the float32 blend (the
// device's forward, bit for bit).
uint64_t pixel_sig_f32(const float* proj32, const uint32_t* payloads, uint32_t begin, uint32_t end, float pxf,
                       float pyf, uint32_t idx, bool* touched) {
  *touched = false;
  float Tr = 1.0f;
  uint64_t h = 1469598103934665603ull;
  for (uint32_t i = begin; i < end; ++i) {
    const float* p = proj32 + 12 * (size_t)payloads[i];
    const float dx = pxf - p[0], dy = pyf - p[1];
    const float power = nts_blend_power(p[4], p[5], p[6], dx, dy);
    if (!(power <= 0.0f)) continue;  // also a NaN power
    const float ar = p[11] * nts_expf(power);
    const bool clamped = !(ar < NTS_ALPHA_MAX);
    const float alpha = clamped ? NTS_ALPHA_MAX : ar;
    if (alpha < NTS_ALPHA_MIN) continue;
    h = (h ^ ((uint64_t)i * 2 + (clamped ? 1 : 0))) * 1099511628211ull;
    if (payloads[i] == idx) *touched = true;
    Tr = Tr * (1.0f - alpha);
    if (Tr < NTS_T_MIN) {
      h = (h ^ 0xABCDull) * 1099511628211ull;
      break;
    }
  }
  return h;
}

const int kWidthOf[6] = {3, 4, 3, 1, 3, 45};

}  // namespace

int orc_fd_local_mse(const orc_splats_f64* s, uint32_t n, const orc_camera* cam, const orc_tiling* t,
                     const double* proj64, const float* proj32, const uint32_t* payloads, const uint32_t* ranges,
                     const float* target, const float* weights, uint32_t idx, int attr, int comp, double h,
                     double* fd) {
  *fd = 0.0;
  if (idx >= n || attr < 0 || attr > 5 || comp < 0 || comp >= kWidthOf[attr]) return -1;
  // this splat's parameters, perturbed
  double pos[3], rot[4], ls[3], lg, dc[3], rest[45];
  double* dst[6] = {pos, rot, ls, &lg, dc, rest};
  const double* src[6] = {s->pos, s->rot, s->log_scale, s->logit, s->sh_dc, s->sh_rest};
  double rec[2][12];
  for (int sign = 0; sign < 2; ++sign) {
    for (int a = 0; a < 6; ++a)
      for (int k = 0; k < kWidthOf[a]; ++k) dst[a][k] = src[a][(size_t)kWidthOf[a] * idx + k];
    dst[attr][comp] += sign == 0 ? h : -h;
    const Splats<double> one{pos, rot, ls, &lg, dc, rest};
    project_one<double>(one, 0, cam, rec[sign]);
  }
  // the fixed binning of this splat (float32 record) and its perturbed bins
  const float* b32 = proj32 + 12 * (size_t)idx;
  int32_t x0, y0, x1, y1;
  const bool vis = tile_bbox<float>(b32, t, &x0, &y0, &x1, &y1);
  int disc = 0;
  for (int sign = 0; sign < 2; ++sign) {
    float r32[12];
    for (int k = 0; k < 12; ++k) r32[k] = (float)rec[sign][k];
    int32_t a0, b0, a1, b1;
    const bool v = tile_bbox<float>(r32, t, &a0, &b0, &a1, &b1);
    if (v != vis || (v && (a0 != x0 || b0 != y0 || a1 != x1 || b1 != y1))) disc = 1;
    if (v && depth16_of(rec[sign][2], t->depth_max) != depth16_of((double)b32[2], t->depth_max)) disc = 1;
  }
  if (!vis) return disc;  // contributes nowhere at theta
  const uint32_t ts = t->tile_size, tiles_x = (t->width + ts - 1) / ts;
  double sum[2] = {0.0, 0.0};
  for (int32_t ty = y0; ty <= y1; ++ty)
    for (int32_t tx = x0; tx <= x1; ++tx) {
      const uint32_t tile = (uint32_t)ty * tiles_x + (uint32_t)tx;
      const uint32_t begin = ranges[2 * tile], end = ranges[2 * tile + 1];
      for (uint32_t ly = 0; ly < ts; ++ly)
        for (uint32_t lx = 0; lx < ts; ++lx) {
          const uint32_t px = tx * ts + lx, py = ty * ts + ly;
          if (px >= t->width || py >= t->height) continue;
          const double pxf = (double)((float)px + 0.5f), pyf = (double)((float)py + 0.5f);
          const float* tg = target + 4 * ((size_t)py * t->width + px);
          const float* wp = weights ? weights + 4 * ((size_t)py * t->width + px) : nullptr;
          uint64_t sg[2];
          bool hit[3];
          for (int sign = 0; sign < 2; ++sign)
            sum[sign] +=
                pixel_sq_err(proj64, payloads, begin, end, idx, rec[sign], pxf, pyf, tg, wp, &sg[sign], &hit[sign]);
          // the float32 forward (the device's) must take the same pattern, or
          // the difference measures another active set than the gradient;
          // only pixels the splat reaches carry its derivative
          const uint64_t s32 = pixel_sig_f32(proj32, payloads, begin, end, (float)px + 0.5f, (float)py + 0.5f, idx,
                                             &hit[2]);
          if (!(hit[0] || hit[1] || hit[2])) continue;
          if (sg[0] != sg[1]) disc = 1;
          if (!disc && s32 != sg[0]) disc = 2;
        }
    }
  *fd = (sum[0] - sum[1]) / (2.0 * h);
  if (!weights) *fd /= 3.0 * (double)t->width * (double)t->height;
  return disc;
}

void orc_fd_local_mse_batch(const orc_splats_f64* s, uint32_t n, const orc_camera* cam, const orc_tiling* t,
                            const double* proj64, const float* proj32, const uint32_t* payloads,
                            const uint32_t* ranges, const float* target, const float* weights, uint32_t count,
                            const uint32_t* idx, const int32_t* attr, const int32_t* comp, const double* h,
                            double* fd, int32_t* disc) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t k = 0; k < (int64_t)count; ++k)
    disc[k] = orc_fd_local_mse(s, n, cam, t, proj64, proj32, payloads, ranges, target, weights, idx[k], attr[k],
                               comp[k], h[k], &fd[k]);
}

// ---- SSIM loss (loss.rs:23-415) ---------------------------------------------
// Per-channel SSIM over an 11-tap separable Gaussian (sigma 1.5, normalised,
// reflect-101 borders), C1 = 1e-4, C2 = 9e-4; loss = 1 - mean over pixels and
// channels; the gradient through the blur adjoint (a scatter with the same
// reflection map).  float32 arithmetic in the reference's loop orders; the
// kernel taps use the shared exponential.
namespace {

const int kSsimWin = 11, kSsimR = 5;

void ssim_kernel(float k[kSsimWin]) {
  const float centre = (float)(kSsimWin - 1) * 0.5f;
  const float two_sigma_sq = 2.0f * 1.5f * 1.5f;
  float sum = 0.0f;
  for (int i = 0; i < kSsimWin; ++i) {
    const float d = (float)i - centre;
    k[i] = nts_expf(-d * d / two_sigma_sq);
    sum += k[i];
  }
  for (int i = 0; i < kSsimWin; ++i) k[i] = k[i] / sum;
}

int reflect101(int i, int n) {
  if (n == 1) return 0;
  const int period = 2 * (n - 1);
  int x = i % period;
  if (x < 0) x += period;
  if (x >= n) x = period - x;
  return x;
}

void blur(const float* src, int w, int h, const float* k, float* dst) {
  std::vector<float> tmp((size_t)w * h);
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      float acc = 0.0f;
      for (int t = 0; t < kSsimWin; ++t) acc += src[(size_t)y * w + reflect101(x + t - kSsimR, w)] * k[t];
      tmp[(size_t)y * w + x] = acc;
    }
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      float acc = 0.0f;
      for (int t = 0; t < kSsimWin; ++t) acc += tmp[(size_t)reflect101(y + t - kSsimR, h) * w + x] * k[t];
      dst[(size_t)y * w + x] = acc;
    }
}

void blur_adjoint(const float* src, int w, int h, const float* k, float* dst) {
  std::vector<float> tmp((size_t)w * h, 0.0f);
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      const float sv = src[(size_t)y * w + x];
      for (int t = 0; t < kSsimWin; ++t) tmp[(size_t)reflect101(y + t - kSsimR, h) * w + x] += sv * k[t];
    }
  for (size_t i = 0; i < (size_t)w * h; ++i) dst[i] = 0.0f;
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      const float sv = tmp[(size_t)y * w + x];
      for (int t = 0; t < kSsimWin; ++t) dst[(size_t)y * w + reflect101(x + t - kSsimR, w)] += sv * k[t];
    }
}

// loss.rs:212-338 for one channel: returns 1 - mean(SSIM), grad = dL/dx
float ssim_channel(const float* x, const float* y, int w, int h, const float* k, float* grad) {
  const size_t n = (size_t)w * h;
  std::vector<float> xx(n), yy(n), xy(n), mx(n), my(n), gxx(n), gyy(n), gxy(n);
  for (size_t i = 0; i < n; ++i) {
    xx[i] = x[i] * x[i];
    yy[i] = y[i] * y[i];
    xy[i] = x[i] * y[i];
  }
  blur(x, w, h, k, mx.data());
  blur(y, w, h, k, my.data());
  blur(xx.data(), w, h, k, gxx.data());
  blur(yy.data(), w, h, k, gyy.data());
  blur(xy.data(), w, h, k, gxy.data());
  std::vector<float> P(n), Q(n), R(n), mxq(n), myr(n);
  double sum = 0.0;
  for (size_t i = 0; i < n; ++i) {
    const float mux = mx[i], muy = my[i];
    const float mux2 = mux * mux, muy2 = muy * muy;
    const float sxx = gxx[i] - mux2, syy = gyy[i] - muy2, sxy = gxy[i] - mux * muy;
    const float av = 2.0f * mux * muy + 1e-4f;
    const float bv = 2.0f * sxy + 9e-4f;
    const float cv = mux2 + muy2 + 1e-4f;
    const float dv = sxx + syy + 9e-4f;
    sum += (double)((av * bv) / (cv * dv));
    const float cd = cv * dv;
    P[i] = 2.0f * bv * (muy * cv - mux * av) / (cv * cd);
    Q[i] = -av * bv / (cv * dv * dv);
    R[i] = 2.0f * av / cd;
    mxq[i] = mux * Q[i];
    myr[i] = muy * R[i];
  }
  std::vector<float> sp(n), sq(n), sr(n), smq(n), smr(n);
  blur_adjoint(P.data(), w, h, k, sp.data());
  blur_adjoint(Q.data(), w, h, k, sq.data());
  blur_adjoint(R.data(), w, h, k, sr.data());
  blur_adjoint(mxq.data(), w, h, k, smq.data());
  blur_adjoint(myr.data(), w, h, k, smr.data());
  const float inv_n = 1.0f / (float)n;
  for (size_t i = 0; i < n; ++i) {
    const float d = sp[i] + 2.0f * (x[i] * sq[i] - smq[i]) + (y[i] * sr[i] - smr[i]);
    grad[i] = -inv_n * d;
  }
  return 1.0f - (float)(sum / (double)n);
}

}  // namespace

void orc_ssim_kernel(float* k) { ssim_kernel(k); }

void orc_ssim_blur(const float* src, int w, int h, float* dst, int adjoint) {
  float k[kSsimWin];
  ssim_kernel(k);
  if (adjoint) blur_adjoint(src, w, h, k, dst);
  else blur(src, w, h, k, dst);
}

// ssim_loss_grad (loss.rs:346-383): channels of float4 images, loss averaged
// over the three channels, grad = float4 (w = 0).
float orc_ssim_loss_grad(const float* pred, const float* target, int w, int h, float* grad) {
  float k[kSsimWin];
  ssim_kernel(k);
  const size_t n = (size_t)w * h;
  std::vector<float> px(n), tx(n), g(n);
  float loss = 0.0f;
  float l[3];
  for (int c = 0; c < 3; ++c) {
    for (size_t i = 0; i < n; ++i) {
      px[i] = pred[4 * i + c];
      tx[i] = target[4 * i + c];
    }
    l[c] = ssim_channel(px.data(), tx.data(), w, h, k, g.data());
    for (size_t i = 0; i < n; ++i) grad[4 * i + c] = g[i] / 3.0f;
  }
  for (size_t i = 0; i < n; ++i) grad[4 * i + 3] = 0.0f;
  loss = (l[0] + l[1] + l[2]) / 3.0f;
  return loss;
}

// tonemap.wgsl:27-41.
void orc_tonemap(const float* image, uint32_t w, uint32_t h, uint8_t* out) {
  for (uint64_t i = 0; i < (uint64_t)w * h; ++i) {
    for (int c = 0; c < 3; ++c) {
      const float x = image[4 * i + c];
      float tm = x / (1.0f + nts_maxf(x, 0.0f));
      tm = nts_minf(nts_maxf(tm, 0.0f), 1.0f);
      const float s = tm > 0.0031308f ? 1.055f * nts_srgb_pow(tm) - 0.055f : tm * 12.92f;
      out[4 * i + c] = (uint8_t)lrintf(nts_minf(nts_maxf(s, 0.0f), 1.0f) * 255.0f);
    }
    const float a = nts_minf(nts_maxf(image[4 * i + 3], 0.0f), 1.0f);
    out[4 * i + 3] = (uint8_t)lrintf(a * 255.0f);
  }
}


// Timing driver of the CPU baseline: the reference step sequence of
// train.rs:155-250 over n_views views of one scene (project, TileBinner::bin
// with its per-call allocation, composite with the saved state, L1 dL/dC,
// recompute backward, project_backward with the full chain rule), the views'
// gradients summed, then one Adam step on every attribute slab +
// constraints.  `warmup` untimed runs, then `reps` timed ones; each view's
// wall time of each timed run (steady_clock) goes to
// view_secs[rep * n_views + v], the Adam step's to adam_secs[rep].  threads >
// 0 sets the OpenMP thread count for the call (restored after).
int orc_time_step_views(const orc_splats_f32* s, uint32_t n, const orc_camera* cams, const orc_tiling* tilings,
                        const float* const* targets, uint32_t n_views, int warmup, int reps, int threads,
                        double* view_secs, double* adam_secs) {
  if (!s || !cams || !tilings || !targets || n_views == 0 || reps <= 0) return -1;
#ifdef _OPENMP
  const int prev_threads = omp_get_max_threads();
  if (threads > 0) omp_set_num_threads(threads);
#endif
  const size_t widths[6] = {3, 4, 3, 1, 3, 45};
  std::vector<float> total[6], g[6], params[6], m[6], v[6];
  const float* src[6] = {s->pos, s->rot, s->log_scale, s->logit, s->sh_dc, s->sh_rest};
  for (int a = 0; a < 6; ++a) {
    total[a].assign(widths[a] * n, 0.0f);
    g[a].assign(widths[a] * n, 0.0f);
    params[a].assign(src[a], src[a] + widths[a] * n);
    m[a].assign(widths[a] * n, 0.0f);
    v[a].assign(widths[a] * n, 0.0f);
  }
  orc_grads_f32 gg{g[0].data(), g[1].data(), g[2].data(), g[3].data(), g[4].data(), g[5].data()};
  std::vector<float> proj(12 * (size_t)n), pgrad(12 * (size_t)n);
  typedef std::chrono::steady_clock clk;
  for (int run = 0; run < warmup + reps; ++run) {
    for (int a = 0; a < 6; ++a) std::fill(total[a].begin(), total[a].end(), 0.0f);
    for (uint32_t vi = 0; vi < n_views; ++vi) {
      const auto t0 = clk::now();
      const orc_tiling& t = tilings[vi];
      const uint64_t np = (uint64_t)t.width * t.height;
      const uint32_t ntiles = ((t.width + t.tile_size - 1) / t.tile_size) * ((t.height + t.tile_size - 1) / t.tile_size);
      orc_project_f32(s, n, &cams[vi], proj.data());
      const uint64_t k = orc_bin(proj.data(), n, &t, nullptr, nullptr, nullptr);
      std::vector<uint32_t> keys(k > 0 ? k : 1), pay(k > 0 ? k : 1), ranges(2 * (size_t)ntiles);
      orc_bin(proj.data(), n, &t, keys.data(), pay.data(), ranges.data());
      std::vector<float> img(4 * np), ft(np), dl(4 * np);
      std::vector<uint32_t> li(np);
      orc_composite(proj.data(), pay.data(), ranges.data(), &t, img.data(), ft.data(), li.data());
      (void)orc_loss_grad(img.data(), targets[vi], t.width, t.height, 0, dl.data());
      orc_composite_backward_recompute(proj.data(), pay.data(), ranges.data(), &t, n, ft.data(), li.data(), dl.data(),
                                       pgrad.data());
      orc_project_backward_full(s, n, &cams[vi], pgrad.data(), &gg);
      for (int a = 0; a < 6; ++a) {
        float* dst = total[a].data();
        const float* gs = g[a].data();
        const int64_t cnt = (int64_t)(widths[a] * n);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < cnt; ++i) dst[i] = dst[i] + gs[i];
      }
      if (run >= warmup) view_secs[(size_t)(run - warmup) * n_views + vi] =
          std::chrono::duration<double>(clk::now() - t0).count();
    }
    const auto t0 = clk::now();
    for (int a = 0; a < 6; ++a) {
      const int64_t cnt = (int64_t)(widths[a] * n), chunk = 1 << 16;
#pragma omp parallel for schedule(static)
      for (int64_t c = 0; c < cnt; c += chunk) {
        const uint64_t len = (uint64_t)std::min<int64_t>(chunk, cnt - c);
        orc_adam_step(params[a].data() + c, total[a].data() + c, m[a].data() + c, v[a].data() + c, len, 1.6e-4f,
                      0.9f, 0.999f, 1e-15f, (uint64_t)run + 1);
      }
    }
    orc_apply_constraints(params[1].data(), params[2].data(), params[3].data(), n, -10.0f, 5.0f, 10.0f);
    if (run >= warmup) adam_secs[run - warmup] = std::chrono::duration<double>(clk::now() - t0).count();
  }
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(prev_threads);
#endif
  return 0;
}

}  // extern "C"
