// Device-side projection maths (EWA splatting with the rasteriser's
// conventions).  Operation order is fixed so that results are bit-identical
// to the host restatement compiled with -ffp-contract=off:
//   a*b*c == (a*b)*c, sums left to right, and a 3x3 product
//   (A*B)[c][r] = A[0][r]*B[c][0] + A[1][r]*B[c][1] + A[2][r]*B[c][2]
// with matrices stored column-major as m[col][row].
#pragma once
#include "nts_internal.h"

struct M3f {
  float c[3][3];
};

__device__ __forceinline__ M3f m3_mul(const M3f& a, const M3f& b) {
  M3f o;
#pragma unroll
  for (int col = 0; col < 3; ++col)
#pragma unroll
    for (int row = 0; row < 3; ++row)
      o.c[col][row] = a.c[0][row] * b.c[col][0] + a.c[1][row] * b.c[col][1] + a.c[2][row] * b.c[col][2];
  return o;
}

__device__ __forceinline__ M3f m3_t(const M3f& a) {
  M3f o;
#pragma unroll
  for (int col = 0; col < 3; ++col)
#pragma unroll
    for (int row = 0; row < 3; ++row) o.c[col][row] = a.c[row][col];
  return o;
}

__device__ __forceinline__ float3 m3_mulv(const M3f& a, float3 v) {
  return make_float3(a.c[0][0] * v.x + a.c[1][0] * v.y + a.c[2][0] * v.z,
                     a.c[0][1] * v.x + a.c[1][1] * v.y + a.c[2][1] * v.z,
                     a.c[0][2] * v.x + a.c[1][2] * v.y + a.c[2][2] * v.z);
}

__device__ __forceinline__ float dot3f(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Rotation from a (w, x, y, z) quaternion, columns as glam's Quat::to_mat3.
__device__ __forceinline__ M3f quat_to_m3(float w, float x, float y, float z) {
  M3f r;
  r.c[0][0] = 1.0f - 2.0f * (y * y + z * z);
  r.c[0][1] = 2.0f * (x * y + w * z);
  r.c[0][2] = 2.0f * (x * z - w * y);
  r.c[1][0] = 2.0f * (x * y - w * z);
  r.c[1][1] = 1.0f - 2.0f * (x * x + z * z);
  r.c[1][2] = 2.0f * (y * z + w * x);
  r.c[2][0] = 2.0f * (x * z + w * y);
  r.c[2][1] = 2.0f * (y * z - w * x);
  r.c[2][2] = 1.0f - 2.0f * (x * x + y * y);
  return r;
}

__device__ __forceinline__ void sh_basis15(float x, float y, float z, float b[15]) {
  const float xx = x * x, yy = y * y, zz = z * z;
  b[0] = -NTS_SH_C1 * y;
  b[1] = NTS_SH_C1 * z;
  b[2] = -NTS_SH_C1 * x;
  b[3] = NTS_SH_C2_0 * x * y;
  b[4] = NTS_SH_C2_1 * y * z;
  b[5] = NTS_SH_C2_2 * (2.0f * zz - xx - yy);
  b[6] = NTS_SH_C2_3 * x * z;
  b[7] = NTS_SH_C2_4 * (xx - yy);
  b[8] = NTS_SH_C3_0 * y * (3.0f * xx - yy);
  b[9] = NTS_SH_C3_1 * x * y * z;
  b[10] = NTS_SH_C3_2 * y * (4.0f * zz - xx - yy);
  b[11] = NTS_SH_C3_3 * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
  b[12] = NTS_SH_C3_4 * x * (4.0f * zz - xx - yy);
  b[13] = NTS_SH_C3_5 * z * (xx - yy);
  b[14] = NTS_SH_C3_6 * x * (xx - 3.0f * yy);
}

// d b_k / d (x, y, z) of the 15 basis functions of sh_basis15.
__device__ __forceinline__ void sh_basis_grad15(float x, float y, float z, float g[15][3]) {
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

__device__ __forceinline__ M3f cam_rot(const DevCamera& cam) {
  M3f w;
#pragma unroll
  for (int col = 0; col < 3; ++col)
#pragma unroll
    for (int row = 0; row < 3; ++row) w.c[col][row] = cam.view[col * 4 + row];
  return w;
}

// view * (p, 1), then divided by w.
__device__ __forceinline__ float3 to_cam(const DevCamera& cam, float px, float py, float pz) {
  const float* v = cam.view;
  const float o0 = v[0] * px + v[4] * py + v[8] * pz + v[12] * 1.0f;
  const float o1 = v[1] * px + v[5] * py + v[9] * pz + v[13] * 1.0f;
  const float o2 = v[2] * px + v[6] * py + v[10] * pz + v[14] * 1.0f;
  const float o3 = v[3] * px + v[7] * py + v[11] * pz + v[15] * 1.0f;
  return make_float3(o0 / o3, o1 / o3, o2 / o3);
}

// Screen-space covariance of one splat.  Returns false for a degenerate
// determinant.  Shared by the forward projection and its backward.
struct Cov2D {
  float s00, s01, s11, det;
  float3 j0, j1;
  float3 r0, r1;  // rows of J * Sigma_cam
  M3f R, W;
  float s2[3];
};

__device__ __forceinline__ bool screen_cov(const DevCamera& cam, float3 t, float qw, float qx, float qy, float qz,
                                           float ls0, float ls1, float ls2, Cov2D& o) {
  o.R = quat_to_m3(qw, qx, qy, qz);
  const float sg0 = nts_expf(ls0), sg1 = nts_expf(ls1), sg2 = nts_expf(ls2);
  o.s2[0] = sg0 * sg0;
  o.s2[1] = sg1 * sg1;
  o.s2[2] = sg2 * sg2;
  M3f Ms;
#pragma unroll
  for (int col = 0; col < 3; ++col)
#pragma unroll
    for (int row = 0; row < 3; ++row) Ms.c[col][row] = o.R.c[col][row] * o.s2[col];
  const M3f cov_w = m3_mul(Ms, m3_t(o.R));
  o.W = cam_rot(cam);
  const M3f cov_c = m3_mul(m3_mul(o.W, cov_w), m3_t(o.W));
  const float f = cam.focal;
  const float inv_z = 1.0f / t.z;
  const float inv_z2 = inv_z * inv_z;
  o.j0 = make_float3(f * inv_z, 0.0f, -f * t.x * inv_z2);
  o.j1 = make_float3(0.0f, -f * inv_z, f * t.y * inv_z2);
  float3& r0 = o.r0;
  float3& r1 = o.r1;
  r0.x = cov_c.c[0][0] * o.j0.x + cov_c.c[1][0] * o.j0.y + cov_c.c[2][0] * o.j0.z;
  r0.y = cov_c.c[0][1] * o.j0.x + cov_c.c[1][1] * o.j0.y + cov_c.c[2][1] * o.j0.z;
  r0.z = cov_c.c[0][2] * o.j0.x + cov_c.c[1][2] * o.j0.y + cov_c.c[2][2] * o.j0.z;
  r1.x = cov_c.c[0][0] * o.j1.x + cov_c.c[1][0] * o.j1.y + cov_c.c[2][0] * o.j1.z;
  r1.y = cov_c.c[0][1] * o.j1.x + cov_c.c[1][1] * o.j1.y + cov_c.c[2][1] * o.j1.z;
  r1.z = cov_c.c[0][2] * o.j1.x + cov_c.c[1][2] * o.j1.y + cov_c.c[2][2] * o.j1.z;
  o.s00 = dot3f(r0, o.j0) + NTS_COV_DILATION;
  o.s01 = dot3f(r0, o.j1);
  o.s11 = dot3f(r1, o.j1) + NTS_COV_DILATION;
  o.det = o.s00 * o.s11 - o.s01 * o.s01;
  return o.det > 0.0f;
}

// Tile bbox (tile_count.wgsl:39-50).  false when the bbox is empty.
__device__ __forceinline__ bool tile_bbox_dev(float mx, float my, float r, float ts, int32_t tiles_x, int32_t tiles_y,
                                              int32_t& x0, int32_t& y0, int32_t& x1, int32_t& y1) {
  x0 = nts_maxi(nts_f2i_sat(floorf(nts_div_tile(mx - r, ts))), 0);
  y0 = nts_maxi(nts_f2i_sat(floorf(nts_div_tile(my - r, ts))), 0);
  x1 = nts_mini(nts_f2i_sat(floorf(nts_div_tile(mx + r, ts))), tiles_x - 1);
  y1 = nts_mini(nts_f2i_sat(floorf(nts_div_tile(my + r, ts))), tiles_y - 1);
  return !(x1 < x0 || y1 < y0);
}

// The tile rectangle of a projected record: the reference's (CULL false) or
// that one cut to the tiles its alpha >= 1/255 ellipse can reach
// (nts_alpha_tile_rect).  false when empty.
template <bool CULL>
__device__ __forceinline__ bool record_tiles(const Projected& p, float ts, int32_t tiles_x, int32_t tiles_y,
                                             int32_t& x0, int32_t& y0, int32_t& x1, int32_t& y1) {
  if (p.b.w == 0.0f || !tile_bbox_dev(p.a.x, p.a.y, p.a.w, ts, tiles_x, tiles_y, x0, y0, x1, y1)) return false;
  return !CULL || nts_alpha_tile_rect(p.a.x, p.a.y, p.b.x, p.b.y, p.b.z, p.c.w, ts, &x0, &y0, &x1, &y1);
}

// 16-bit depth bucket (tile_emit.wgsl:109-110).
__device__ __forceinline__ uint32_t depth16(float depth, float depth_max) {
  float dn = depth / depth_max;
  dn = dn < 0.0f ? 0.0f : (dn > 1.0f ? 1.0f : dn);
  return (uint32_t)(dn * 65535.0f);
}
