// Per-splat projection and its backward as device functions over a splat's
// 59 parameters held in registers (shared by the single-view and the
// multi-view kernels).  Arithmetic is fixed to match the host restatement
// bit for bit (see nts_geom.h).
#pragma once
#include "nts_geom.h"

struct SplatP {
  float v[P_COUNT];
};

__device__ __forceinline__ void load_splat(const float* __restrict__ P, uint32_t cap, uint32_t i, SplatP& s) {
#pragma unroll
  for (int k = 0; k < P_COUNT; ++k) s.v[k] = P[(size_t)k * cap + i];
}

// Forward projection of one splat (project_gaussians.wgsl:156-259).  Returns
// an all-zero record for culled splats.
__device__ __forceinline__ Projected project_p(const SplatP& s, const DevCamera& cam) {
  Projected out;
  out.a = make_float4(0.f, 0.f, 0.f, 0.f);
  out.b = out.a;
  out.c = out.a;
  const float px = s.v[P_POS + 0];
  const float py = s.v[P_POS + 1];
  const float pz = s.v[P_POS + 2];
  const float3 cx = to_cam(cam, px, py, pz);
  const float depth = -cx.z;
  if (depth < cam.near_) return out;
  const float3 t = make_float3(cx.x, cx.y, depth);
  const float W = cam.width, H = cam.height, f = cam.focal;
  const float mx = W * 0.5f + (f * t.x) / t.z;
  const float my = H * 0.5f - (f * t.y) / t.z;
  Cov2D cv;
  if (!screen_cov(cam, t, s.v[P_ROT + 0], s.v[P_ROT + 1],
                  s.v[P_ROT + 2], s.v[P_ROT + 3],
                  s.v[P_SCALE + 0], s.v[P_SCALE + 1],
                  s.v[P_SCALE + 2], cv))
    return out;
  const float inv_det = 1.0f / cv.det;
  const float ca = cv.s11 * inv_det, cb = -cv.s01 * inv_det, cc = cv.s00 * inv_det;
  const float mid = 0.5f * (cv.s00 + cv.s11);
  const float disc = mid * mid - cv.det;
  const float lambda = mid + sqrtf(disc > 0.0f ? disc : 0.0f);
  const float radius = ceilf(3.0f * sqrtf(lambda));
  const float maxpix = (W > H ? W : H) + radius;
  if (mx < -radius || mx >= W + radius || my < -radius || my >= H + radius || radius >= maxpix) return out;

  float dx = cam.cam_pos[0] - px, dy = cam.cam_pos[1] - py, dz = cam.cam_pos[2] - pz;
  const float inv_len = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
  dx = dx * inv_len;
  dy = dy * inv_len;
  dz = dz * inv_len;
  float b[15];
  sh_basis15(dx, dy, dz, b);
  float rgb[3], acc[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) rgb[ch] = NTS_SH_C0 * s.v[P_DC + ch];
#pragma unroll
  for (int k = 0; k < 15; ++k)
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) acc[ch] = acc[ch] + s.v[P_REST + 15 * ch + k] * b[k];
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) {
    const float c = (rgb[ch] + acc[ch]) + 0.5f;
    rgb[ch] = c > 0.0f ? c : 0.0f;
  }
  const float alpha = nts_sigmoidf(s.v[P_LOGIT]);
  out.a = make_float4(mx, my, depth, radius);
  out.b = make_float4(ca, cb, cc, 1.0f);
  out.c = make_float4(rgb[0], rgb[1], rgb[2], alpha);
  return out;
}

// project_backward.wgsl:79-304: chain the 2-D gradients back to the splat
// parameters.  Writes every gradient plane of splat i.  FULL adds the terms
// the reference drops (project_backward.wgsl:17-24): the colour clamp gate,
// the view direction's dependence on the position and the Jacobian's
// dependence on the camera-space mean.
// ACC adds each term to g[k] (g[k] = g[k] + term) instead of storing it:
// a sum over views then needs no per-view temporary of 59 registers.
template <bool FULL, bool ACC = false>
__device__ __forceinline__ void project_backward_p(const SplatP& s, const DevCamera& cam, const ProjectedGrad& pg,
                                                   float (&g)[P_COUNT]) {
  if (!ACC) {
#pragma unroll
    for (int k = 0; k < P_COUNT; ++k) g[k] = 0.0f;
  }
  // store or accumulate one gradient term
  auto put = [&](int k, float v) { g[k] = ACC ? g[k] + v : v; };
  const float px = s.v[P_POS + 0];
  const float py = s.v[P_POS + 1];
  const float pz = s.v[P_POS + 2];
  const float3 cx = to_cam(cam, px, py, pz);
  const float depth = -cx.z;
  bool live = !(depth < cam.near_);
  const float3 t = make_float3(cx.x, cx.y, depth);
  const float qw = s.v[P_ROT + 0], qx = s.v[P_ROT + 1];
  const float qy = s.v[P_ROT + 2], qz = s.v[P_ROT + 3];
  Cov2D cv;
  if (live)
    live = screen_cov(cam, t, qw, qx, qy, qz, s.v[P_SCALE + 0], s.v[P_SCALE + 1],
                      s.v[P_SCALE + 2], cv);
  if (live) {
    const float f = cam.focal;
    const float inv_z = 1.0f / t.z;
    const float inv_z2 = inv_z * inv_z;
    const float inv_det = 1.0f / cv.det;
    const float ca = cv.s11 * inv_det, cb = -cv.s01 * inv_det, cc = cv.s00 * inv_det;
    const float dmx = pg.a.x, dmy = pg.a.y, dop = pg.a.z;
    const float dca = pg.b.x, dcb = pg.b.y, dcc = pg.b.z;
    float dcol[3] = {pg.c.x, pg.c.y, pg.c.z};

    const float os = nts_sigmoidf(s.v[P_LOGIT]);
    put(P_LOGIT, dop * os * (1.0f - os));

    float dx = cam.cam_pos[0] - px, dy = cam.cam_pos[1] - py, dz = cam.cam_pos[2] - pz;
    const float inv_len = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
    dx = dx * inv_len;
    dy = dy * inv_len;
    dz = dz * inv_len;
    float b[15];
    sh_basis15(dx, dy, dz, b);
    if (FULL) {  // max(SH + 0.5, 0) passes no gradient where it clamps
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) {
        float acc = 0.0f;
#pragma unroll
        for (int k = 0; k < 15; ++k) acc = acc + s.v[P_REST + 15 * ch + k] * b[k];
        const float raw = (NTS_SH_C0 * s.v[P_DC + ch] + acc) + 0.5f;
        if (!(raw > 0.0f)) dcol[ch] = 0.0f;
      }
    }
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) put(P_DC + ch, dcol[ch] * NTS_SH_C0);
#pragma unroll
    for (int k = 0; k < 15; ++k) {
      put(P_REST + k, dcol[0] * b[k]);
      put(P_REST + 15 + k, dcol[1] * b[k]);
      put(P_REST + 30 + k, dcol[2] * b[k]);
    }

    // dSigma2d = -(C dconic C) with the off-diagonal halved / summed.
    const float cm[2][2] = {{ca, cb}, {cb, cc}};
    const float dm[2][2] = {{dca, dcb * 0.5f}, {dcb * 0.5f, dcc}};
    float t1[2][2], full[2][2];
#pragma unroll
    for (int col = 0; col < 2; ++col)
#pragma unroll
      for (int row = 0; row < 2; ++row) t1[col][row] = cm[0][row] * dm[col][0] + cm[1][row] * dm[col][1];
#pragma unroll
    for (int col = 0; col < 2; ++col)
#pragma unroll
      for (int row = 0; row < 2; ++row)
        full[col][row] = -1.0f * (t1[0][row] * cm[col][0] + t1[1][row] * cm[col][1]);
    const float ds00 = full[0][0], ds11 = full[1][1];
    const float ds01 = full[0][1] + full[1][0];
    const float dsig[2][2] = {{ds00, ds01 * 0.5f}, {ds01 * 0.5f, ds11}};
    float d00 = 0.0f, d01 = 0.0f, d02 = 0.0f, d11 = 0.0f, d12 = 0.0f, d22 = 0.0f;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int bb = 0; bb < 2; ++bb) {
        const float coef = dsig[bb][a];
        const float3 ja = (a == 0) ? cv.j0 : cv.j1;
        const float3 jb = (bb == 0) ? cv.j0 : cv.j1;
        d00 = d00 + ja.x * coef * jb.x;
        d01 = d01 + ja.x * coef * jb.y;
        d02 = d02 + ja.x * coef * jb.z;
        d11 = d11 + ja.y * coef * jb.y;
        d12 = d12 + ja.y * coef * jb.z;
        d22 = d22 + ja.z * coef * jb.z;
      }
    M3f dcov_cam;
    dcov_cam.c[0][0] = d00; dcov_cam.c[0][1] = d01; dcov_cam.c[0][2] = d02;
    dcov_cam.c[1][0] = d01; dcov_cam.c[1][1] = d11; dcov_cam.c[1][2] = d12;
    dcov_cam.c[2][0] = d02; dcov_cam.c[2][1] = d12; dcov_cam.c[2][2] = d22;
    const M3f dcw = m3_mul(m3_mul(m3_t(cv.W), dcov_cam), cv.W);
    M3f RS2, sym;
    const M3f dcwT = m3_t(dcw);
#pragma unroll
    for (int col = 0; col < 3; ++col)
#pragma unroll
      for (int row = 0; row < 3; ++row) {
        RS2.c[col][row] = cv.R.c[col][row] * cv.s2[col];
        sym.c[col][row] = dcw.c[col][row] + dcwT.c[col][row];
      }
    const M3f dR = m3_mul(sym, RS2);
    const float G00 = dR.c[0][0], G10 = dR.c[0][1], G20 = dR.c[0][2];
    const float G01 = dR.c[1][0], G11 = dR.c[1][1], G21 = dR.c[1][2];
    const float G02 = dR.c[2][0], G12 = dR.c[2][1], G22 = dR.c[2][2];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float3 rk = make_float3(cv.R.c[k][0], cv.R.c[k][1], cv.R.c[k][2]);
      const float ds2 = dot3f(rk, m3_mulv(dcw, rk));
      put(P_SCALE + k, ds2 * 2.0f * cv.s2[k]);
    }
    put(P_ROT + 0, 2.0f * (qz * (G10 - G01) + qy * (G02 - G20) + qx * (G21 - G12)));
    put(P_ROT + 1, 2.0f * (qy * (G01 + G10) + qz * (G02 + G20) + qw * (G21 - G12)) - 4.0f * qx * (G11 + G22));
    put(P_ROT + 2, -4.0f * qy * (G00 + G22) + 2.0f * (qx * (G01 + G10) + qw * (G02 - G20) + qz * (G12 + G21)));
    put(P_ROT + 3, -4.0f * qz * (G00 + G11) + 2.0f * (qw * (G10 - G01) + qx * (G02 + G20) + qy * (G12 + G21)));

    const float3 dtx = make_float3(f * inv_z, 0.0f, -f * t.x * inv_z2);
    const float3 dty = make_float3(0.0f, -f * inv_z, f * t.y * inv_z2);
    float3 dLdt = make_float3(dtx.x * dmx + dty.x * dmy, dtx.y * dmx + dty.y * dmy, dtx.z * dmx + dty.z * dmy);
    if (FULL) {  // dL/dJ = 2 dSigma2d J Sigma_cam, then dJ/dt
      const float h01 = ds01 * 0.5f;
      const float3 r0 = cv.r0, r1 = cv.r1;
      const float3 dJ0 = make_float3(2.0f * (ds00 * r0.x + h01 * r1.x), 2.0f * (ds00 * r0.y + h01 * r1.y),
                                     2.0f * (ds00 * r0.z + h01 * r1.z));
      const float3 dJ1 = make_float3(2.0f * (h01 * r0.x + ds11 * r1.x), 2.0f * (h01 * r0.y + ds11 * r1.y),
                                     2.0f * (h01 * r0.z + ds11 * r1.z));
      const float fz2 = f * inv_z2;
      const float fz3 = fz2 * inv_z;
      const float jtx = dJ0.z * -fz2;
      const float jty = dJ1.z * fz2;
      const float jtz = dJ0.x * -fz2 + dJ0.z * (2.0f * t.x * fz3) + dJ1.y * fz2 + dJ1.z * (-2.0f * t.y * fz3);
      dLdt.x = dLdt.x + jtx;
      dLdt.y = dLdt.y + jty;
      dLdt.z = dLdt.z + jtz;
    }
    const float3 dLdcam = make_float3(dLdt.x, dLdt.y, -dLdt.z);
    float3 dp = m3_mulv(m3_t(cv.W), dLdcam);
    if (FULL) {  // colour through the view direction (cam - p) / |cam - p|
      float db[15][3];
      sh_basis_grad15(dx, dy, dz, db);
      float vx = 0.0f, vy = 0.0f, vz = 0.0f;
#pragma unroll
      for (int k = 0; k < 15; ++k) {
        const float wk = dcol[0] * s.v[P_REST + k] + dcol[1] * s.v[P_REST + 15 + k] +
                         dcol[2] * s.v[P_REST + 30 + k];
        vx = vx + wk * db[k][0];
        vy = vy + wk * db[k][1];
        vz = vz + wk * db[k][2];
      }
      const float pr = dx * vx + dy * vy + dz * vz;
      dp.x = dp.x - (vx - dx * pr) * inv_len;
      dp.y = dp.y - (vy - dy * pr) * inv_len;
      dp.z = dp.z - (vz - dz * pr) * inv_len;
    }
    put(P_POS + 0, dp.x);
    put(P_POS + 1, dp.y);
    put(P_POS + 2, dp.z);
  }
}


// project_backward_p split in two for the multi-view kernel, which sums a
// splat's gradient over up to 32 views: holding all 59 parameters and 59
// sums across that loop leaves one wave per SIMD.  Same formulas; the sums
// meet in another order (position: geometric part over views, then each
// colour channel's view-direction part).
//
// Geometry half: the gradients through the 2-D mean, conic and opacity into
// g = {pos x, y, z, rot w, x, y, z, log-scale 0, 1, 2, logit} (added).
// Returns whether the view keeps the splat (the forward's tests).
template <bool FULL>
__device__ __forceinline__ bool project_backward_geo(const float (&q)[11], const DevCamera& cam, float4 pga,
                                                     float4 pgb, float (&g)[11]) {
  const float px = q[0], py = q[1], pz = q[2];
  const float3 cx = to_cam(cam, px, py, pz);
  const float depth = -cx.z;
  if (depth < cam.near_) return false;
  const float3 t = make_float3(cx.x, cx.y, depth);
  const float qw = q[3], qx = q[4], qy = q[5], qz = q[6];
  Cov2D cv;
  if (!screen_cov(cam, t, qw, qx, qy, qz, q[7], q[8], q[9], cv)) return false;
  const float f = cam.focal;
  const float inv_z = 1.0f / t.z;
  const float inv_z2 = inv_z * inv_z;
  const float inv_det = 1.0f / cv.det;
  const float ca = cv.s11 * inv_det, cb = -cv.s01 * inv_det, cc = cv.s00 * inv_det;
  const float dmx = pga.x, dmy = pga.y, dop = pga.z;
  const float dca = pgb.x, dcb = pgb.y, dcc = pgb.z;
  const float os = nts_sigmoidf(q[10]);
  g[10] = g[10] + dop * os * (1.0f - os);
  const float cm[2][2] = {{ca, cb}, {cb, cc}};
  const float dm[2][2] = {{dca, dcb * 0.5f}, {dcb * 0.5f, dcc}};
  float t1[2][2], full[2][2];
#pragma unroll
  for (int col = 0; col < 2; ++col)
#pragma unroll
    for (int row = 0; row < 2; ++row) t1[col][row] = cm[0][row] * dm[col][0] + cm[1][row] * dm[col][1];
#pragma unroll
  for (int col = 0; col < 2; ++col)
#pragma unroll
    for (int row = 0; row < 2; ++row)
      full[col][row] = -1.0f * (t1[0][row] * cm[col][0] + t1[1][row] * cm[col][1]);
  const float ds00 = full[0][0], ds11 = full[1][1];
  const float ds01 = full[0][1] + full[1][0];
  const float dsig[2][2] = {{ds00, ds01 * 0.5f}, {ds01 * 0.5f, ds11}};
  float d00 = 0.0f, d01 = 0.0f, d02 = 0.0f, d11 = 0.0f, d12 = 0.0f, d22 = 0.0f;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int bb = 0; bb < 2; ++bb) {
      const float coef = dsig[bb][a];
      const float3 ja = (a == 0) ? cv.j0 : cv.j1;
      const float3 jb = (bb == 0) ? cv.j0 : cv.j1;
      d00 = d00 + ja.x * coef * jb.x;
      d01 = d01 + ja.x * coef * jb.y;
      d02 = d02 + ja.x * coef * jb.z;
      d11 = d11 + ja.y * coef * jb.y;
      d12 = d12 + ja.y * coef * jb.z;
      d22 = d22 + ja.z * coef * jb.z;
    }
  M3f dcov_cam;
  dcov_cam.c[0][0] = d00; dcov_cam.c[0][1] = d01; dcov_cam.c[0][2] = d02;
  dcov_cam.c[1][0] = d01; dcov_cam.c[1][1] = d11; dcov_cam.c[1][2] = d12;
  dcov_cam.c[2][0] = d02; dcov_cam.c[2][1] = d12; dcov_cam.c[2][2] = d22;
  const M3f dcw = m3_mul(m3_mul(m3_t(cv.W), dcov_cam), cv.W);
  M3f RS2, sym;
  const M3f dcwT = m3_t(dcw);
#pragma unroll
  for (int col = 0; col < 3; ++col)
#pragma unroll
    for (int row = 0; row < 3; ++row) {
      RS2.c[col][row] = cv.R.c[col][row] * cv.s2[col];
      sym.c[col][row] = dcw.c[col][row] + dcwT.c[col][row];
    }
  const M3f dR = m3_mul(sym, RS2);
  const float G00 = dR.c[0][0], G10 = dR.c[0][1], G20 = dR.c[0][2];
  const float G01 = dR.c[1][0], G11 = dR.c[1][1], G21 = dR.c[1][2];
  const float G02 = dR.c[2][0], G12 = dR.c[2][1], G22 = dR.c[2][2];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float3 rk = make_float3(cv.R.c[k][0], cv.R.c[k][1], cv.R.c[k][2]);
    const float ds2 = dot3f(rk, m3_mulv(dcw, rk));
    g[7 + k] = g[7 + k] + ds2 * 2.0f * cv.s2[k];
  }
  g[3] = g[3] + 2.0f * (qz * (G10 - G01) + qy * (G02 - G20) + qx * (G21 - G12));
  g[4] = g[4] + (2.0f * (qy * (G01 + G10) + qz * (G02 + G20) + qw * (G21 - G12)) - 4.0f * qx * (G11 + G22));
  g[5] = g[5] + (-4.0f * qy * (G00 + G22) + 2.0f * (qx * (G01 + G10) + qw * (G02 - G20) + qz * (G12 + G21)));
  g[6] = g[6] + (-4.0f * qz * (G00 + G11) + 2.0f * (qw * (G10 - G01) + qx * (G02 + G20) + qy * (G12 + G21)));
  const float3 dtx = make_float3(f * inv_z, 0.0f, -f * t.x * inv_z2);
  const float3 dty = make_float3(0.0f, -f * inv_z, f * t.y * inv_z2);
  float3 dLdt = make_float3(dtx.x * dmx + dty.x * dmy, dtx.y * dmx + dty.y * dmy, dtx.z * dmx + dty.z * dmy);
  if (FULL) {  // dL/dJ = 2 dSigma2d J Sigma_cam, then dJ/dt
    const float h01 = ds01 * 0.5f;
    const float3 r0 = cv.r0, r1 = cv.r1;
    const float3 dJ0 = make_float3(2.0f * (ds00 * r0.x + h01 * r1.x), 2.0f * (ds00 * r0.y + h01 * r1.y),
                                   2.0f * (ds00 * r0.z + h01 * r1.z));
    const float3 dJ1 = make_float3(2.0f * (h01 * r0.x + ds11 * r1.x), 2.0f * (h01 * r0.y + ds11 * r1.y),
                                   2.0f * (h01 * r0.z + ds11 * r1.z));
    const float fz2 = f * inv_z2;
    const float fz3 = fz2 * inv_z;
    const float jtx = dJ0.z * -fz2;
    const float jty = dJ1.z * fz2;
    const float jtz = dJ0.x * -fz2 + dJ0.z * (2.0f * t.x * fz3) + dJ1.y * fz2 + dJ1.z * (-2.0f * t.y * fz3);
    dLdt.x = dLdt.x + jtx;
    dLdt.y = dLdt.y + jty;
    dLdt.z = dLdt.z + jtz;
  }
  const float3 dLdcam = make_float3(dLdt.x, dLdt.y, -dLdt.z);
  const float3 dp = m3_mulv(m3_t(cv.W), dLdcam);
  g[0] = g[0] + dp.x;
  g[1] = g[1] + dp.y;
  g[2] = g[2] + dp.z;
  return true;
}

// Colour half for one channel of a kept view: the channel's SH DC and rest
// gradients (added to gdc, grest) and, FULL, the gate of max(SH + 0.5, 0) and
// the channel's share of the position gradient through the view direction
// (added to psh, to be subtracted from the position gradient).
template <bool FULL>
__device__ __forceinline__ void project_backward_colour(float px, float py, float pz, float dc,
                                                        const float (&rest)[15], const DevCamera& cam, float dcol,
                                                        float& gdc, float (&grest)[15], float3& psh) {
  float dx = cam.cam_pos[0] - px, dy = cam.cam_pos[1] - py, dz = cam.cam_pos[2] - pz;
  const float inv_len = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
  dx = dx * inv_len;
  dy = dy * inv_len;
  dz = dz * inv_len;
  float b[15];
  sh_basis15(dx, dy, dz, b);
  if (FULL) {
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < 15; ++k) acc = acc + rest[k] * b[k];
    const float raw = (NTS_SH_C0 * dc + acc) + 0.5f;
    if (!(raw > 0.0f)) dcol = 0.0f;
  }
  gdc = gdc + dcol * NTS_SH_C0;
#pragma unroll
  for (int k = 0; k < 15; ++k) grest[k] = grest[k] + dcol * b[k];
  if (FULL) {
    float db[15][3];
    sh_basis_grad15(dx, dy, dz, db);
    float vx = 0.0f, vy = 0.0f, vz = 0.0f;
#pragma unroll
    for (int k = 0; k < 15; ++k) {
      const float wk = dcol * rest[k];
      vx = vx + wk * db[k][0];
      vy = vy + wk * db[k][1];
      vz = vz + wk * db[k][2];
    }
    const float pr = dx * vx + dy * vy + dz * vz;
    psh.x = psh.x + (vx - dx * pr) * inv_len;
    psh.y = psh.y + (vy - dy * pr) * inv_len;
    psh.z = psh.z + (vz - dz * pr) * inv_len;
  }
}
