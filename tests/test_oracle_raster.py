"""Pins the checker's projection, forward blend and backward:
closed forms, an independent float64 EWA restatement, brute-force per-pixel
blending, and float64 central finite differences."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from nanotracer_amd import synth

from .helpers import projected_record


def _unit_splat():
    return synth.HostSplats(np.array([[0, 0, -5]], np.float32), np.array([[1, 0, 0, 0]], np.float32),
                            np.zeros((1, 3), np.float32), np.array([10.0], np.float32),
                            np.array([[0.5, 0.3, 0.1]], np.float32), np.zeros((1, 45), np.float32))


def test_cpu_oracle_centers_on_axis():
    # raster.rs:999-1026
    s = _unit_splat()
    cam = O.camera_from_pose((0, 0, 0), (0, 0, -1), (0, 1, 0), math.pi / 2, 256, 256, 1)
    p = O.project(s, cam)[0]
    assert p[7] == 1.0 and abs(p[0] - 128) < 1e-3 and abs(p[1] - 128) < 1e-3 and abs(p[2] - 5) < 1e-3
    assert 50 < p[3] < 100 and abs(p[5]) < 1e-3 and abs(p[4] - p[6]) < 1e-3


def _ewa_f64(pos, q, ls, eye, target, up, fov, W, H):
    """Independent float64 EWA: explicit look-at, the pinhole Jacobian
    obtained by differentiating the projection numerically, conic and radius
    by numpy linear algebra.

    Convention of the reference (project_gaussians.wgsl:203-219): the
    Jacobian is taken with respect to t = (x, y, depth) with depth = -z_cam,
    and applied to the camera-space covariance in (x, y, z_cam) coordinates
    without the diag(1, 1, -1) change of variables.  Off-axis this flips the
    sign of the xz / yz covariance terms relative to textbook EWA; it is
    reproduced here because parity is with the reference."""
    eye, target, up = map(np.asarray, (eye, target, up))
    f = target - eye
    f = f / np.linalg.norm(f)
    s = np.cross(f, up)
    s = s / np.linalg.norm(s)
    u = np.cross(s, f)
    Rv = np.stack([s, u, -f])  # world -> camera rotation (rows)
    focal = H / (2 * math.tan(fov / 2))

    def proj_t(t):  # t = (x, y, depth)
        return np.array([W / 2 + focal * t[0] / t[2], H / 2 - focal * t[1] / t[2]])

    c = Rv @ (pos - eye)
    t0 = np.array([c[0], c[1], -c[2]])
    mean = proj_t(t0)
    J = np.zeros((2, 3))
    h = 1e-6
    for k in range(3):
        d = np.zeros(3)
        d[k] = h
        J[:, k] = (proj_t(t0 + d) - proj_t(t0 - d)) / (2 * h)
    w, x, y, z = q / np.linalg.norm(q)
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    S = np.diag(np.exp(ls))
    cov_cam = Rv @ (R @ S @ S @ R.T) @ Rv.T
    cov2 = J @ cov_cam @ J.T + 0.3 * np.eye(2)
    conic = np.linalg.inv(cov2)
    lam = np.linalg.eigvalsh(cov2).max()
    return mean, t0[2], conic, math.ceil(3 * math.sqrt(lam))


def test_projection_matches_independent_ewa():
    s = synth.make_scene(400, 7, 1.0, log_scale_mu=-2.5)
    eye, target, up = (0.3, 0.4, 3.0), (0.0, 0.1, 0.0), (0.0, 1.0, 0.0)
    W, H = 320, 240
    cam = O.camera_from_pose(eye, target, up, math.pi / 3, W, H, s.n)
    p = O.project(s, cam, np.float64)
    checked = 0
    for i in range(s.n):
        if p[i, 7] == 0:
            continue
        mean, depth, conic, rad = _ewa_f64(s.pos[i].astype(np.float64), s.rot[i].astype(np.float64),
                                           s.log_scale[i].astype(np.float64), eye, target, up, math.pi / 3, W, H)
        assert np.allclose(p[i, :2], mean, rtol=1e-5, atol=1e-4)
        assert abs(p[i, 2] - depth) < 1e-5
        assert np.allclose(p[i, [4, 5, 6]], [conic[0, 0], conic[0, 1], conic[1, 1]], rtol=1e-4, atol=1e-7)
        assert abs(p[i, 3] - rad) <= 1
        checked += 1
    assert checked > 100
    # fp32 path agrees with fp64 to float precision
    p32 = O.project(s, cam, np.float32)
    vis = p[:, 7] != 0
    assert np.array_equal(p32[:, 7] != 0, vis)
    assert np.allclose(p32[vis][:, [0, 1, 2]], p[vis][:, [0, 1, 2]], rtol=1e-5, atol=1e-3)


def test_near_and_frustum_cull_write_zero_records():
    s = synth.make_scene(3, 1, 1.0)
    s.pos[:] = [[0, 0, 1.0], [100.0, 0, -2.0], [0, 0, -2.0]]  # behind, far off-screen, visible
    cam = O.camera_from_pose((0, 0, 0), (0, 0, -1), (0, 1, 0), math.pi / 3, 64, 64, 3)
    p = O.project(s, cam)
    assert not p[0].any() and not p[1].any() and p[2, 7] == 1.0


def _fma32(a, b, c):
    """float32 fused multiply-add, exactly: a b + c in rationals, rounded
    once to the nearest float32 (ties to even)."""
    from fractions import Fraction

    x = Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c))
    if x == 0:
        return np.float32(0.0) if (np.float32(a) * np.float32(b) + np.float32(c)) >= 0 else np.float32(-0.0)
    sgn, x = (1, x) if x > 0 else (-1, -x)
    e = x.numerator.bit_length() - x.denominator.bit_length()
    if Fraction(2) ** e > x:
        e -= 1
    q = max(e, -126) - 23  # the float32 ulp at x
    m = x / Fraction(2) ** q
    r = m.numerator // m.denominator
    rem = m - r
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and r % 2 == 1):
        r += 1
    return np.float32(sgn * float(r * Fraction(2) ** q))


def test_fma32_rounds_once():
    from fractions import Fraction

    f = np.float32
    # a product-then-add loses 2^-24 of (1 + 2^-12)^2 - 1; the fused form keeps it
    a = f(1 + 2.0 ** -12)
    assert _fma32(a, a, f(-1)) == f(2.0 ** -11 + 2.0 ** -24)
    assert f(a * a) - f(1) == f(2.0 ** -11)
    # against the exact value: no float32 neighbour is closer (ties to even)
    rng = np.random.default_rng(11)
    for a, b, c in rng.uniform(-2, 2, (300, 3)).astype(np.float32):
        x = Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c))
        r = _fma32(a, b, c)
        d = abs(Fraction(float(r)) - x)
        for nb in (np.nextafter(r, f(np.inf)), np.nextafter(r, f(-np.inf))):
            dn = abs(Fraction(float(nb)) - x)
            assert d < dn or (d == dn and (r.view(np.uint32) & 1) == 0)


def _brute_force_pixel(proj, px, py, t):
    """Front-to-back over every visible splat whose tile box covers the
    pixel's tile, ordered by (depth16, index) - no ranges, no sort."""
    ts = t.tile_size
    tx, ty = px // ts, py // ts
    cand = []
    for i in range(proj.shape[0]):
        p = proj[i]
        if p[7] == 0:
            continue
        x0 = max(math.floor(np.float32(p[0] - p[3]) / np.float32(ts)), 0)
        x1 = math.floor(np.float32(p[0] + p[3]) / np.float32(ts))
        y0 = max(math.floor(np.float32(p[1] - p[3]) / np.float32(ts)), 0)
        y1 = math.floor(np.float32(p[1] + p[3]) / np.float32(ts))
        if x0 <= tx <= x1 and y0 <= ty <= y1:
            dn = min(max(np.float32(p[2]) / np.float32(t.depth_max), np.float32(0)), np.float32(1))
            cand.append((int(np.float32(dn) * np.float32(65535.0)), i))
    cand.sort()
    f = np.float32
    C = [f(0), f(0), f(0)]
    T = f(1)
    for _, i in cand:
        p = proj[i]
        dx, dy = f(f(px) + f(0.5)) - p[0], f(f(py) + f(0.5)) - p[1]
        power = f(-0.5) * _fma32(f(p[6] * dy), dy, _fma32(f(p[4] * dx), dx, f(f(f(2) * p[5]) * dx) * dy))
        if power > 0:
            continue
        g = O.expf(np.array([power], np.float32))[0]
        alpha = min(p[11] * g, f(0.99))
        if alpha < f(1.0 / 255.0):
            continue
        for c in range(3):
            C[c] = _fma32(f(T * alpha), p[8 + c], C[c])
        T = f(T * f(f(1) - alpha))
        if T < f(1e-4):
            break
    return np.array(C + [f(1) - T], np.float32)


def test_tiled_blend_equals_brute_force_per_pixel():
    rng = np.random.default_rng(4)
    n = 60
    proj = np.zeros((n, 12), np.float32)
    proj[:, 0] = rng.uniform(-5, 37, n)
    proj[:, 1] = rng.uniform(-5, 29, n)
    proj[:, 2] = rng.uniform(0.5, 9, n)
    sig = rng.uniform(1.0, 6.0, n)
    proj[:, 3] = np.ceil(3 * sig)
    proj[:, 4] = 1 / sig ** 2
    proj[:, 5] = rng.uniform(-0.2, 0.2, n) / sig ** 2
    proj[:, 6] = 1 / sig ** 2
    proj[:, 7] = 1
    proj[:, 8:11] = rng.uniform(0, 1, (n, 3))
    proj[:, 11] = rng.uniform(0.2, 1.0, n)
    t = O.tiling(32, 24, 8, 10.0)
    keys, pay, rg = O.bin(proj, t)
    img = O.composite(proj, pay, rg, t)
    for py in range(0, 24, 3):
        for px in range(0, 32, 3):
            assert np.array_equal(img[py, px], _brute_force_pixel(proj, px, py, t)), (px, py)


def test_single_splat_alpha_is_analytic():
    # alpha = o * exp(-r^2 / (2 sigma^2)) for an isotropic splat
    proj = projected_record(32.5, 32.5, 5.0, 30.0, conic=(0.04, 0.0, 0.04), opacity=0.8)[None]
    t = O.tiling(64, 64, 16, 100.0)
    keys, pay, rg = O.bin(proj, t)
    img = O.composite(proj, pay, rg, t)
    for (x, y) in [(32, 32), (35, 30), (40, 40), (20, 33)]:
        r2 = (x + 0.5 - 32.5) ** 2 + (y + 0.5 - 32.5) ** 2
        a = min(0.8 * math.exp(-0.5 * 0.04 * r2), 0.99)
        if a < 1 / 255:
            a = 0.0
        assert abs(img[y, x, 3] - a) < 1e-6 and abs(img[y, x, 0] - a) < 1e-6


def _smooth_scene():
    """Low-opacity, well-separated-depth splats: no pixel reaches T < 1e-4,
    so the reference-parity backward is the exact chain rule."""
    rng = np.random.default_rng(8)
    n = 12
    proj = np.zeros((n, 12), np.float64)
    proj[:, 0] = rng.uniform(6, 26, n)
    proj[:, 1] = rng.uniform(6, 26, n)
    proj[:, 2] = np.linspace(1, 9, n)
    sig = rng.uniform(2.0, 4.0, n)
    proj[:, 3] = 40  # box covers the image: no clipping discontinuity
    proj[:, 4] = 1 / sig ** 2
    proj[:, 5] = rng.uniform(-0.1, 0.1, n) / sig ** 2
    proj[:, 6] = 1 / (sig * 1.3) ** 2
    proj[:, 7] = 1
    proj[:, 8:11] = rng.uniform(0.2, 1, (n, 3))
    proj[:, 11] = rng.uniform(0.2, 0.5, n)
    return proj


def test_reference_parity_backward_matches_finite_differences():
    proj = _smooth_scene()
    t = O.tiling(32, 32, 16, 10.0)
    keys, pay, rg = O.bin(proj.astype(np.float32), t)
    rng = np.random.default_rng(9)
    wpix = rng.normal(0, 1, (32, 32, 4))
    wpix[..., 3] = 0

    def loss(pr):
        return float((O.composite_f64(pr, pay, rg, t) * wpix).sum())

    img = O.composite(proj.astype(np.float32), pay, rg, t)
    g = O.composite_backward_ref(proj.astype(np.float32), pay, rg, t, img, wpix.astype(np.float32))
    h = 1e-5
    for i in range(proj.shape[0]):
        for col, scale in ((0, 1), (1, 1), (11, 1), (4, 1e-2), (5, 1e-2), (6, 1e-2), (8, 1), (9, 1), (10, 1)):
            hp = h * scale
            a, b = proj.copy(), proj.copy()
            a[i, col] += hp
            b[i, col] -= hp
            fd = (loss(a) - loss(b)) / (2 * hp)
            gcol = {0: 0, 1: 1, 11: 2, 4: 4, 5: 5, 6: 6, 8: 8, 9: 9, 10: 10}[col]
            an = g[i, gcol]
            assert abs(an - fd) <= 2e-3 * max(abs(fd), abs(an)) + 2e-4, (i, col, an, fd)


def _fd_scene():
    # backward_finite_diff.rs:29-49
    rest = np.zeros(45, np.float32)
    rest[0] = 0.05
    rest[20] = -0.03
    return synth.HostSplats(np.array([[0, 0, -5]], np.float32), np.array([[1, 0, 0, 0]], np.float32),
                            np.array([[-0.5, -0.6, -0.7]], np.float32), np.array([2.0], np.float32),
                            np.array([[0.4, 0.3, 0.2]], np.float32), rest[None].copy())


def test_finite_diff_matches_analytic():
    """tests/backward_finite_diff.rs:174-246 with its tolerances, L = sum rgb."""
    s = _fd_scene()
    W = H = 32
    cam = O.camera_from_pose((0, 0, 0), (0, 0, -1), (0, 1, 0), math.pi / 3, W, H, 1)
    t = O.tiling(W, H, 16, 50.0)

    def forward_loss(sp):
        p = O.project(sp, cam)
        keys, pay, rg = O.bin(p, t)
        return float(O.composite(p, pay, rg, t)[..., :3].astype(np.float64).sum())

    p = O.project(s, cam)
    keys, pay, rg = O.bin(p, t)
    img = O.composite(p, pay, rg, t)
    dl = np.zeros((H, W, 4), np.float32)
    dl[..., :3] = 1
    pg = O.composite_backward_ref(p, pay, rg, t, img, dl)
    g = O.project_backward(s, cam, pg)

    def numerical(mut):
        a, b = s.copy(), s.copy()
        mut(a, 1e-3)
        mut(b, -1e-3)
        return (forward_loss(a) - forward_loss(b)) / 2e-3

    def close(a, n, rel, ab):
        d = abs(a - n)
        return d <= ab or d <= rel * (abs(a) + abs(n))

    def add(attr, idx, e):
        def f(sp, h):
            getattr(sp, attr)[idx] += h
        return f

    assert close(g.sh_dc[0, 0], numerical(add("sh_dc", (0, 0), 0)), 0.02, 0.05)
    assert close(g.logit[0], numerical(add("logit", 0, 0)), 0.05, 0.05)
    assert close(g.sh_rest[0, 0], numerical(add("sh_rest", (0, 0), 0)), 0.05, 0.05)
    assert close(g.pos[0, 0], numerical(add("pos", (0, 0), 0)), 0.10, 0.5)
    assert close(g.log_scale[0, 0], numerical(add("log_scale", (0, 0), 0)), 0.10, 0.5)
