"""Pins the checker's scalar maths: exp against float64, powi against the
closed form, and the SH basis against the real spherical harmonics."""
import math

import numpy as np
import pytest
from scipy import special

from oracle import oracle as O


def test_expf_within_one_ulp():
    x = np.concatenate([np.linspace(-87.3, 88.7, 3_000_001), np.linspace(-1, 1, 200_001)]).astype(np.float32)
    y = O.expf(x).astype(np.float64)
    ref = np.exp(x.astype(np.float64))
    ulp = np.spacing(ref.astype(np.float32)).astype(np.float64)
    assert (np.abs(y - ref) / ulp).max() <= 1.0


def test_expf_special_values():
    x = np.array([0.0, -0.0, -200.0, -104.5, 89.0, np.inf, -np.inf, np.nan, 1.0, -1.0], np.float32)
    y = O.expf(x)
    assert y[0] == 1.0 and y[1] == 1.0
    assert y[2] == 0.0 and y[3] == 0.0
    assert np.isinf(y[4]) and np.isinf(y[5]) and y[6] == 0.0 and np.isnan(y[7])
    assert abs(y[8] - math.e) <= 2.5e-7 and abs(y[9] - 1 / math.e) <= 1e-7


def test_expf_subnormal_range():
    x = np.linspace(-103.9, -87.4, 20001).astype(np.float32)
    y = O.expf(x).astype(np.float64)
    ref = np.exp(x.astype(np.float64))
    # one rounding into the subnormal grid (spacing 2^-149)
    assert np.abs(y - ref).max() <= 2.0 ** -149


@pytest.mark.parametrize("b,t", [(0.9, 1), (0.9, 7), (0.999, 1), (0.999, 100), (0.999, 12345), (0.5, 30)])
def test_powi_matches_closed_form(b, t):
    assert abs(O.powi(b, t) - float(np.float32(b)) ** t) <= 1e-6 * max(1.0, float(np.float32(b)) ** t) + 1e-12


def _real_sh(l, m, theta, phi):
    """Real SH with the Condon-Shortley phase, from scipy's complex Y."""
    if hasattr(special, "sph_harm_y"):
        Y = lambda mm: special.sph_harm_y(l, mm, theta, phi)
    else:
        Y = lambda mm: special.sph_harm(mm, l, phi, theta)
    if m > 0:
        return math.sqrt(2) * (-1) ** m * np.real(Y(m))
    if m < 0:
        return math.sqrt(2) * (-1) ** m * np.imag(Y(-m))
    return np.real(Y(0))


def test_sh_basis_is_the_real_spherical_harmonics():
    """Colours from a splat with one unit SH coefficient equal the basis
    function (+0.5 offset) of the real SH for that (l, m), up to the sign
    convention of the 3DGS basis and the 6-digit constants."""
    from nanotracer_amd import synth

    rng = np.random.default_rng(0)
    dirs = rng.normal(size=(64, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    # (index in sh_rest band order) -> (l, m); 3DGS order per band is m = -l..l,
    # with the basis negated for odd m in band 1 and matching CS phase in others
    lm = [(1, -1), (1, 0), (1, 1), (2, -2), (2, -1), (2, 0), (2, 1), (2, 2),
          (3, -3), (3, -2), (3, -1), (3, 0), (3, 1), (3, 2), (3, 3)]
    cam = O.camera_from_pose((0, 0, 0), (0, 0, -1), (0, 1, 0), math.pi / 2, 64, 64, 1)
    for k, (l, m) in enumerate(lm):
        vals = []
        refs = []
        for d in dirs:
            # splat placed so that the direction splat -> camera is d, in front
            # of the camera; use a tiny splat far away along -d to keep it visible
            pos = -np.asarray(d) * 5.0
            if pos[2] > -1.0:  # must lie in front (z < 0)
                continue
            s = synth.HostSplats(pos[None].astype(np.float32), np.array([[1, 0, 0, 0]], np.float32),
                                 np.full((1, 3), -3.0, np.float32), np.array([0.0], np.float32),
                                 np.zeros((1, 3), np.float32), np.zeros((1, 45), np.float32))
            s.sh_rest[0, k] = 1.0  # red channel, coefficient k
            s.sh_dc[0] = -0.5 / 0.2820948  # cancel the +0.5 offset
            p = O.project(s, cam, np.float64)
            if p[0, 7] == 0:
                continue
            theta = math.acos(np.clip(d[2], -1, 1))
            phi = math.atan2(d[1], d[0])
            vals.append(p[0, 8])
            refs.append(_real_sh(l, m, theta, phi))
        vals, refs = np.array(vals), np.array(refs)
        assert len(vals) > 10
        # the renderer clamps at 0: compare where the reference is positive
        sgn = 1.0 if np.dot(np.maximum(refs, 0), vals) >= np.dot(np.maximum(-refs, 0), vals) else -1.0
        pos = sgn * refs > 1e-3
        assert pos.sum() > 3
        assert np.allclose(vals[pos], sgn * refs[pos], rtol=2e-5, atol=2e-6), (k, l, m)


def test_logf_within_two_ulp():
    x = np.concatenate([np.geomspace(1e-8, 1.0, 2_000_001), np.linspace(0.5, 2.0, 400_001),
                        np.geomspace(1e-30, 1e30, 200_001)]).astype(np.float32)
    y = O.logf(x).astype(np.float64)
    ref = np.log(x.astype(np.float64))
    # near x = 1 the result is tiny: measure in ulps of the result, with an
    # absolute floor of one ulp of 1e-7 there
    ulp = np.maximum(np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64), 2.0 ** -47)
    assert (np.abs(y - ref) / ulp).max() <= 2.0
    assert O.logf(np.array([1.0], np.float32))[0] == 0.0


def test_cos2pi_accuracy_and_symmetry():
    u = np.concatenate([np.linspace(0.0, 1.0, 2_000_001, endpoint=False),
                        np.array([0.0, 0.125, 0.25, 0.375, 0.5, 0.625, 0.75, 0.875])]).astype(np.float32)
    y = O.cos2pi(u).astype(np.float64)
    ref = np.cos(2.0 * np.pi * u.astype(np.float64))
    assert np.abs(y - ref).max() <= 1.5e-7
    # closed forms at the octant boundaries
    c = O.cos2pi(np.array([0.0, 0.25, 0.5, 0.75], np.float32))
    assert c[0] == 1.0 and abs(c[1]) < 1e-7 and c[2] == -1.0 and abs(c[3]) < 1e-7
    # cos(2 pi u) = cos(2 pi (1 - u)), on multiples of 2^-20 where 1 - u is exact
    v = (np.arange(1, 2 ** 19, 37) / 2.0 ** 20).astype(np.float32)
    assert np.array_equal(O.cos2pi(v), O.cos2pi((1.0 - v.astype(np.float64)).astype(np.float32)))


def test_lcg_jump_equals_stepping():
    # train.rs:552-556: s' = 1664525 s + 1013904223 mod 2^32
    s0 = 7 * 1664525 + 1013904223 & 0xFFFFFFFF
    s = s0
    for k in range(1, 600):
        s = (s * 1664525 + 1013904223) & 0xFFFFFFFF
        if k in (1, 2, 3, 12, 13, 24, 255, 256, 599):
            assert O.lcg_jump(s0, k) == s
    assert O.lcg_jump(s0, 0) == s0


def test_box_muller_is_the_transform_of_its_uniforms():
    out, uni = O.box_muller_stream(123456789, 200_000)
    u1 = np.maximum(uni[:, 0].astype(np.float64), 1e-8)
    ref = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * uni[:, 1].astype(np.float64))
    assert np.abs(out - ref).max() <= 2e-6 * np.maximum(1.0, np.abs(ref)).max()
    # the uniforms are the LCG's state / 2^32
    s = 123456789
    for i in range(4):
        s = (s * 1664525 + 1013904223) & 0xFFFFFFFF
        assert uni[i // 2, i % 2] == np.float32(s / 2.0 ** 32)
        s2 = s
    assert abs(out.mean()) < 0.01 and abs(out.std() - 1.0) < 0.01


def test_tile_division_by_a_power_of_two_is_exact():
    # nts_div_tile multiplies by the reciprocal for power-of-two tile sizes;
    # the quotient must equal the correctly rounded division bit for bit
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.uniform(-5000, 5000, 200000), rng.normal(0, 3, 1000)]).astype(np.float32)
    for ts in (1.0, 8.0, 16.0, 32.0, 64.0):
        q = x / np.float32(ts)
        m = x * np.float32(1.0 / ts)
        assert np.array_equal(q.view(np.uint32), m.view(np.uint32)), ts
