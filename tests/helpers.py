"""Shared helpers of the test suite (scenes, cameras, comparisons)."""
from __future__ import annotations

import math

import numpy as np

from nanotracer_amd import synth


def config_a_scene():
    """5,000 splats, means U[-1,1]^3 (the CPU-runnable configuration)."""
    return synth.make_scene(5000, 0x5EED0000, 1.0)


def config_b_scene(n=200_000):
    return synth.make_scene(n, 0x5EED0001, 3.0)


def config_a_camera():
    return dict(eye=(0.0, 0.0, 3.0), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov_y=math.pi / 3, width=256,
                height=256, depth_max=6.0)


def config_b_camera():
    u = synth.uniform((0x5EED0001 << 8) + 31, 1)
    eye = synth.orbit_eye(6.0, 2.0 * math.pi * float(u[0]), 0.3)
    return dict(eye=eye, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov_y=math.pi / 3, width=1280, height=720,
                depth_max=12.0)


def rel_norm_err(a: np.ndarray, b: np.ndarray) -> float:
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    d = np.linalg.norm(a - b)
    n = np.linalg.norm(b)
    return float(d / n) if n > 0 else float(d)


def lcg_u32(seed: int, n: int, mod: int | None = None) -> np.ndarray:
    """The reference tests' LCG (x * 1664525 + 1013904223 mod 2^32)."""
    out = np.empty(n, np.uint32)
    s = seed & 0xFFFFFFFF
    for i in range(n):
        s = (s * 1664525 + 1013904223) & 0xFFFFFFFF
        out[i] = s % mod if mod else s
    return out


def projected_record(x, y, depth, r, conic=(1.0, 0.0, 1.0), color=(1.0, 1.0, 1.0), opacity=1.0, visible=1.0):
    return np.array([x, y, depth, r, conic[0], conic[1], conic[2], visible, color[0], color[1], color[2], opacity],
                    np.float32)


def edge_records(n: int, seed: int, w: int, h: int) -> np.ndarray:
    """Projected records that stress the culling bounds: needles at every
    angle with conic condition kappa log-uniform in [1, 1e6] (the culling
    limit is 2e4), opacities at and around 1/255, a few NaN and infinite
    opacities and degenerate conics, means on and off screen."""
    rng = np.random.default_rng(seed)
    theta = rng.uniform(0.0, np.pi, n)
    s_minor = np.sqrt(0.3 + rng.uniform(0.0, 0.5, n))
    ratio = np.exp(rng.uniform(0.0, np.log(1e3), n))  # kappa up to ~ratio^2
    s_major = np.minimum(s_minor * ratio, 400.0)
    c, s = np.cos(theta), np.sin(theta)
    sxx = c * c * s_major**2 + s * s * s_minor**2
    syy = s * s * s_major**2 + c * c * s_minor**2
    sxy = c * s * (s_major**2 - s_minor**2)
    det = sxx * syy - sxy * sxy
    qa, qb, qc = syy / det, -sxy / det, sxx / det
    op = rng.uniform(0.01, 1.0, n)
    k = rng.integers(0, 6, n)
    op[k == 0] = (1.0 / 255.0) * (1.0 + rng.uniform(-2e-4, 2e-4, int((k == 0).sum())))
    op[k == 1] = 0.999 / 255.0 * (1.0 + rng.uniform(-1e-6, 1e-6, int((k == 1).sum())))
    op[k == 2] = 1.0
    recs = np.zeros((n, 12), np.float32)
    recs[:, 0] = rng.uniform(-30.0, w + 30.0, n)
    recs[:, 1] = rng.uniform(-30.0, h + 30.0, n)
    recs[:, 2] = rng.permutation(n) * (50.0 / n) + 0.5  # distinct depths in (0, 50)
    recs[:, 3] = np.ceil(np.minimum(3.0 * s_major, 300.0))
    recs[:, 4], recs[:, 5], recs[:, 6] = qa, qb, qc
    recs[:, 7] = 1.0
    recs[:, 8:11] = rng.uniform(0.0, 1.0, (n, 3))
    recs[:, 11] = op
    special = rng.choice(n, 8, replace=False)
    recs[special[0:2], 11] = np.nan
    recs[special[2], 11] = np.inf
    recs[special[3:5], 5] = np.sqrt(recs[special[3:5], 4] * recs[special[3:5], 6]) * 1.5  # det < 0
    recs[special[5], 4] = recs[special[5], 6] = 0.0  # det = 0 with qb != 0 -> det < 0
    return recs


def conic_kappa(recs: np.ndarray) -> np.ndarray:
    """kappa = (sqrt(qa qc) + |qb|)^2 / det per record, in float32 as the device computes it."""
    qa, qb, qc = recs[:, 4], recs[:, 5], recs[:, 6]
    with np.errstate(all="ignore"):
        det = qa * qc - qb * qb
        s = np.sqrt(qa * qc) + np.abs(qb)
        return (s * s / det).astype(np.float32)
