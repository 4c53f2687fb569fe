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
