"""Deterministic synthetic scenes and cameras (no datasets: the workloads are
drawn from stated distributions).

RNG: SplitMix64 (Steele, Lea & Flood, "Fast splittable pseudorandom number
generators", OOPSLA 2014) -> float64 uniforms in [0, 1) from the top 53 bits
-> standard normals by Box-Muller (cosine branch).  Each attribute draws from
its own stream, seeded (seed << 8) + attribute index.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n outputs of SplitMix64 started from `seed` (uint64)."""
    i = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, n: int) -> np.ndarray:
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, n: int) -> np.ndarray:
    u = uniform(seed, 2 * n)
    u1 = 1.0 - u[0::2]  # (0, 1]
    u2 = u[1::2]
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


@dataclass
class HostSplats:
    """Host SoA in the SplatBuffer layout (crates/nano-optimize/src/splat_store.rs:21-36)."""

    pos: np.ndarray        # (n, 3)
    rot: np.ndarray        # (n, 4) w, x, y, z
    log_scale: np.ndarray  # (n, 3)
    logit: np.ndarray      # (n,)
    sh_dc: np.ndarray      # (n, 3)
    sh_rest: np.ndarray    # (n, 45) planar R[15] G[15] B[15]

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])

    def copy(self) -> "HostSplats":
        return HostSplats(*(np.array(a, copy=True) for a in self.arrays()))

    def arrays(self):
        return (self.pos, self.rot, self.log_scale, self.logit, self.sh_dc, self.sh_rest)

    @staticmethod
    def zeros(n: int) -> "HostSplats":
        z = lambda *s: np.zeros(s, np.float32)
        return HostSplats(z(n, 3), z(n, 4), z(n, 3), z(n), z(n, 3), z(n, 45))

    def flat(self) -> np.ndarray:
        """All 59 parameters of every splat, attribute-major."""
        return np.concatenate([a.reshape(-1) for a in self.arrays()])

    @staticmethod
    def from_flat(flat: np.ndarray, n: int) -> "HostSplats":
        """Inverse of flat()."""
        out, o = [], 0
        for w in (3, 4, 3, 1, 3, 45):
            a = np.ascontiguousarray(flat[o:o + w * n], np.float32)
            out.append(a if w == 1 else a.reshape(n, w))
            o += w * n
        return HostSplats(*out)

    def subset(self, idx) -> "HostSplats":
        return HostSplats(*(np.ascontiguousarray(a[idx]) for a in self.arrays()))


def make_scene(n: int, seed: int, mean_half_extent: float = 1.0, log_scale_mu: float = -4.0,
               log_scale_sd: float = 0.5, logit_mu: float = 0.0, logit_sd: float = 1.0,
               rest_sd: float = 0.05) -> HostSplats:
    """Splats with the benchmark distributions: means U[-e, e]^3, log-scales
    N(mu, sd) per axis, quaternions normalise(N(0, I4)) stored (w, x, y, z),
    opacity logits N(mu, sd), SH DC U(0, 1), higher bands N(0, rest_sd)."""
    s = seed << 8
    pos = (uniform(s + 0, 3 * n) * 2.0 - 1.0) * mean_half_extent
    ls = normal(s + 1, 3 * n) * log_scale_sd + log_scale_mu
    q = normal(s + 2, 4 * n).reshape(n, 4)
    q /= np.maximum(np.linalg.norm(q, axis=1, keepdims=True), 1e-12)
    lg = normal(s + 3, n) * logit_sd + logit_mu
    dc = uniform(s + 4, 3 * n)
    rest = normal(s + 5, 45 * n) * rest_sd
    f = lambda a, *shape: np.ascontiguousarray(a.reshape(*shape).astype(np.float32))
    return HostSplats(f(pos, n, 3), f(q, n, 4), f(ls, n, 3), f(lg, n), f(dc, n, 3), f(rest, n, 45))


def fast_scene(n: int, seed: int, mean_half_extent: float = 3.0, log_scale_mu: float = -4.0,
               log_scale_sd: float = 0.5, logit_mu: float = 0.0, logit_sd: float = 1.0,
               rest_sd: float = 0.05) -> HostSplats:
    """The make_scene distributions drawn with numpy's PCG64 in float32 (the
    higher SH bands as rows of a random pool): for large resident serving
    scenes, where only the distribution matters."""
    rng = np.random.default_rng(seed)
    f32 = np.float32
    pos = (rng.random((n, 3), dtype=f32) * f32(2.0) - f32(1.0)) * f32(mean_half_extent)
    q = rng.standard_normal((n, 4), dtype=f32)
    q /= np.maximum(np.linalg.norm(q, axis=1, keepdims=True), f32(1e-12))
    ls = rng.standard_normal((n, 3), dtype=f32) * f32(log_scale_sd) + f32(log_scale_mu)
    lg = rng.standard_normal(n, dtype=f32) * f32(logit_sd) + f32(logit_mu)
    dc = rng.random((n, 3), dtype=f32)
    # higher SH bands from a pool of 65536 N(0, rest_sd) rows
    pool = rng.standard_normal((65536, 45), dtype=f32) * f32(rest_sd)
    rest = pool[rng.integers(0, 65536, n)]
    return HostSplats(pos, q, ls, lg, dc, rest)


def perturb_means(s: HostSplats, seed: int, sd: float = 0.02) -> HostSplats:
    out = s.copy()
    out.pos = np.ascontiguousarray((s.pos.astype(np.float64) + normal((seed << 8) + 9, 3 * s.n).reshape(-1, 3) * sd)
                                   .astype(np.float32))
    return out


def orbit_eye(radius: float, azimuth: float, elevation: float):
    ce = math.cos(elevation)
    return (radius * ce * math.sin(azimuth), radius * math.sin(elevation), radius * ce * math.cos(azimuth))


RESOLUTIONS_C = ((640, 480), (1280, 720), (1920, 1080), (2560, 1440))


def orbit_views(n_views: int, seed: int, radius: float, resolutions=RESOLUTIONS_C):
    """Seeded orbit cameras: azimuth U[0, 2pi), elevation U[-0.4, 0.6] rad,
    resolution drawn uniformly from `resolutions`.  Returns a list of
    (eye, width, height)."""
    u = uniform((seed << 8) + 17, 3 * n_views).reshape(n_views, 3)
    views = []
    for k in range(n_views):
        az = 2.0 * math.pi * u[k, 0]
        el = -0.4 + 1.0 * u[k, 1]
        w, h = resolutions[min(int(u[k, 2] * len(resolutions)), len(resolutions) - 1)]
        views.append((orbit_eye(radius, az, el), w, h))
    return views


RESOLUTIONS_E = ((512, 512), (1280, 720), (1920, 1080), (3840, 2160))


def serving_scene_sizes(count: int, seed: int, lo: int = 50_000, hi: int = 3_000_000) -> np.ndarray:
    """Splat counts log-uniform in [lo, hi] (the resident serving scenes)."""
    rng = np.random.default_rng(seed + 1)
    return np.exp(rng.uniform(math.log(lo), math.log(hi), count)).astype(np.int64)


def serving_scenes(count: int, seed: int, lo: int = 50_000, hi: int = 3_000_000):
    """`count` resident scenes with splat counts log-uniform in [lo, hi]:
    windows of one `hi`-splat pool (a per-scene start and scale), so that
    setting up tens of millions of resident splats costs seconds.  Yields
    HostSplats in order."""
    pool = fast_scene(hi, seed, 3.0)
    counts = serving_scene_sizes(count, seed, lo, hi)
    rng = np.random.default_rng(seed + 1)
    rng.uniform(size=count)  # the draws of serving_scene_sizes
    for n in counts:
        start = int(rng.integers(0, pool.n))
        idx = (start + np.arange(n)) % pool.n
        sc = float(rng.uniform(0.8, 1.2))
        arrs = [np.ascontiguousarray(a[idx]) for a in pool.arrays()]
        arrs[0] *= np.float32(sc)
        yield HostSplats(*arrs)
