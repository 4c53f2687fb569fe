"""The batched and serving paths bin only the tiles of a splat's alpha >=
1/255 box (nts_alpha_tile_rect), and mark the pairs of those tiles its
alpha >= 1/255 region cannot reach (nts_tile_cull_hits) with the all-ones
key.  On the host restatement: the live culled pairs are reference pairs in
the reference's order, nothing with opacity below 1/255 is binned, and the
image, final transmittance and last contributing splat of every pixel equal
those of the reference binning."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from nanotracer_amd import synth


def _scene(n, seed):
    rng = np.random.default_rng(seed)
    s = synth.make_scene(n, seed, 1.2, log_scale_mu=-3.2, log_scale_sd=0.9)
    elong = rng.random(n) < 0.3
    s.log_scale[elong, 0] += rng.uniform(1.0, 3.0, int(elong.sum()))
    faint = rng.random(n) < 0.15
    s.logit[faint] = rng.uniform(-9.0, -5.0, int(faint.sum())).astype(np.float32)
    return s


def _last_splat(pay, rg, li, ts, w, h):
    """Splat index of each pixel's last contributor (or -1) from slice-relative indices."""
    tx = -(-w // ts)
    out = np.full(li.shape, -1, np.int64)
    ys, xs = np.nonzero(li)
    tiles = (ys // ts) * tx + xs // ts
    out[ys, xs] = pay[rg[2 * tiles] + li[ys, xs] - 1]
    return out


@pytest.mark.parametrize("tile", [8, 16, 32])
def test_culled_binning_keeps_every_pixel(tile):
    s = _scene(4000, 0xC011 + tile)
    w, h = 160, 120
    cam = O.camera_from_pose((0.3, 0.2, 2.6), (0, 0, 0), (0, 1, 0), math.pi / 3, w, h, s.n)
    t = O.tiling(w, h, tile, 6.0)
    p = O.project(s, cam)
    k0, p0, r0 = O.bin(p, t)
    k1, p1, r1 = O.bin(p, t, cull=True)
    assert 0 < len(k1) < len(k0)
    # the pairs of tiles the alpha >= 1/255 region cannot reach carry the
    # all-ones key: they close the list, and no tile range reaches them
    live = k1 != 0xFFFFFFFF
    nl = int(live.sum())
    assert 0 < nl < len(k1) and live[:nl].all() and not live[nl:].any()
    assert r1.reshape(-1, 2)[:, 1].max() == nl
    # the live pairs: an order-preserving subset of the reference pairs
    pos = {}
    for i, pair in enumerate(zip(k0.tolist(), p0.tolist())):
        pos[pair] = i
    idx = np.array([pos[pair] for pair in zip(k1[:nl].tolist(), p1[:nl].tolist())])
    assert np.all(np.diff(idx) > 0)
    # splats that cannot reach alpha 1/255 anywhere are not binned
    faint = np.nonzero(p[:, 11] * np.float32(255.0) < np.float32(0.999))[0]
    assert faint.size and not np.isin(p1, faint).any()
    img0, ft0, li0 = O.composite(p, p0, r0, t, with_state=True)
    img1, ft1, li1 = O.composite(p, p1, r1, t, with_state=True)
    assert np.array_equal(img0.view(np.uint32), img1.view(np.uint32))
    assert np.array_equal(ft0.view(np.uint32), ft1.view(np.uint32))
    assert np.array_equal(_last_splat(p0, r0, li0, tile, w, h), _last_splat(p1, r1, li1, tile, w, h))


@pytest.mark.parametrize("tile", [8, 32])
def test_culled_binning_edge_records(tile):
    """Needles across the kappa limit, opacities at 1/255, NaN and infinite
    opacities (a NaN alpha clamps to 0.99 in the blend, so such a splat is
    drawn), degenerate conics: the culled binning still renders every pixel
    exactly as the reference binning, keeps the reference tiles of every
    record it may not cull, and culls some of the others."""
    from .helpers import conic_kappa, edge_records
    w, h = 192, 128
    p = edge_records(3000, 0xED6E + tile, w, h)
    t = O.tiling(w, h, tile, 60.0)
    k0, p0, r0 = O.bin(p, t)
    k1, p1, r1 = O.bin(p, t, cull=True)
    assert 0 < len(k1) < len(k0)
    img0, ft0, li0 = O.composite(p, p0, r0, t, with_state=True)
    img1, ft1, li1 = O.composite(p, p1, r1, t, with_state=True)
    assert np.array_equal(img0.view(np.uint32), img1.view(np.uint32))
    assert np.array_equal(ft0.view(np.uint32), ft1.view(np.uint32))
    assert np.array_equal(_last_splat(p0, r0, li0, tile, w, h), _last_splat(p1, r1, li1, tile, w, h))
    c0 = np.bincount(p0, minlength=len(p))
    c1 = np.bincount(p1[k1 != 0xFFFFFFFF], minlength=len(p))  # live pairs
    kap = conic_kappa(p)
    with np.errstate(invalid="ignore"):
        keep = ~(kap < 2e4) | ~(p[:, 11] * np.float32(255.0) <= np.float32(3.4e38))
    keep &= ~(p[:, 11] * np.float32(255.0) < np.float32(0.999))
    assert keep.sum() >= 10
    assert np.array_equal(c0[keep], c1[keep])
    assert (c1[~keep] <= c0[~keep]).all() and (c1[~keep] < c0[~keep]).any()
