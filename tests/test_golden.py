"""The committed golden fixtures (tests/golden, written by make_golden.py):
the checker must reproduce them bit for bit, and the device path must match
them (keys, ranges, images bit-exact; gradients within 1e-4 norm-wise)."""
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
from nanotracer_amd import synth

from .helpers import rel_norm_err

G = Path(__file__).resolve().parent / "golden"
NAMES = ("pos", "rot", "log_scale", "logit", "sh_dc", "sh_rest")


def _scene64():
    d = np.load(G / "scene64_32x32.npz")
    s = synth.HostSplats(*[d[n] for n in NAMES])
    cam = O.camera_from_pose((0.0, 0.0, 2.5), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), math.pi / 3, 32, 32, s.n)
    return d, s, cam, O.tiling(32, 32, 16, 5.0)


def test_binner_fixture_holds_the_reference_assertions():
    # tile_binner.rs:477-499 and 527-539
    d = np.load(G / "binner_cases.npz")
    assert list(d["four_keys"] >> 16) == [0, 2, 12, 15]
    assert list(d["four_payloads"]) == [0, 1, 2, 3]
    r = d["four_ranges"].reshape(-1, 2)
    assert [(t, b, e) for t, (b, e) in enumerate(r) if e > b] == [(0, 0, 1), (2, 1, 2), (12, 2, 3), (15, 3, 4)]
    assert list(d["two_keys"] >> 16) == [5, 5] and list(d["two_payloads"]) == [1, 0]
    assert list(d["two_ranges"][10:12]) == [0, 2]


@pytest.mark.parametrize("case", ["four", "two"])
def test_oracle_reproduces_binner_fixture(case):
    d = np.load(G / "binner_cases.npz")
    t = O.tiling(64, 64, 16, 100.0)
    k, p, r = O.bin(d[f"{case}_projected"], t)
    assert np.array_equal(k, d[f"{case}_keys"]) and np.array_equal(p, d[f"{case}_payloads"])
    assert np.array_equal(r, d[f"{case}_ranges"])


def test_oracle_reproduces_scene64_fixture():
    d, s, cam, t = _scene64()
    assert np.array_equal(np.array(cam.view, np.float32), d["view"])
    proj = O.project(s, cam)
    assert np.array_equal(proj.view(np.uint32), d["projected"].view(np.uint32))
    k, p, r = O.bin(proj, t)
    assert np.array_equal(k, d["keys"]) and np.array_equal(p, d["payloads"]) and np.array_equal(r, d["ranges"])
    img, ft, li = O.composite(proj, p, r, t, with_state=True)
    assert np.array_equal(img.view(np.uint32), d["image"].view(np.uint32))
    assert np.array_equal(ft.view(np.uint32), d["final_t"].view(np.uint32)) and np.array_equal(li, d["last_idx"])
    loss, dl = O.loss_grad(img, d["target"], 0)
    assert loss == float(d["loss"]) and np.array_equal(dl, d["dl_dc"])
    pg = O.composite_backward_recompute(proj, p, r, t, ft, li, dl)
    assert np.array_equal(pg, d["projected_grad"])
    g = O.project_backward(s, cam, pg, full=True)
    for n, a in zip(NAMES, g.arrays()):
        assert np.array_equal(a, d[f"g_{n}"]), n


def test_oracle_reproduces_config_a_step():
    d = np.load(G / "config_a_step.npz")
    s = synth.HostSplats(*[d[f"before_{n}"] for n in NAMES])
    assert np.array_equal(s.pos, synth.make_scene(5000, 0x5EED0000, 1.0).pos)  # the generator is stable too
    cam = O.camera_from_pose((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), math.pi / 3, 256, 256, s.n)
    t = O.tiling(256, 256, 16, 6.0)
    op = O.project(s, cam)
    k, pay, rg = O.bin(op, t)
    img = O.composite(op, pay, rg, t)
    loss, dl = O.loss_grad(img, d["target"], 0)
    assert loss == float(d["loss"])
    g = O.project_backward(s, cam, O.composite_backward_ref(op, pay, rg, t, img, dl))
    for n, a in zip(NAMES, g.arrays()):
        assert np.array_equal(a, d[f"grad_{n}"]), n


@pytest.mark.gpu
def test_device_matches_scene64_fixture(ctx):
    from nanotracer_amd import _lib as L
    from nanotracer_amd.api import CameraUniform, GpuSplatBuffer, Rasterizer, TileBinner, TilingParams

    d, s, cam, t = _scene64()
    g = GpuSplatBuffer.upload(ctx, s)
    r = Rasterizer()
    gcam = CameraUniform.from_pose((0.0, 0.0, 2.5), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), math.pi / 3, 32, 32, s.n)
    proj = r.alloc_projected(ctx, s.n)
    r.project(ctx, g, gcam, proj)
    assert np.array_equal(proj.download(np.float32, s.n * 12).reshape(-1, 12).view(np.uint32),
                          d["projected"].view(np.uint32))
    params = TilingParams(32, 32, 16, 5.0)
    bins = TileBinner().bin(ctx, proj, s.n, params)
    assert np.array_equal(bins.sorted_keys(), d["keys"]) and np.array_equal(bins.sorted_payloads(), d["payloads"])
    assert np.array_equal(bins.tile_ranges(), d["ranges"])
    img = r.alloc_image(ctx, 32, 32)
    ft = ctx.storage_buffer_zeroed(32 * 32 * 4)
    li = ctx.storage_buffer_zeroed(32 * 32 * 4)
    r.composite_ex(ctx, proj, bins, img, params, ft, li)
    assert np.array_equal(img.download(np.float32, 32 * 32 * 4).reshape(32, 32, 4).view(np.uint32),
                          d["image"].view(np.uint32))
    dl = ctx.storage_buffer(d["dl_dc"])
    pg = r.alloc_projected_grad(ctx, s.n)
    r.composite_backward_recompute(ctx, proj, bins, ft, li, dl, pg, params)
    got = pg.download(np.float32, s.n * 12).reshape(-1, 12)
    assert rel_norm_err(got, d["projected_grad"]) < 1e-4
    r.project_backward(ctx, g, pg, gcam, mode=L.BWD_RECOMPUTE)
    for n, a in zip(NAMES, g.grads().arrays()):
        assert rel_norm_err(a, d[f"g_{n}"]) < 1e-4, n
