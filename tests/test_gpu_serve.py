"""The render server: requests of several resident scenes batched as they
arrive, several batches in flight.  Each result must be bit-identical to
the same request rendered alone, and float results to the host
restatement of the forward."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from nanotracer_amd import _lib as L
from nanotracer_amd import synth
from nanotracer_amd.api import CameraUniform, RenderServer, ViewSpec

pytestmark = pytest.mark.gpu

RES = ((160, 120), (320, 240), (256, 256), (480, 270))


def _scenes():
    return [synth.make_scene(n, 0x5EED0100 + k, 2.0) for k, n in enumerate((5_000, 12_000, 30_000, 8_000))]


def _requests(scenes, count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        sid = int(rng.integers(len(scenes)))
        w, h = RES[int(rng.integers(len(RES)))]
        az, el = float(rng.uniform(0, 2 * math.pi)), float(rng.uniform(-0.4, 0.4))
        eye = synth.orbit_eye(5.0, az, el)
        cam = CameraUniform.from_pose(eye, (0, 0, 0), (0, 1, 0), math.pi / 3, w, h, scenes[sid].n)
        out.append((sid, ViewSpec(cam, w, h, 10.0)))
    return out


def _serve(srv, reqs, fmt):
    ids = [srv.submit(sid, v, fmt)[0] for sid, v in reqs]
    res = [srv.wait(i) for i in ids]
    return [r[0] for r in res], [r[1] for r in res]


@pytest.mark.parametrize("fmt", [L.OUT_RGBA8, L.OUT_F32])
def test_mixed_batches_equal_solo_renders(ctx, fmt):
    scenes = _scenes()
    srv = RenderServer(slots=2, max_views=8, max_pixels=1 << 20, max_records=200_000, pair_capacity=4_000_000)
    ids = [srv.add_scene(s) for s in scenes]
    assert ids == [0, 1, 2, 3]
    reqs = _requests(scenes, 40, 1)
    mixed, infos = _serve(srv, reqs, fmt)
    st = srv.stats()
    assert st["completed"] == 40 and st["failed"] == 0
    assert st["batches"] < 40 and max(i["batch_views"] for i in infos) > 1  # requests were batched
    for (sid, v), img in zip(reqs, mixed):
        solo, info = srv.wait(srv.submit(sid, v, fmt)[0])
        assert info["batch_views"] == 1
        assert np.array_equal(img.view(np.uint8), solo.view(np.uint8))
    srv.close()


def test_mixed_formats_in_one_batch(ctx):
    # rgba8 (4 B per pixel) and float (16 B per pixel) results of one batch
    # share the pinned staging buffer: each must come back whole
    scenes = _scenes()
    srv = RenderServer(slots=1, max_views=16, max_pixels=1 << 21, max_records=200_000, pair_capacity=8_000_000)
    for s in scenes:
        srv.add_scene(s)
    reqs = _requests(scenes, 16, 4)
    fmts = [L.OUT_F32 if k % 3 == 0 else L.OUT_RGBA8 for k in range(16)]
    ids = [srv.submit(sid, v, f)[0] for (sid, v), f in zip(reqs, fmts)]
    res = [srv.wait(i) for i in ids]
    assert max(info["batch_views"] for _, info in res) > 1
    for (sid, v), f, (img, _) in zip(reqs, fmts, res):
        solo, info = srv.wait(srv.submit(sid, v, f)[0])
        assert info["batch_views"] == 1
        assert np.array_equal(img.view(np.uint8), solo.view(np.uint8))
    srv.close()


@pytest.mark.parametrize("tile,lds", [(16, 128), (32, 512)])
def test_float_results_match_the_host_forward(ctx, tile, lds):
    # images depend on the tile size (the reference bins by tile rectangles):
    # the server and the oracle use the same one
    scenes = _scenes()
    srv = RenderServer(slots=3, max_views=6, max_pixels=1 << 20, max_records=200_000, pair_capacity=4_000_000,
                       tile_size=tile, lds_batch=lds)
    for s in scenes:
        srv.add_scene(s)
    reqs = _requests(scenes, 12, 2)
    imgs, _ = _serve(srv, reqs, L.OUT_F32)
    for (sid, v), img in zip(reqs, imgs):
        ocam = O.camera_like(v.camera)
        t = O.tiling(v.width, v.height, tile, v.depth_max)
        op = O.project(scenes[sid], ocam)
        k, pay, rg = O.bin(op, t)
        ref = O.composite(op, pay, rg, t)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))
        assert np.array_equal(O.tonemap(img), O.tonemap(ref))
    srv.close()


def test_overflowing_batches_rerun_one_by_one(ctx):
    scenes = _scenes()
    # room for a few views' pairs at a time, not for a full batch
    srv = RenderServer(slots=2, max_views=8, max_pixels=1 << 20, max_records=200_000, pair_capacity=150_000)
    for s in scenes:
        srv.add_scene(s)
    reqs = _requests(scenes, 24, 3)
    out = []
    for (sid, v) in reqs:
        out.append(srv.submit(sid, v, L.OUT_RGBA8)[0])
    results = []
    for rid in out:
        try:
            results.append(srv.wait(rid)[0])
        except L.NtsError as e:
            assert e.code == L.NTS_E_CAPACITY
            results.append(None)
    st = srv.stats()
    assert st["completed"] + st["failed"] == 24
    big = RenderServer(slots=1, max_views=1, max_pixels=1 << 20, max_records=200_000, pair_capacity=4_000_000)
    for s in scenes:
        big.add_scene(s)
    for (sid, v), img in zip(reqs, results):
        if img is not None:
            solo, _ = big.wait(big.submit(sid, v, L.OUT_RGBA8)[0])
            assert np.array_equal(img, solo)
    srv.close()
    big.close()


def test_admission_limits(ctx):
    scenes = _scenes()
    srv = RenderServer(slots=1, max_views=4, max_pixels=100_000, max_records=40_000, pair_capacity=1_000_000)
    sid = srv.add_scene(scenes[0])
    cam = CameraUniform.from_pose((0, 0, 5), (0, 0, 0), (0, 1, 0), math.pi / 3, 480, 270, scenes[0].n)
    rid, _ = srv.submit(sid, ViewSpec(cam, 480, 270, 10.0))  # 129,600 px > 100,000
    with pytest.raises(L.NtsError):
        srv.wait(rid)
    big = synth.make_scene(50_000, 7, 2.0)  # more records than a batch holds
    with pytest.raises(L.NtsError):
        srv.add_scene(big)
    srv.close()
    tight = RenderServer(slots=1, max_views=4, max_pixels=100_000, max_records=40_000, pair_capacity=1_000_000,
                         hbm_budget=200 << 20)
    with pytest.raises(L.NtsError):
        for _ in range(100):  # 30k splats = 7 MB each: stops at the budget
            tight.add_scene(synth.make_scene(30_000, 9, 2.0))
    assert tight.stats()["hbm_in_use"] <= 200 << 20
    tight.close()


def test_scenes_load_from_ply(ctx, tmp_path):
    from nanotracer_amd.api import write_ply

    scenes = _scenes()[:2]
    paths = []
    for k, s in enumerate(scenes):
        p = tmp_path / f"scene{k}.ply"
        write_ply(p, s)
        paths.append(p)
    srv = RenderServer(slots=1, max_views=4, max_pixels=1 << 20, max_records=200_000, pair_capacity=4_000_000)
    a = [srv.add_scene(s) for s in scenes]
    b = [srv.add_scene_ply(p) for p in paths]
    assert a == [0, 1] and b == [2, 3]
    for sid, v in _requests(scenes, 6, 3):
        x, _ = srv.wait(srv.submit(sid, v, L.OUT_F32)[0])
        y, _ = srv.wait(srv.submit(sid + 2, v, L.OUT_F32)[0])
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    with pytest.raises(L.NtsError):
        srv.add_scene_ply(tmp_path / "missing.ply")
    srv.close()
