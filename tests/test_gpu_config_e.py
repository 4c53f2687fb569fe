"""The render server at its full size against solo renders and the oracle.

64 resident scenes with splat counts log-uniform in [50k, 3M] (SH order 3),
a 24 GiB HBM budget (the scenes, then three slot workspaces that each
batch's working set must fit), up to 32 views per launch, resolutions drawn from
{512^2, 1280x720, 1920x1080, 3840x2160}: the bench's serving workload.  A
burst of requests arrives at once, so batches form while earlier ones are
on the device.  Then:

- batches hold more than one view on average, and at least one holds 8+;
- eight sampled requests, one of them 3840x2160 on the largest (3M-splat
  class) scene, equal their solo render bit for bit, and their float images
  equal the host restatement of the forward (project_gaussians,
  TileBinner::bin, rasterize: nano-view/src/lib.rs:761-796 per frame) bit
  for bit."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from nanotracer_amd import _lib as L
from nanotracer_amd import synth
from nanotracer_amd.api import CameraUniform, RenderServer, ViewSpec, device_memory

pytestmark = pytest.mark.gpu

SEED = 0x5EED0002 + 5
TILE, LDS = 32, 512


def _request(rng, sizes, sid=None, res=None):
    sid = int(rng.integers(len(sizes))) if sid is None else sid
    w, h = synth.RESOLUTIONS_E[int(rng.integers(4))] if res is None else res
    eye = synth.orbit_eye(float(rng.uniform(6.0, 9.0)), float(rng.uniform(0, 2 * math.pi)),
                          float(rng.uniform(-0.5, 0.5)))
    cam = CameraUniform.from_pose(eye, (0, 0, 0), (0, 1, 0), math.pi / 3, w, h, int(sizes[sid]))
    return sid, ViewSpec(cam, w, h, 24.0)


def test_config_e_full_size_burst(ctx):
    scenes = list(synth.serving_scenes(64, SEED))
    sizes = [s.n for s in scenes]
    resident = sum(((n + 63) // 64) * 64 * 4 * 59 for n in sizes)
    # the serving budget: 24 GiB of what the device reports free; the scenes
    # first, the rest split into three slot workspaces that bound each batch
    free, _ = device_memory(0)
    budget = min(24 << 30, int(0.9 * free))
    slot_bytes = (budget - resident) // 3 - (1 << 20)
    srv = RenderServer(hbm_budget=budget, slots=3, max_views=32, max_pixels=96 << 20, max_records=0,
                       pair_capacity=0, tile_size=TILE, lds_batch=LDS, slot_bytes=slot_bytes)
    for s in scenes:
        srv.add_scene(s)
    st = srv.stats()
    assert st["scenes"] == 64 and st["hbm_in_use"] <= budget
    print(f"{sum(sizes)} resident splats, {st['hbm_in_use'] / 1e9:.1f} GB of a {budget / 1e9:.1f} GB budget "
          f"({slot_bytes / 1e9:.2f} GB per slot)", flush=True)

    rng = np.random.default_rng(SEED + 3)
    big = int(np.argmax(sizes))
    reqs = [_request(rng, sizes) for _ in range(160)]
    reqs[37] = _request(rng, sizes, big, (3840, 2160))
    sample = [37, 0, 11, 52, 80, 101, 123, 159]
    ids = []
    for k, (sid, v) in enumerate(reqs):
        ids.append(srv.submit(sid, v, L.OUT_F32 if k in sample else L.OUT_RGBA8)[0])
    results = [srv.wait(i) for i in ids]
    st = srv.stats()
    views = [info["batch_views"] for _, info in results]
    print(f"burst of {len(reqs)}: {st['batches']} batches, mean {np.mean(views):.2f} views per batch "
          f"(max {max(views)}), max in flight {st['max_in_flight']}, {st['overflow_reruns']} reruns, "
          f"{st['memory_limited']} batches closed by memory", flush=True)
    assert st["completed"] >= len(reqs) and st["failed"] == 0
    assert np.mean(views) > 1.0 and max(views) >= 8
    assert st["memory_limited"] > 0  # the budget, not the view cap, bounds batches
    for k in sample:
        sid, v = reqs[k]
        img = results[k][0]
        solo, info = srv.wait(srv.submit(sid, v, L.OUT_F32)[0])
        assert info["batch_views"] == 1
        assert np.array_equal(img.view(np.uint32), solo.view(np.uint32)), k
        ocam = O.camera_like(v.camera)
        t = O.tiling(v.width, v.height, TILE, v.depth_max)
        op = O.project(scenes[sid], ocam)
        _, pay, rg = O.bin(op, t)
        ref = O.composite(op, pay, rg, t)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), k
        print(f"request {k}: scene {sid} ({sizes[sid]} splats) {v.width}x{v.height} in a batch of "
              f"{results[k][1]['batch_views']}: equal to its solo render and the host forward", flush=True)
    srv.close()
