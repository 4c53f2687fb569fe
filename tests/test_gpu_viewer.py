"""The viewer's per-frame render sequence (State::frame, nano-view
lib.rs:754-796) on the device against the oracle: composite and rgba8
bit-exact, across splat swaps (lib.rs:712-731) and a viewport resize, and
the live-training channel (run_with_training, lib.rs:76-139) rendering the
published snapshots."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from nanotracer_amd import synth
from nanotracer_amd.train import TrainConfig, View
from nanotracer_amd.viewer import FrameRenderer, OrbitCamera, run_with_training

pytestmark = pytest.mark.gpu


def _oracle_frame(s, orbit, w, h):
    cam = O.camera_from_pose(orbit.position(), orbit.target, (0.0, 1.0, 0.0), orbit.fov_y, w, h, s.n)
    t = O.tiling(w, h, 16, orbit.depth_max())
    p = O.project(s, cam)
    k, pay, rg = O.bin(p, t)
    img = O.composite(p, pay, rg, t)
    return img, O.tonemap(img)


def _check(r, s, orbit):
    r.frame(orbit)
    img, rgba = _oracle_frame(s, orbit, r.image_w, r.image_h)
    assert np.array_equal(r.composite().view(np.uint32), img.view(np.uint32))
    assert np.array_equal(r.rgba8(), rgba)


def test_frames_match_oracle_across_swaps_and_resize(ctx):
    s = synth.make_scene(20_000, 0x71E3, 2.0)
    orbit = OrbitCamera(distance=6.0, azimuth=0.4, elevation=0.3)
    r = FrameRenderer(ctx, s, 320, 200)
    _check(r, s, orbit)
    orbit.azimuth = 1.9
    _check(r, s, orbit)
    # fewer splats: in place; more than the capacity: a new scene
    fewer = synth.make_scene(12_000, 0x71E4, 2.0)
    r.update_splats(fewer)
    _check(r, fewer, orbit)
    more = synth.make_scene(30_000, 0x71E5, 2.0)
    r.update_splats(more)
    assert r.gpu_splats.capacity >= 30_000
    _check(r, more, orbit)
    r.resize_viewport(256, 144)
    _check(r, more, OrbitCamera(distance=30.0, azimuth=-0.7, elevation=-0.2))  # depth_max 120
    r.free()


def test_live_training_renders_published_snapshots(ctx):
    s = synth.make_scene(3000, 0x71E6, 1.0)
    views = []
    for k in range(2):
        eye = synth.orbit_eye(3.0, math.pi * k, 0.2)
        cam = O.camera_from_pose(eye, (0, 0, 0), (0, 1, 0), math.pi / 3, 96, 64, s.n)
        t = O.tiling(96, 64, 16, 6.0)
        tp = O.project(synth.perturb_means(s, 0xF00D + k, 0.05), cam)
        kk, pay, rg = O.bin(tp, t)
        views.append(View(eye, 96, 64, O.composite(tp, pay, rg, t)))
    cfg = TrainConfig(iterations=8, max_splats=4000, diagnostics_every=0)
    orbit = OrbitCamera(distance=3.5, azimuth=0.3, elevation=0.2)
    frames = []
    r, ch = run_with_training(ctx, s, views, 6.0, orbit, cfg, 160, 96, publish_every=2,
                              on_frame=lambda rr, st, snap: frames.append((st.total_pairs, snap.version)))
    assert frames and ch.error is None
    its = [h[0] for h in ch.history]
    assert its == sorted(its) and its[-1] == 7 and all((i + 1) % 2 == 0 for i in its)
    # the last frame shows the final snapshot
    final = ch.latest()
    assert final.version == len(range(1, 9, 2)) and frames[-1][1] == final.version
    _check(r, final.splats, orbit)
    r.free()
