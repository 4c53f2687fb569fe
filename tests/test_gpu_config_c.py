"""The batched training iteration at its full size against the oracle.

One scene of 1,000,000 splats (SH order 3), 16 cameras on a seeded
radius-6 orbit with resolutions drawn from {640x480, 1280x720, 1920x1080,
2560x1440}, per-view L1 losses summed: the bench's workload.  Steps run
through hipGraph replays of nts_train_step_views, at tile 32 with the LDS
batches the autotuner picks for this workload (forward 512, backward 128),
then at tile 16.  Against the oracle run view by view
(train.rs:155-250 once per view, gradients summed over views):

- keys, payloads and ranges of three views, one of them 2560x1440, bit-exact
  against the oracle's culled binning, and with the reference's binning
  (BIN_REFERENCE) against TileBinner::bin's;
- those views' image, final T and last contributor bit-exact;
- all 16 per-view losses to 1e-9 relative;
- the summed parameter gradient within 1e-4 norm-wise per attribute;
- parameters after Adam within the 2 lr bound of a sign flip."""
import math
import time

import numpy as np
import pytest

from oracle import oracle as O
from nanotracer_amd import _lib as L
from nanotracer_amd import synth
from nanotracer_amd.api import AdamConfig, CameraUniform, GpuSplatBuffer, ViewBatch, ViewSpec

from .helpers import rel_norm_err
from .test_gpu_batch import _oracle_batch_step

pytestmark = pytest.mark.gpu

N, V, SEED = 1_000_000, 16, 0x5EED0002


def _workload():
    s = synth.make_scene(N, SEED, 3.0)
    views = []
    for eye, w, h in synth.orbit_views(V, SEED, 6.0):
        cam = CameraUniform.from_pose(eye, (0, 0, 0), (0, 1, 0), math.pi / 3, w, h, N)
        views.append((ViewSpec(cam, w, h, 12.0), O.camera_like(cam), eye))
    return s, views


def _pick_views(views):
    """The first 2560x1440 view and the first two views of other sizes."""
    big = next(k for k, v in enumerate(views) if (v[0].width, v[0].height) == (2560, 1440))
    others = [k for k, v in enumerate(views) if (v[0].width, v[0].height) != (2560, 1440)][:2]
    return [big] + others


def _state(g):
    mg, vg = g.moments()
    return (g.readback(), [a.reshape(-1).copy() for a in mg.arrays()], [a.reshape(-1).copy() for a in vg.arrays()],
            g.grad_acc())


def _check_step(b, g, views, targets, tile, before, t, picked):
    s0, m, v, acc = before
    cfg = AdamConfig()
    t0 = time.time()
    ref_s, ref_g, ref_losses = _oracle_batch_step(s0, views, targets, tile, L.BWD_RECOMPUTE, m, v, acc, t, cfg)
    print(f"tile {tile}: oracle step {time.time() - t0:.1f} s", flush=True)
    assert np.allclose(b.losses(), ref_losses, rtol=1e-9, atol=0), (b.losses(), ref_losses)
    for name, x, y in zip(("pos", "rot", "scale", "logit", "dc", "rest"), g.grads().arrays(), ref_g.arrays()):
        err = rel_norm_err(x, y)
        print(f"tile {tile}: {name} gradient norm-wise {err:.2e}", flush=True)
        assert err < 1e-4, (tile, name, err)
    for x, y in zip(g.readback().arrays(), ref_s.arrays()):
        d = np.abs(x.astype(np.float64) - y)
        assert d.max() <= 2.0 * cfg.lr + 1e-6
        assert (d > 1e-6).mean() < 2e-3
    for k in picked:
        spec, ocam, _ = views[k]
        tt = O.tiling(spec.width, spec.height, tile, spec.depth_max)
        op = O.project(s0, ocam)
        ok, opay, orr = O.bin(op, tt, cull=True)
        keys, pay, rng = b.bins(k, tile)
        assert np.array_equal(keys, ok) and np.array_equal(pay, opay) and np.array_equal(rng, orr), (tile, k)
        img, ft, li = O.composite(op, opay, orr, tt, with_state=True)
        assert np.array_equal(b.image(k).view(np.uint32), img.view(np.uint32)), (tile, k)
        gft, gli = b.state(k)
        assert np.array_equal(gft.view(np.uint32), ft.view(np.uint32)) and np.array_equal(gli, li), (tile, k)
        print(f"tile {tile}: view {k} ({spec.width}x{spec.height}, {len(keys)} pairs) bit-exact", flush=True)


def test_config_c_full_size_graph_replay(ctx):
    s, views = _workload()
    specs = [v[0] for v in views]
    pixels = sum(v.width * v.height for v in specs)
    picked = _pick_views(views)
    # targets: the perturbed scene rendered on the device (input data to both sides)
    tg = GpuSplatBuffer.upload(ctx, synth.perturb_means(s, SEED + 1), trainable=False)
    probe = ViewBatch(ctx, V, N, pixels, 400_000_000)
    probe.render(tg, specs, 16)
    targets = [probe.image(k) for k in range(V)]
    probe.render(tg, specs, 16)
    pairs16, _, ov = probe.stats()
    assert ov == 0
    probe.free()
    tg.free()
    tbuf = ctx.storage_buffer(np.concatenate([x.reshape(-1) for x in targets]))
    g = GpuSplatBuffer.upload(ctx, s)
    b = ViewBatch(ctx, V, N, pixels, int(pairs16 * 1.3) + 1_000_000)
    # tile 32, the tuned LDS batches: capture at t = 1, compare the replay at t = 2
    b.set_lds_batch(512, 128)
    b.train_step(g, specs, tbuf, 1, 32, L.LOSS_L1, L.BWD_RECOMPUTE, graph=True, write_grads=True)
    before = _state(g)
    b.train_step(g, specs, tbuf, 2, 32, L.LOSS_L1, L.BWD_RECOMPUTE, graph=True, write_grads=True)
    ctx.sync()
    pairs, visible, ov = b.stats()
    assert ov == 0
    print(f"tile 32: {pairs} pairs, {visible} visible records", flush=True)
    _check_step(b, g, views, targets, 32, before, 2, picked)
    # the reference's binning at tile 32: the picked views' keys, payloads and
    # ranges bit-exact against TileBinner::bin, images unchanged
    s_now = g.readback()
    b.render(g, specs, 32)
    culled = {k: b.image(k) for k in picked}
    b.set_binning(L.BIN_REFERENCE)
    b.render(g, specs, 32)
    rpairs, _, ov = b.stats()
    assert ov == 0 and rpairs > pairs
    for k in picked:
        spec, ocam, _ = views[k]
        tt = O.tiling(spec.width, spec.height, 32, spec.depth_max)
        op = O.project(s_now, ocam)
        rk, rp, rr = O.bin(op, tt)
        keys, pay, rng = b.bins(k, 32)
        assert np.array_equal(keys, rk) and np.array_equal(pay, rp) and np.array_equal(rng, rr), k
        assert np.array_equal(b.image(k).view(np.uint32), culled[k].view(np.uint32)), k
        print(f"reference binning: view {k} ({spec.width}x{spec.height}, {len(rk)} pairs) keys bit-exact", flush=True)
    b.set_binning(L.BIN_CULLED)
    # tile 16: a new capture, then a replay
    b.set_lds_batch(256, 256)
    b.train_step(g, specs, tbuf, 3, 16, L.LOSS_L1, L.BWD_RECOMPUTE, graph=True, write_grads=True)
    before = _state(g)
    b.train_step(g, specs, tbuf, 4, 16, L.LOSS_L1, L.BWD_RECOMPUTE, graph=True, write_grads=True)
    ctx.sync()
    assert b.stats()[2] == 0
    _check_step(b, g, views, targets, 16, before, 4, picked[:1])
    b.free()
    g.free()
