"""Reference semantics of the scene and optimiser entry points, and the
batched step's handling of a pair overflow.

- AdamGpu's three dispatches called one by one (accumulate_grad,
  step_all, apply_constraints: train.rs:228-245) equal the oracle's
  sequence and the fused kernel, bit for bit.
- GpuSplatBuffer::sync_from (splat_gpu.rs:124-150) rewrites parameters and
  keeps moments, gradients and grad_acc; upload starts them at zero.
- A scene created without NTS_SCENE_TRAINABLE renders like a trainable one
  and refuses the calls that need gradients or moments.
- A batched step whose pairs overflow the capacity leaves the scene
  untouched and is reported by the next stats call (once)."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from nanotracer_amd import _lib as L
from nanotracer_amd import synth
from nanotracer_amd.api import (AdamConfig, AdamGpu, CameraUniform, GpuSplatBuffer, Rasterizer, ViewBatch,
                                ViewSpec)

from .test_gpu_optim import _oracle_fused_step

pytestmark = pytest.mark.gpu


def _scene_with_state(ctx, n=3000, seed=99):
    s = synth.make_scene(n, seed, 1.0)
    s.rot *= 1.3
    s.log_scale[::5] = -11.0
    s.logit[::7] = 12.0
    rng = np.random.default_rng(seed + 1)
    g = synth.HostSplats(*[rng.normal(0, 0.05, a.shape).astype(np.float32) for a in s.arrays()])
    m = synth.HostSplats(*[rng.normal(0, 0.01, a.shape).astype(np.float32) for a in s.arrays()])
    v = synth.HostSplats(*[rng.uniform(0, 1e-3, a.shape).astype(np.float32) for a in s.arrays()])
    acc = rng.uniform(0, 1, n).astype(np.float32)
    gb = GpuSplatBuffer.upload(ctx, s)
    gb.set_grads(g)
    gb.set_moments(m, v)
    gb.set_grad_acc(acc)
    return s, g, m, v, acc, gb


def test_separate_optimiser_dispatches_match_oracle_and_fused(ctx):
    cfg = AdamConfig()
    t = 7
    s, g, m, v, acc, gb = _scene_with_state(ctx)
    adam = AdamGpu()
    adam.accumulate_grad(ctx, gb)
    adam.step_all(ctx, gb, t, cfg, cfg)
    adam.apply_constraints(ctx, gb)
    mf = [a.reshape(-1).copy() for a in m.arrays()]
    vf = [a.reshape(-1).copy() for a in v.arrays()]
    acc_ref = acc.copy()
    ref = _oracle_fused_step(s, g, mf, vf, acc_ref, t, cfg, cfg)
    for x, y in zip(gb.readback().arrays(), ref.arrays()):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    gm, gv = gb.moments()
    for x, y in zip(gm.arrays(), mf):
        assert np.array_equal(x.reshape(-1), y)
    for x, y in zip(gv.arrays(), vf):
        assert np.array_equal(x.reshape(-1), y)
    assert np.array_equal(gb.grad_acc(), acc_ref)
    # the fused kernel gives the same bits
    _, _, _, _, _, fb = _scene_with_state(ctx)
    adam.step_all_fused(ctx, fb, t, cfg, cfg)
    for x, y in zip(fb.readback().arrays(), ref.arrays()):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    assert np.array_equal(fb.grad_acc(), acc_ref)
    # step_all alone leaves the constraints undone and grad_acc untouched
    _, _, _, _, _, sb = _scene_with_state(ctx)
    adam.step_all(ctx, sb, t, cfg, cfg)
    assert np.array_equal(sb.grad_acc(), acc)
    assert np.abs(np.linalg.norm(sb.readback().rot, axis=1) - 1.0).max() > 0.1
    for b in (gb, fb, sb):
        b.free()


def test_sync_from_keeps_optimiser_state(ctx):
    s, g, m, v, acc, gb = _scene_with_state(ctx)
    s2 = synth.make_scene(s.n, 1234, 1.0)
    gb.sync_from(s2)
    for x, y in zip(gb.readback().arrays(), s2.arrays()):
        assert np.array_equal(x, y)
    gm, gv = gb.moments()
    for x, y in zip(gm.arrays(), m.arrays()):
        assert np.array_equal(x, y)
    for x, y in zip(gv.arrays(), v.arrays()):
        assert np.array_equal(x, y)
    for x, y in zip(gb.grads().arrays(), g.arrays()):
        assert np.array_equal(x, y)
    assert np.array_equal(gb.grad_acc(), acc)
    # another count is not a sync
    with pytest.raises(L.NtsError) as e:
        gb.sync_from(synth.make_scene(s.n - 1, 5, 1.0))
    assert e.value.code == L.NTS_E_INVALID
    # upload: fresh optimiser state
    gb.upload_from(s2)
    gm, gv = gb.moments()
    assert all(not x.any() for x in gm.arrays()) and all(not x.any() for x in gv.arrays())
    assert not gb.grad_acc().any()
    gb.free()


def test_render_only_scene(ctx):
    s = synth.make_scene(4000, 0x5EED0000, 1.0)
    cam = CameraUniform.from_pose((0, 0, 3), (0, 0, 0), (0, 1, 0), math.pi / 3, 128, 96, s.n)
    r = Rasterizer()
    out = []
    for trainable in (True, False):
        gb = GpuSplatBuffer.upload(ctx, s, trainable=trainable)
        proj = r.alloc_projected(ctx, s.n)
        r.project(ctx, gb, cam, proj)
        out.append(proj.download(np.float32, s.n * 12))
        if not trainable:
            for call in (gb.grads, gb.moments, gb.grad_acc, gb.zero_grads, gb.zero_grad_acc, gb.prune,
                         lambda: AdamGpu().step_all_fused(ctx, gb, 1, AdamConfig(), AdamConfig()),
                         lambda: AdamGpu().accumulate_grad(ctx, gb)):
                with pytest.raises(L.NtsError) as e:
                    call()
                assert e.value.code == L.NTS_E_STATE
            AdamGpu().apply_constraints(ctx, gb)  # parameters only: allowed
            gb.sync_from(s)
        proj.free()
        gb.free()
    assert np.array_equal(out[0], out[1])


def test_overflowing_step_leaves_scene_untouched(ctx):
    s = synth.make_scene(20_000, 0x5EED0002, 3.0)
    res = ((320, 240), (256, 256))
    specs = []
    for eye, w, h in synth.orbit_views(2, 5, 6.0, res):
        specs.append(ViewSpec(CameraUniform.from_pose(eye, (0, 0, 0), (0, 1, 0), math.pi / 3, w, h, s.n), w, h,
                              12.0))
    pixels = sum(w * h for w, h in res)
    tbuf = ctx.storage_buffer(np.full(pixels * 4, 0.25, np.float32))
    gb = GpuSplatBuffer.upload(ctx, s)
    before = gb.readback()
    m0, v0 = gb.moments()
    small = ViewBatch(ctx, 2, s.n, pixels, 1000)  # far below the pairs of these views
    for t, graph in ((1, False), (2, True), (3, True)):
        small.train_step(gb, specs, tbuf, t, 16, L.LOSS_L1, L.BWD_RECOMPUTE, graph=graph, write_grads=True)
    ctx.sync()
    for x, y in zip(gb.readback().arrays(), before.arrays()):
        assert np.array_equal(x, y)
    m1, v1 = gb.moments()
    assert all(np.array_equal(x, y) for x, y in zip(m1.arrays(), m0.arrays()))
    assert all(np.array_equal(x, y) for x, y in zip(v1.arrays(), v0.arrays()))
    assert not gb.grad_acc().any()
    pairs, _, overflow = small.stats()
    assert pairs > 1000 and overflow == 3  # three skipped updates, reported once
    # a batch with room trains the same scene; the latch was cleared by the read
    big = ViewBatch(ctx, 2, s.n, pixels, 4_000_000)
    big.train_step(gb, specs, tbuf, 4, 16, L.LOSS_L1, L.BWD_RECOMPUTE)
    assert big.stats()[2] == 0
    assert any(not np.array_equal(x, y) for x, y in zip(gb.readback().arrays(), before.arrays()))
    small.train_step(gb, specs, tbuf, 5, 16, L.LOSS_L1, L.BWD_RECOMPUTE)
    assert small.stats()[2] == 1 and small.stats()[2] == 0
    for b in (small, big, gb):
        b.free()
