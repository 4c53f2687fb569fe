"""Adam, the fused optimiser step and the single-view training step on the
GPU against the oracle."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from nanotracer_amd import synth
from nanotracer_amd.api import AdamConfig, AdamGpu, GpuSplatBuffer
from nanotracer_amd.train import Trainer, View

from .helpers import config_a_camera, config_a_scene, rel_norm_err

pytestmark = pytest.mark.gpu


def _run_parity(ctx, n, steps, grad_fn):
    # tests/adam_gpu_parity.rs:20-98
    cfg = AdamConfig()
    cpu_p = np.full(n, 0.5, np.float32)
    cpu_m = np.zeros(n, np.float32)
    cpu_v = np.zeros(n, np.float32)
    pb, gb = ctx.storage_buffer(cpu_p), ctx.storage_buffer_zeroed(n * 4)
    mb, vb = ctx.storage_buffer_zeroed(n * 4), ctx.storage_buffer_zeroed(n * 4)
    for t in range(1, steps + 1):
        g = np.array([grad_fn(t, i) for i in range(n)], np.float32)
        O.adam_step(cpu_p, g, cpu_m, cpu_v, cfg.lr, cfg.beta1, cfg.beta2, cfg.eps, t)
        gb.upload(g)
        AdamGpu().step(ctx, pb, gb, mb, vb, n, cfg.lr, t, cfg)
    gp, gm, gv = pb.download(np.float32, n), mb.download(np.float32, n), vb.download(np.float32, n)
    assert np.abs(gp - cpu_p).max() < 1e-4 and np.abs(gm - cpu_m).max() < 1e-5 and np.abs(gv - cpu_v).max() < 1e-5
    # same arithmetic on both sides: bit-identical
    assert np.array_equal(gp, cpu_p) and np.array_equal(gm, cpu_m) and np.array_equal(gv, cpu_v)


def test_adam_parity_positive_grad(ctx):
    _run_parity(ctx, 256, 10, lambda t, i: 0.1)


def test_adam_parity_negative_grad(ctx):
    _run_parity(ctx, 256, 10, lambda t, i: -0.2)


def test_adam_parity_mixed_signs(ctx):
    _run_parity(ctx, 1024, 20, lambda t, i: (1.0 if i % 2 == 0 else -1.0) * (0.05 + 0.01 * t))


def test_adam_parity_single_step_matches_moments(ctx):
    _run_parity(ctx, 64, 1, lambda t, i: 0.01 * (i + 1.0))


def test_adam_parity_large_n(ctx):
    _run_parity(ctx, 8192, 5, lambda t, i: 0.05 * math.sin(i * 0.01 + t * 0.1))


def _oracle_fused_step(s, g, m, v, acc, t, cfg_pos, cfg_attr):
    """accumulate_grad + six per-attribute Adam steps + constraints
    (train.rs:228-245) on host arrays; m, v are lists of 6 flat slabs."""
    s = s.copy()
    O.accumulate_grad(g.pos, acc)
    arrs = list(s.arrays())
    for k, (p, gr) in enumerate(zip(arrs, g.arrays())):
        cfg = cfg_pos if k == 0 else cfg_attr
        flat = p.reshape(-1)
        O.adam_step(flat, gr.reshape(-1), m[k], v[k], cfg.lr, cfg.beta1, cfg.beta2, cfg.eps, t)
    O.apply_constraints(s.rot, s.log_scale, s.logit)
    return s


def test_fused_step_bit_exact(ctx):
    # adam_gpu_parity.rs:161-376 (step_all on splats with distinct
    # per-attribute gradients), extended to constraints and grad_acc
    n = 3000
    s = synth.make_scene(n, 99, 1.0)
    s.rot *= 1.3  # off the unit sphere: renormalisation must act
    s.log_scale[::5] = -11.0  # below the clamp
    s.logit[::7] = 12.0
    rng = np.random.default_rng(1)
    g = synth.HostSplats(*[rng.normal(0, sc, a.shape).astype(np.float32)
                           for a, sc in zip(s.arrays(), (1e-3, 2e-3, 3e-3, 4e-2, 5e-3, 6e-4))])
    gpu = GpuSplatBuffer.upload(ctx, s)
    gpu.set_grads(g)
    acc0 = rng.uniform(0, 1, n).astype(np.float32)
    gpu.set_grad_acc(acc0)
    cfg_pos, cfg_attr = AdamConfig(lr=1.6e-4), AdamConfig(lr=2.5e-3)
    m = [np.zeros(a.size, np.float32) for a in s.arrays()]
    v = [np.zeros(a.size, np.float32) for a in s.arrays()]
    acc = acc0.copy()
    ref = s
    for t in (1, 2, 3):
        AdamGpu().step_all_fused(ctx, gpu, t, cfg_pos, cfg_attr)
        ref = _oracle_fused_step(ref, g, m, v, acc, t, cfg_pos, cfg_attr)
        gpu.set_grads(g)
    out = gpu.readback()
    for a, b in zip(out.arrays(), ref.arrays()):
        assert np.array_equal(a, b)
    gm, gv = gpu.moments()
    for k, (a, b) in enumerate(zip(gm.arrays(), gv.arrays())):
        assert np.array_equal(a.reshape(-1), m[k]) and np.array_equal(b.reshape(-1), v[k])
    assert np.array_equal(gpu.grad_acc(), acc)


def _oracle_step(s, ocam, t_params, target, m, v, acc, step, cfg):
    """One full step of the reference sequence on the oracle."""
    t = t_params
    op = O.project(s, ocam)
    keys, pay, rg = O.bin(op, t)
    img = O.composite(op, pay, rg, t)
    loss, dl = O.loss_grad(img, target, 0)
    pg = O.composite_backward_ref(op, pay, rg, t, img, dl)
    g = O.project_backward(s, ocam, pg)
    s2 = _oracle_fused_step(s, g, m, v, acc, step, cfg, cfg)
    return s2, g, img, loss


def test_config_a_training_per_step_parity(ctx):
    """Config A (5,000 splats, 256^2, L1, Adam): 100 steps on the GPU, each
    step checked against the oracle started from the GPU's state."""
    s = config_a_scene()
    c = config_a_camera()
    target_s = synth.perturb_means(s, 0x7A46E7)
    ocam = O.camera_from_pose(c["eye"], c["target"], c["up"], c["fov_y"], c["width"], c["height"], s.n)
    t = O.tiling(c["width"], c["height"], 16, c["depth_max"])
    tp = O.project(target_s, ocam)
    tk, tpay, trg = O.bin(tp, t)
    target = O.composite(tp, tpay, trg, t)
    gpu = GpuSplatBuffer.upload(ctx, s)
    view = View(c["eye"], c["width"], c["height"], target)
    tr = Trainer(ctx, gpu, [view], c["depth_max"])
    cfg = AdamConfig()
    losses = []
    for it in range(100):
        check = it < 5 or it % 10 == 0
        if check:
            before = gpu.readback()
            mg, vg = gpu.moments()
            m = [a.reshape(-1).copy() for a in mg.arrays()]
            v = [a.reshape(-1).copy() for a in vg.arrays()]
            acc = gpu.grad_acc()
        res = tr.step(it)
        losses.append(res.loss)
        if check:
            ref_s, ref_g, ref_img, ref_loss = _oracle_step(before, ocam, t, target, m, v, acc, it + 1, cfg)
            assert abs(res.loss - ref_loss) <= 1e-9, (it, res.loss, ref_loss)
            gg = gpu.grads()
            for name, a, b in zip(("pos", "rot", "scale", "logit", "dc", "rest"), gg.arrays(), ref_g.arrays()):
                assert rel_norm_err(a, b) < 1e-4, (it, name, rel_norm_err(a, b))
            after = gpu.readback()
            lr = cfg.lr
            for a, b in zip(after.arrays(), ref_s.arrays()):
                d = np.abs(a.astype(np.float64) - b)
                # identical except where a tiny gradient's sign differs
                # under the atomic sum order: Adam then moves by <= 2 lr
                assert d.max() <= 2.0 * lr + 1e-6
                assert (d > 1e-6).mean() < 2e-3
    assert losses[-1] < losses[0]


def test_recompute_mode_training_per_step_parity(ctx):
    """Config A in the Differentiation form (recompute backward, exact chain
    rule): each checked step matches the oracle's recompute path."""
    from nanotracer_amd import _lib as L

    s = config_a_scene()
    c = config_a_camera()
    ocam = O.camera_from_pose(c["eye"], c["target"], c["up"], c["fov_y"], c["width"], c["height"], s.n)
    t = O.tiling(c["width"], c["height"], 16, c["depth_max"])
    tp = O.project(synth.perturb_means(s, 0x7A46E7), ocam)
    tk, tpay, trg = O.bin(tp, t)
    target = O.composite(tp, tpay, trg, t)
    gpu = GpuSplatBuffer.upload(ctx, s)
    tr = Trainer(ctx, gpu, [View(c["eye"], c["width"], c["height"], target)], c["depth_max"],
                 backward_mode=L.BWD_RECOMPUTE)
    cfg = AdamConfig()
    losses = []
    for it in range(30):
        check = it in (0, 1, 2, 10, 29)
        if check:
            before = gpu.readback()
            mg, vg = gpu.moments()
            m = [a.reshape(-1).copy() for a in mg.arrays()]
            v = [a.reshape(-1).copy() for a in vg.arrays()]
            acc = gpu.grad_acc()
        res = tr.step(it)
        losses.append(res.loss)
        if check:
            op = O.project(before, ocam)
            keys, pay, rg = O.bin(op, t)
            img, ft, li = O.composite(op, pay, rg, t, with_state=True)
            loss, dl = O.loss_grad(img, target, 0)
            pg = O.composite_backward_recompute(op, pay, rg, t, ft, li, dl)
            ref_g = O.project_backward(before, ocam, pg, full=True)
            ref_s = _oracle_fused_step(before, ref_g, m, v, acc, it + 1, cfg, cfg)
            assert abs(res.loss - loss) <= 1e-9
            gg = gpu.grads()
            for name, a, b in zip(("pos", "rot", "scale", "logit", "dc", "rest"), gg.arrays(), ref_g.arrays()):
                assert rel_norm_err(a, b) < 1e-4, (it, name, rel_norm_err(a, b))
            after = gpu.readback()
            for a, b in zip(after.arrays(), ref_s.arrays()):
                d = np.abs(a.astype(np.float64) - b)
                assert d.max() <= 2.0 * cfg.lr + 1e-6
                assert (d > 1e-6).mean() < 2e-3
    assert losses[-1] < losses[0]
