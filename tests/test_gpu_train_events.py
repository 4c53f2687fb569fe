"""train() across prune and densify events with the reference's constants
(train.rs:103-114: prune every 100 from 50 at logit < -5.3, densify every
200 from 100 above an average |dpos|_1 of 2e-4, split above log-scale
-2.3) against an oracle-driven loop, 201 iterations on a 5,000-splat
256x256 scene (four views cycled as train.rs:155-165 does).

- The events happen at exactly the reference's iterations: a prune at 100,
  a prune then a densify at 200, nothing anywhere else.
- Each event's result equals the oracle's prune / densify
  (train.rs:259-362, 414-602) applied to the device's own state just
  before it, bit for bit: every parameter, both Adam moments and grad_acc
  (zero after a densify).
- Between events the oracle runs the reference step sequence
  (train.rs:155-250) from the device state of the segment's start; every
  iteration's loss agrees to 1e-3 relative.  (Float atomics sum the
  device's gradients in another order, and Adam turns a near-zero gradient
  of the other sign into a step of the other sign, so the two trajectories
  are compared by loss, not by bits.)"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from nanotracer_amd import synth
from nanotracer_amd.api import AdamConfig
from nanotracer_amd.train import TrainConfig, View, train

from .test_gpu_optim import _oracle_fused_step

pytestmark = pytest.mark.gpu

W = H = 256
DEPTH_MAX = 6.0


def _views(s, n=4):
    out = []
    for k in range(n):
        eye = synth.orbit_eye(3.0, 2 * math.pi * k / n, 0.2)
        cam = O.camera_from_pose(eye, (0, 0, 0), (0, 1, 0), math.pi / 3, W, H, s.n)
        t = O.tiling(W, H, 16, DEPTH_MAX)
        tp = O.project(synth.perturb_means(s, 0xC0FFEE + k, 0.05), cam)
        _, pay, rg = O.bin(tp, t)
        out.append(View(eye, W, H, O.composite(tp, pay, rg, t)))
    return out


def _download(g):
    m, v = g.moments()
    mo = []
    for a, b in zip(m.arrays(), v.arrays()):
        mo += [a.reshape(-1).copy(), b.reshape(-1).copy()]
    return g.readback(), mo, g.grad_acc()


def _oracle_step(s, view, mo, acc, t, cfg):
    cam = O.camera_from_pose(view.eye, view.target, view.up, view.fov_y, W, H, s.n)
    tt = O.tiling(W, H, 16, DEPTH_MAX)
    op = O.project(s, cam)
    _, pay, rg = O.bin(op, tt)
    img = O.composite(op, pay, rg, tt)
    loss, dl = O.loss_grad(img, view.target_image, 0)
    pg = O.composite_backward_ref(op, pay, rg, tt, img, dl)
    g = O.project_backward(s, cam, pg)
    m, v = mo[0::2], mo[1::2]
    s2 = _oracle_fused_step(s, g, m, v, acc, t, cfg, cfg)
    return s2, [x for pair in zip(m, v) for x in pair], loss


def test_train_events_match_the_oracle(ctx):
    s = synth.make_scene(5000, 0x5EED0000, 1.0, logit_mu=-2.0, logit_sd=2.0)
    assert (s.logit < -5.3).sum() > 50  # splats for the first prune
    views = _views(s)
    cfg = TrainConfig(iterations=201, max_splats=6000, diagnostics_every=0)
    losses, counts, events = [], [], {}
    train(ctx, s, views, DEPTH_MAX, cfg, on_iter=lambda it, loss, n: (losses.append(loss), counts.append(n)),
          log=lambda *_: None, on_event=lambda it, stage, g: events.setdefault(it, {}).update({stage: _download(g)}))
    assert sorted(events) == [100, 200]
    acfg = AdamConfig()
    # each event from the device's own state
    for it, ev in events.items():
        (bs, bmo, bacc), (as_, amo, aacc) = ev["before"], ev["after"]
        rs, rmo, racc = O.prune(bs, -5.3, bmo, bacc)
        removed = bs.n - rs.n
        if it == 200:
            # grad_count: the 201 steps 0..200 since the start (train.rs:251, 327-331)
            rs, rmo, racc, (cl, sp, par) = O.densify(rs, it + 1, 2e-4, -2.3, cfg.max_splats, it, rmo, racc)
            assert cl + sp > 0 and rs.n <= cfg.max_splats
            racc = np.zeros_like(racc)
        assert removed > 0 or it == 200
        assert as_.n == rs.n, (it, as_.n, rs.n)
        for a, b in zip(as_.arrays(), rs.arrays()):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), it
        for a, b in zip(amo, rmo):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), it
        assert np.array_equal(aacc.view(np.uint32), racc.view(np.uint32)), it
        print(f"iter {it}: {bs.n} -> {as_.n} splats, bit-exact against the oracle's events", flush=True)
    # the count moves only at the events
    for k in range(1, len(counts)):
        if counts[k] != counts[k - 1]:
            assert k in events, k
    # segments of the oracle-driven loop, each from the device state at its start
    seg_start = {0: (s.copy(), [np.zeros(a.size, np.float32) for a in s.arrays() for _ in (0, 1)],
                     np.zeros(s.n, np.float32))}
    for it, ev in events.items():
        seg_start[it + 1] = ev["after"]
    starts = sorted(seg_start)
    ref_losses = []
    for a, b in zip(starts, starts[1:] + [cfg.iterations]):
        cur, mo, acc = seg_start[a]
        cur, mo, acc = cur.copy(), [x.copy() for x in mo], acc.copy()
        for it in range(a, b):
            cur, mo, loss = _oracle_step(cur, views[it % len(views)], mo, acc, it + 1, acfg)
            ref_losses.append(loss)
            if it in events:  # the events themselves were checked above
                break
    dev, ref = np.array(losses), np.array(ref_losses)
    rel = np.abs(dev - ref) / np.abs(ref)
    print(f"loss {dev[0]:.5f} -> {dev[-1]:.5f}; oracle loop {ref[0]:.5f} -> {ref[-1]:.5f}; max rel {rel.max():.2e}")
    assert rel.max() < 1e-3
    assert dev[-1] < dev[0]
