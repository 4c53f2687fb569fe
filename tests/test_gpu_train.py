"""The train() driver: view cycling, the prune / densify cadence of
train.rs:259-331 (grad_acc and grad_count reset after densify), and
checkpoints that restore the whole optimiser state."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from nanotracer_amd import synth
from nanotracer_amd.api import GpuSplatBuffer
from nanotracer_amd.train import TrainConfig, TrainState, View, load_checkpoint, save_checkpoint, train

pytestmark = pytest.mark.gpu


def _views(s, n=3, w=128, h=96):
    views = []
    for k in range(n):
        eye = synth.orbit_eye(3.0, 2 * math.pi * k / n, 0.2)
        cam = O.camera_from_pose(eye, (0, 0, 0), (0, 1, 0), math.pi / 3, w, h, s.n)
        t = O.tiling(w, h, 16, 6.0)
        tp = O.project(synth.perturb_means(s, 0xC0FFEE + k, 0.05), cam)
        kk, pay, rg = O.bin(tp, t)
        views.append(View(eye, w, h, O.composite(tp, pay, rg, t)))
    return views


def test_cadence_and_resets(ctx):
    s = synth.make_scene(3000, 0x7EA1, 1.0, log_scale_mu=-2.6, log_scale_sd=0.5, logit_mu=-2.0, logit_sd=2.5)
    views = _views(s)
    cfg = TrainConfig(iterations=61, max_splats=4000, prune_every=10, prune_warmup=5, densify_every=20,
                      densify_warmup=10, densify_grad_threshold=2e-6, diagnostics_every=30)
    seen, logs = [], []
    out = train(ctx, s, views, 6.0, cfg, on_iter=lambda it, loss, n: seen.append((it, loss, n)), log=logs.append)
    assert [x[0] for x in seen] == list(range(61))
    counts = [x[2] for x in seen]
    changes = [seen[k][0] for k in range(len(seen)) if counts[k] != (counts[k - 1] if k else s.n)]
    # the count only moves at a prune (10, 20, ...) or densify (20, 40, 60) iteration
    assert changes and all(it >= 10 and it % 10 == 0 for it in changes)
    assert any("pruned" in l for l in logs) and any("densify" in l for l in logs)
    assert out.n == counts[-1] <= cfg.max_splats
    assert all(np.isfinite(x[1]) for x in seen)


def test_checkpoint_restores_the_optimiser_state(ctx, tmp_path):
    s = synth.make_scene(2000, 0x7EA2, 1.0)
    views = _views(s, 2)
    base = str(tmp_path / "ckpt")
    cfg = TrainConfig(iterations=12, max_splats=3000, checkpoint_every=12, checkpoint_path=base,
                      diagnostics_every=0)
    train(ctx, s, views, 6.0, cfg, log=lambda *_: None)
    g, st = load_checkpoint(ctx, base, 6000)
    assert (st.iteration, st.t, st.grad_count) == (12, 12, 12)
    # save again from the restored buffer: identical files
    save_checkpoint(str(tmp_path / "again"), g, st)
    a, b = np.load(base + ".npz"), np.load(str(tmp_path / "again.npz"))
    for k in ("m", "v", "grad_acc", "state"):
        assert np.array_equal(a[k], b[k]), k
    assert (tmp_path / "ckpt.ply").read_bytes() == (tmp_path / "again.ply").read_bytes()
    g.free()
    # resume continues from iteration 12 with the saved step counter
    seen = []
    cfg2 = TrainConfig(iterations=16, max_splats=3000, diagnostics_every=0)
    train(ctx, None, views, 6.0, cfg2, on_iter=lambda it, loss, n: seen.append(it), resume=base, log=lambda *_: None)
    assert seen == [12, 13, 14, 15]
