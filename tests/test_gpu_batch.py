"""The batched multi-view step (one launch per stage across all views)
against the oracle run view by view: keys, ranges and images bit-exact
(by default the batched path bins only the tiles a splat's alpha >= 1/255
ellipse can reach: its keys are the oracle's with that cull, its images
those of the reference binning; with BIN_REFERENCE its keys are the
reference's),
per-view losses, summed gradients within 1e-4 norm-wise, parameters after
Adam; and hipGraph replays against eager steps."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from nanotracer_amd import _lib as L
from nanotracer_amd import synth
from nanotracer_amd.api import AdamConfig, CameraUniform, GpuSplatBuffer, ViewBatch, ViewSpec

from .helpers import rel_norm_err

pytestmark = pytest.mark.gpu

RES = ((320, 240), (480, 270), (256, 256), (400, 300))


def _views(n_splats, n_views=4, seed=5, res=RES):
    out = []
    for k, (eye, w, h) in enumerate(synth.orbit_views(n_views, seed, 6.0, res)):
        cam = CameraUniform.from_pose(eye, (0, 0, 0), (0, 1, 0), math.pi / 3, w, h, n_splats)
        out.append((ViewSpec(cam, w, h, 12.0), O.camera_like(cam), eye))
    return out


def _targets(scene_t, views, tile):
    imgs = []
    for spec, ocam, _ in views:
        t = O.tiling(spec.width, spec.height, tile, spec.depth_max)
        p = O.project(scene_t, ocam)
        k, pay, rg = O.bin(p, t)
        imgs.append(O.composite(p, pay, rg, t))
    return imgs


def _is_subsequence(sub_k, sub_p, k, p):
    j = 0
    for a, b in zip(sub_k, sub_p):
        while j < len(k) and (k[j] != a or p[j] != b):
            j += 1
        if j == len(k):
            return False
        j += 1
    return True


@pytest.mark.parametrize("tile", [16, 8, 32])
def test_render_views_bit_exact_per_view(ctx, tile):
    s = synth.make_scene(20_000, 0x5EED0002, 3.0)
    views = _views(s.n)
    g = GpuSplatBuffer.upload(ctx, s)
    b = ViewBatch(ctx, 8, s.n, sum(w * h for w, h in RES), 2_000_000)
    b.render(g, [v[0] for v in views], tile)
    tp, vis, ov = b.stats()
    assert ov == 0 and tp > 0
    for k, (spec, ocam, _) in enumerate(views):
        t = O.tiling(spec.width, spec.height, tile, spec.depth_max)
        op = O.project(s, ocam)
        ok, opay, orr = O.bin(op, t, cull=True)
        keys, pay, rng = b.bins(k, tile)
        assert np.array_equal(keys, ok) and np.array_equal(pay, opay) and np.array_equal(rng, orr)
        # the culled pairs are reference pairs, in the reference's order
        rk, rp, rr = O.bin(op, t)
        live = ok != 0xFFFFFFFF  # pairs of tiles the alpha >= 1/255 region misses close the list
        assert live.sum() < len(rk) and _is_subsequence(ok[live], opay[live], rk, rp)
        img, ft, li = O.composite(op, opay, orr, t, with_state=True)
        assert np.array_equal(img.view(np.uint32), O.composite(op, rp, rr, t).view(np.uint32))
        gimg = b.image(k)
        assert np.array_equal(gimg.view(np.uint32), img.view(np.uint32))
        gft, gli = b.state(k)
        assert np.array_equal(gft.view(np.uint32), ft.view(np.uint32)) and np.array_equal(gli, li)
    b.free()


@pytest.mark.parametrize("tile", [32, 16, 8])
def test_render_views_count_table_sort_with_an_empty_view(ctx, tile):
    # a pair capacity of 16M or more: the binning sort takes its offsets from
    # count tables (tile bits 8, 9 and 11 at tiles 32, 16 and 8: one pass by
    # digit, then 1 + 8 and 3 + 8 bits); one view looks away and has no pair
    s = synth.make_scene(20_000, 0x5EED0002, 3.0)
    views = _views(s.n)
    w, h = 400, 300
    cam = CameraUniform.from_pose((0, 0, 8), (0, 0, 16), (0, 1, 0), math.pi / 3, w, h, s.n)
    views.insert(2, (ViewSpec(cam, w, h, 12.0), O.camera_like(cam), (0, 0, 8)))
    g = GpuSplatBuffer.upload(ctx, s)
    b = ViewBatch(ctx, 8, s.n, sum(v[0].width * v[0].height for v in views), 20_000_000)
    b.render(g, [v[0] for v in views], tile)
    tp, vis, ov = b.stats()
    assert ov == 0 and tp > 0
    for k, (spec, ocam, _) in enumerate(views):
        t = O.tiling(spec.width, spec.height, tile, spec.depth_max)
        op = O.project(s, ocam)
        ok, opay, orr = O.bin(op, t, cull=True)
        keys, pay, rng = b.bins(k, tile)
        assert np.array_equal(keys, ok) and np.array_equal(pay, opay) and np.array_equal(rng, orr)
        if k == 2:
            assert len(ok) == 0 and not rng.any()
        img, ft, li = O.composite(op, opay, orr, t, with_state=True)
        assert np.array_equal(b.image(k).view(np.uint32), img.view(np.uint32))
        gft, gli = b.state(k)
        assert np.array_equal(gft.view(np.uint32), ft.view(np.uint32)) and np.array_equal(gli, li)
    b.free()


@pytest.mark.parametrize("tile", [16, 8, 32])
def test_reference_binning_keys_are_the_references(ctx, tile):
    """With BIN_REFERENCE the batched path emits TileBinner::bin's pairs: every
    view's keys, payloads and ranges bit-exact against the oracle's reference
    binning (tile_binner.rs:181-382), the images those of the culled rule."""
    s = synth.make_scene(20_000, 0x5EED0002, 3.0)
    views = _views(s.n)
    g = GpuSplatBuffer.upload(ctx, s)
    b = ViewBatch(ctx, 8, s.n, sum(w * h for w, h in RES), 4_000_000)
    b.render(g, [v[0] for v in views], tile)
    culled = [b.image(k) for k in range(len(views))]
    culled_pairs = b.stats()[0]
    b.set_binning(L.BIN_REFERENCE)
    b.render(g, [v[0] for v in views], tile)
    tp, vis, ov = b.stats()
    assert ov == 0 and tp > culled_pairs
    total = 0
    for k, (spec, ocam, _) in enumerate(views):
        t = O.tiling(spec.width, spec.height, tile, spec.depth_max)
        op = O.project(s, ocam)
        rk, rp, rr = O.bin(op, t)
        keys, pay, rng = b.bins(k, tile)
        assert np.array_equal(keys, rk) and np.array_equal(pay, rp) and np.array_equal(rng, rr), (tile, k)
        total += len(rk)
        img = O.composite(op, rp, rr, t)
        assert np.array_equal(b.image(k).view(np.uint32), img.view(np.uint32)), (tile, k)
        assert np.array_equal(culled[k].view(np.uint32), img.view(np.uint32)), (tile, k)
    assert total == tp
    # back to the default rule
    b.set_binning(L.BIN_CULLED)
    b.render(g, [v[0] for v in views], tile)
    assert b.stats()[0] == culled_pairs
    b.free()


def test_reference_binning_step_matches_oracle(ctx):
    """A training step with the reference's binning: the oracle step view by
    view (train.rs:155-250), as for the culled rule, eager and graph-replayed."""
    s = synth.make_scene(20_000, 0x5EED0002, 3.0)
    views = _views(s.n)
    tile = 16
    targets = _targets(synth.perturb_means(s, 77), views, tile)
    tbuf = ctx.storage_buffer(np.concatenate([x.reshape(-1) for x in targets]))
    g = GpuSplatBuffer.upload(ctx, s)
    b = ViewBatch(ctx, 8, s.n, sum(w * h for w, h in RES), 4_000_000)
    b.set_binning(L.BIN_REFERENCE)
    cfg = AdamConfig()
    for step in (1, 2, 3):
        before = g.readback()
        mg, vg = g.moments()
        m = [a.reshape(-1).copy() for a in mg.arrays()]
        v = [a.reshape(-1).copy() for a in vg.arrays()]
        acc = g.grad_acc()
        b.train_step(g, [x[0] for x in views], tbuf, step, tile, L.LOSS_L1, L.BWD_RECOMPUTE, graph=step > 1,
                     write_grads=True)
        ref_s, ref_g, ref_losses = _oracle_batch_step(before, views, targets, tile, L.BWD_RECOMPUTE, m, v, acc, step,
                                                      cfg)
        assert np.allclose(b.losses(), ref_losses, rtol=1e-9, atol=0)
        for name, x, y in zip(("pos", "rot", "scale", "logit", "dc", "rest"), g.grads().arrays(), ref_g.arrays()):
            assert rel_norm_err(x, y) < 1e-4, (step, name, rel_norm_err(x, y))
        for x, y in zip(g.readback().arrays(), ref_s.arrays()):
            d = np.abs(x.astype(np.float64) - y)
            assert d.max() <= 2.0 * cfg.lr + 1e-6
            assert (d > 1e-6).mean() < 2e-3
        spec, ocam, _ = views[0]
        tt = O.tiling(spec.width, spec.height, tile, spec.depth_max)
        rk, rp, rr = O.bin(O.project(before, ocam), tt)
        keys, pay, rng = b.bins(0, tile)
        assert np.array_equal(keys, rk) and np.array_equal(pay, rp) and np.array_equal(rng, rr), step
    b.free()


def _oracle_batch_step(s, views, targets, tile, mode, m, v, acc, t, cfg, ssim_lambda=None):
    """The reference sequence view by view; gradients summed over views in
    view order (float32), then accumulate_grad + Adam + constraints.  The
    loss is L1, or the combined MSE + SSIM loss with ssim_lambda."""
    from .test_gpu_optim import _oracle_fused_step

    total = None
    losses = []
    for (spec, ocam, _), tgt in zip(views, targets):
        tt = O.tiling(spec.width, spec.height, tile, spec.depth_max)
        op = O.project(s, ocam)
        k, pay, rg = O.bin(op, tt)
        img, ft, li = O.composite(op, pay, rg, tt, with_state=True)
        if ssim_lambda is None:
            loss, dl = O.loss_grad(img, tgt, 0)
        else:
            loss, _, _, dl = O.combined_loss_grad(img, tgt, ssim_lambda)
        losses.append(loss)
        if mode == L.BWD_RECOMPUTE:
            pg = O.composite_backward_recompute(op, pay, rg, tt, ft, li, dl)
        else:
            pg = O.composite_backward_ref(op, pay, rg, tt, img, dl)
        g = O.project_backward(s, ocam, pg, full=(mode == L.BWD_RECOMPUTE))
        total = g if total is None else synth.HostSplats(*[a + b for a, b in zip(total.arrays(), g.arrays())])
    return _oracle_fused_step(s, total, m, v, acc, t, cfg, cfg), total, losses


# a pair capacity of 16M or more takes the binning sort through its count tables
@pytest.mark.parametrize("mode,cap", [(L.BWD_RECOMPUTE, 2_000_000), (L.BWD_REF_PARITY, 2_000_000),
                                      (L.BWD_RECOMPUTE, 20_000_000)])
def test_train_step_views_matches_oracle(ctx, mode, cap):
    s = synth.make_scene(20_000, 0x5EED0002, 3.0)
    views = _views(s.n)
    tile = 16
    targets = _targets(synth.perturb_means(s, 77), views, tile)
    tbuf = ctx.storage_buffer(np.concatenate([x.reshape(-1) for x in targets]))
    g = GpuSplatBuffer.upload(ctx, s)
    b = ViewBatch(ctx, 8, s.n, sum(w * h for w, h in RES), cap)
    cfg = AdamConfig()
    for step in (1, 2, 3):
        before = g.readback()
        mg, vg = g.moments()
        m = [a.reshape(-1).copy() for a in mg.arrays()]
        v = [a.reshape(-1).copy() for a in vg.arrays()]
        acc = g.grad_acc()
        b.train_step(g, [x[0] for x in views], tbuf, step, tile, L.LOSS_L1, mode, write_grads=True)
        ref_s, ref_g, ref_losses = _oracle_batch_step(before, views, targets, tile, mode, m, v, acc, step, cfg)
        assert np.allclose(b.losses(), ref_losses, rtol=1e-9, atol=0)
        gg = g.grads()
        for name, x, y in zip(("pos", "rot", "scale", "logit", "dc", "rest"), gg.arrays(), ref_g.arrays()):
            assert rel_norm_err(x, y) < 1e-4, (step, name, rel_norm_err(x, y))
        after = g.readback()
        for x, y in zip(after.arrays(), ref_s.arrays()):
            d = np.abs(x.astype(np.float64) - y)
            assert d.max() <= 2.0 * cfg.lr + 1e-6
            assert (d > 1e-6).mean() < 2e-3
    b.free()


@pytest.mark.parametrize("lam,mode", [(0.2, L.BWD_RECOMPUTE), (1.0, L.BWD_REF_PARITY), (0.0, L.BWD_RECOMPUTE)])
def test_train_step_views_ssim_matches_oracle(ctx, lam, mode):
    s = synth.make_scene(20_000, 0x5EED0003, 3.0)
    views = _views(s.n)
    tile = 16
    targets = _targets(synth.perturb_means(s, 79), views, tile)
    tbuf = ctx.storage_buffer(np.concatenate([x.reshape(-1) for x in targets]))
    g = GpuSplatBuffer.upload(ctx, s)
    b = ViewBatch(ctx, 8, s.n, sum(w * h for w, h in RES), 2_000_000)
    b.set_ssim_lambda(lam)
    cfg = AdamConfig()
    for step in (1, 2):
        before = g.readback()
        mg, vg = g.moments()
        m = [a.reshape(-1).copy() for a in mg.arrays()]
        v = [a.reshape(-1).copy() for a in vg.arrays()]
        acc = g.grad_acc()
        b.train_step(g, [x[0] for x in views], tbuf, step, tile, L.LOSS_SSIM, mode, write_grads=True,
                     graph=(step == 2))
        ref_s, ref_g, ref_losses = _oracle_batch_step(before, views, targets, tile, mode, m, v, acc, step, cfg,
                                                      ssim_lambda=lam)
        assert np.allclose(b.losses(), ref_losses, rtol=1e-6, atol=0), (b.losses(), ref_losses)
        gg = g.grads()
        for name, x, y in zip(("pos", "rot", "scale", "logit", "dc", "rest"), gg.arrays(), ref_g.arrays()):
            assert rel_norm_err(x, y) < 1e-4, (step, name, rel_norm_err(x, y))
        after = g.readback()
        for x, y in zip(after.arrays(), ref_s.arrays()):
            d = np.abs(x.astype(np.float64) - y)
            assert d.max() <= 2.0 * cfg.lr + 1e-6
            assert (d > 1e-6).mean() < 2e-3
    b.free()


def test_ssim_step_needs_the_setter(ctx):
    s = synth.make_scene(1000, 0x5EED0004, 3.0)
    views = _views(s.n, 1)
    g = GpuSplatBuffer.upload(ctx, s)
    b = ViewBatch(ctx, 2, s.n, 320 * 240, 200_000)
    tbuf = ctx.storage_buffer(np.zeros((240, 320, 4), np.float32))
    with pytest.raises(L.NtsError, match="nts_batch_set_ssim_lambda"):
        b.train_step(g, [views[0][0]], tbuf, 1, 16, L.LOSS_SSIM, L.BWD_RECOMPUTE)
    b.free()


def test_graph_replay_matches_eager_steps(ctx):
    s = synth.make_scene(20_000, 0x5EED0002, 3.0)
    views = _views(s.n)
    tile = 16
    targets = _targets(synth.perturb_means(s, 78), views, tile)
    tbuf = ctx.storage_buffer(np.concatenate([x.reshape(-1) for x in targets]))
    specs = [x[0] for x in views]
    runs = []
    for graph in (False, True):
        g = GpuSplatBuffer.upload(ctx, s)
        b = ViewBatch(ctx, 8, s.n, sum(w * h for w, h in RES), 2_000_000)
        losses = []
        for step in range(1, 6):
            b.train_step(g, specs, tbuf, step, tile, graph=graph)
            ctx.sync()
            losses.append(b.losses().sum())
        runs.append((losses, g.readback()))
        b.free()
        g.free()
    (le, pe), (lg, pg) = runs
    # identical inputs; only the float atomics' order differs between runs
    assert np.allclose(le, lg, rtol=1e-6)
    assert le[-1] < le[0]
    for x, y in zip(pe.arrays(), pg.arrays()):
        assert np.abs(x - y).max() <= 5 * 1.6e-4


def test_single_view_batch_equals_plain_path(ctx):
    from nanotracer_amd.train import Trainer, View

    s = synth.make_scene(20_000, 0x5EED0002, 3.0)
    (spec, ocam, eye), = _views(s.n, 1)
    g = GpuSplatBuffer.upload(ctx, s)
    tr = Trainer(ctx, g, [View(eye, spec.width, spec.height)], spec.depth_max)
    proj, bins, img, cam, params = tr.render(tr.views[0])
    plain = img.download(np.float32, spec.width * spec.height * 4).reshape(spec.height, spec.width, 4)
    b = ViewBatch(ctx, 1, s.n, spec.width * spec.height, 1_000_000)
    b.render(g, [spec], 16)
    assert np.array_equal(b.image(0).view(np.uint32), plain.view(np.uint32))
    keys, pay, rng = b.bins(0, 16)
    live = keys != 0xFFFFFFFF
    assert _is_subsequence(keys[live], pay[live], bins.sorted_keys(), bins.sorted_payloads())
    b.free()


def test_overflow_is_flagged_not_written_past(ctx):
    s = synth.make_scene(20_000, 0x5EED0002, 3.0)
    views = _views(s.n)
    g = GpuSplatBuffer.upload(ctx, s)
    b = ViewBatch(ctx, 8, s.n, sum(w * h for w, h in RES), 10_000)
    b.render(g, [v[0] for v in views], 16)
    tp, vis, ov = b.stats()
    assert ov == 1 and tp > 10_000
    b.free()


@pytest.mark.parametrize("tile,fwd_batch", [(8, 64), (16, 512), (32, 128), (32, 512)])
def test_lds_batch_keeps_images_bit_exact(ctx, tile, fwd_batch):
    s = synth.make_scene(20_000, 0x5EED0002, 3.0)
    views = _views(s.n)
    g = GpuSplatBuffer.upload(ctx, s)
    b = ViewBatch(ctx, 8, s.n, sum(w * h for w, h in RES), 2_000_000)
    b.set_lds_batch(fwd_batch, 256)
    b.render(g, [v[0] for v in views], tile)
    for k, (spec, ocam, _) in enumerate(views):
        t = O.tiling(spec.width, spec.height, tile, spec.depth_max)
        op = O.project(s, ocam)
        ok, opay, orr = O.bin(op, t, cull=True)
        img, ft, li = O.composite(op, opay, orr, t, with_state=True)
        assert np.array_equal(b.image(k).view(np.uint32), img.view(np.uint32)), k
        gft, gli = b.state(k)
        assert np.array_equal(gli, li)
    b.free()


@pytest.mark.parametrize("tile,bwd_batch", [(32, 64), (32, 128), (32, 256), (8, 128), (16, 64)])
def test_lds_batch_backward_matches_oracle(ctx, tile, bwd_batch):
    s = synth.make_scene(20_000, 0x5EED0002, 3.0)
    views = _views(s.n)
    targets = _targets(synth.perturb_means(s, 79), views, tile)
    tbuf = ctx.storage_buffer(np.concatenate([x.reshape(-1) for x in targets]))
    g = GpuSplatBuffer.upload(ctx, s)
    b = ViewBatch(ctx, 8, s.n, sum(w * h for w, h in RES), 2_000_000)
    b.set_lds_batch(128, bwd_batch)
    cfg = AdamConfig()
    m = [np.zeros(a.size, np.float32) for a in s.arrays()]
    v = [np.zeros(a.size, np.float32) for a in s.arrays()]
    acc = np.zeros(s.n, np.float32)
    b.train_step(g, [x[0] for x in views], tbuf, 1, tile, L.LOSS_L1, L.BWD_RECOMPUTE, write_grads=True)
    _, ref_g, ref_losses = _oracle_batch_step(s, views, targets, tile, L.BWD_RECOMPUTE, m, v, acc, 1, cfg)
    assert np.allclose(b.losses(), ref_losses, rtol=1e-9, atol=0)
    for name, x, y in zip(("pos", "rot", "scale", "logit", "dc", "rest"), g.grads().arrays(), ref_g.arrays()):
        assert rel_norm_err(x, y) < 1e-4, (name, rel_norm_err(x, y))
    b.free()


def test_wave_culling_exact_on_elongated_and_faint_splats(ctx):
    """Needle-like splats (axis ratios up to e^7), splats whose peak alpha
    sits just above and below 1/255, and large flat ones: the wave and pixel
    culls must not change a single bit of the image or the saved state."""
    rng = np.random.default_rng(11)
    n = 6000
    s = synth.make_scene(n, 0x5EED0003, 1.5)
    ls = np.full((n, 3), -4.0, np.float32)
    ls[:, 0] = rng.uniform(-1.5, -0.5, n)
    ls[:, 1:] = rng.uniform(-8.0, -5.0, (n, 2))
    big = rng.random(n) < 0.1
    ls[big] = rng.uniform(-2.0, -1.0, (int(big.sum()), 3))
    s.log_scale[:] = ls
    thr = math.log((1.0 / 255.0) / (1.0 - 1.0 / 255.0))
    s.logit[:] = np.where(rng.random(n) < 0.3, thr + rng.normal(0, 0.01, n),
                             rng.normal(0.5, 1.5, n)).astype(np.float32)
    res = ((320, 240), (256, 256))
    views = _views(s.n, 2, seed=9, res=res)
    g = GpuSplatBuffer.upload(ctx, s)
    for tile in (8, 16, 32):
        b = ViewBatch(ctx, 2, s.n, 2 * max(w * h for w, h in res), 4_000_000)
        b.render(g, [v[0] for v in views], tile)
        for k, (spec, ocam, _) in enumerate(views):
            t = O.tiling(spec.width, spec.height, tile, spec.depth_max)
            op = O.project(s, ocam)
            ok, opay, orr = O.bin(op, t, cull=True)
            img, ft, li = O.composite(op, opay, orr, t, with_state=True)
            assert np.array_equal(img.view(np.uint32), O.composite(op, *O.bin(op, t)[1:], t).view(np.uint32))
            assert np.array_equal(b.image(k).view(np.uint32), img.view(np.uint32)), (tile, k)
            gft, gli = b.state(k)
            assert np.array_equal(gft.view(np.uint32), ft.view(np.uint32)) and np.array_equal(gli, li)
        b.free()


def test_autotuner_picks_a_candidate_and_trains(ctx):
    from nanotracer_amd.autotune import autotune_batch

    s = synth.make_scene(20_000, 0x5EED0002, 3.0)
    views = _views(s.n)
    specs = [x[0] for x in views]
    targets = _targets(synth.perturb_means(s, 80), views, 16)
    tbuf = ctx.storage_buffer(np.concatenate([x.reshape(-1) for x in targets]))
    g = GpuSplatBuffer.upload(ctx, s)
    b = ViewBatch(ctx, 8, s.n, sum(w * h for w, h in RES), 4_000_000)
    r = autotune_batch(ctx, b, g, specs, tbuf, 1, tiles=(16, 32), fwd_batches=(128, 256), bwd_batches=(64, 128))
    assert r.tile in (16, 32) and r.fwd_batch in (128, 256) and r.bwd_batch in (64, 128)
    # two steps per timed configuration, then two timed pairs per scheduler
    assert r.next_t == 1 + 2 * 2 * 2 + 2 * 2 * 2 and r.step_ms > 0
    assert len([row for row in r.table if "fwd_ms" in row and "scheduler" not in row]) == 4
    assert r.scheduler in ("grid", "persistent") and set(r.blend_ms) == {"grid", "persistent"}
    b.train_step(g, specs, tbuf, r.next_t, r.tile, graph=True)
    ctx.sync()
    assert np.isfinite(b.losses()).all()
    b.free()


@pytest.mark.parametrize("tile,fb,bb", [(8, 64, 64), (16, 256, 128), (32, 512, 128), (32, 128, 256)])
def test_persistent_scheduler_matches_grid(ctx, tile, fb, bb):
    """The persistent tile scheduler (resident workgroups taking tiles from a
    counter) against one workgroup per tile: identical images, final T and
    last contributors; on the same scene the same losses and gradients (the
    float atomics sum in another order); after graph-replayed steps the
    same parameters within Adam's step (a near-zero gradient summed to the
    other sign moves a parameter by lr the other way)."""
    s = synth.make_scene(20_000, 0x5EED0002, 3.0)
    views = _views(s.n)
    specs = [x[0] for x in views]
    targets = _targets(synth.perturb_means(s, 81), views, tile)
    tbuf = ctx.storage_buffer(np.concatenate([x.reshape(-1) for x in targets]))
    out = {}
    for sched in (L.SCHED_GRID, L.SCHED_PERSISTENT):
        g = GpuSplatBuffer.upload(ctx, s)
        b = ViewBatch(ctx, 8, s.n, sum(w * h for w, h in RES), 2_000_000)
        b.set_lds_batch(fb, bb)
        b.set_scheduler(sched)
        b.render(g, specs, tile)
        imgs = [b.image(k) for k in range(len(specs))]
        states = [b.state(k) for k in range(len(specs))]
        b.train_step(g, specs, tbuf, 1, tile, L.LOSS_L1, L.BWD_RECOMPUTE, write_grads=True)
        first = (b.losses(), g.grads())
        for t in (2, 3, 4):
            b.train_step(g, specs, tbuf, t, tile, L.LOSS_L1, L.BWD_RECOMPUTE, graph=True)
        ctx.sync()
        assert b.stats()[2] == 0
        out[sched] = (imgs, states, first, b.losses(), g.readback())
        b.free()
        g.free()
    (ia, sa, (la, ga), la4, pa), (ib, sb, (lb, gb), lb4, pb) = out[L.SCHED_GRID], out[L.SCHED_PERSISTENT]
    for x, y in zip(ia, ib):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    for (fa, lia), (fb_, lib_) in zip(sa, sb):
        assert np.array_equal(fa.view(np.uint32), fb_.view(np.uint32)) and np.array_equal(lia, lib_)
    assert np.allclose(la, lb, rtol=1e-9, atol=0)
    for name, x, y in zip(("pos", "rot", "scale", "logit", "dc", "rest"), ga.arrays(), gb.arrays()):
        assert rel_norm_err(x, y) < 1e-5, (name, rel_norm_err(x, y))
    assert np.allclose(la4, lb4, rtol=1e-4, atol=0)
    for x, y in zip(pa.arrays(), pb.arrays()):
        assert np.abs(x.astype(np.float64) - y).max() <= 2 * 3 * 1.6e-4 + 1e-6
