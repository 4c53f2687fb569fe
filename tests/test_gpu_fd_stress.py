"""Differentiation stress (config D): 3,000,000 splats with log-scales
N(-5, 1) and opacity logits N(1, 1.5) at 1920x1080.  The device gradient of
the recompute backward (exact chain rule) is checked by float64 central
differences on 2,048 seeded parameters spanning positions, log-scales,
rotations, opacities and SH coefficients.

The loss is the mean-square error against a render of a perturbed copy.
The rasteriser's backward receives dL/dC and returns d<dL/dC, C>/dtheta, so
the asserted check differentiates exactly that functional, sum <w, C(theta)>
with w = the device's dL/dC held fixed; the central differences of the MSE
itself are reported beside it (there the float32 residual C - target of
near-equal pixels adds its own rounding to the gradient, the same in the
float32 host restatement).  Differences re-blend only the tiles of the
perturbed splat over the fixed float32 binning (orc_fd_local_mse).  A
parameter whose +-h evaluations change, on a pixel the splat reaches, the
contributing set (an alpha crossing 1/255 or the 0.99 clamp, an early
termination), or change the splat's tiles or depth bucket, is reported as
discontinuous and left out: the analytic gradient has no term for those
jumps.  So is one whose float64 contribution pattern at theta differs from
the float32 forward's (the device's) on such a pixel.  Values are compared
on the summed scale (x 3WH) with |g - fd| <= 1e-4 max(|g|, |fd|) + 1e-6.
At h = 1e-3 (with one Richardson step) nearly every position of these
1-pixel splats moves a cutoff contour across a pixel centre, so the
asserted comparison uses h = 1e-5 (float64 keeps the rounding below 1e-9 of
the gradient); the h = 1e-3 counts are reported alongside."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from nanotracer_amd import _lib as L
from nanotracer_amd import synth
from nanotracer_amd.api import CameraUniform, GpuSplatBuffer, Rasterizer, TileBinner, TilingParams, loss_grad

pytestmark = pytest.mark.gpu

W, H, N = 1920, 1080, 3_000_000
SPLIT = (("pos", 0, 3, 512), ("log_scale", 2, 3, 384), ("rot", 1, 4, 384), ("logit", 3, 1, 256),
         ("sh_dc", 4, 3, 256), ("sh_rest", 5, 45, 256))


def test_config_d_gradient_by_central_differences(ctx):
    s = synth.fast_scene(N, 0x5EED0003, 3.0, log_scale_mu=-5.0, log_scale_sd=1.0, logit_mu=1.0, logit_sd=1.5)
    u = synth.uniform((0x5EED0003 << 8) + 31, 1)
    eye = synth.orbit_eye(6.0, 2.0 * math.pi * float(u[0]), 0.3)
    cam = CameraUniform.from_pose(eye, (0, 0, 0), (0, 1, 0), math.pi / 3, W, H, N)
    params = TilingParams(W, H, 16, 12.0)
    r, binner = Rasterizer(), TileBinner()

    def render(scene):
        proj = r.alloc_projected(ctx, N)
        r.project(ctx, scene, cam, proj)
        bins = binner.bin(ctx, proj, N, params)
        img, ft, li = r.alloc_image(ctx, W, H), ctx.storage_buffer_zeroed(W * H * 4), ctx.storage_buffer_zeroed(W * H * 4)
        r.composite_ex(ctx, proj, bins, img, params, ft, li)
        return proj, bins, img, ft, li

    tgt_scene = GpuSplatBuffer.upload(ctx, synth.perturb_means(s, 0x5EED0004, 0.01))
    _, tb, timg, _, _ = render(tgt_scene)
    target = timg.download(np.float32, W * H * 4).reshape(H, W, 4)
    tb.free()
    tgt_scene.free()
    g = GpuSplatBuffer.upload(ctx, s)
    proj, bins, img, ft, li = render(g)
    dl = ctx.storage_buffer_zeroed(W * H * 16)
    tbuf = ctx.storage_buffer(target)
    loss_grad(ctx, img, tbuf, dl, W, H, L.LOSS_MSE)
    weights = dl.download(np.float32, W * H * 4).reshape(H, W, 4)
    pg = r.alloc_projected_grad(ctx, N)
    r.composite_backward_recompute(ctx, proj, bins, ft, li, dl, pg, params)
    r.project_backward(ctx, g, pg, cam, mode=L.BWD_RECOMPUTE)
    grads = g.grads()

    ocam = O.camera_like(cam)
    fdc = O.FdContext(s, ocam, O.tiling(W, H, 16, 12.0), target, weights)
    mse = O.FdContext.__new__(O.FdContext)  # the same context differentiating the MSE
    mse.__dict__.update(fdc.__dict__)
    mse.weights = None
    assert np.array_equal(fdc.keys, bins.sorted_keys())  # the differences use the device's binning
    scale = 3.0 * W * H  # compare on the summed loss
    rng = np.random.default_rng(0x5EED0005)
    vis = np.nonzero(fdc.proj32[:, 7] != 0)[0]
    report, outliers, samples = {}, {}, {}
    for name, attr, width, count in SPLIT:
        ga = getattr(grads, name)
        idx = rng.choice(vis, count, replace=False).astype(np.uint32)
        comp = rng.integers(width, size=count).astype(np.int32)
        gv = (ga[idx] if ga.ndim == 1 else ga[idx, comp]).astype(np.float64) * scale
        fd5, d5 = fdc.fd_many(idx, attr, comp, 1e-5)
        f3a, d3a = fdc.fd_many(idx, attr, comp, 1e-3)
        f3b, d3b = fdc.fd_many(idx, attr, comp, 5e-4)
        fm, dm = mse.fd_many(idx, attr, comp, 1e-5)
        fd5, fd3, fm = fd5 * scale, (4 * f3b - f3a) / 3 * scale, fm * scale
        c5, c3, cm = d5 == 0, (d3a == 0) & (d3b == 0), dm == 0
        tol = lambda a, b: np.abs(a - b) <= 1e-4 * np.maximum(np.abs(a), np.abs(b)) + 1e-6
        ok5, ok3, okm = tol(gv, fd5), tol(gv, fd3), tol(gv, fm)
        nrm = float(np.linalg.norm((gv - fd5)[c5]) / max(np.linalg.norm(fd5[c5]), 1e-300)) if c5.any() else 0.0
        report[name] = dict(params=count, continuous=int(c5.sum()), f32_f64_pattern=int((d5 == 2).sum()),
                            pass_vjp=int((ok5 & c5).sum()), normwise_vjp=nrm, continuous_h1e3=int(c3.sum()),
                            pass_h1e3=int((ok3 & c3).sum()), pass_mse=int((okm & cm).sum()),
                            continuous_mse=int(cm.sum()))
        bad = np.nonzero(c5 & ~ok5)[0]
        outliers[name] = [int(k) for k in bad]
        samples[name] = (idx, comp, gv, fd5, c5)
        if bad.size:
            print(name, "outside the tolerance:", [(int(idx[k]), int(comp[k]), float(gv[k]), float(fd5[k]))
                                                  for k in bad[:8]])
    print("\nconfig D central differences:", report)
    # A continuous parameter outside the per-parameter acceptance is accepted
    # only where the float32 host restatement of the same backward (the
    # oracle: recompute backward + full chain rule over the same binning and
    # dL/dC) misses the float64 differences too, by at least half the
    # acceptance (the method in float32 cannot meet 1e-4 there: dL/dalpha =
    # G (T c - S / (1 - alpha)) and the conic's moment sums cancel where
    # a splat's colour matches the colour behind it), and the device stays
    # within 5e-4 of the differences.  Over the remaining continuous
    # parameters the norm-wise error per attribute is below 1e-4.
    if any(outliers.values()):
        _, ft32, li32 = O.composite(fdc.proj32, fdc.pay, fdc.rg, fdc.t, with_state=True)
        pg32 = O.composite_backward_recompute(fdc.proj32, fdc.pay, fdc.rg, fdc.t, ft32, li32, weights)
        g32 = O.project_backward(s, ocam, pg32, full=True)
    for name, attr, width, count in SPLIT:
        rep, (idx, comp, gv, fd5, c5) = report[name], samples[name]
        assert rep["continuous"] >= 0.5 * rep["params"], (name, rep)
        keep = c5.copy()
        for k in outliers[name]:
            a32 = getattr(g32, name)
            g_host = float(a32[idx[k]] if a32.ndim == 1 else a32[idx[k], comp[k]]) * scale
            dev_err, host_err = abs(gv[k] - fd5[k]) / abs(fd5[k]), abs(g_host - fd5[k]) / abs(fd5[k])
            print(f"{name} outlier {int(idx[k])}.{int(comp[k])}: device {gv[k]:.9g} ({dev_err:.1e}), float32 host "
                  f"{g_host:.9g} ({host_err:.1e}), float64 differences {fd5[k]:.9g}")
            assert host_err >= 0.5e-4 and dev_err <= 5e-4, (name, int(idx[k]), int(comp[k]), dev_err, host_err)
            keep[k] = False
        nrm = float(np.linalg.norm((gv - fd5)[keep]) / max(np.linalg.norm(fd5[keep]), 1e-300))
        assert nrm < 1e-4, (name, nrm, rep)
