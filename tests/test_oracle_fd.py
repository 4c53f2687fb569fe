"""Float64 central finite differences pin the recompute-mode backward and the
full chain rule (the Differentiation form) on a scene with dense occlusion
and saturated pixels.  Parameters whose +-h evaluations change the active
(splat, pixel) set, a tile box, a depth bucket or a saturation point are
reported as discontinuous and excluded from the tolerance."""
import math

import numpy as np

from oracle import oracle as O
from nanotracer_amd import synth

W, H = 48, 40


def _scene():
    s = synth.make_scene(40, 21, 0.3, log_scale_mu=-1.8, log_scale_sd=0.3, logit_mu=3.0, logit_sd=1.5, rest_sd=0.2)
    s.pos[:, 2] -= 0.2
    return s


def _cam(n):
    return O.camera_from_pose((0.2, 0.1, 2.5), (0, 0, 0), (0, 1, 0), math.pi / 3, W, H, n)


def _render64(s, cam, t):
    s64 = synth.HostSplats(*[a.astype(np.float64) for a in s.arrays()])
    p64 = O.project(s64, cam, np.float64)
    keys, pay, rg = O.bin(p64.astype(np.float32), t)
    img, li, nu = O.composite_f64(p64, pay, rg, t, with_state=True)
    sig = (keys.tobytes(), pay.tobytes(), li.tobytes(), nu.tobytes())
    return img, sig


def _fd_check(full: bool, h: float = 1e-5):
    s = _scene()
    cam = _cam(s.n)
    t = O.tiling(W, H, 16, 6.0)
    rng = np.random.default_rng(2)
    wpix = rng.normal(0, 1, (H, W, 4))
    wpix[..., 3] = 0
    p = O.project(s, cam)
    keys, pay, rg = O.bin(p, t)
    img, ft, li = O.composite(p, pay, rg, t, with_state=True)
    assert (ft < 1e-4).any(), "the scene must contain saturated pixels"
    pg = O.composite_backward_recompute(p, pay, rg, t, ft, li, wpix.astype(np.float32))
    g = O.project_backward(s, cam, pg, full=full)
    _, sig0 = _render64(s, cam, t)
    ok, bad, disc = {}, {}, 0
    worst = []
    for i in range(s.n):
        if p[i, 7] == 0:
            continue
        for attr, width in (("pos", 3), ("rot", 4), ("log_scale", 3), ("logit", 1), ("sh_dc", 3), ("sh_rest", 45)):
            for c in range(width):
                if attr == "sh_rest" and c % 7 != 0:
                    continue
                a, b = s.copy(), s.copy()
                arr_a, arr_b = getattr(a, attr), getattr(b, attr)
                if width == 1:
                    arr_a[i] += h
                    arr_b[i] -= h
                else:
                    arr_a[i, c] += h
                    arr_b[i, c] -= h
                ia, sa = _render64(a, cam, t)
                ib, sb = _render64(b, cam, t)
                if sa != sig0 or sb != sig0:
                    disc += 1
                    continue
                fd = float(((ia - ib) * wpix).sum() / (2 * h))
                an = float(getattr(g, attr)[i] if width == 1 else getattr(g, attr)[i, c])
                err = abs(an - fd)
                tol = 2e-3 * max(abs(an), abs(fd)) + 2e-5
                if err <= tol:
                    ok[attr] = ok.get(attr, 0) + 1
                else:
                    bad[attr] = bad.get(attr, 0) + 1
                    worst.append((attr, i, c, an, fd))
    return ok, bad, disc, worst


def test_recompute_backward_full_chain_matches_finite_differences():
    ok, bad, disc, worst = _fd_check(full=True)
    for attr in ("pos", "rot", "log_scale", "logit", "sh_dc", "sh_rest"):
        assert ok.get(attr, 0) >= 20, (attr, ok, disc)
    assert not bad, worst[:10]


def test_reference_chain_rule_is_caught_by_the_same_check():
    """The reference's project_backward drops dJ/dt and dSH/dp: the same
    finite-difference check must flag its position gradients."""
    ok, bad, disc, worst = _fd_check(full=False)
    assert bad.get("pos", 0) >= 10, (ok, bad)


def test_local_fd_equals_the_whole_image_difference():
    """orc_fd_local_mse re-blends only the tiles of the perturbed splat; it
    must equal the central difference of the whole float64 image over the
    same fixed binning (the means of two full renders)."""
    from nanotracer_amd import synth

    s = synth.make_scene(400, 0xFD10CA1, 0.8, log_scale_mu=-3.0, log_scale_sd=0.6, logit_mu=0.5)
    cam = O.camera_from_pose((0.0, 0.3, 2.5), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), math.pi / 3, 80, 64, s.n)
    t = O.tiling(80, 64, 16, 5.0)
    tp = O.project(synth.perturb_means(s, 77, 0.03), cam)
    tk, tpay, trg = O.bin(tp, t)
    target = O.composite(tp, tpay, trg, t)
    ctx = O.FdContext(s, cam, t, target)
    rng = np.random.default_rng(5)
    vis = np.nonzero(ctx.proj32[:, 7] != 0)[0]
    checked = 0
    for k in range(24):
        idx = int(rng.choice(vis))
        attr = k % 6
        comp = int(rng.integers((3, 4, 3, 1, 3, 45)[attr]))
        h = 1e-4
        fd, disc = ctx.fd(idx, attr, comp, h)
        ims = []
        for sign in (1, -1):
            s64 = synth.HostSplats(*[np.array(a, np.float64) for a in s.arrays()])
            a = s64.arrays()[attr]
            if a.ndim == 1:
                a[idx] += sign * h
            else:
                a[idx, comp] += sign * h
            p64 = O.project(s64, cam, np.float64)
            ims.append(O.composite_f64(p64, ctx.pay, ctx.rg, t))
        sq = [np.sum((im[..., :3] - target[..., :3].astype(np.float64)) ** 2) for im in ims]
        ref = (sq[0] - sq[1]) / (2 * h) / (3.0 * 80 * 64)
        if disc:
            continue  # discontinuous, or float32 / float64 patterns differ
        # the whole-image sums cancel to ~1e-16 of the image loss: 1e-7 relative
        assert abs(fd - ref) <= 1e-7 * abs(ref) + 1e-13, (idx, attr, comp, fd, ref)
        checked += 1
    assert checked >= 12
