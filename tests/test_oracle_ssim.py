"""Pins the checker's SSIM loss (loss.rs:23-415): the kernel, the reflect-101
blur against scipy.ndimage (mode 'mirror' is reflect-101), the blur adjoint
by the inner-product identity, an independent float64 SSIM, and the
reference's own tests (identity, noise, finite differences, the combined
loss)."""
import numpy as np
import pytest
from scipy import ndimage

from oracle import oracle as O


def test_kernel_is_the_normalised_gaussian():
    k = O.ssim_kernel().astype(np.float64)
    d = np.arange(11) - 5.0
    ref = np.exp(-d * d / 4.5)
    ref /= ref.sum()
    assert np.abs(k - ref).max() < 1e-7 and abs(k.sum() - 1.0) < 1e-6
    assert np.array_equal(O.ssim_kernel(), O.ssim_kernel()[::-1])


@pytest.mark.parametrize("h,w", [(8, 8), (13, 6), (40, 57), (3, 5), (1, 7), (2, 2), (1, 1)])
def test_blur_is_scipy_mirror_gaussian(h, w):
    x = np.random.default_rng(h * w).random((h, w)).astype(np.float32)
    ref = ndimage.gaussian_filter(x.astype(np.float64), 1.5, mode="mirror", truncate=5 / 1.5)
    assert np.abs(O.ssim_blur(x) - ref).max() < 2e-6


@pytest.mark.parametrize("h,w", [(8, 8), (13, 6), (40, 57)])
def test_adjoint_identity(h, w):
    rng = np.random.default_rng(h + w)
    x = rng.random((h, w)).astype(np.float32)
    y = rng.random((h, w)).astype(np.float32)
    lhs = np.sum(O.ssim_blur(x).astype(np.float64) * y)
    rhs = np.sum(x.astype(np.float64) * O.ssim_blur(y, adjoint=True))
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
    if h < 12:  # near the border the blur is not self-adjoint
        assert not np.allclose(O.ssim_blur(y, adjoint=True), O.ssim_blur(y), atol=1e-6)


def _ssim64(x, y):
    f = lambda a: ndimage.gaussian_filter(a, 1.5, mode="mirror", truncate=5 / 1.5)
    mx, my = f(x), f(y)
    sxx, syy, sxy = f(x * x) - mx * mx, f(y * y) - my * my, f(x * y) - mx * my
    s = (2 * mx * my + 1e-4) * (2 * sxy + 9e-4) / ((mx * mx + my * my + 1e-4) * (sxx + syy + 9e-4))
    return 1.0 - s.mean()


def test_ssim_matches_an_independent_float64_ssim():
    rng = np.random.default_rng(3)
    p = rng.random((24, 31, 4)).astype(np.float32)
    t = np.clip(p + rng.normal(0, 0.1, p.shape), 0, 1).astype(np.float32)
    loss, _ = O.ssim_loss_grad(p, t)
    ref = np.mean([_ssim64(p[..., c].astype(np.float64), t[..., c].astype(np.float64)) for c in range(3)])
    assert abs(loss - ref) < 1e-5


def _checkerboard(w, h, scale):
    yy, xx = np.mgrid[0:h, 0:w]
    v = np.where((xx + yy) % 2 == 0, scale, 0.0).astype(np.float32)
    return np.stack([v, v * 0.5, v * 0.25, np.zeros_like(v)], -1).astype(np.float32)


def test_ssim_is_one_for_identical_images():
    # loss.rs:433-445
    img = _checkerboard(8, 8, 0.7)
    loss, g = O.ssim_loss_grad(img, img)
    assert abs(loss) < 1e-5 and np.linalg.norm(g[..., :3], axis=-1).max() < 1e-4


def test_ssim_grows_with_noise():
    # loss.rs:447-466 (the reference's LCG noise)
    img = _checkerboard(8, 8, 0.7)
    noisy = img.copy()
    s = 1
    for y in range(8):
        for x in range(8):
            s = (s * 1103515245 + 12345) & 0xFFFFFFFF
            noisy[y, x, :3] += np.float32(((s >> 16) / 65535.0 - 0.5) * 0.4)
    assert O.ssim_loss_grad(noisy, img)[0] > O.ssim_loss_grad(img, img)[0]


def test_ssim_gradient_matches_finite_differences():
    # loss.rs:469-513, and a float64 check of the same gradient
    n = 64
    f = (np.arange(n) / n).astype(np.float32)
    pred = np.stack([0.3 + 0.4 * f, 0.5 - 0.3 * f, 0.2 + 0.6 * f, 0 * f], -1).reshape(8, 8, 4).astype(np.float32)
    tgt = np.stack([0.4 + 0.3 * f, 0.6 - 0.4 * f, 0.3 + 0.5 * f, 0 * f], -1).reshape(8, 8, 4).astype(np.float32)
    _, g = O.ssim_loss_grad(pred, tgt)
    worst = 0.0
    for i in range(n):
        for c in range(3):
            y, x = divmod(i, 8)
            pp, pm = pred.copy(), pred.copy()
            pp[y, x, c] += np.float32(1e-3)
            pm[y, x, c] -= np.float32(1e-3)
            fd = (O.ssim_loss_grad(pp, tgt)[0] - O.ssim_loss_grad(pm, tgt)[0]) / 2e-3
            assert abs(fd - g[y, x, c]) < 5e-3
            # float64 reference of the same loss
            p64p, p64m = pred.astype(np.float64), pred.astype(np.float64)
            p64p[y, x, c] += 1e-6
            p64m[y, x, c] -= 1e-6
            l64 = lambda q: np.mean([_ssim64(q[..., k], tgt[..., k].astype(np.float64)) for k in range(3)])
            fd64 = (l64(p64p) - l64(p64m)) / 2e-6
            worst = max(worst, abs(fd64 - g[y, x, c]) / (abs(fd64) + 1e-3))
    assert worst < 2e-3


def _pair(seed):
    n = 64
    f = (np.arange(n) / n).astype(np.float32)
    pred = np.stack([f, 1.0 - f, np.full(n, 0.3, np.float32), 0 * f], -1).reshape(8, 8, 4).astype(np.float32)
    tgt = np.stack([f * 0.9, 1.0 - f * 1.1, np.full(n, 0.4, np.float32), 0 * f], -1).reshape(8, 8, 4).astype(np.float32)
    return pred, tgt


def test_combined_loss_limits_and_weighting():
    # loss.rs:515-583
    pred, tgt = _pair(0)
    mse, mg = O.loss_grad(pred, tgt, 1)
    ssim, sg = O.ssim_loss_grad(pred, tgt)
    t0 = O.combined_loss_grad(pred, tgt, 0.0)
    assert t0[0] == mse and t0[2] == 0.0 and np.array_equal(t0[3], mg)
    t1 = O.combined_loss_grad(pred, tgt, 1.0)
    assert t1[0] == ssim and t1[1] == 0.0 and np.array_equal(t1[3], sg)
    t2 = O.combined_loss_grad(pred, tgt, 0.2)
    assert np.abs(t2[3] - (mg * 0.8 + sg * 0.2)).max() < 1e-6
