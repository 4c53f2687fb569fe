"""Combined MSE + SSIM loss on the device (loss.rs:23-415) against the host
restatement: dL/dC bit-exact (the device gathers the blur transpose in the
scatter's summation order and shares the host's float32 operation order),
the loss components within 1e-6 relative (float64 sums in another order)."""
import numpy as np
import pytest

from oracle import oracle as O
from nanotracer_amd.api import combined_loss_grad, loss_grad
from nanotracer_amd import _lib as L

pytestmark = pytest.mark.gpu


def _pair(h, w, seed, noise=0.08):
    rng = np.random.default_rng(seed)
    tgt = np.zeros((h, w, 4), np.float32)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    for c in range(3):
        tgt[..., c] = 0.5 + 0.4 * np.sin(xx * (0.11 + 0.05 * c) + yy * 0.07 * (c + 1))
    tgt[..., 3] = 1.0
    pred = tgt + rng.normal(0, noise, tgt.shape).astype(np.float32)
    pred[..., 3] = rng.random((h, w)).astype(np.float32)
    return np.ascontiguousarray(pred), tgt


def _device(ctx, pred, tgt, lam):
    h, w = pred.shape[:2]
    pb, tb = ctx.storage_buffer(pred), ctx.storage_buffer(tgt)
    db = ctx.storage_buffer_zeroed(h * w * 16)
    total, mse, ssim = combined_loss_grad(ctx, pb, tb, db, w, h, lam)
    return total, mse, ssim, db.download(np.float32, h * w * 4).reshape(h, w, 4)


def _check(ctx, pred, tgt, lam):
    total, mse, ssim, grad = _device(ctx, pred, tgt, lam)
    rt, rm, rs, rg = O.combined_loss_grad(pred, tgt, lam)
    bad = np.argwhere(grad != rg)
    assert bad.size == 0, f"lam={lam} {pred.shape}: {len(bad)} gradient values differ, first {bad[:3].tolist()}"
    for a, b in ((total, rt), (mse, rm), (ssim, rs)):
        assert abs(a - b) <= 1e-6 * max(abs(b), 1e-6), (lam, a, b)


@pytest.mark.parametrize("lam", [0.0, 0.2, 0.5, 1.0])
@pytest.mark.parametrize("h,w", [(6, 6), (9, 14), (64, 48), (131, 257)])
def test_combined_matches_oracle(ctx, h, w, lam):
    pred, tgt = _pair(h, w, h * 1000 + w)
    _check(ctx, pred, tgt, lam)


@pytest.mark.parametrize("h,w", [(1, 1), (3, 5), (5, 2), (1, 17)])
def test_images_below_one_window(ctx, h, w):
    # axes shorter than the window reflect more than once (periodic reflect-101)
    pred, tgt = _pair(h, w, 7 + h * w)
    _check(ctx, pred, tgt, 0.2)


def test_full_hd(ctx):
    pred, tgt = _pair(1080, 1920, 11, noise=0.03)
    _check(ctx, pred, tgt, 0.2)


def test_identity_and_clamp(ctx):
    pred, tgt = _pair(32, 40, 3)
    total, mse, ssim, grad = _device(ctx, tgt, tgt, 1.0)
    assert abs(ssim) < 1e-6 and mse == 0.0 and np.abs(grad).max() < 1e-6
    # lambda outside [0, 1] is clamped
    a = _device(ctx, pred, tgt, 1.7)
    b = _device(ctx, pred, tgt, 1.0)
    assert a[:3] == b[:3] and np.array_equal(a[3], b[3])
    c = _device(ctx, pred, tgt, -0.5)
    d = ctx.storage_buffer_zeroed(32 * 40 * 16)
    m = loss_grad(ctx, ctx.storage_buffer(pred), ctx.storage_buffer(tgt), d, 40, 32, L.LOSS_MSE)
    assert c[0] == pytest.approx(m, rel=1e-6) and c[2] == 0.0
    assert np.array_equal(c[3], d.download(np.float32, 32 * 40 * 4).reshape(32, 40, 4))
