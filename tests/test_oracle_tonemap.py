"""The oracle's tonemap (tonemap.wgsl:18-41) pinned by its closed form.

Per channel: Reinhard x / (1 + max(x, 0)), clamped to [0, 1], then the sRGB
transfer (12.92 t below the cutoff 0.0031308, 1.055 t^(1/2.4) - 0.055
above it), then the rgba8unorm store: clamp to [0, 1], scale by 255, round
to nearest.  Alpha is the coverage w = 1 - T through the same store.

The closed form is evaluated here in float64, independently of the
oracle's float32 code.  The two may differ by one level only where the
exact value sits within 1e-3 of a rounding boundary (float32 powf and
division round differently from float64 there); everywhere else the
levels are equal.  Planted mistakes (no Reinhard, Reinhard without the
max(x, 0), exponent 2.4 instead of 1/2.4, no 0.055 offset, truncation
instead of rounding) must each fail this check."""
import numpy as np

from oracle import oracle as O


def _srgb_u8_f64(x, reinhard=True, expo=1.0 / 2.4, offset=0.055, rnd=np.rint, keep_max=True):
    x = np.asarray(x, np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        tm = x / (1.0 + (np.maximum(x, 0.0) if keep_max else x)) if reinhard else x
    tm = np.clip(np.nan_to_num(tm, nan=0.0), 0.0, 1.0)
    s = np.where(tm > 0.0031308, 1.055 * np.power(tm, expo) - offset, tm * 12.92)
    v = np.clip(s, 0.0, 1.0) * 255.0
    return rnd(v), v


def _inputs():
    rng = np.random.default_rng(0x70E)
    c = 0.0031308
    x_cut = c / (1.0 - c)  # Reinhard maps x_cut to the sRGB cutoff
    xs = np.concatenate([
        np.array([0.0, -0.0, -1.0, -1e6, 1.0, 3.0, 1e6, np.float32(3.4e38), 1e-30, x_cut]),
        np.linspace(-2.0, 2.0, 4001),
        np.geomspace(1e-8, 1e4, 20000),
        x_cut * (1.0 + np.linspace(-1e-4, 1e-4, 201)),
        rng.uniform(0.0, 1.0, 20000),
        rng.exponential(1.0, 20000),
    ]).astype(np.float32)
    n = xs.size
    img = np.zeros((n, 4), np.float32)
    img[:, 0] = xs
    img[:, 1] = xs[::-1]
    img[:, 2] = rng.permutation(xs)
    img[:, 3] = np.concatenate([np.linspace(-0.5, 1.5, n // 2), rng.uniform(0.0, 1.0, n - n // 2)])
    return img.reshape(1, n, 4)


def _check(out, img, **variant):
    """Count of channel values that disagree with the closed form beyond the
    allowed one level at a rounding boundary."""
    bad = 0
    for ch in range(3):
        want, v = _srgb_u8_f64(img[0, :, ch], **variant)
        got = out[0, :, ch].astype(np.float64)
        near = np.abs((v - np.floor(v)) - 0.5) < 1e-3
        bad += int(((got != want) & ~(near & (np.abs(got - want) <= 1))).sum())
    return bad


def test_tonemap_closed_form():
    img = _inputs()
    out = O.tonemap(img)
    assert out.dtype == np.uint8 and out.shape == img.shape
    assert _check(out, img) == 0
    # alpha: the coverage through the unorm store (the float32 product can
    # land exactly on a half level, which then rounds to even)
    v = np.clip(img[0, :, 3].astype(np.float64), 0.0, 1.0) * 255.0
    got = out[0, :, 3].astype(np.float64)
    near = np.abs((v - np.floor(v)) - 0.5) < 1e-3
    assert not ((got != np.rint(v)) & ~(near & (np.abs(got - np.rint(v)) <= 1))).any()
    assert ((got != np.rint(v)) & near).sum() <= 8


def test_tonemap_special_values():
    img = np.array([[[0.0, -5.0, 1e30, 1.0], [1.0, 0.0, 0.0, 0.0], [np.float32(0.0031308 / (1 - 0.0031308)), 0.0,
                                                                   0.0, 0.5]]], np.float32)
    out = O.tonemap(img)
    assert out[0, 0, :3].tolist() == [0, 0, 255] and out[0, 0, 3] == 255
    # x = 1: Reinhard 0.5, sRGB 1.055 * 0.5^(1/2.4) - 0.055 = 0.73536, 187.52 -> 188
    assert out[0, 1, 0] == 188 and out[0, 1, 3] == 0
    # at the cutoff both branches give 0.04045 -> 10.31 -> 10; alpha 0.5 -> 127.5 -> 128 (to even: 128)
    assert out[0, 2, 0] == 10 and out[0, 2, 3] == 128


def test_tonemap_pin_rejects_planted_mistakes():
    img = _inputs()
    out = O.tonemap(img)
    for variant in (dict(reinhard=False), dict(expo=2.4), dict(offset=0.0), dict(rnd=np.floor),
                    dict(keep_max=False)):
        assert _check(out, img, **variant) > 0, variant
