"""Prune and densify on the device (train.rs:414-602) against the host
restatement: every parameter, both Adam moments and grad_acc bit-exact, in
the order the reference's descending swap_remove leaves them."""
import numpy as np
import pytest

from oracle import oracle as O
from nanotracer_amd import synth
from nanotracer_amd.api import GpuSplatBuffer

pytestmark = pytest.mark.gpu

WS = (3, 4, 3, 1, 3, 45)


def _state(n, seed):
    rng = np.random.default_rng(seed)
    s = synth.make_scene(n, seed, 1.0, log_scale_mu=-2.6, log_scale_sd=0.6, logit_mu=-3.0, logit_sd=2.5)
    m = synth.HostSplats(*[rng.normal(0, 1e-3, a.shape).astype(np.float32) for a in s.arrays()])
    v = synth.HostSplats(*[rng.uniform(0, 1e-6, a.shape).astype(np.float32) for a in s.arrays()])
    acc = rng.exponential(2e-4 * 50, n).astype(np.float32)
    return s, m, v, acc


def _upload(ctx, s, m, v, acc, cap):
    g = GpuSplatBuffer.upload(ctx, s, capacity=cap)
    g.set_moments(m, v)
    g.set_grad_acc(acc)
    return g


def _flat_moments(m, v):
    out = []
    for a, b in zip(m.arrays(), v.arrays()):
        out += [a.reshape(-1).copy(), b.reshape(-1).copy()]
    return out


def _check(g, ref_s, ref_mo, ref_acc):
    out = g.readback()
    assert out.n == ref_s.n
    for a, b in zip(out.arrays(), ref_s.arrays()):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    gm, gv = g.moments()
    for k, (a, b) in enumerate(zip(gm.arrays(), gv.arrays())):
        assert np.array_equal(a.reshape(-1), ref_mo[2 * k]) and np.array_equal(b.reshape(-1), ref_mo[2 * k + 1])
    assert np.array_equal(g.grad_acc(), ref_acc)


@pytest.mark.parametrize("n,seed", [(1, 3), (37, 4), (5000, 5), (200_003, 6)])
def test_prune_matches_swap_remove_order(ctx, n, seed):
    s, m, v, acc = _state(n, seed)
    g = _upload(ctx, s, m, v, acc, n)
    removed = g.prune(-5.3)
    ref_s, ref_mo, ref_acc = O.prune(s, -5.3, _flat_moments(m, v), acc)
    assert removed == n - ref_s.n
    _check(g, ref_s, ref_mo, ref_acc)


def test_prune_everything_and_nothing(ctx):
    s, m, v, acc = _state(100, 8)
    s.logit[:] = -9.0
    g = _upload(ctx, s, m, v, acc, 100)
    assert g.prune(-5.3) == 100 and g.n == 0
    s.logit[:] = 1.0
    g2 = _upload(ctx, s, m, v, acc, 100)
    assert g2.prune(-5.3) == 0 and g2.n == 100
    _check(g2, s, _flat_moments(m, v), acc)


@pytest.mark.parametrize("n,max_splats,grad_count,it", [
    (40, 1000, 10, 200),        # everything accepted
    (5000, 5600, 50, 400),      # budget binds mid-list
    (100_000, 180_000, 100, 1200),
    (3000, 3000, 20, 600),      # no room: nothing happens
])
def test_densify_matches_host(ctx, n, max_splats, grad_count, it):
    s, m, v, acc = _state(n, n + it)
    g = _upload(ctx, s, m, v, acc, n + 2 * max(0, max_splats - n) + 8)
    counts = g.densify(grad_count, max_splats, it)
    ref_s, ref_mo, ref_acc, ref_counts = O.densify(s, grad_count, 2e-4, -2.3, max_splats, it,
                                                    _flat_moments(m, v), acc)
    assert counts == tuple(ref_counts)
    if max_splats > n:
        assert counts[0] > 0 and counts[1] > 0  # both kinds exercised
    _check(g, ref_s, ref_mo, ref_acc)


def test_densify_then_prune_sequence(ctx):
    """The reference's cadence at an iteration that does both: prune first,
    then densify, then grad_acc = 0 (train.rs:259-331)."""
    s, m, v, acc = _state(20_000, 21)
    g = _upload(ctx, s, m, v, acc, 60_000)
    g.prune(-5.3)
    counts = g.densify(100, 30_000, 200)
    g.zero_grad_acc()
    r_s, r_mo, r_acc = O.prune(s, -5.3, _flat_moments(m, v), acc)
    mo_m = synth.HostSplats(*[r_mo[2 * k].reshape(-1, w) if w > 1 else r_mo[2 * k] for k, w in enumerate(WS)])
    mo_v = synth.HostSplats(*[r_mo[2 * k + 1].reshape(-1, w) if w > 1 else r_mo[2 * k + 1] for k, w in enumerate(WS)])
    r_s, r_mo, r_acc, r_counts = O.densify(r_s, 100, 2e-4, -2.3, 30_000, 200, _flat_moments(mo_m, mo_v), r_acc)
    assert counts == tuple(r_counts)
    _check(g, r_s, r_mo, np.zeros(r_s.n, np.float32))


def test_densify_needs_room(ctx):
    from nanotracer_amd._lib import NtsError

    s, m, v, acc = _state(1000, 9)
    g = _upload(ctx, s, m, v, acc, 1000)
    with pytest.raises(NtsError):
        g.densify(10, 2000, 200)
    assert g.n == 1000
