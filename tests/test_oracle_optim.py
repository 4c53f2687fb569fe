"""Pins the checker's Adam, constraints, grad accumulator, loss, prune and
densify: closed forms and the reference's known-answer tests (adam.rs,
train.rs:626-743, loss.rs)."""
import math

import numpy as np

from oracle import oracle as O
from nanotracer_amd import synth


def _adam_f64(p, gs, lr=1.6e-4, b1=0.9, b2=0.999, eps=1e-15):
    """Textbook Adam (Kingma & Ba 2015, Algorithm 1) in float64."""
    p = p.astype(np.float64)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for t, g in enumerate(gs, start=1):
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        p = p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, m, v


def test_adam_matches_textbook_sequences():
    # the five sequences of tests/adam_gpu_parity.rs
    seqs = [(256, 10, lambda t, i: 0.1), (256, 10, lambda t, i: -0.2),
            (1024, 20, lambda t, i: (1.0 if i % 2 == 0 else -1.0) * (0.05 + 0.01 * t)),
            (64, 1, lambda t, i: 0.01 * (i + 1.0)), (8192, 5, lambda t, i: 0.05 * math.sin(i * 0.01 + t * 0.1))]
    for n, steps, fn in seqs:
        p = np.full(n, 0.5, np.float32)
        m = np.zeros(n, np.float32)
        v = np.zeros(n, np.float32)
        gs = []
        for t in range(1, steps + 1):
            g = np.array([fn(t, i) for i in range(n)], np.float32)
            gs.append(g.astype(np.float64))
            O.adam_step(p, g, m, v, 1.6e-4, 0.9, 0.999, 1e-15, t)
        rp, rm, rv = _adam_f64(np.full(n, 0.5), gs)
        assert np.abs(p - rp).max() < 1e-6
        assert np.abs(m - rm).max() < 1e-6 and np.abs(v - rv).max() < 1e-7


def test_adam_first_step_moves_by_lr_against_the_gradient():
    # adam.rs:119-130: first step update is lr * sign(g) when eps << |g|
    p = np.array([1.0, 2.0, 3.0], np.float32)
    m, v = np.zeros(3, np.float32), np.zeros(3, np.float32)
    O.adam_step(p, np.array([1.0, -1.0, 1e-3], np.float32), m, v, 1.6e-4, 0.9, 0.999, 1e-15, 1)
    assert np.allclose(p, [1.0 - 1.6e-4, 2.0 + 1.6e-4, 3.0 - 1.6e-4], atol=2e-7)


def test_zero_gradient_lanes_are_inert():
    p = np.array([1.0, 0.25], np.float32)
    m, v = np.zeros(2, np.float32), np.zeros(2, np.float32)
    for t in range(1, 5):
        O.adam_step(p, np.zeros(2, np.float32), m, v, 1.6e-4, 0.9, 0.999, 1e-15, t)
    assert list(p) == [1.0, 0.25] and not m.any() and not v.any()


def test_constraints_closed_form():
    rot = np.array([[2.0, 0, 0, 0], [0, 0, 0, 0], [1, 1, 1, 1]], np.float32)
    ls = np.array([[-11, 0, 6], [-10, 5, 1], [0, 0, 0]], np.float32)
    lg = np.array([-12, 3, 12], np.float32)
    O.apply_constraints(rot, ls, lg)
    assert np.allclose(rot[0], [1, 0, 0, 0]) and not rot[1].any() and np.allclose(rot[2], [0.5] * 4)
    assert np.array_equal(ls, np.array([[-10, 0, 5], [-10, 5, 1], [0, 0, 0]], np.float32))
    assert list(lg) == [-10, 3, 10]


def test_accumulate_grad_is_l1_of_dpos():
    dpos = np.array([[1, -2, 3], [-0.5, 0, 0.25]], np.float32)
    acc = np.array([10, 0], np.float32)
    O.accumulate_grad(dpos, acc)
    assert list(acc) == [16.0, 0.75]


def test_loss_closed_forms():
    pred = np.zeros((2, 3, 4), np.float32)
    tgt = np.zeros((2, 3, 4), np.float32)
    pred[0, 0, :3] = [1, -2, 0]
    loss, dl = O.loss_grad(pred, tgt, 1)  # MSE: sum d^2 / 3WH, grad 2d / 3WH
    assert abs(loss - 5 / 18) < 1e-9 and np.allclose(dl[0, 0, :3], [2 / 18, -4 / 18, 0])
    loss, dl = O.loss_grad(pred, tgt, 0)  # L1: sum |d| / 3WH, grad sign(d) / 3WH
    assert abs(loss - 3 / 18) < 1e-9 and np.allclose(dl[0, 0, :3], [1 / 18, -1 / 18, 0]) and dl[0, 0, 2] == 0


def _mk(pos, scale, logit, dc):
    n = len(pos)
    return synth.HostSplats(np.array(pos, np.float32).reshape(n, 3), np.tile(np.array([1, 0, 0, 0], np.float32), (n, 1)),
                            np.array(scale, np.float32).reshape(n, 3), np.array(logit, np.float32),
                            np.array(dc, np.float32).reshape(n, 3), np.zeros((n, 45), np.float32))


def test_densify_clones_small_and_splits_large():
    # train.rs:626-704
    s = _mk([[0, 0, 0], [5, 5, 5], [-1, -1, -1]], [[-3] * 3, [0] * 3, [-1] * 3], [2.0, 2.0, 2.0],
            [[0.5] * 3, [0.3] * 3, [0.1, 0.2, 0.3]])
    mom = [np.zeros(3 * w, np.float32) for w in (3, 3, 4, 4, 3, 3, 1, 1, 3, 3, 45, 45)]
    out, mo, acc, counts = O.densify(s, 1, 0.5, -2.3, 100, 42, mom, np.array([1.0, 1.0, 0.0001], np.float32))
    assert counts == (1, 2, 1)
    assert out.n == 5 and acc.size == 5
    assert [m.size for m in mo] == [5 * w for w in (3, 3, 4, 4, 3, 3, 1, 1, 3, 3, 45, 45)]
    # clone: parent keeps its place with logit + ln(1/2); the clone is appended
    assert abs(out.logit[0] - (2.0 - math.log(2))) < 1e-6
    # the split parent (index 1) is gone: its slot holds the last child
    lsplit = -math.log(1.6)
    assert np.allclose(out.log_scale[1], [lsplit] * 3, atol=1e-6)
    children = [i for i in range(out.n) if np.allclose(out.log_scale[i], [lsplit] * 3, atol=1e-6)]
    assert len(children) == 2
    for i in children:  # sampled around the parent with sigma = 1
        assert np.linalg.norm(out.pos[i] - 5.0) < 6.0


def test_prune_drops_low_opacity_splats():
    # train.rs:706-743
    s = _mk([[i] * 3 for i in range(5)], [[-1] * 3] * 5, [2, -10, 2, -10, 2], [[0] * 3] * 5)
    mom = [np.zeros(5 * w, np.float32) for w in (3, 3, 4, 4, 3, 3, 1, 1, 3, 3, 45, 45)]
    out, mo, acc = O.prune(s, -5.3, mom, np.array([10, 20, 30, 40, 50], np.float32))
    assert out.n == 3 and acc.size == 3 and [m.size for m in mo] == [3 * w for w in (3, 3, 4, 4, 3, 3, 1, 1, 3, 3, 45, 45)]
    assert np.allclose(out.logit, 2.0)
    # swap_remove applied high-to-low: remove 3 (last=4 moves in), then 1 (last=3 -> 4's data moves)
    assert list(acc) == [10, 50, 30]
    assert np.allclose(out.pos[:, 0], [0, 4, 2])
