"""Writes the golden fixtures of the test suite from the host restatement.

The reference commits no golden vectors for this path (its tests compute
everything at run time), so these fixtures freeze the checker's own outputs
on small seeded inputs: a later change to the checker or the device path
that moves a bit shows up against them.

    python tests/golden/make_golden.py

Fixtures (numpy .npz, no pickles):
  scene64_32x32.npz   64 splats in front of a 32x32 camera, tile 16:
                      inputs, projected records, sorted keys / payloads,
                      ranges, image, final T, last index, dL/dC of an L1 loss
                      against a perturbed copy, the recompute backward's
                      projected gradient and the full-chain parameter
                      gradient.
  binner_cases.npz    the inputs of the four-splat and two-splat binner
                      tests of tile_binner.rs:440-541 with the keys,
                      payloads and ranges they produce.
  config_a_step.npz   config A (5,000 splats, 256x256, L1, Adam 1.6e-4):
                      parameters before and after one reference-parity
                      training step, its loss and parameter gradient.
"""
from __future__ import annotations

import math
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

from oracle import oracle as O  # noqa: E402
from nanotracer_amd import synth  # noqa: E402


def scene64():
    s = synth.make_scene(64, 0x601DE4, 0.6, log_scale_mu=-2.2, log_scale_sd=0.4, logit_mu=0.5)
    cam = O.camera_from_pose((0.0, 0.0, 2.5), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), math.pi / 3, 32, 32, s.n)
    t = O.tiling(32, 32, 16, 5.0)
    return s, cam, t


def make_scene64():
    s, cam, t = scene64()
    proj = O.project(s, cam)
    keys, pay, rg = O.bin(proj, t)
    img, ft, li = O.composite(proj, pay, rg, t, with_state=True)
    tgt_s = synth.perturb_means(s, 0x601DE5, 0.05)
    tp = O.project(tgt_s, cam)
    tk, tpay, trg = O.bin(tp, t)
    target = O.composite(tp, tpay, trg, t)
    loss, dl = O.loss_grad(img, target, 0)
    pg = O.composite_backward_recompute(proj, pay, rg, t, ft, li, dl)
    g = O.project_backward(s, cam, pg, full=True)
    np.savez_compressed(HERE / "scene64_32x32.npz", pos=s.pos, rot=s.rot, log_scale=s.log_scale, logit=s.logit,
                        sh_dc=s.sh_dc, sh_rest=s.sh_rest, view=np.array(cam.view, np.float32),
                        cam_pos=np.array(cam.cam_pos, np.float32), focal=np.float32(cam.focal),
                        projected=proj, keys=keys, payloads=pay, ranges=rg, image=img, final_t=ft, last_idx=li,
                        target=target, loss=np.float64(loss), dl_dc=dl, projected_grad=pg,
                        g_pos=g.pos, g_rot=g.rot, g_log_scale=g.log_scale, g_logit=g.logit, g_sh_dc=g.sh_dc,
                        g_sh_rest=g.sh_rest)


def _rec(x, y, depth, r):
    return np.array([x, y, depth, r, 1.0, 0.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0], np.float32)


def make_binner_cases():
    # the inputs of tile_binner.rs:440-541: four splats at tile centres of a
    # 64x64 image (depths 10, 20, 30, 5), two splats in tile (1, 1) at depths
    # 80 and 5; tile 16, depth_max 100
    four = np.stack([_rec(8, 8, 10.0, 4), _rec(40, 8, 20.0, 4), _rec(8, 56, 30.0, 4), _rec(56, 56, 5.0, 4)])
    two = np.stack([_rec(24, 24, 80.0, 4), _rec(24, 24, 5.0, 4)])
    out = {}
    for name, proj in (("four", four), ("two", two)):
        t = O.tiling(64, 64, 16, 100.0)
        k, p, r = O.bin(proj, t)
        out.update({f"{name}_projected": proj, f"{name}_tiling": np.array([64, 64, 16], np.uint32),
                    f"{name}_depth_max": np.float32(100.0), f"{name}_keys": k, f"{name}_payloads": p,
                    f"{name}_ranges": r})
    np.savez_compressed(HERE / "binner_cases.npz", **out)


def make_config_a_step():
    s = synth.make_scene(5000, 0x5EED0000, 1.0)
    cam = O.camera_from_pose((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), math.pi / 3, 256, 256, s.n)
    t = O.tiling(256, 256, 16, 6.0)
    tp = O.project(synth.perturb_means(s, 0x7A46E7), cam)
    tk, tpay, trg = O.bin(tp, t)
    target = O.composite(tp, tpay, trg, t)
    op = O.project(s, cam)
    k, pay, rg = O.bin(op, t)
    img = O.composite(op, pay, rg, t)
    loss, dl = O.loss_grad(img, target, 0)
    pg = O.composite_backward_ref(op, pay, rg, t, img, dl)
    g = O.project_backward(s, cam, pg)
    after = s.copy()
    acc = np.zeros(s.n, np.float32)
    O.accumulate_grad(g.pos, acc)
    for p, gr in zip(after.arrays(), g.arrays()):
        flat = p.reshape(-1)
        O.adam_step(flat, gr.reshape(-1), np.zeros_like(flat), np.zeros_like(flat), 1.6e-4, 0.9, 0.999, 1e-15, 1)
    O.apply_constraints(after.rot, after.log_scale, after.logit)
    np.savez_compressed(HERE / "config_a_step.npz", target=target, loss=np.float64(loss), grad_acc=acc,
                        **{f"before_{n}": a for n, a in zip(NAMES, s.arrays())},
                        **{f"after_{n}": a for n, a in zip(NAMES, after.arrays())},
                        **{f"grad_{n}": a for n, a in zip(NAMES, g.arrays())})


NAMES = ("pos", "rot", "log_scale", "logit", "sh_dc", "sh_rest")

if __name__ == "__main__":
    make_scene64()
    make_binner_cases()
    make_config_a_step()
    for f in sorted(HERE.glob("*.npz")):
        print(f.name, f.stat().st_size)
