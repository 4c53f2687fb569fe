"""Benchmark: differentiable splat rasteriser training step + render on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints ONE JSON line.  The timed region holds K full training steps
(project, bin, forward blend, device loss, backward blend, project_backward,
fused Adam) with every input resident in HBM; it is bracketed by device
synchronisation.  Afterwards a few steps run with per-kernel HIP-event timing
(on the library's own stream) to report the dominant kernel against the HBM
roofline, and the host restatement of the reference path (the checker) is
timed on a bounded sample as the CPU baseline.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
N_SPLATS = 1_000_000
WIDTH, HEIGHT = 1920, 1080
SEED = 0x5EED0002


def build_workload(ctx, backward_mode):
    from nanotracer_amd import synth
    from nanotracer_amd.api import GpuSplatBuffer
    from nanotracer_amd.train import Trainer, View

    s = synth.make_scene(N_SPLATS, SEED, 3.0)
    u = synth.uniform((SEED << 8) + 31, 1)
    eye = synth.orbit_eye(6.0, 2.0 * math.pi * float(u[0]), 0.3)
    # target: the same scene with perturbed means, rendered on the device
    tgt_scene = GpuSplatBuffer.upload(ctx, synth.perturb_means(s, SEED + 1))
    tmp = Trainer(ctx, tgt_scene, [View(eye, WIDTH, HEIGHT)], 12.0)
    _, bins, img, _, _ = tmp.render(tmp.views[0])
    target = img.download(np.float32, WIDTH * HEIGHT * 4).reshape(HEIGHT, WIDTH, 4)
    bins.free()
    tgt_scene.free()
    g = GpuSplatBuffer.upload(ctx, s)
    tr = Trainer(ctx, g, [View(eye, WIDTH, HEIGHT, target)], 12.0, backward_mode=backward_mode)
    return s, eye, target, g, tr


def alg_bytes(name: str, N: int, Nv: int, K: int, P: int, T: int, kb: int = 4) -> float:
    """Algorithmic (compulsory) HBM bytes of one launch of each kernel."""
    table = {
        "project": 236 * N + 48 * N,
        "tile_count": 20 * N + 4 * N,
        "scan_reduce": 4 * N,
        "scan_down": 8 * N,
        "tile_emit": 24 * N + (kb + 4) * K,
        "sort_hist": kb * K,
        "sort_onesweep": 2 * (kb + 4) * K,
        "tile_ranges": kb * K + 8 * T,
        "blend_fwd": 8 * T + 40 * K + 16 * P,
        "loss": 48 * P,
        "blend_bwd": 8 * T + 40 * K + 32 * P + 36 * Nv,
        "blend_bwd_recompute": 8 * T + 40 * K + 24 * P + 36 * Nv,
        "project_backward_full": (236 + 48 + 236) * N,
        "project_backward": (236 + 48 + 236) * N,
        "adam_step": 7 * 236 * N + 8 * N,
    }
    return float(table.get(name, 0))


def cpu_baseline(s, eye, target):
    """The checker (host restatement of the reference path) on one full
    training step of the same workload; returns iterations/s."""
    from oracle import oracle as O

    t0 = time.perf_counter()
    cam = O.camera_from_pose(eye, (0, 0, 0), (0, 1, 0), math.pi / 3, WIDTH, HEIGHT, s.n)
    t = O.tiling(WIDTH, HEIGHT, 16, 12.0)
    op = O.project(s, cam)
    keys, pay, rg = O.bin(op, t)
    img = O.composite(op, pay, rg, t)
    loss, dl = O.loss_grad(img, target, 0)
    pg = O.composite_backward_ref(op, pay, rg, t, img, dl)
    g = O.project_backward(s, cam, pg)
    flat_p = [a.reshape(-1).copy() for a in s.arrays()]
    for p, gr in zip(flat_p, g.arrays()):
        m = np.zeros_like(p)
        v = np.zeros_like(p)
        O.adam_step(p, gr.reshape(-1), m, v, 1.6e-4, 0.9, 0.999, 1e-15, 1)
    dt = time.perf_counter() - t0
    return dict(value=1.0 / dt, unit="iters/s", cores=O.num_threads(), kind="port",
                sample=f"one full training step (project, bin, blend fwd, L1, blend bwd, project_backward, Adam) "
                       f"of the {N_SPLATS}-splat {WIDTH}x{HEIGHT} workload, {dt:.2f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--backward", choices=("ref", "recompute"), default="ref",
                    help="reference-parity backward or the recompute (final-T, last-index) backward")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(see the sharding design); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from nanotracer_amd import _lib as L
    from nanotracer_amd.api import GpuCtx

    ctx = GpuCtx(0)
    mode = L.BWD_RECOMPUTE if args.backward == "recompute" else L.BWD_REF_PARITY
    s, eye, target, g, tr = build_workload(ctx, mode)
    for it in range(args.warmup):
        tr.step(it)
    ctx.sync()
    t0 = time.perf_counter()
    pairs = 0
    for it in range(args.steps):
        pairs = tr.step(args.warmup + it).total_pairs
    ctx.sync()
    dt = time.perf_counter() - t0
    ms_per_step = dt * 1e3 / args.steps
    iters_s = args.steps / dt

    # render-only latency (forward: project, bin, blend) at the same size
    lat = []
    for _ in range(30):
        ctx.sync()
        a = time.perf_counter()
        _, bins, _, _, _ = tr.render(tr.views[0])
        ctx.sync()
        lat.append((time.perf_counter() - a) * 1e3)
        bins.free()
    lat = np.array(lat[5:])
    render_mpx_s = WIDTH * HEIGHT / (lat.mean() * 1e-3) / 1e6

    # per-kernel timing on the library stream
    ctx.prof_reset()
    ctx.prof_enable(True)
    nprof = 3
    for it in range(nprof):
        tr.step(1000 + it)
    ctx.prof_enable(False)
    rep = ctx.prof_report()
    proj = tr._bufs["proj"].download(np.float32, N_SPLATS * 12).reshape(-1, 12)
    Nv = int((proj[:, 7] != 0).sum())
    T = ((WIDTH + 15) // 16) * ((HEIGHT + 15) // 16)
    kernels = {}
    for name, (launches, ms) in rep.items():
        per = alg_bytes(name, N_SPLATS, Nv, pairs, WIDTH * HEIGHT, T)
        kernels[name] = dict(launches=launches, avg_ms=ms / launches, alg_bytes=per,
                             gbs=(per / (ms / launches * 1e-3) / 1e9) if per else None)
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"] * kernels[k]["launches"])
    dk = kernels[dom]
    achieved = dk["gbs"] or 0.0
    roofline = dict(bound="hbm", kernel=dom, achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None, avg_ms=round(dk["avg_ms"], 4),
                    alg_bytes_per_launch=dk["alg_bytes"])

    out = dict(
        metric="train iters/s + render Mpx/s at 1M splats 1080p; p50/p99 render latency",
        value=round(iters_s, 3), unit="iters/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=round(ms_per_step, 3), higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f32",
        data="synthetic",
        config=dict(workload="plain single-view training step (project, bin, blend fwd, L1, blend bwd, "
                             "project_backward, fused Adam)", splats=N_SPLATS, width=WIDTH, height=HEIGHT,
                    tile=16, pairs=pairs, visible=Nv, backward=args.backward),
        render_mpx_s=round(render_mpx_s, 2), render_p50_ms=round(float(np.percentile(lat, 50)), 3),
        render_p99_ms=round(float(np.percentile(lat, 99)), 3),
        roofline=roofline,
        kernels={k: dict(avg_ms=round(v["avg_ms"], 4), launches=v["launches"],
                         gbs=None if v["gbs"] is None else round(v["gbs"], 1)) for k, v in kernels.items()},
    )
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(s, eye, target)
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
