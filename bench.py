"""Benchmark: differentiable splat rasteriser on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload batched|plain]

Prints ONE JSON line.

batched (default): one scene of 1,000,000 splats (SH order 3, fp32); each
  training iteration renders 16 cameras on a seeded radius-6 orbit with
  resolutions drawn from {640x480, 1280x720, 1920x1080, 2560x1440}, sums the
  per-view L1 losses and runs backward + fused Adam: one launch per stage
  for all views, the whole iteration captured once in a hipGraph and
  replayed.  The blend tile size (8/16/32) and LDS batches (64-512) are
  autotuned on the workload during warm-up (--tile / --lds-batch fix
  them).  `value` = training iterations per second.
plain: the single-view path (one 1920x1080 view per iteration, reference
  call sequence).

Also reported: single-view 1080p render throughput and p50/p99 latency at
1M splats; the plain single-view training path's views/s; the render server
(64 resident scenes of 50k-3M splats under an HBM budget from hipMemGetInfo,
batches bounded by the slot workspaces, up to 32 views per batch, 3 batches
in flight) swept over Poisson rates from 200 req/s until p99 passes 20 ms,
then at 3200 req/s for its throughput ceiling, with p50/p99
arrival-to-host latency, Mpx/s and views per batch; per-kernel HIP-event
timings with the dominant kernel and the whole step against the HBM
roofline; peaks measured on the device (copy bandwidth, LDS read bandwidth,
shader clock); and the CPU baseline (the host restatement of the reference
path, timed in C++ on this host's cores on a bounded sample of the same
workload).
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
N_VIEWS = 16
SEED = 0x5EED0002
TILE = "auto"
LDS_BATCH = "auto"


def make_views(n):
    from nanotracer_amd import synth
    from nanotracer_amd.api import CameraUniform, ViewSpec

    views = []
    for eye, w, h in synth.orbit_views(N_VIEWS, SEED, 6.0):
        cam = CameraUniform.from_pose(eye, (0, 0, 0), (0, 1, 0), math.pi / 3, w, h, n)
        views.append((ViewSpec(cam, w, h, 12.0), eye))
    return views


def alg_bytes(name: str, N: int, Nv: int, K: int, P: int, T: int, V: int, kb: int = 4) -> float:
    """Algorithmic (compulsory) HBM bytes of one launch of each kernel.
    N splats, Nv visible (splat, view) records, K pairs, P pixels, T tiles,
    V views (all summed over the launch's views)."""
    table = {
        "project": 236 * N + 48 * N,
        "project_views": 236 * N + 52 * N * V,
        "tile_count": 20 * N + 4 * N,
        "scan_reduce": 4 * N * V,
        "scan_down": 8 * N * V,
        "tile_emit": 24 * N + (kb + 4) * K,
        "emit_views": 24 * N * V + (kb + 4) * K,
        "sort_hist": kb * K,
        "sort_part_count": kb * K,
        "sort_onesweep": 2 * (kb + 4) * K,
        "tile_ranges": kb * K + 8 * T,
        "blend_fwd": 8 * T + 40 * K + 16 * P,
        # the batched step's forward also clears the V x N projected-gradient records (36 B)
        "blend_fwd_loss": 8 * T + 40 * K + 16 * P + 16 * P + 16 * P + 8 * P + 36 * N * V,
        "loss": 48 * P,
        "blend_bwd": 8 * T + 40 * K + 32 * P + 36 * Nv,
        "blend_bwd_recompute": 8 * T + 40 * K + 24 * P + 36 * Nv,
        "project_backward": (236 + 48 + 236) * N,
        "project_backward_full": (236 + 48 + 236) * N,
        "adam_step": 7 * 236 * N + 8 * N,
        "backward_adam_views": 236 * N + 36 * N * V + 6 * 236 * N + 8 * N,
        "ssim_moments": 32 * P + 60 * P,
        "ssim_grad": 60 * P + 32 * P + 16 * P,
    }
    return float(table.get(name, 0))


def kernel_table(rep, N, Nv, K, P, T, V):
    kernels = {}
    for name, (launches, ms) in rep.items():
        per = alg_bytes(name, N, Nv, K, P, T, V)
        kernels[name] = dict(launches=launches, avg_ms=ms / launches, alg_bytes=per,
                             gbs=(per / (ms / launches * 1e-3) / 1e9) if per else None)
    return kernels


# HBM bytes per launch measured with rocprofv3 (2.0 x FETCH_SIZE + 1.0 x WRITE_SIZE, the factors
# read on the 4 GiB copy kernel of the peak probe) at the (kernel, tile size, that kernel's LDS batch)
# it was profiled at, with the pair count of that run (the steps with the default culled binning);
# profiles/hbm_counters_train_step_1M_16views.md.
PROFILED_TRAFFIC = {("blend_bwd_recompute", 32, 128): (4.8937e+09, 45_319_957),
                    ("blend_fwd_loss", 32, 512): (4.4507e+09, 45_319_957)}
# Wave-level VALU instructions per launch (rocprofv3 SQ_INSTS_VALU, same profile and pair count).
PROFILED_VALU = {("blend_bwd_recompute", 32, 128): (3.2419e9, 45_319_957),
                 ("blend_fwd_loss", 32, 512): (1.9617e9, 45_319_957)}
# LDS bytes per launch moved by the lanes (rocprofv3 SQ_INSTS_LDS_LOAD_BANDWIDTH + _STORE_BANDWIDTH +
# _ATOMIC_BANDWIDTH, 64-byte units) and the share of the kernel's cycles the LDS spends on indexed
# operations (SQ_LDS_IDX_ACTIVE over 256 CUs against GRBM_GUI_ACTIVE over 8 XCDs), same profile;
# profiles/hbm_counters_train_step_1M_16views.md.
PROFILED_LDS = {("blend_bwd_recompute", 32, 128): ((543362000.0 + 15983100.0 + 27815900.0) * 64, 45_319_957, 0.670),
                ("blend_fwd_loss", 32, 512): ((1957190000.0 + 13700600.0) * 64, 45_319_957, 0.405)}
N_SIMDS = 1024  # 256 CUs x 4 SIMDs


def blend_evals_of(ctx, evals, kernels, cfg, pairs):
    """The blend kernels' work counted on the device (one counted step) and the recompute
    backward against the ceiling of its own arithmetic: a probe replays the backward's
    per-(splat, wave) code (bwd_hit, the nine-term wave reduce, the LDS slot atomic) at the
    kernel's occupancy with every pixel taking part and no memory traffic or barriers; the
    counted evaluations at the probe's rate give the time the arithmetic alone needs.  The
    probe's evaluation runs all four pixel branches, the dearest path, so that time is an upper
    estimate of the arithmetic's share."""
    tile, fb, bb = cfg
    lanes_px_bwd = 64 * (4 if tile == 32 else 1)  # pixels a backward wave covers (2x2 per lane at tile 32)
    out = dict(evals)
    out["fwd_evaluated_pixel_splats"] = evals["fwd_wave_evals"] * 64
    out["bwd_evaluated_pixel_splats"] = evals["bwd_wave_evals"] * lanes_px_bwd
    for name, key, ev in (("blend_fwd_loss", (tile, fb), "fwd_wave_evals"),
                          ("blend_bwd_recompute", (tile, bb), "bwd_wave_evals")):
        pk = (name,) + key
        if pk in PROFILED_VALU and evals[ev] and pairs:
            instr = PROFILED_VALU[pk][0] * pairs / PROFILED_VALU[pk][1]
            out[name.split("_")[1] + "_valu_per_wave_eval"] = round(instr / evals[ev], 1)
    if "blend_bwd_recompute" in kernels and evals["bwd_wave_evals"]:
        rate = ctx.probe_blend_mix(tile, bb)
        meas = kernels["blend_bwd_recompute"]["avg_ms"]
        ceil_ms = evals["bwd_wave_evals"] / rate * 1e3
        out["bwd_mix_probe"] = dict(wave_evals_per_s=round(rate, 1), arithmetic_ms=round(ceil_ms, 4),
                                    measured_ms=round(meas, 4), frac=round(ceil_ms / meas, 3))
    return out


def roofline_of(kernels, cfg=None, pairs=None):
    """cfg = (tile, forward LDS batch, backward LDS batch) of the run."""
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"] * kernels[k]["launches"])
    dk = kernels[dom]
    achieved = dk["gbs"] or 0.0
    traffic, note = None, None
    if cfg is not None and pairs:
        tile, fb, bb = cfg
        key = (dom, tile, bb if dom.startswith("blend_bwd") else fb)
        if key in PROFILED_TRAFFIC:
            t_bytes, t_pairs = PROFILED_TRAFFIC[key]
            traffic = round(t_bytes * pairs / t_pairs)
            note = "rocprofv3 FETCH_SIZE/WRITE_SIZE of this kernel at this tile and LDS batch, scaled by pairs"
    out = dict(bound="hbm", kernel=dom, achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
               frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic, avg_ms=round(dk["avg_ms"], 4),
               alg_bytes_per_launch=dk["alg_bytes"])
    if note:
        out["traffic_note"] = note
    return out


def valu_of(kernels, cfg, pairs, peaks):
    """VALU issue rate and LDS bandwidth of the blend kernels: the profiled instruction count and LDS
    bytes, scaled by this run's pairs, over this run's HIP-event kernel time; the issue rate (at the
    shader clock measured on this device) beside the rate of independent FMA chains measured on this
    device (the VALU's issue ceiling for FP32 work), the LDS rate beside the LDS read bandwidth
    measured on this device."""
    CLOCK_HZ = peaks["clock_ghz"] * 1e9
    LDS_PEAK_TBS = peaks["lds_read_tbs"]
    out = {}
    tile, fb, bb = cfg
    for name in ("blend_bwd_recompute", "blend_fwd_loss"):
        key = (name, tile, bb if name.startswith("blend_bwd") else fb)
        if name not in kernels or key not in PROFILED_VALU or not pairs:
            continue
        sec = kernels[name]["avg_ms"] * 1e-3
        instr = PROFILED_VALU[key][0] * pairs / PROFILED_VALU[key][1]
        rate = instr / N_SIMDS / (sec * CLOCK_HZ)
        fma = peaks.get("fma_per_simd_cycle") or None
        out[name] = dict(valu_instructions=round(instr), valu_issue_per_simd_cycle=round(rate, 4),
                         fma_probe_rate=round(fma, 4) if fma else None,
                         vs_fma_probe=round(rate / fma, 3) if fma else None)
        if key in PROFILED_LDS:
            lb, lp, busy = PROFILED_LDS[key]
            tbs = lb * pairs / lp / sec / 1e12
            out[name].update(lds_tb_s=round(tbs, 2), lds_peak_tb_s=round(LDS_PEAK_TBS, 2),
                             lds_frac=round(tbs / LDS_PEAK_TBS, 3), lds_busy_profiled=busy)
    return out


def render_latency(ctx, g, n, tile, lds=None):
    """Single-view 1920x1080 forward (project, bin, blend) latency, at the
    training step's tile and (when given) its forward LDS batch."""
    from nanotracer_amd import synth
    from nanotracer_amd.api import CameraUniform, ViewBatch, ViewSpec

    u = synth.uniform((SEED << 8) + 31, 1)
    eye = synth.orbit_eye(6.0, 2.0 * math.pi * float(u[0]), 0.3)
    spec = ViewSpec(CameraUniform.from_pose(eye, (0, 0, 0), (0, 1, 0), math.pi / 3, 1920, 1080, n), 1920, 1080, 12.0)
    b = ViewBatch(ctx, 1, n, 1920 * 1080, 20_000_000)
    if lds is not None:
        b.set_lds_batch(*lds)
    lat = []
    for _ in range(40):
        ctx.sync()
        a = time.perf_counter()
        b.render(g, [spec], tile)
        ctx.sync()
        lat.append((time.perf_counter() - a) * 1e3)
    b.free()
    lat = np.array(lat[10:])
    return dict(render_mpx_s=round(1920 * 1080 / (lat.mean() * 1e-3) / 1e6, 2),
                render_p50_ms=round(float(np.percentile(lat, 50)), 3),
                render_p99_ms=round(float(np.percentile(lat, 99)), 3))


def cpu_baseline_batched(s, views, targets, tile):
    """The host restatement of the reference path (oracle/nts_oracle.cpp,
    steady_clock inside C++) on a bounded sample of the same 16-view
    workload: one view of each resolution class (project, bin, blend with
    the saved state, L1, recompute backward, project_backward), median of 5
    runs after 1 warm-up on every core, plus the iteration's one Adam step.
    The 16-view iteration is the sum, over its views, of the median of the
    sampled view of the same resolution, plus Adam.  Single-thread: the
    smallest class's view, median of 3 after the all-core runs."""
    import platform
    from oracle import oracle as O

    classes = {}
    for k, (spec, _) in enumerate(views):
        classes.setdefault((spec.width, spec.height), k)
    picked = sorted(classes.values())
    cams = [O.camera_like(views[k][0].camera) for k in picked]
    tils = [O.tiling(views[k][0].width, views[k][0].height, tile, views[k][0].depth_max) for k in picked]
    tg = [targets[k] for k in picked]
    t0 = time.perf_counter()
    vs, ad = O.time_step_views(s, cams, tils, tg, warmup=1, reps=5, threads=0)
    med = np.median(vs, axis=0)
    per_class = {(views[k][0].width, views[k][0].height): float(m) for k, m in zip(picked, med)}
    t_iter = sum(per_class[(spec.width, spec.height)] for spec, _ in views) + float(np.median(ad))
    small = min(per_class, key=lambda wh: wh[0] * wh[1])
    ks = picked[[views[k][0].width * views[k][0].height for k in picked].index(small[0] * small[1])]
    vs1, _ = O.time_step_views(s, [O.camera_like(views[ks][0].camera)],
                               [O.tiling(small[0], small[1], tile, views[ks][0].depth_max)], [targets[ks]],
                               warmup=0, reps=3, threads=1)
    single = float(np.median(vs1[:, 0]))
    cpu_wall = time.perf_counter() - t0
    pixels = sum(spec.width * spec.height for spec, _ in views)
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return dict(value=round(1.0 / t_iter, 5), unit="iters/s", cores=O.num_threads(), kind="port",
                sample=f"one view of each resolution class of the 16-view iteration ({', '.join(f'{w}x{h}' for w, h in sorted(per_class))}), "
                       f"median of 5 after 1 warm-up on all cores, + one Adam step; iteration = sum over the 16 "
                       f"views of their class's median + Adam ({cpu_wall:.1f} s of CPU work in this run)",
                ms_per_iter=round(t_iter * 1e3, 1), views_per_s=round(len(views) / t_iter, 3),
                mpx_s=round(pixels / t_iter / 1e6, 3),
                view_ms_by_class={f"{w}x{h}": round(v * 1e3, 1) for (w, h), v in sorted(per_class.items())},
                adam_ms=round(float(np.median(ad)) * 1e3, 1),
                single_thread=dict(view=f"{small[0]}x{small[1]}", ms=round(single * 1e3, 1),
                                   all_core_ms=round(per_class[small] * 1e3, 1),
                                   speedup=round(single / per_class[small], 2)),
                cpu_model=model or platform.processor(), hardware_concurrency=os.cpu_count(),
                omp_num_threads=os.environ.get("OMP_NUM_THREADS"))


def plain_views_per_s(ctx, s, steps=10, warmup=3):
    """The plain single-view path (the reference's call sequence, one
    1920x1080 view per training step, reference-parity backward) on the same
    scene: GPU views per second."""
    from nanotracer_amd import synth
    from nanotracer_amd.api import GpuSplatBuffer
    from nanotracer_amd.train import Trainer, View

    u = synth.uniform((SEED << 8) + 31, 1)
    eye = synth.orbit_eye(6.0, 2.0 * math.pi * float(u[0]), 0.3)
    g = GpuSplatBuffer.upload(ctx, s)
    tr = Trainer(ctx, g, [View(eye, 1920, 1080, np.zeros((1080, 1920, 4), np.float32))], 12.0)
    for it in range(warmup):
        tr.step(it)
    ctx.sync()
    t0 = time.perf_counter()
    for it in range(steps):
        tr.step(warmup + it)
    ctx.sync()
    dt = time.perf_counter() - t0
    g.free()
    return round(steps / dt, 2)


SERVE_SCENES = 64
SERVE_RES = ((512, 512), (1280, 720), (1920, 1080), (3840, 2160))


SERVE_RATES = (200.0, 400.0, 600.0, 800.0, 1200.0, 1600.0, 1800.0, 2000.0, 2400.0)
SERVE_SATURATION_RATE = 3200.0  # past what the server sustains: its throughput ceiling
SERVE_BUDGET_CAP = 24 << 30  # the server's share of the device


def _serve_rate(srv, sizes, rate, n_req, rng):
    """Poisson arrivals at `rate` req/s, each one camera of a random resident
    scene at a random resolution; latency = arrival to the tonemapped image
    in host memory."""
    from nanotracer_amd import synth
    from nanotracer_amd.api import CameraUniform, ViewSpec

    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
    reqs = []
    for k in range(n_req):
        sid = int(rng.integers(len(sizes)))
        w, h = SERVE_RES[int(rng.integers(len(SERVE_RES)))]
        eye = synth.orbit_eye(float(rng.uniform(6.0, 9.0)), float(rng.uniform(0, 2 * math.pi)),
                              float(rng.uniform(-0.5, 0.5)))
        cam = CameraUniform.from_pose(eye, (0, 0, 0), (0, 1, 0), math.pi / 3, w, h, sizes[sid])
        reqs.append((sid, ViewSpec(cam, w, h, 24.0)))
    before = srv.stats()
    lat, pix, views = [], 0, []
    inflight = {}
    t0 = time.perf_counter()
    i = 0
    while i < n_req or inflight:
        now = time.perf_counter() - t0
        while i < n_req and arrivals[i] <= now:
            rid, _ = srv.submit(*reqs[i])
            inflight[rid] = (now - arrivals[i], reqs[i][1].width * reqs[i][1].height)
            i += 1
        done = []
        for rid, (lag, px) in inflight.items():
            r = srv.poll(rid)
            if r is not None:
                lat.append(lag * 1e3 + r[1]["latency_ms"])
                views.append(r[1]["batch_views"])
                pix += px
                done.append(rid)
        for rid in done:
            del inflight[rid]
        if not done and (i >= n_req or arrivals[i] > time.perf_counter() - t0):
            time.sleep(0.0002)
    wall = time.perf_counter() - t0
    st = srv.stats()
    d = {k: st[k] - before[k] for k in ("batches", "overflow_reruns", "failed", "memory_limited")}
    lat = np.array(lat)
    return dict(rate_req_s=rate, requests=n_req, p50_ms=round(float(np.percentile(lat, 50)), 3),
                p99_ms=round(float(np.percentile(lat, 99)), 3), mean_ms=round(float(lat.mean()), 3),
                mpx_s=round(pix / wall / 1e6, 1), mean_views_per_batch=round(float(np.mean(views)), 2),
                max_views_per_batch=int(max(views)), batches=int(d["batches"]),
                batches_closed_by_memory=int(d["memory_limited"]), overflow_reruns=int(d["overflow_reruns"]),
                failed=int(d["failed"]), max_in_flight_so_far=int(st["max_in_flight"]))


def run_serving(rates=SERVE_RATES, n_req=1500, seed=SEED + 5, tile=32, lds=512):
    """The render server over 64 resident scenes of 50k-3M splats, swept over
    Poisson arrival rates until p99 passes 20 ms.  HBM budget: the smaller of
    24 GiB and 90% of the free memory hipMemGetInfo reports; the scenes take
    their share and the rest is three slot workspaces, which bound every
    batch's working set (admission by memory)."""
    from nanotracer_amd import synth
    from nanotracer_amd.api import CameraUniform, RenderServer, ViewSpec, device_memory

    sizes = [int(n) for n in synth.serving_scene_sizes(SERVE_SCENES, seed)]
    resident = sum(((n + 63) // 64) * 64 * 4 * 59 for n in sizes)
    free, total = device_memory(0)
    budget = min(SERVE_BUDGET_CAP, int(0.9 * free))
    slot_bytes = (budget - resident) // 3 - (1 << 20)
    srv = RenderServer(hbm_budget=budget, slots=3, max_views=32, max_pixels=96 << 20, max_records=0,
                       pair_capacity=0, tile_size=tile, lds_batch=lds, slot_bytes=slot_bytes)
    t_setup = time.perf_counter()
    for sc in synth.serving_scenes(SERVE_SCENES, seed):
        srv.add_scene(sc)
    t_setup = time.perf_counter() - t_setup
    # warm-up: one request per resolution
    for w, h in SERVE_RES:
        cam = CameraUniform.from_pose((0, 0, 8), (0, 0, 0), (0, 1, 0), math.pi / 3, w, h, sizes[0])
        srv.wait(srv.submit(0, ViewSpec(cam, w, h, 24.0))[0])
    rng = np.random.default_rng(seed + 2)
    sweep = []
    for rate in rates:
        sweep.append(_serve_rate(srv, sizes, rate, n_req, rng))
        print(f"[serving] {rate:.0f} req/s: p50 {sweep[-1]['p50_ms']} ms, p99 {sweep[-1]['p99_ms']} ms, "
              f"{sweep[-1]['mpx_s']} Mpx/s, {sweep[-1]['mean_views_per_batch']} views/batch",
              file=sys.stderr, flush=True)
        if sweep[-1]["p99_ms"] > 20.0:
            break
    # one run past what the server sustains gives its throughput ceiling
    sweep.append(_serve_rate(srv, sizes, SERVE_SATURATION_RATE, n_req, rng))
    print(f"[serving] {SERVE_SATURATION_RATE:.0f} req/s: p50 {sweep[-1]['p50_ms']} ms, "
          f"p99 {sweep[-1]['p99_ms']} ms, {sweep[-1]['mpx_s']} Mpx/s (saturation)",
          file=sys.stderr, flush=True)
    st = srv.stats()
    srv.close()
    ok = [r["rate_req_s"] for r in sweep if r["p99_ms"] < 20.0 and r["failed"] == 0]
    return dict(scenes=len(sizes), resident_splats=int(sum(sizes)), hbm_free_gb=round(free / 1e9, 1),
                hbm_budget_gb=round(budget / 1e9, 1), resident_gb=round(resident / 1e9, 1),
                slot_gb=round(slot_bytes / 1e9, 2), hbm_in_use_gb=round(st["hbm_in_use"] / 1e9, 1),
                setup_s=round(t_setup, 1), sweep=sweep, max_rate_p99_under_20ms=max(ok) if ok else None)


def run_batched(args, ctx):
    from nanotracer_amd import _lib as L
    from nanotracer_amd import synth
    from nanotracer_amd.api import GpuSplatBuffer, ViewBatch

    s = synth.make_scene(N_SPLATS, SEED, 3.0)
    views = make_views(s.n)
    specs = [v[0] for v in views]
    pixels = sum(v.width * v.height for v in specs)
    # targets: the scene with perturbed means, rendered on the device at the
    # reference's tile size
    tgt_scene = GpuSplatBuffer.upload(ctx, synth.perturb_means(s, SEED + 1))
    probe = ViewBatch(ctx, N_VIEWS, s.n, pixels, 600_000_000)
    probe.render(tgt_scene, specs, 16)
    targets = [probe.image(k) for k in range(N_VIEWS)]
    # pair capacity for the smallest tile size the run may use
    probe.render(tgt_scene, specs, 8 if TILE == "auto" else TILE)
    pairs_t, _, _ = probe.stats()
    probe.free()
    tgt_scene.free()
    tbuf = ctx.storage_buffer(np.concatenate([x.reshape(-1) for x in targets]))
    g = GpuSplatBuffer.upload(ctx, s)
    cap = int(max(pairs_t, 1) * 1.3) + 1_000_000
    b = ViewBatch(ctx, N_VIEWS, s.n, pixels, cap)
    mode = L.BWD_RECOMPUTE if args.backward == "recompute" else L.BWD_REF_PARITY
    t = 1
    tune = None
    if TILE == "auto" or LDS_BATCH == "auto":
        # the tuning steps are training steps: they count as warm-up
        from nanotracer_amd.autotune import TILES, autotune_batch

        tune = autotune_batch(ctx, b, g, specs, tbuf, t, L.LOSS_L1, mode,
                              tiles=TILES if TILE == "auto" else (TILE,))
        t = tune.next_t
        tile, lds = tune.tile, (tune.fwd_batch, tune.bwd_batch)
    else:
        tile, lds = TILE, LDS_BATCH
        b.set_lds_batch(*lds)
    for _ in range(max(args.warmup, 1)):
        b.train_step(g, specs, tbuf, t, tile, L.LOSS_L1, mode, graph=True)
        t += 1
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.train_step(g, specs, tbuf, t, tile, L.LOSS_L1, mode, graph=True)
        t += 1
    ctx.sync()
    dt = time.perf_counter() - t0
    pairs, visible, overflow = b.stats()
    if overflow:
        raise RuntimeError("pair capacity overflowed during the timed run")
    losses = b.losses()
    iters_s = args.steps / dt
    # per-kernel timing (eager launches with HIP events on the library stream)
    ctx.prof_reset()
    ctx.prof_enable(True)
    for _ in range(3):
        b.train_step(g, specs, tbuf, t, tile, L.LOSS_L1, mode, graph=False)
        t += 1
    ctx.prof_enable(False)
    n_tiles = sum((-(-v.width // tile)) * (-(-v.height // tile)) for v in specs)
    kernels = kernel_table(ctx.prof_report(), s.n, visible, pairs, pixels, n_tiles, N_VIEWS)
    # evaluation counts of one more step, taken by the counting form of the
    # blend kernels (not timed; the grid scheduler, which the form requires)
    if tune and tune.scheduler != "grid":
        b.set_scheduler(L.SCHED_GRID)
    b.set_eval_counting(True)
    b.train_step(g, specs, tbuf, t, tile, L.LOSS_L1, mode, graph=False)
    t += 1
    evals = b.eval_counts()
    b.set_eval_counting(False)
    if tune and tune.scheduler != "grid":
        b.set_scheduler(L.SCHED_PERSISTENT)
    ssim = ssim_step(ctx, b, g, specs, tbuf, t, tile, mode, min(args.steps, 10), s.n, visible, pairs, pixels,
                     n_tiles, args.ssim_lambda) if args.ssim_lambda > 0 else None
    # (the SSIM leg advanced the Adam step by fewer than 100)
    refbin = reference_binning_step(ctx, b, g, specs, tbuf, t + 100, tile, mode, min(args.steps, 10))
    b.free()
    out = dict(
        metric="train iters/s + render Mpx/s at 1M splats 1080p; p50/p99 render latency",
        value=round(iters_s, 3), unit="iters/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=round(dt * 1e3 / args.steps, 3), higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype="f32", data="synthetic",
        config=dict(workload="batched training iteration: 1M splats, 16 views of mixed resolution per launch, "
                             "hipGraph replay", splats=s.n, views=N_VIEWS, pixels=pixels, tile=tile,
                    lds_batch=dict(forward=lds[0], backward=lds[1]),
                    tuning="autotuned during warm-up" if tune else "fixed",
                    scheduler=tune.scheduler if tune else "grid", pairs=pairs,
                    visible_records=visible, backward=args.backward, loss="L1 summed over views"),
        views_per_s=round(iters_s * N_VIEWS, 2), train_mpx_s=round(iters_s * pixels / 1e6, 2),
        final_loss=float(losses.sum()),
    )
    out.update(render_latency(ctx, g, s.n, tile, lds))
    peaks = ctx.probe_peaks(4 << 30)
    peaks.update(ctx.probe_valu())
    out["roofline"] = roofline_of(kernels, (tile, lds[0], lds[1]), pairs)
    out["roofline"]["frac_of_measured_copy"] = round(out["roofline"]["achieved"] / peaks["hbm_copy_gbs"], 4)
    out["peaks"] = dict(hbm_copy_gbs=round(peaks["hbm_copy_gbs"], 1), lds_read_tbs=round(peaks["lds_read_tbs"], 2),
                        clock_ghz=round(peaks["clock_ghz"], 3), hbm_vendor_gbs=HBM_PEAK_GBS,
                        valu_fp32_fma_tflops=round(peaks["fma_tflops"], 1),
                        valu_fma_per_simd_cycle=round(peaks["fma_per_simd_cycle"], 4),
                        how="copy of 4 GiB (read + write bytes); conflict-free ds_read_b128 over all CUs; "
                            "shader cycle counter over the wall clock during the LDS run; independent "
                            "v_fma_f32 chains at 8 waves per SIMD")
    # the whole step against HBM: compulsory bytes of the step (params, Adam, records, pairs, pixels, tiles)
    b_step = 2368.0 * s.n + 132.0 * visible + 116.0 * pairs + 76.0 * pixels + 8.0 * n_tiles
    ach = b_step / (out["ms_per_step"] * 1e-3) / 1e9
    out["step_roofline"] = dict(alg_bytes=b_step, achieved_gbs=round(ach, 1), peak_gbs=HBM_PEAK_GBS,
                                frac=round(ach / HBM_PEAK_GBS, 4),
                                frac_of_measured_copy=round(ach / peaks["hbm_copy_gbs"], 4),
                                formula="2368 N + sum_v (132 N_v + 116 K_v + 76 P_v + 8 T_v)",
                                counters=dict(N=s.n, N_v=visible, K=pairs, P=pixels, T=n_tiles))
    out["valu"] = valu_of(kernels, (tile, lds[0], lds[1]), pairs, peaks)
    out["blend_evals"] = blend_evals_of(ctx, evals, kernels, (tile, lds[0], lds[1]), pairs)
    if ssim:
        out["ssim_step"] = ssim
    out["reference_binning_step"] = refbin
    if tune:
        out["autotune"] = dict(predicted_step_ms=round(tune.step_ms, 3), blend_ms_by_scheduler=tune.blend_ms,
                               table=tune.table)
    out["kernels"] = {k: dict(avg_ms=round(v["avg_ms"], 4), launches=v["launches"],
                              gbs=None if v["gbs"] is None else round(v["gbs"], 1)) for k, v in kernels.items()}
    out["plain_single_view"] = dict(views_per_s=plain_views_per_s(ctx, s), view="1920x1080, tile 16",
                                    sequence="reference call sequence, reference-parity backward")
    if args.serve_rate > 0:
        out["serving"] = run_serving(tile=tile, lds=lds[0])
    if not args.no_cpu_baseline:
        cb = cpu_baseline_batched(s, views, targets, tile)
        out["cpu_baseline"] = cb
        out["vs_cpu"] = dict(batched_views_per_s=out["views_per_s"], cpu_views_per_s=cb["views_per_s"],
                             batched_over_cpu=round(out["views_per_s"] / cb["views_per_s"], 1),
                             plain_1080p_views_per_s=out["plain_single_view"]["views_per_s"],
                             cpu_1080p_views_per_s=round(1e3 / cb["view_ms_by_class"]["1920x1080"], 3)
                             if "1920x1080" in cb["view_ms_by_class"] else None)
        if out["vs_cpu"]["cpu_1080p_views_per_s"]:
            out["vs_cpu"]["plain_over_cpu_1080p"] = round(out["vs_cpu"]["plain_1080p_views_per_s"]
                                                          / out["vs_cpu"]["cpu_1080p_views_per_s"], 1)
    return out


def reference_binning_step(ctx, b, g, specs, tbuf, t, tile, mode, steps):
    """The same batched iteration with the reference's binning (BIN_REFERENCE:
    TileBinner::bin's pairs, so every view's keys are the reference's), graph
    replays timed as in the headline run; the batch goes back to the default
    culled rule afterwards."""
    from nanotracer_amd import _lib as L

    b.set_binning(L.BIN_REFERENCE)
    try:
        for _ in range(2):
            b.train_step(g, specs, tbuf, t, tile, L.LOSS_L1, mode, graph=True)
            t += 1
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            b.train_step(g, specs, tbuf, t, tile, L.LOSS_L1, mode, graph=True)
            t += 1
        ctx.sync()
        dt = time.perf_counter() - t0
        pairs, _, overflow = b.stats()
        if overflow:
            raise RuntimeError("pair capacity overflowed in the reference-binning steps")
    finally:
        b.set_binning(L.BIN_CULLED)
    return dict(binning="reference (TileBinner::bin pairs; keys bit-exact)", steps=steps,
                ms_per_step=round(dt * 1e3 / steps, 3), iters_s=round(steps / dt, 3), pairs=pairs)


def ssim_step(ctx, b, g, specs, tbuf, t, tile, mode, steps, N, Nv, K, P, T, lam=0.2):
    """The same batched iteration with the reference's default training loss,
    (1 - lam) MSE + lam (1 - SSIM) (train.rs:48-52, main.rs:189-190): graph
    replays timed after the headline run, then per-kernel HIP-event times."""
    from nanotracer_amd import _lib as L

    b.set_ssim_lambda(lam)
    b.train_step(g, specs, tbuf, t, tile, L.LOSS_SSIM, mode, graph=True)
    t += 1
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.train_step(g, specs, tbuf, t, tile, L.LOSS_SSIM, mode, graph=True)
        t += 1
    ctx.sync()
    dt = time.perf_counter() - t0
    ctx.prof_reset()
    ctx.prof_enable(True)
    for _ in range(2):
        b.train_step(g, specs, tbuf, t, tile, L.LOSS_SSIM, mode, graph=False)
        t += 1
    ctx.prof_enable(False)
    ks = kernel_table(ctx.prof_report(), N, Nv, K, P, T, len(specs))
    return dict(ssim_lambda=lam, steps=steps, ms_per_step=round(dt * 1e3 / steps, 3),
                iters_s=round(steps / dt, 3), final_loss=float(b.losses().sum()),
                kernels={k: dict(avg_ms=round(v["avg_ms"], 4), gbs=None if v["gbs"] is None else round(v["gbs"], 1))
                         for k, v in ks.items() if k.startswith("ssim") or k.startswith("blend")})


def run_plain(args, ctx):
    from nanotracer_amd import _lib as L
    from nanotracer_amd import synth
    from nanotracer_amd.api import GpuSplatBuffer
    from nanotracer_amd.train import Trainer, View

    W, H = 1920, 1080
    s = synth.make_scene(N_SPLATS, SEED, 3.0)
    u = synth.uniform((SEED << 8) + 31, 1)
    eye = synth.orbit_eye(6.0, 2.0 * math.pi * float(u[0]), 0.3)
    tgt_scene = GpuSplatBuffer.upload(ctx, synth.perturb_means(s, SEED + 1))
    tmp = Trainer(ctx, tgt_scene, [View(eye, W, H)], 12.0)
    _, bins, img, _, _ = tmp.render(tmp.views[0])
    target = img.download(np.float32, W * H * 4).reshape(H, W, 4)
    bins.free()
    tgt_scene.free()
    g = GpuSplatBuffer.upload(ctx, s)
    mode = L.BWD_RECOMPUTE if args.backward == "recompute" else L.BWD_REF_PARITY
    tr = Trainer(ctx, g, [View(eye, W, H, target)], 12.0, backward_mode=mode)
    for it in range(args.warmup):
        tr.step(it)
    ctx.sync()
    t0 = time.perf_counter()
    pairs = 0
    for it in range(args.steps):
        pairs = tr.step(args.warmup + it).total_pairs
    ctx.sync()
    dt = time.perf_counter() - t0
    ctx.prof_reset()
    ctx.prof_enable(True)
    for it in range(3):
        tr.step(1000 + it)
    ctx.prof_enable(False)
    proj = tr._bufs["proj"].download(np.float32, N_SPLATS * 12).reshape(-1, 12)
    Nv = int((proj[:, 7] != 0).sum())
    kernels = kernel_table(ctx.prof_report(), N_SPLATS, Nv, pairs, W * H, ((W + 15) // 16) * ((H + 15) // 16), 1)
    out = dict(
        metric="train iters/s + render Mpx/s at 1M splats 1080p; p50/p99 render latency",
        value=round(args.steps / dt, 3), unit="iters/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=round(dt * 1e3 / args.steps, 3), higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype="f32", data="synthetic",
        config=dict(workload="plain single-view training step (reference call sequence)", splats=N_SPLATS,
                    width=W, height=H, tile=16, pairs=pairs, visible=Nv, backward=args.backward),
    )
    out.update(render_latency(ctx, g, s.n, 16))
    out["roofline"] = roofline_of(kernels)
    out["kernels"] = {k: dict(avg_ms=round(v["avg_ms"], 4), launches=v["launches"],
                              gbs=None if v["gbs"] is None else round(v["gbs"], 1)) for k, v in kernels.items()}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=("batched", "plain"), default="batched")
    ap.add_argument("--backward", choices=("ref", "recompute"), default="recompute",
                    help="reference-parity backward or the recompute (final-T, last-index) backward")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tile", choices=("auto", "8", "16", "32"), default="auto",
                    help="blend tile size, or autotuned during warm-up")
    ap.add_argument("--lds-batch", default="auto", help="blend forward,backward LDS batch, or auto")
    ap.add_argument("--serve-rate", type=float, default=1.0,
                    help="0 skips the render-server sweep (Poisson rates from 200 req/s until p99 > 20 ms, then 3200 req/s)")
    ap.add_argument("--ssim-lambda", type=float, default=0.2,
                    help="also time the iteration with the combined MSE + SSIM loss; 0 skips it")
    args = ap.parse_args()
    global TILE, LDS_BATCH
    TILE = args.tile if args.tile == "auto" else int(args.tile)
    LDS_BATCH = args.lds_batch if args.lds_batch == "auto" else tuple(int(x) for x in args.lds_batch.split(","))
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the sharding design is view-parallel with one gradient reduce-scatter); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    from nanotracer_amd.api import GpuCtx

    ctx = GpuCtx(0)
    out = run_batched(args, ctx) if args.workload == "batched" else run_plain(args, ctx)
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
