"""Autotuner of the batched step's blend configuration: tile size (8, 16 or
32 pixels), splats per LDS batch (forward 64-512, backward 64-256) and the
tile scheduler (one workgroup per tile, or persistent workgroups taking
tiles from a device counter).

It times real training steps of the caller's own workload (scene, views,
targets) with per-kernel HIP-event timing, so the choice follows the data:
splat footprints, resolutions and depth complexity all move the optimum.
The forward and backward blend are separate launches, so one step times
one forward batch and one backward batch together: a tile size needs
max(|forward batches|, |backward batches|) timed steps, each after one
untimed step of the same configuration.  The winner minimises the summed
kernel time of a step: the rest of the step (projection, binning, sort,
Adam) for that tile size plus the fastest forward and fastest backward.

Tuning steps are ordinary training steps: they advance the scene and the
Adam step counter, like warm-up iterations.
"""
from __future__ import annotations

from dataclasses import dataclass, field

from . import _lib as L

TILES = (8, 16, 32)
FWD_BATCHES = (64, 128, 256, 512)
BWD_BATCHES = (64, 128, 256)
SCHED_NAMES = {L.SCHED_GRID: "grid", L.SCHED_PERSISTENT: "persistent"}


@dataclass
class TuneResult:
    tile: int
    fwd_batch: int
    bwd_batch: int
    step_ms: float          # predicted summed kernel time of one step
    next_t: int             # Adam step to use next
    table: list = field(default_factory=list)  # one row per timed step
    scheduler: str = "grid"  # the blend kernels' tile scheduler picked
    blend_ms: dict = field(default_factory=dict)  # forward + backward blend time per scheduler


def _split(rep: dict):
    fwd = sum(ms for name, (_, ms) in rep.items() if name.startswith("blend_fwd"))
    bwd = sum(ms for name, (_, ms) in rep.items() if name.startswith("blend_bwd"))
    total = sum(ms for _, (_, ms) in rep.items())
    return fwd, bwd, total - fwd - bwd


def autotune_batch(ctx, batch, splats, views, targets, t: int, loss_kind: int = L.LOSS_L1,
                   backward_mode: int = L.BWD_RECOMPUTE, tiles=TILES, fwd_batches=FWD_BATCHES,
                   bwd_batches=BWD_BATCHES, **step_kw) -> TuneResult:
    """Pick (tile, forward batch, backward batch) for `batch` on this
    workload and leave the batch set to it.  `t` is the next Adam step."""
    rows = []
    best = None
    for tile in tiles:
        fwd_t, bwd_t, rest_t = {}, {}, []
        overflow = False
        for i in range(max(len(fwd_batches), len(bwd_batches))):
            fb, bb = fwd_batches[i % len(fwd_batches)], bwd_batches[i % len(bwd_batches)]
            batch.set_lds_batch(fb, bb)
            batch.train_step(splats, views, targets, t, tile, loss_kind, backward_mode, **step_kw)
            t += 1
            ctx.sync()
            ctx.prof_reset()
            ctx.prof_enable(True)
            try:
                batch.train_step(splats, views, targets, t, tile, loss_kind, backward_mode, **step_kw)
                t += 1
                ctx.sync()
                rep = ctx.prof_report()
            finally:
                ctx.prof_enable(False)
            if batch.stats()[2]:
                overflow = True  # the pair capacity cannot hold this tile size
                break
            f, b, r = _split(rep)
            fwd_t[fb] = min(f, fwd_t.get(fb, f))
            bwd_t[bb] = min(b, bwd_t.get(bb, b))
            rest_t.append(r)
            rows.append(dict(tile=tile, fwd_batch=fb, bwd_batch=bb, fwd_ms=round(f, 4), bwd_ms=round(b, 4),
                             rest_ms=round(r, 4)))
        if overflow or not fwd_t:
            rows.append(dict(tile=tile, skipped="pair capacity"))
            continue
        fb = min(fwd_t, key=fwd_t.get)
        bb = min(bwd_t, key=bwd_t.get)
        total = min(rest_t) + fwd_t[fb] + bwd_t[bb]
        if best is None or total < best[0]:
            best = (total, tile, fb, bb, min(rest_t))
    if best is None:
        raise RuntimeError("no tile size fits the batch's pair capacity")
    total, tile, fb, bb, rest = best
    batch.set_lds_batch(fb, bb)
    # the tile scheduler at the chosen configuration: one workgroup per tile
    # against resident workgroups taking tiles from a counter
    blend = {}
    for sched in (L.SCHED_GRID, L.SCHED_PERSISTENT):
        batch.set_scheduler(sched)
        best_ms = None
        for _ in range(2):
            batch.train_step(splats, views, targets, t, tile, loss_kind, backward_mode, **step_kw)
            t += 1
            ctx.sync()
            ctx.prof_reset()
            ctx.prof_enable(True)
            try:
                batch.train_step(splats, views, targets, t, tile, loss_kind, backward_mode, **step_kw)
                t += 1
                ctx.sync()
                rep = ctx.prof_report()
            finally:
                ctx.prof_enable(False)
            f, b, _ = _split(rep)
            best_ms = (f, b) if best_ms is None or f + b < sum(best_ms) else best_ms
        blend[sched] = best_ms
        rows.append(dict(tile=tile, fwd_batch=fb, bwd_batch=bb, scheduler=SCHED_NAMES[sched],
                         fwd_ms=round(best_ms[0], 4), bwd_ms=round(best_ms[1], 4)))
    sched = min(blend, key=lambda k: sum(blend[k]))
    batch.set_scheduler(sched)
    total = rest + sum(blend[sched])
    return TuneResult(tile=tile, fwd_batch=fb, bwd_batch=bb, step_ms=total, next_t=t, table=rows,
                      scheduler=SCHED_NAMES[sched], blend_ms={SCHED_NAMES[k]: round(sum(v), 4)
                                                              for k, v in blend.items()})
