"""Single-view training step and loop over the C ABI (the sequence of
train(), crates/nano-optimize/src/train.rs:155-250): project -> bin ->
composite -> dL/dC -> composite_backward -> project_backward ->
accumulate_grad + Adam + constraints.  The loss is computed on the device
(the reference reads the image back and computes it on the CPU)."""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .api import (AdamConfig, AdamGpu, CameraUniform, GpuCtx, GpuSplatBuffer, Rasterizer, TileBinner,
                  TilingParams, combined_loss_grad, loss_grad)
from .synth import HostSplats


@dataclass
class View:
    eye: tuple
    width: int
    height: int
    target_image: np.ndarray | None = None   # (H, W, 4) float32
    target: tuple = (0.0, 0.0, 0.0)
    up: tuple = (0.0, 1.0, 0.0)
    fov_y: float = 1.0471975511965976


@dataclass
class StepResult:
    loss: float
    total_pairs: int


@dataclass
class Trainer:
    ctx: GpuCtx
    splats: GpuSplatBuffer
    views: list
    depth_max: float
    tile_size: int = 16
    loss_kind: int = L.LOSS_L1
    backward_mode: int = L.BWD_REF_PARITY
    adam_pos: AdamConfig = field(default_factory=AdamConfig)
    adam_attr: AdamConfig = field(default_factory=AdamConfig)
    t: int = 0
    ssim_lambda: float = 0.2   # LOSS_SSIM only: (1 - lambda) MSE + lambda (1 - SSIM)

    def __post_init__(self):
        self.r = Rasterizer()
        self.binner = TileBinner()
        self.adam = AdamGpu()
        self._bufs = {}
        self._targets = {}

    def _buf(self, key, nbytes):
        b = self._bufs.get(key)
        if b is None or b.nbytes < nbytes:
            b = self.ctx.storage_buffer_zeroed(nbytes)
            self._bufs[key] = b
        return b

    def camera(self, v: View):
        return CameraUniform.from_pose(v.eye, v.target, v.up, v.fov_y, v.width, v.height, self.splats.n)

    def render(self, v: View):
        """Forward only: returns (projected, bins, image buffer, camera, params)."""
        n = self.splats.n
        cam = self.camera(v)
        proj = self._buf("proj", max(n, 1) * 48)
        self.r.project(self.ctx, self.splats, cam, proj)
        params = TilingParams(v.width, v.height, self.tile_size, self.depth_max)
        bins = self.binner.bin(self.ctx, proj, n, params)
        img = self._buf(("img", v.width, v.height), v.width * v.height * 16)
        if self.backward_mode == L.BWD_RECOMPUTE:
            ft = self._buf(("ft", v.width, v.height), v.width * v.height * 4)
            li = self._buf(("li", v.width, v.height), v.width * v.height * 4)
            self.r.composite_ex(self.ctx, proj, bins, img, params, ft, li)
        else:
            self.r.composite(self.ctx, proj, bins, img, params)
        return proj, bins, img, cam, params

    def step(self, iteration: int) -> StepResult:
        v = self.views[iteration % len(self.views)]
        n = self.splats.n
        proj, bins, img, cam, params = self.render(v)
        key = id(v)
        tgt = self._targets.get(key)
        if tgt is None:
            tgt = self.ctx.storage_buffer(np.ascontiguousarray(v.target_image, np.float32))
            self._targets[key] = tgt
        dl = self._buf(("dl", v.width, v.height), v.width * v.height * 16)
        if self.loss_kind == L.LOSS_SSIM:
            loss = combined_loss_grad(self.ctx, img, tgt, dl, v.width, v.height, self.ssim_lambda)[0]
        else:
            loss = loss_grad(self.ctx, img, tgt, dl, v.width, v.height, self.loss_kind)
        pg = self._buf("pgrad", max(n, 1) * 48)
        pg.zero()
        if self.backward_mode == L.BWD_RECOMPUTE:
            self.r.composite_backward_recompute(self.ctx, proj, bins, self._bufs[("ft", v.width, v.height)],
                                                self._bufs[("li", v.width, v.height)], dl, pg, params)
        else:
            self.r.composite_backward(self.ctx, proj, bins, img, dl, pg, params)
        self.r.project_backward(self.ctx, self.splats, pg, cam, self.backward_mode)
        self.t += 1
        self.adam.step_all_fused(self.ctx, self.splats, self.t, self.adam_pos, self.adam_attr)
        total = bins.total_pairs
        bins.free()
        return StepResult(loss, total)


# ---------------------------------------------------------------------------
# train() driver (train.rs:65-401)
# ---------------------------------------------------------------------------
@dataclass
class TrainConfig:
    """TrainConfig (src/main.rs:418-437) and the cadence constants of
    train.rs:103-114."""

    iterations: int = 1000
    max_splats: int = 1_000_000
    tile_size: int = 16
    loss_kind: int = L.LOSS_L1
    ssim_lambda: float = 0.2           # LOSS_SSIM: TrainConfig::ssim_lambda (train.rs:48-52)
    backward_mode: int = L.BWD_REF_PARITY
    adam_pos: AdamConfig = field(default_factory=AdamConfig)
    adam_attr: AdamConfig = field(default_factory=AdamConfig)
    prune_every: int = 100
    prune_warmup: int = 50
    prune_opacity_logit: float = -5.3
    densify_every: int = 200
    densify_warmup: int = 100
    densify_grad_threshold: float = 2e-4
    split_scale_threshold: float = -2.3
    diagnostics_every: int = 100
    checkpoint_every: int = 0          # 0: no periodic checkpoints
    checkpoint_path: str | None = None  # base path: <base>.ply + <base>.npz


@dataclass
class TrainState:
    iteration: int = 0   # next iteration to run
    t: int = 0           # Adam step (adam_pos.t)
    grad_count: int = 0  # steps accumulated into grad_acc since the last densify


def save_checkpoint(base, splats: GpuSplatBuffer, state: TrainState) -> None:
    """Parameters as a 3DGS PLY (base.ply) and the optimiser state the
    reference does not persist (base.npz): both Adam moments, grad_acc,
    the Adam step, grad_count and the next iteration."""
    from .api import write_ply

    write_ply(f"{base}.ply", splats.readback())
    m, v = splats.moments()
    np.savez(f"{base}.npz", m=m.flat(), v=v.flat(), grad_acc=splats.grad_acc(),
             state=np.array([state.iteration, state.t, state.grad_count], np.int64))


def load_checkpoint(ctx: GpuCtx, base, capacity: int):
    """Inverse of save_checkpoint: (GpuSplatBuffer, TrainState)."""
    from .api import read_ply

    host = read_ply(f"{base}.ply")
    d = np.load(f"{base}.npz")
    g = GpuSplatBuffer.upload(ctx, host, capacity=max(capacity, host.n))
    g.set_moments(HostSplats.from_flat(d["m"], host.n), HostSplats.from_flat(d["v"], host.n))
    g.set_grad_acc(d["grad_acc"])
    it, t, gc = (int(x) for x in d["state"])
    return g, TrainState(iteration=it, t=t, grad_count=gc)


def _norm(a) -> float:
    return float(np.sqrt(np.sum(np.asarray(a, np.float64) ** 2)))


def train(ctx: GpuCtx, splats, views, depth_max: float, cfg: TrainConfig | None = None, on_iter=None,
          resume: str | None = None, log=print, on_snapshot=None, snapshot_every: int = 1, on_event=None):
    """The training loop of train.rs:155-401 on the device: view iter %
    len(views) each iteration, one fused Adam step, prune every
    `prune_every` iterations from `prune_warmup`, densify every
    `densify_every` from `densify_warmup` (then grad_acc = 0 and
    grad_count = 0), gradient-norm diagnostics every `diagnostics_every`
    and at the end, optional checkpoints.  `splats` is a HostSplats (fresh
    start) or ignored when `resume` names a checkpoint.  on_iter(iter, loss,
    n) is called after every iteration; on_snapshot(iter, HostSplats, loss),
    the reference's on_iter with the splats (train.rs:55-64), every
    `snapshot_every` iterations and after the last.  on_event(iter,
    stage, scene) is called with stage "before" and "after" around each
    prune / densify checkpoint (train.rs:259-362), the device scene in
    hand.  Returns the final HostSplats."""
    cfg = cfg or TrainConfig()
    cap = max(cfg.max_splats, splats.n if splats is not None else 0) + cfg.max_splats
    if resume:
        g, st = load_checkpoint(ctx, resume, cap)
    else:
        g, st = GpuSplatBuffer.upload(ctx, splats, capacity=cap), TrainState()
    tr = Trainer(ctx, g, list(views), depth_max, cfg.tile_size, cfg.loss_kind, cfg.backward_mode, cfg.adam_pos,
                 cfg.adam_attr, st.t, cfg.ssim_lambda)
    for it in range(st.iteration, cfg.iterations):
        res = tr.step(it)
        st.t = tr.t
        st.grad_count += 1
        do_prune = it >= cfg.prune_warmup and it % cfg.prune_every == 0
        do_densify = it >= cfg.densify_warmup and it % cfg.densify_every == 0
        if on_event is not None and (do_prune or do_densify):
            on_event(it, "before", g)
        if do_prune:
            before = g.n
            removed = g.prune(cfg.prune_opacity_logit)
            if removed:
                log(f"[train] iter {it}: pruned {removed} low-opacity splats ({before}->{g.n})")
        if do_densify and st.grad_count > 0:
            cloned, split, parents = g.densify(st.grad_count, cfg.max_splats, it, cfg.densify_grad_threshold,
                                               cfg.split_scale_threshold)
            if cloned + split + parents:
                log(f"[train] iter {it}: densify cloned {cloned} split {split} (removed {parents} parents) "
                    f"-> {g.n} splats")
            g.zero_grad_acc()
            st.grad_count = 0
        if on_event is not None and (do_prune or do_densify):
            on_event(it, "after", g)
        st.iteration = it + 1
        if on_iter is not None:
            on_iter(it, res.loss, g.n)
        if on_snapshot is not None and ((it + 1) % max(1, snapshot_every) == 0 or it + 1 == cfg.iterations):
            on_snapshot(it, g.readback(), res.loss)
        if cfg.diagnostics_every and (it % cfg.diagnostics_every == 0 or it + 1 == cfg.iterations):
            gr = g.grads()
            log(f"[train] iter {it:>5} | view {it % len(tr.views)}/{len(tr.views)} | loss {res.loss:.4f} | "
                f"grad |pos|={_norm(gr.pos):.2e} |op|={_norm(gr.logit):.2e} |dc|={_norm(gr.sh_dc):.2e} "
                f"|scale|={_norm(gr.log_scale):.2e} | {g.n} splats")
        if cfg.checkpoint_every and cfg.checkpoint_path and (it + 1) % cfg.checkpoint_every == 0:
            save_checkpoint(cfg.checkpoint_path, g, st)
    out = g.readback()
    g.free()
    return out
