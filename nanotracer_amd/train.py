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
                  TilingParams, loss_grad)
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
