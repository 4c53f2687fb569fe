"""Host-side mirror of the public surface of crates/nano-optimize
(src/lib.rs:21-45) over the C ABI.  Names follow the reference: GpuCtx for
WgpuCtx, GpuSplatBuffer, Rasterizer, TileBinner, PrefixScan, RadixSort,
AdamGpu, CameraUniform.from_pose, TilingParams.  Every call runs on the
gfx950 kernels of libnts.so; there is no host fallback.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .synth import HostSplats

PROJECTED_BYTES = 48        # ProjectedSplat, raster.rs:27-39
PROJECTED_GRAD_BYTES = 48   # ProjectedGrad, raster.rs:46-57


class Buffer:
    """A device allocation owned by the library (wgpu::Buffer stand-in)."""

    def __init__(self, ctx: "GpuCtx", nbytes: int):
        self.ctx = ctx
        h = C.c_void_p()
        L.check(L.lib().nts_buffer_create(ctx.h, nbytes, C.byref(h)))
        self.h = h
        self.nbytes = nbytes

    def upload(self, a: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(a)
        L.check(L.lib().nts_buffer_upload(self.ctx.h, self.h, offset, L.vptr(a), a.nbytes))

    def download(self, dtype, count: int, offset: int = 0) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        L.check(L.lib().nts_buffer_download(self.ctx.h, self.h, offset, L.vptr(out), out.nbytes))
        return out

    def zero(self) -> None:
        L.check(L.lib().nts_buffer_zero(self.ctx.h, self.h))

    def free(self) -> None:
        if self.h:
            L.lib().nts_buffer_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                self.free()
        except Exception:
            pass


# prune / densify constants (train.rs:103-114)
PRUNE_OPACITY_LOGIT = -5.3
DENSIFY_GRAD_THRESHOLD = 2e-4
SPLIT_SCALE_THRESHOLD = -2.3


class GpuCtx:
    """WgpuCtx (gpu.rs:21-152) over a HIP device and stream."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        h = C.c_void_p()
        L.check(L.lib().nts_ctx_create(device, hbm_budget, C.byref(h)))
        self.h = h
        self.device = device

    def close(self) -> None:
        if self.h:
            L.lib().nts_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def storage_buffer(self, a: np.ndarray) -> Buffer:
        a = np.ascontiguousarray(a)
        b = Buffer(self, max(a.nbytes, 4))
        b.upload(a)
        return b

    def storage_buffer_zeroed(self, nbytes: int) -> Buffer:
        return Buffer(self, max(int(nbytes), 4))

    def readback(self, buf: Buffer, dtype, count: int) -> np.ndarray:
        return buf.download(dtype, count)

    def sync(self) -> None:
        L.check(L.lib().nts_sync(self.h))

    # per-kernel HIP-event timing
    def prof_enable(self, on: bool = True) -> None:
        L.check(L.lib().nts_prof_enable(self.h, int(on)))

    def prof_reset(self) -> None:
        L.check(L.lib().nts_prof_reset(self.h))

    def prof_query(self, prefix: str):
        ms, n = C.c_double(), C.c_uint64()
        L.check(L.lib().nts_prof_query(self.h, prefix.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def probe_peaks(self, copy_bytes: int = 4 << 30) -> dict:
        """Measured HBM copy bandwidth (GB/s), LDS 128-bit read bandwidth
        (TB/s) and shader clock (GHz) of this device."""
        h, l, c = C.c_double(), C.c_double(), C.c_double()
        L.check(L.lib().nts_probe_peaks(self.h, int(copy_bytes), C.byref(h), C.byref(l), C.byref(c)))
        return dict(hbm_copy_gbs=h.value, lds_read_tbs=l.value, clock_ghz=c.value)

    def probe_valu(self) -> dict:
        """Measured FP32 FMA ceiling of this device: TFLOP/s and wave
        instructions per SIMD per shader cycle."""
        t, r = C.c_double(), C.c_double()
        L.check(L.lib().nts_probe_valu(self.h, C.byref(t), C.byref(r)))
        return dict(fma_tflops=t.value, fma_per_simd_cycle=r.value)

    def probe_blend_mix(self, tile_size: int = 32, bwd_batch: int = 128) -> float:
        """(splat, wave) evaluations per second of the recompute backward's
        own per-evaluation code at its occupancy, with no memory traffic or
        barriers: the ceiling of its arithmetic."""
        r = C.c_double()
        L.check(L.lib().nts_probe_blend_mix(self.h, int(tile_size), int(bwd_batch), C.byref(r)))
        return r.value

    def prof_report(self) -> dict:
        buf = C.create_string_buffer(1 << 16)
        L.check(L.lib().nts_prof_report(self.h, buf, len(buf)))
        out = {}
        for line in buf.value.decode().splitlines():
            name, n, ms = line.rsplit(" ", 2)
            out[name] = (int(n), float(ms))
        return out


def _host_struct(s: HostSplats) -> L.SplatsHost:
    hs = L.SplatsHost()
    hs.n = s.n
    hs.pos, hs.rot, hs.log_scale = L.fptr(s.pos), L.fptr(s.rot), L.fptr(s.log_scale)
    hs.logit, hs.sh_dc, hs.sh_rest = L.fptr(s.logit), L.fptr(s.sh_dc), L.fptr(s.sh_rest)
    return hs


def read_ply(path) -> HostSplats:
    """read_ply (nano-splat/src/ply.rs:116-134) through the library."""
    n = C.c_uint32()
    L.check(L.lib().nts_ply_count(str(path).encode(), C.byref(n)))
    out = HostSplats.zeros(n.value)
    hs = _host_struct(out)
    L.check(L.lib().nts_ply_read(str(path).encode(), C.byref(hs)))
    return out


def write_ply(path, splats: HostSplats) -> None:
    """write_ply (nano-splat/src/ply.rs:17-28) through the library."""
    L.check(L.lib().nts_ply_write(str(path).encode(), C.byref(_host_struct(splats))))


class GpuSplatBuffer:
    """GpuSplatBuffer + GradSplatBuffers + AdamMomentBuffers + grad_acc
    (splat_gpu.rs:36-384) as one device scene of attribute planes.  With
    trainable=False only the parameters are held (a viewer's scene), and the
    calls that need gradients, moments or grad_acc raise NTS_E_STATE."""

    def __init__(self, ctx: GpuCtx, capacity: int, trainable: bool = True):
        self.ctx = ctx
        h = C.c_void_p()
        flags = L.SCENE_TRAINABLE if trainable else 0
        L.check(L.lib().nts_scene_create(ctx.h, max(1, int(capacity)), flags, C.byref(h)))
        self.h = h
        self.capacity = capacity
        self.trainable = trainable

    @classmethod
    def upload(cls, ctx: GpuCtx, splats: HostSplats, capacity: int | None = None,
               trainable: bool = True) -> "GpuSplatBuffer":
        """GpuSplatBuffer::upload (splat_gpu.rs:94-120): a new scene, fresh
        optimiser state."""
        g = cls(ctx, capacity or max(1, splats.n), trainable)
        g.upload_from(splats)
        return g

    @property
    def n(self) -> int:
        return int(L.lib().nts_scene_count(self.h))

    def upload_from(self, splats: HostSplats) -> None:
        """Set the splat count and parameters of this scene; gradients,
        moments and grad_acc restart at zero (the state of a fresh upload)."""
        L.check(L.lib().nts_scene_upload(self.ctx.h, self.h, C.byref(_host_struct(splats))))

    def sync_from(self, splats: HostSplats) -> None:
        """GpuSplatBuffer::sync_from (splat_gpu.rs:124-150): parameters of
        the same splat count rewritten in place; moments, gradients and
        grad_acc are kept."""
        L.check(L.lib().nts_scene_sync_params(self.ctx.h, self.h, C.byref(_host_struct(splats))))

    def readback(self) -> HostSplats:
        out = HostSplats.zeros(self.n)
        L.check(L.lib().nts_scene_download(self.ctx.h, self.h, C.byref(_host_struct(out))))
        return out

    def grads(self) -> HostSplats:
        out = HostSplats.zeros(self.n)
        L.check(L.lib().nts_scene_download_grads(self.ctx.h, self.h, C.byref(_host_struct(out))))
        return out

    def set_grads(self, g: HostSplats) -> None:
        L.check(L.lib().nts_scene_upload_grads(self.ctx.h, self.h, C.byref(_host_struct(g))))

    def zero_grads(self) -> None:
        L.check(L.lib().nts_scene_zero_grads(self.ctx.h, self.h))

    def moments(self):
        m, v = HostSplats.zeros(self.n), HostSplats.zeros(self.n)
        L.check(L.lib().nts_scene_download_moments(self.ctx.h, self.h, C.byref(_host_struct(m)),
                                                   C.byref(_host_struct(v))))
        return m, v

    def set_moments(self, m: HostSplats, v: HostSplats) -> None:
        L.check(L.lib().nts_scene_upload_moments(self.ctx.h, self.h, C.byref(_host_struct(m)),
                                                 C.byref(_host_struct(v))))

    def grad_acc(self) -> np.ndarray:
        out = np.zeros(self.n, np.float32)
        L.check(L.lib().nts_scene_download_grad_acc(self.ctx.h, self.h, L.vptr(out)))
        return out

    def set_grad_acc(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.float32)
        L.check(L.lib().nts_scene_upload_grad_acc(self.ctx.h, self.h, L.vptr(a)))

    def zero_grad_acc(self) -> None:
        """AdamGpu::zero_grad_acc (adam_gpu.rs:492-541)."""
        L.check(L.lib().nts_scene_zero_grad_acc(self.ctx.h, self.h))

    def prune(self, logit_threshold: float = PRUNE_OPACITY_LOGIT) -> int:
        """prune_low_opacity (train.rs:572-602) on the device; returns the
        number removed."""
        removed = C.c_uint32()
        L.check(L.lib().nts_prune(self.ctx.h, self.h, C.c_float(logit_threshold), C.byref(removed)))
        return int(removed.value)

    def densify(self, grad_count: int, max_splats: int, iteration: int,
                grad_threshold: float = DENSIFY_GRAD_THRESHOLD,
                split_log_scale: float = SPLIT_SCALE_THRESHOLD):
        """densify (train.rs:414-549) on the device; returns (cloned, split
        children, parents removed).  The scene needs room for the clones and
        both children of every split."""
        counts = (C.c_uint32 * 3)()
        L.check(L.lib().nts_densify(self.ctx.h, self.h, int(grad_count), C.c_float(grad_threshold),
                                    C.c_float(split_log_scale), int(max_splats), int(iteration), counts))
        return tuple(int(c) for c in counts)

    def free(self) -> None:
        if self.h:
            L.lib().nts_scene_destroy(self.ctx.h, self.h)
            self.h = None


class CameraUniform:
    """CameraUniform::from_pose (raster.rs:79-99)."""

    @staticmethod
    def from_pose(eye, target, up, fov_y: float, width: int, height: int, n_splats: int) -> L.Camera:
        f3 = lambda v: (C.c_float * 3)(*[float(x) for x in v])
        cam = L.Camera()
        L.check(L.lib().nts_camera_from_pose(f3(eye), f3(target), f3(up), float(fov_y), int(width), int(height),
                                             int(n_splats), C.byref(cam)))
        return cam


@dataclass
class TilingParams:
    """TilingParams (tile_binner.rs:21-42)."""

    width: int
    height: int
    tile_size: int = 16
    depth_max: float = 100.0

    def tiles_x(self) -> int:
        return -(-self.width // self.tile_size)

    def tiles_y(self) -> int:
        return -(-self.height // self.tile_size)

    def num_tiles(self) -> int:
        return self.tiles_x() * self.tiles_y()

    def c(self) -> L.Tiling:
        return L.Tiling(self.width, self.height, self.tile_size, self.depth_max)


class TileBinningResult:
    """TileBinningResult (tile_binner.rs:77-83); device-resident."""

    def __init__(self, ctx: GpuCtx, h):
        self.ctx = ctx
        self.h = h
        tp, nt = C.c_uint64(), C.c_uint32()
        L.check(L.lib().nts_bins_info(h, C.byref(tp), C.byref(nt)))
        self.total_pairs = tp.value
        self.num_tiles = nt.value

    def sorted_keys(self) -> np.ndarray:
        out = np.zeros(self.total_pairs, np.uint32)
        L.check(L.lib().nts_bins_download_keys_u32(self.ctx.h, self.h, L.vptr(out)))
        return out

    def sorted_keys_u64(self) -> np.ndarray:
        out = np.zeros(self.total_pairs, np.uint64)
        L.check(L.lib().nts_bins_download_keys_u64(self.ctx.h, self.h, L.vptr(out)))
        return out

    def sorted_payloads(self) -> np.ndarray:
        out = np.zeros(self.total_pairs, np.uint32)
        L.check(L.lib().nts_bins_download_payloads(self.ctx.h, self.h, L.vptr(out)))
        return out

    def tile_ranges(self) -> np.ndarray:
        out = np.zeros(2 * self.num_tiles, np.uint32)
        L.check(L.lib().nts_bins_download_ranges(self.ctx.h, self.h, L.vptr(out)))
        return out

    def free(self) -> None:
        if self.h:
            L.lib().nts_bins_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                self.free()
        except Exception:
            pass


class TileBinner:
    """TileBinner::bin (tile_binner.rs:181-382)."""

    def __init__(self, ctx: GpuCtx | None = None):
        pass

    def bin(self, ctx: GpuCtx, projected: Buffer, n_splats: int, params: TilingParams) -> TileBinningResult:
        h = C.c_void_p()
        L.check(L.lib().nts_bin(ctx.h, projected.h, int(n_splats), C.byref(params.c()), C.byref(h)))
        return TileBinningResult(ctx, h)


class PrefixScan:
    """PrefixScan::scan (prefix_scan.rs:127-186): exclusive u32, in place."""

    def __init__(self, ctx: GpuCtx | None = None):
        pass

    def scan(self, ctx: GpuCtx, buf: Buffer, n: int) -> None:
        L.check(L.lib().nts_scan_u32(ctx.h, buf.h, int(n)))


class RadixSort:
    """RadixSort::sort (radix_sort.rs:144-303): stable, in place."""

    def __init__(self, ctx: GpuCtx | None = None):
        pass

    def sort(self, ctx: GpuCtx, keys: Buffer, vals: Buffer, n: int, key_bits: int = 32) -> None:
        L.check(L.lib().nts_sort_pairs_u32(ctx.h, keys.h, vals.h, int(n), int(key_bits)))

    def sort_u64(self, ctx: GpuCtx, keys: Buffer, vals: Buffer, n: int, key_bits: int = 64) -> None:
        L.check(L.lib().nts_sort_pairs_u64(ctx.h, keys.h, vals.h, int(n), int(key_bits)))

    def sort_noalloc(self, ctx: GpuCtx, keys: Buffer, vals: Buffer, n: int, key_bits: int = 32) -> None:
        """sort() with no device allocation: the library's static per-GPU
        workspace, at most SORT_NOALLOC_MAX_PAIRS pairs; concurrent calls on
        one GPU race."""
        L.check(L.lib().nts_sort_pairs_u32_noalloc(ctx.h, keys.h, vals.h, int(n), int(key_bits)))

    def sort_tiles(self, ctx: GpuCtx, keys: Buffer, vals: Buffer, n: int, tile_bits: int, ranges: Buffer,
                   num_tiles: int, key_bytes: int = 4) -> None:
        """The binning sort: sort() over the low 16 + tile_bits bits of
        (tile << 16 | depth16) keys, plus ranges (2 u32 per tile, [begin,
        end)) of the tiles below num_tiles.  Keys are below
        2^(16 + tile_bits) or all ones (a dropped pair); with any all-ones
        key, tile ids stay below 2^tile_bits - 1."""
        L.check(L.lib().nts_sort_tile_pairs(ctx.h, keys.h, vals.h, int(n), int(key_bytes), int(tile_bits),
                                            ranges.h, int(num_tiles)))


class Rasterizer:
    """Rasterizer (raster.rs:114-804)."""

    def __init__(self, ctx: GpuCtx | None = None):
        pass

    def alloc_image(self, ctx: GpuCtx, width: int, height: int) -> Buffer:
        return ctx.storage_buffer_zeroed(width * height * 16)

    def alloc_projected(self, ctx: GpuCtx, n: int) -> Buffer:
        return ctx.storage_buffer_zeroed(max(n, 1) * PROJECTED_BYTES)

    def alloc_projected_grad(self, ctx: GpuCtx, n: int) -> Buffer:
        return ctx.storage_buffer_zeroed(max(n, 1) * PROJECTED_GRAD_BYTES)

    def zero_projected_grad(self, ctx: GpuCtx, buf: Buffer, n: int | None = None) -> None:
        buf.zero()

    def project(self, ctx: GpuCtx, splats: GpuSplatBuffer, camera: L.Camera, projected: Buffer) -> None:
        L.check(L.lib().nts_project(ctx.h, splats.h, C.byref(camera), projected.h))

    def composite(self, ctx: GpuCtx, projected: Buffer, bins: TileBinningResult, image: Buffer,
                  params: TilingParams) -> None:
        L.check(L.lib().nts_composite(ctx.h, projected.h, bins.h, image.h, C.byref(params.c())))

    def composite_ex(self, ctx: GpuCtx, projected: Buffer, bins: TileBinningResult, image: Buffer,
                     params: TilingParams, final_t: Buffer | None, last_idx: Buffer | None) -> None:
        """composite that also saves the final transmittance and last contributor per pixel."""
        L.check(L.lib().nts_composite_ex(ctx.h, projected.h, bins.h, image.h, C.byref(params.c()),
                                         final_t.h if final_t else None, last_idx.h if last_idx else None))

    def composite_backward(self, ctx: GpuCtx, projected: Buffer, bins: TileBinningResult, forward_out: Buffer,
                           dl_dc: Buffer, projected_grad: Buffer, params: TilingParams) -> None:
        L.check(L.lib().nts_composite_backward(ctx.h, projected.h, bins.h, forward_out.h, dl_dc.h,
                                               projected_grad.h, C.byref(params.c())))

    def composite_backward_recompute(self, ctx: GpuCtx, projected: Buffer, bins: TileBinningResult,
                                     final_t: Buffer, last_idx: Buffer, dl_dc: Buffer, projected_grad: Buffer,
                                     params: TilingParams) -> None:
        L.check(L.lib().nts_composite_backward_ex(ctx.h, projected.h, bins.h, None, final_t.h, last_idx.h, dl_dc.h,
                                                  projected_grad.h, C.byref(params.c()), L.BWD_RECOMPUTE))

    def project_backward(self, ctx: GpuCtx, splats: GpuSplatBuffer, projected_grad: Buffer,
                         camera: L.Camera, mode: int = L.BWD_REF_PARITY) -> None:
        L.check(L.lib().nts_project_backward_ex(ctx.h, splats.h, projected_grad.h, C.byref(camera), int(mode)))

    def tonemap(self, ctx: GpuCtx, composite: Buffer, target: Buffer, width: int, height: int) -> None:
        L.check(L.lib().nts_tonemap(ctx.h, composite.h, target.h, width, height))


def loss_grad(ctx: GpuCtx, image: Buffer, target: Buffer, dl_dc: Buffer, width: int, height: int,
              kind: int = L.LOSS_L1) -> float:
    """dL/dC on the device (the reference computes it on the CPU, train.rs:180-196)."""
    out = C.c_double()
    L.check(L.lib().nts_loss_grad(ctx.h, image.h, target.h, dl_dc.h, width, height, kind, C.byref(out)))
    return out.value


def combined_loss_grad(ctx: GpuCtx, image: Buffer, target: Buffer, dl_dc: Buffer, width: int, height: int,
                       lam: float) -> tuple:
    """(1 - lam) MSE + lam (1 - SSIM) and its dL/dC on the device
    (combined_loss_grad, loss.rs:390-415): returns (total, mse, ssim)."""
    t, m, s = C.c_double(), C.c_double(), C.c_double()
    L.check(L.lib().nts_loss_grad_ssim(ctx.h, image.h, target.h, dl_dc.h, width, height, float(lam), C.byref(t),
                                       C.byref(m), C.byref(s)))
    return t.value, m.value, s.value


@dataclass
class AdamConfig:
    """AdamConfig (adam.rs:22-39); defaults as the reference."""

    lr: float = 1.6e-4
    beta1: float = 0.9
    beta2: float = 0.999
    eps: float = 1e-15

    def c(self) -> L.AdamCfg:
        return L.AdamCfg(self.lr, self.beta1, self.beta2, self.eps)


LOG_SCALE_MIN, LOG_SCALE_MAX, OPACITY_LOGIT_MAX = -10.0, 5.0, 10.0  # train.rs:37-39


class AdamGpu:
    """AdamGpu (adam_gpu.rs)."""

    def __init__(self, ctx: GpuCtx | None = None):
        pass

    def step(self, ctx: GpuCtx, params: Buffer, grads: Buffer, m: Buffer, v: Buffer, n: int, lr: float, t: int,
             cfg: AdamConfig) -> None:
        c = AdamConfig(lr, cfg.beta1, cfg.beta2, cfg.eps).c()
        L.check(L.lib().nts_adam_step_flat(ctx.h, params.h, grads.h, m.h, v.h, int(n), C.byref(c), int(t)))

    def step_all(self, ctx: GpuCtx, splats: GpuSplatBuffer, t: int, cfg_pos: AdamConfig,
                 cfg_attr: AdamConfig) -> None:
        """AdamGpu::step_all (adam_gpu.rs:334-427): Adam on every attribute
        of the scene from its gradients."""
        L.check(L.lib().nts_adam_step_all(ctx.h, splats.h, C.byref(cfg_pos.c()), C.byref(cfg_attr.c()), int(t)))

    def apply_constraints(self, ctx: GpuCtx, splats: GpuSplatBuffer, log_scale_min: float = LOG_SCALE_MIN,
                          log_scale_max: float = LOG_SCALE_MAX,
                          opacity_logit_max: float = OPACITY_LOGIT_MAX) -> None:
        """AdamGpu::apply_constraints (adam_gpu.rs:431-489)."""
        cons = L.Constraints(log_scale_min, log_scale_max, opacity_logit_max)
        L.check(L.lib().nts_apply_constraints(ctx.h, splats.h, C.byref(cons)))

    def accumulate_grad(self, ctx: GpuCtx, splats: GpuSplatBuffer) -> None:
        """AdamGpu::accumulate_grad (adam_gpu.rs:492-541): grad_acc +=
        |dpos|_1 of the scene's position gradient."""
        L.check(L.lib().nts_accumulate_grad(ctx.h, splats.h))

    def step_all_fused(self, ctx: GpuCtx, splats: GpuSplatBuffer, t: int, cfg_pos: AdamConfig,
                       cfg_attr: AdamConfig, log_scale_min: float = LOG_SCALE_MIN,
                       log_scale_max: float = LOG_SCALE_MAX, opacity_logit_max: float = OPACITY_LOGIT_MAX) -> None:
        """accumulate_grad + step_all + apply_constraints in one kernel."""
        cons = L.Constraints(log_scale_min, log_scale_max, opacity_logit_max)
        L.check(L.lib().nts_adam_step(ctx.h, splats.h, C.byref(cfg_pos.c()), C.byref(cfg_attr.c()), int(t),
                                      C.byref(cons)))


@dataclass
class ViewSpec:
    """One view of a batch (nts_view_desc)."""

    camera: L.Camera
    width: int
    height: int
    depth_max: float

    def c(self) -> L.ViewDesc:
        return L.ViewDesc(self.camera, float(self.depth_max), int(self.width), int(self.height), 0)


def views_array(views):
    arr = (L.ViewDesc * len(views))()
    for k, v in enumerate(views):
        arr[k] = v.c()
    return arr


class ViewBatch:
    """Workspace of the batched multi-view step (nts_batch): all views of an
    iteration go through each stage in one launch; optional hipGraph replay."""

    def __init__(self, ctx: GpuCtx, max_views: int, max_splats: int, max_pixels: int, pair_capacity: int):
        self.ctx = ctx
        h = C.c_void_p()
        L.check(L.lib().nts_batch_create(ctx.h, int(max_views), int(max_splats), int(max_pixels), int(pair_capacity),
                                         C.byref(h)))
        self.h = h
        self.sizes = []

    def free(self) -> None:
        if self.h:
            L.lib().nts_batch_destroy(self.ctx.h, self.h)
            self.h = None

    def train_step(self, splats: GpuSplatBuffer, views, targets: Buffer, t: int, tile_size: int = 16,
                   loss_kind: int = L.LOSS_L1, backward_mode: int = L.BWD_RECOMPUTE,
                   cfg_pos: AdamConfig | None = None, cfg_attr: AdamConfig | None = None, graph: bool = False,
                   write_grads: bool = False, log_scale_min: float = LOG_SCALE_MIN,
                   log_scale_max: float = LOG_SCALE_MAX, opacity_logit_max: float = OPACITY_LOGIT_MAX) -> None:
        cfg_pos = cfg_pos or AdamConfig()
        cfg_attr = cfg_attr or AdamConfig()
        arr = views_array(views)
        self.sizes = [(v.width, v.height) for v in views]
        flags = (L.STEP_GRAPH if graph else 0) | (L.STEP_WRITE_GRADS if write_grads else 0)
        cons = L.Constraints(log_scale_min, log_scale_max, opacity_logit_max)
        L.check(L.lib().nts_train_step_views(self.ctx.h, splats.h, self.h, arr, len(views), targets.h,
                                             int(tile_size), int(loss_kind), int(backward_mode),
                                             C.byref(cfg_pos.c()), C.byref(cfg_attr.c()), int(t), C.byref(cons),
                                             flags))

    def set_lds_batch(self, fwd_batch: int = 256, bwd_batch: int = 128) -> None:
        """Splats staged per LDS batch by the blend forward / backward."""
        L.check(L.lib().nts_batch_set_lds_batch(self.ctx.h, self.h, int(fwd_batch), int(bwd_batch)))

    def set_scheduler(self, scheduler: int = L.SCHED_GRID) -> None:
        """Tile scheduler of the blend kernels: SCHED_GRID (one workgroup per
        tile) or SCHED_PERSISTENT (resident workgroups take tiles from a
        device counter)."""
        L.check(L.lib().nts_batch_set_scheduler(self.ctx.h, self.h, int(scheduler)))

    def set_binning(self, binning: int = L.BIN_CULLED) -> None:
        """Binning rule of the batch's steps: BIN_CULLED (pairs only for the
        tiles a splat's alpha >= 1/255 region reaches; the default) or
        BIN_REFERENCE (TileBinner::bin's tile rectangle, tile_binner.rs:181-382:
        every view's keys, payloads and ranges are the reference's).  Images
        are the same under both."""
        L.check(L.lib().nts_batch_set_binning(self.ctx.h, self.h, int(binning)))

    def set_eval_counting(self, on: bool = True) -> None:
        """Run the blend kernels in their counting form (a measurement aid:
        slower, same results; grid scheduler only)."""
        L.check(L.lib().nts_batch_set_eval_counting(self.ctx.h, self.h, int(bool(on))))

    def eval_counts(self) -> dict:
        """The last step's counts while counting is on: (splat, wave)
        evaluations and blended / contributing (pixel, splat) pairs of the
        forward and the backward."""
        c = (C.c_uint64 * 4)()
        L.check(L.lib().nts_batch_eval_counts(self.ctx.h, self.h, c))
        return dict(fwd_wave_evals=c[0], fwd_blended_pairs=c[1], bwd_wave_evals=c[2], bwd_contributing_pairs=c[3])

    def set_ssim_lambda(self, lam: float = 0.2) -> None:
        """Blend factor of LOSS_SSIM steps: (1 - lam) MSE + lam (1 - SSIM)
        (train.rs:48-52); allocates the SSIM planes on first use."""
        L.check(L.lib().nts_batch_set_ssim_lambda(self.ctx.h, self.h, float(lam)))

    def render(self, splats: GpuSplatBuffer, views, tile_size: int = 16) -> None:
        arr = views_array(views)
        self.sizes = [(v.width, v.height) for v in views]
        L.check(L.lib().nts_render_views(self.ctx.h, splats.h, self.h, arr, len(views), int(tile_size)))

    def losses(self) -> np.ndarray:
        out = np.zeros(len(self.sizes), np.float64)
        L.check(L.lib().nts_batch_losses(self.ctx.h, self.h, L.vptr(out)))
        return out

    def stats(self):
        tp, vis, ov = C.c_uint64(), C.c_uint32(), C.c_uint32()
        L.check(L.lib().nts_batch_stats(self.ctx.h, self.h, C.byref(tp), C.byref(vis), C.byref(ov)))
        return tp.value, vis.value, ov.value

    def image(self, v: int) -> np.ndarray:
        w, h = self.sizes[v]
        out = np.zeros((h, w, 4), np.float32)
        L.check(L.lib().nts_batch_download_image(self.ctx.h, self.h, v, L.vptr(out)))
        return out

    def state(self, v: int):
        w, h = self.sizes[v]
        ft = np.zeros((h, w), np.float32)
        li = np.zeros((h, w), np.uint32)
        L.check(L.lib().nts_batch_download_state(self.ctx.h, self.h, v, L.vptr(ft), L.vptr(li)))
        return ft, li

    def projected_grad(self, v: int, n: int) -> np.ndarray:
        out = np.zeros((n, 12), np.float32)
        L.check(L.lib().nts_batch_download_projected_grad(self.ctx.h, self.h, v, n, L.vptr(out)))
        return out

    def bins(self, v: int, tile_size: int):
        cnt = C.c_uint64()
        L.check(L.lib().nts_batch_download_bins(self.ctx.h, self.h, v, C.byref(cnt), None, None, None))
        k = cnt.value
        w, h = self.sizes[v]
        ntiles = (-(-w // tile_size)) * (-(-h // tile_size))
        keys = np.zeros(max(k, 1), np.uint32)
        pay = np.zeros(max(k, 1), np.uint32)
        rng = np.zeros(2 * ntiles, np.uint32)
        L.check(L.lib().nts_batch_download_bins(self.ctx.h, self.h, v, C.byref(cnt), L.vptr(keys), L.vptr(pay),
                                                L.vptr(rng)))
        return keys[:k], pay[:k], rng


def device_memory(device: int = 0):
    """(free, total) bytes of a device (hipMemGetInfo)."""
    f, t = C.c_uint64(), C.c_uint64()
    L.check(L.lib().nts_device_memory(int(device), C.byref(f), C.byref(t)))
    return f.value, t.value


class RenderServer:
    """The render server (nts_server): scenes resident in HBM, requests
    batched as they arrive, `slots` batches in flight on their own streams.

    submit() returns a request id at once; the image lands in the array it
    returns once poll() reports the request done (or wait() returns)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0, slots: int = 2, max_views: int = 32,
                 max_pixels: int = 16 << 20, max_records: int = 16 << 20, pair_capacity: int = 120_000_000,
                 tile_size: int = 32, lds_batch: int = 512, slot_bytes: int = 0):
        """slot_bytes > 0: each slot's device workspace; a batch's working set
        must fit it (admission by memory), and 0 for max_records or
        pair_capacity means no cap.  slot_bytes = 0: the workspace holds every
        cap at once."""
        cfg = L.ServerCfg(int(hbm_budget), int(slots), int(max_views), int(max_pixels), int(max_records),
                          int(pair_capacity), int(tile_size), int(lds_batch), int(slot_bytes))
        h = C.c_void_p()
        L.check(L.lib().nts_server_create(int(device), C.byref(cfg), C.byref(h)))
        self.h = h
        self._out = {}  # request id -> output array (kept alive until done)

    def close(self) -> None:
        if self.h:
            L.lib().nts_server_destroy(self.h)
            self.h = None

    def add_scene(self, splats: HostSplats) -> int:
        sid = C.c_uint32()
        L.check(L.lib().nts_server_add_scene(self.h, C.byref(_host_struct(splats)), C.byref(sid)))
        return int(sid.value)

    def add_scene_ply(self, path) -> int:
        """A resident scene read from a 3DGS PLY file."""
        sid = C.c_uint32()
        L.check(L.lib().nts_server_add_scene_ply(self.h, str(path).encode(), C.byref(sid)))
        return int(sid.value)

    def submit(self, scene: int, view: ViewSpec, fmt: int = L.OUT_RGBA8):
        if fmt == L.OUT_RGBA8:
            out = np.zeros((view.height, view.width, 4), np.uint8)
        else:
            out = np.zeros((view.height, view.width, 4), np.float32)
        rid = C.c_uint64()
        vd = view.c()
        L.check(L.lib().nts_server_submit(self.h, int(scene), C.byref(vd), int(fmt), L.vptr(out), C.byref(rid)))
        self._out[rid.value] = out
        return int(rid.value), out

    def poll(self, rid: int):
        """None while pending, else (image, info dict)."""
        done = C.c_int()
        info = L.RequestInfo()
        try:
            L.check(L.lib().nts_server_poll(self.h, int(rid), C.byref(done), C.byref(info)))
        except L.NtsError:
            self._out.pop(rid, None)
            raise
        if not done.value:
            return None
        return self._out.pop(rid), dict(latency_ms=info.latency_ms, pairs=info.pairs, batch_views=info.batch_views)

    def wait(self, rid: int):
        info = L.RequestInfo()
        try:
            L.check(L.lib().nts_server_wait(self.h, int(rid), C.byref(info)))
        finally:
            out = self._out.pop(rid, None)
        return out, dict(latency_ms=info.latency_ms, pairs=info.pairs, batch_views=info.batch_views)

    def stats(self) -> dict:
        st = L.ServeStats()
        L.check(L.lib().nts_server_stats(self.h, C.byref(st)))
        return {name: getattr(st, name) for name, _ in L.ServeStats._fields_}
