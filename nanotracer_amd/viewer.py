"""The viewer's render caller without the window: the per-frame sequence of
State::frame (crates/nano-view/src/lib.rs:754-796) -- project, bin at tile
16 with depth_max = max(4 * orbit distance, 50), composite, tonemap to
rgba8 -- over the C ABI, the orbit camera (lib.rs:152-169), the splat swap
of update_splats (lib.rs:712-731), the viewport resize, the frame-time
statistics, and the live-training channel of run_with_training
(lib.rs:76-139): a worker thread trains and publishes a versioned snapshot
of the splats after every iteration, the render loop picks up the newest.

Window, UI, input handling and the dock layout are not part of this
module: a caller presents the rgba8 frame however it likes."""
from __future__ import annotations

import threading
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from .api import CameraUniform, GpuCtx, GpuSplatBuffer, Rasterizer, TileBinner, TilingParams
from .synth import HostSplats

DEFAULT_VIEWPORT = (1280, 720)


@dataclass
class OrbitCamera:
    """OrbitCamera (lib.rs:152-169): spherical coordinates around a target,
    evaluated in float32 like the reference's glam Vec3."""

    target: tuple = (0.0, 0.0, 0.0)
    distance: float = 5.0
    azimuth: float = 0.0
    elevation: float = 0.0
    fov_y: float = float(np.float32(np.pi / 3))

    def position(self) -> tuple:
        f = np.float32
        d, az, el = f(self.distance), f(self.azimuth), f(self.elevation)
        cos_e = np.cos(el, dtype=np.float32)
        t = [f(x) for x in self.target]
        return (float(t[0] + d * cos_e * np.cos(az, dtype=np.float32)),
                float(t[1] + d * np.sin(el, dtype=np.float32)),
                float(t[2] + d * cos_e * np.sin(az, dtype=np.float32)))

    def depth_max(self) -> float:
        """lib.rs:777: (distance * 4).max(50)."""
        return float(max(np.float32(self.distance) * np.float32(4.0), np.float32(50.0)))


@dataclass
class FrameStats:
    total_pairs: int
    frame_ms: float
    fps: float


class FrameRenderer:
    """The viewer's render state (lib.rs:642-652): the device scene, the
    projected records, the float composite and the rgba8 target, reused
    frame to frame; frame() runs one State::frame compute pass."""

    TILE = 16

    def __init__(self, ctx: GpuCtx, splats: HostSplats, width: int = DEFAULT_VIEWPORT[0],
                 height: int = DEFAULT_VIEWPORT[1]):
        self.ctx = ctx
        self.r = Rasterizer()
        self.binner = TileBinner()
        self.gpu_splats = GpuSplatBuffer.upload(ctx, splats, trainable=False)
        self.last_cpu_splats = splats
        self.image_w, self.image_h = int(width), int(height)
        self.projected = self.r.alloc_projected(ctx, self.gpu_splats.capacity)
        self._alloc_targets()
        self.frame_times: deque = deque()

    def _alloc_targets(self):
        self.composite_buf = self.r.alloc_image(self.ctx, self.image_w, self.image_h)
        self.splat_view = self.ctx.storage_buffer_zeroed(self.image_w * self.image_h * 4)

    def update_splats(self, splats: HostSplats) -> None:
        """lib.rs:712-731: same count -> sync_from in place; another count ->
        a new device scene and projected buffer.  A count within the
        current capacity reuses both."""
        if splats.n == self.gpu_splats.n:
            self.gpu_splats.sync_from(splats)
        elif splats.n <= self.gpu_splats.capacity:
            self.gpu_splats.upload_from(splats)
        else:
            self.gpu_splats.free()
            self.gpu_splats = GpuSplatBuffer.upload(self.ctx, splats, trainable=False)
            self.projected.free()
            self.projected = self.r.alloc_projected(self.ctx, self.gpu_splats.capacity)
        self.last_cpu_splats = splats

    def resize_viewport(self, new_w: int, new_h: int) -> None:
        """Reallocate the render targets for a new viewport size."""
        if (new_w, new_h) == (self.image_w, self.image_h) or new_w == 0 or new_h == 0:
            return
        self.composite_buf.free()
        self.splat_view.free()
        self.image_w, self.image_h = int(new_w), int(new_h)
        self._alloc_targets()

    def frame(self, orbit: OrbitCamera) -> FrameStats:
        t0 = time.perf_counter()
        n = self.gpu_splats.n
        cam = CameraUniform.from_pose(orbit.position(), orbit.target, (0.0, 1.0, 0.0), orbit.fov_y, self.image_w,
                                      self.image_h, n)
        self.r.project(self.ctx, self.gpu_splats, cam, self.projected)
        params = TilingParams(self.image_w, self.image_h, self.TILE, orbit.depth_max())
        bins = self.binner.bin(self.ctx, self.projected, n, params)
        self.r.composite(self.ctx, self.projected, bins, self.composite_buf, params)
        self.r.tonemap(self.ctx, self.composite_buf, self.splat_view, self.image_w, self.image_h)
        pairs = bins.total_pairs
        bins.free()
        self.ctx.sync()
        now = time.perf_counter()
        # frame statistics over the last second (lib.rs:798-803)
        self.frame_times.append(now)
        while self.frame_times and now - self.frame_times[0] > 1.0:
            self.frame_times.popleft()
        return FrameStats(pairs, (now - t0) * 1e3, float(len(self.frame_times)))

    def rgba8(self) -> np.ndarray:
        """The tonemapped frame, (H, W, 4) uint8."""
        return self.splat_view.download(np.uint8, self.image_w * self.image_h * 4).reshape(
            self.image_h, self.image_w, 4)

    def composite(self) -> np.ndarray:
        """The float composite (rgb, 1 - T), (H, W, 4) float32."""
        return self.composite_buf.download(np.float32, self.image_w * self.image_h * 4).reshape(
            self.image_h, self.image_w, 4)

    def free(self) -> None:
        for b in (self.composite_buf, self.splat_view, self.projected):
            b.free()
        self.gpu_splats.free()


@dataclass
class TrainSnapshot:
    """TrainSnapshot (lib.rs:90-99): the splats after an iteration, its
    loss and splat count, and a version that increases per publish."""

    splats: HostSplats
    iteration: int
    loss: float
    splat_count: int
    version: int


@dataclass
class TrainingChannel:
    """The shared slot between the training worker and the render loop
    (SharedSnapshot + paused flag, lib.rs:76-86)."""

    lock: threading.Lock = field(default_factory=threading.Lock)
    snapshot: TrainSnapshot | None = None
    paused: threading.Event = field(default_factory=threading.Event)
    done: threading.Event = field(default_factory=threading.Event)
    error: BaseException | None = None
    history: list = field(default_factory=list)
    last_seen_version: int = 0

    def publish(self, snap: TrainSnapshot) -> None:
        with self.lock:
            self.snapshot = snap

    def latest(self) -> TrainSnapshot | None:
        with self.lock:
            return self.snapshot

    def poll(self) -> TrainSnapshot | None:
        """The newest snapshot if it is newer than the last one seen."""
        s = self.latest()
        if s is None or s.version == self.last_seen_version:
            return None
        self.last_seen_version = s.version
        self.history.append((s.iteration, s.loss, s.splat_count))
        return s


def start_training(device: int, splats: HostSplats, views, depth_max: float, cfg=None,
                   publish_every: int = 1) -> tuple[TrainingChannel, threading.Thread]:
    """run_with_training's worker (lib.rs:87-110): train() on its own
    context (one context per host thread) publishing a snapshot every
    `publish_every` iterations and at the end, sleeping while paused."""
    from .train import train

    ch = TrainingChannel()

    def worker():
        ctx = GpuCtx(device)
        version = 0
        try:
            def on_snapshot(it, host, loss):
                nonlocal version
                version += 1
                ch.publish(TrainSnapshot(host, it, loss, host.n, version))
                while ch.paused.is_set():
                    time.sleep(0.01)

            train(ctx, splats, views, depth_max, cfg, on_snapshot=on_snapshot, snapshot_every=publish_every,
                  log=lambda *_: None)
        except BaseException as e:  # reported to the render side, like the reference's eprintln
            ch.error = e
        finally:
            ctx.close()
            ch.done.set()

    th = threading.Thread(target=worker, name="nts-train", daemon=True)
    th.start()
    return ch, th


def run_with_training(ctx: GpuCtx, splats: HostSplats, views, depth_max: float, orbit: OrbitCamera, cfg=None,
                      width: int = DEFAULT_VIEWPORT[0], height: int = DEFAULT_VIEWPORT[1], on_frame=None,
                      publish_every: int = 1, max_frames: int | None = None):
    """Headless run_with_training: wait for the first snapshot, then render
    frames of the newest snapshot until training ends (or max_frames).
    on_frame(renderer, stats, snapshot) is called after every frame.
    Returns (renderer, channel); the caller frees the renderer."""
    ch, th = start_training(ctx.device, splats, views, depth_max, cfg, publish_every)
    while ch.latest() is None and not ch.done.is_set():
        time.sleep(0.005)
    if ch.error is not None:
        th.join()
        raise ch.error
    first = ch.poll()
    if first is None:
        th.join()
        raise RuntimeError("training ended without publishing a snapshot")
    renderer = FrameRenderer(ctx, first.splats, width, height)
    frames = 0
    snap = first
    while True:
        stats = renderer.frame(orbit)
        frames += 1
        if on_frame is not None:
            on_frame(renderer, stats, snap)
        if max_frames is not None and frames >= max_frames:
            break
        finished = ch.done.is_set()  # read before polling: a final publish precedes done
        newer = ch.poll()
        if newer is not None:
            renderer.update_splats(newer.splats)
            snap = newer
        elif finished:
            break
    th.join()
    if ch.error is not None:
        renderer.free()
        raise ch.error
    return renderer, ch
