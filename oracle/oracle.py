"""ctypes wrapper of the host restatement (oracle/lib/libnts_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker.  Imported by tests/, by the smoke
check and by bench.py's CPU-baseline leg; never by the product path.

Each function restates a piece of crates/nano-optimize (file:line in
nts_oracle.h / nts_oracle.cpp).  All parts are pinned by tests:
  - exp: against float64 exp (<= 1 ulp);
  - scan / sort / count / emit / ranges: the reference's known-answer tests
    (prefix_scan.rs:291-333, radix_sort.rs:357-408, tile_binner.rs:443-541),
    closed forms and brute force;
  - projection: the closed form of raster.rs:999-1026 and a float64 EWA
    restatement written independently in tests/;
  - forward blend: brute force per pixel and the analytic single splat;
  - backward: fp64 central finite differences of the forward;
  - Adam / constraints / grad_acc: closed forms and the adam_gpu_parity.rs
    sequences;
  - prune / densify: the train.rs:626-743 known-answer tests.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "lib" / "libnts_oracle.so"


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


class Camera(C.Structure):
    _fields_ = [
        ("view", C.c_float * 16),
        ("cam_pos", C.c_float * 3),
        ("near_", C.c_float),
        ("focal", C.c_float),
        ("width", C.c_float),
        ("height", C.c_float),
        ("n_splats", C.c_uint32),
    ]


class Tiling(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("tile_size", C.c_uint32), ("depth_max", C.c_float)]


class _Splats32(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("pos", "rot", "log_scale", "logit", "sh_dc", "sh_rest")]


_h = None


def lib():
    global _h
    if _h is None:
        if not LIB.exists():
            build()
        _h = C.CDLL(str(LIB))
        for name in ("orc_tile_count", "orc_bin", "orc_bin_ex"):
            getattr(_h, name).restype = C.c_uint64
        _h.orc_loss_grad.restype = C.c_double
        _h.orc_powi_f32.restype = C.c_float
        _h.orc_powi_f32.argtypes = [C.c_float, C.c_int32]
        _h.orc_prune.restype = C.c_uint32
        _h.orc_densify.restype = C.c_uint32
    return _h


def _p(a: np.ndarray):
    assert a.flags.c_contiguous, "array must be contiguous"
    return C.c_void_p(a.ctypes.data)


def num_threads() -> int:
    return lib().orc_num_threads()


def expf(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty_like(x)
    lib().orc_expf_array(_p(x), _p(y), C.c_uint64(x.size))
    return y


def logf(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty_like(x)
    lib().orc_logf_array(_p(x), _p(y), C.c_uint64(x.size))
    return y


def cos2pi(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty_like(x)
    lib().orc_cos2pi_array(_p(x), _p(y), C.c_uint64(x.size))
    return y


def lcg_jump(state: int, k: int) -> int:
    lib().orc_lcg_jump.restype = C.c_uint32
    return int(lib().orc_lcg_jump(C.c_uint32(state), C.c_uint64(k)))


def box_muller_stream(state: int, n: int):
    out = np.empty(n, np.float32)
    uni = np.empty(2 * n, np.float32)
    lib().orc_box_muller_stream(C.c_uint32(state), C.c_uint64(n), _p(out), _p(uni))
    return out, uni.reshape(n, 2)


def camera_from_pose(eye, target, up, fov_y, width, height, n_splats) -> Camera:
    f3 = lambda v: (C.c_float * 3)(*[float(x) for x in v])
    cam = Camera()
    lib().orc_camera_from_pose(f3(eye), f3(target), f3(up), C.c_float(fov_y), C.c_uint32(width),
                               C.c_uint32(height), C.c_uint32(n_splats), C.byref(cam))
    return cam


def camera_like(cam) -> Camera:
    """Copy any 96-byte camera structure into the oracle's type."""
    out = Camera()
    C.memmove(C.byref(out), C.byref(cam), C.sizeof(Camera))
    return out


def tiling(width, height, tile_size=16, depth_max=100.0) -> Tiling:
    return Tiling(width, height, tile_size, depth_max)


def _splats(s, dtype):
    arrs = [np.ascontiguousarray(a, dtype) for a in (s.pos, s.rot, s.log_scale, s.logit, s.sh_dc, s.sh_rest)]
    st = _Splats32(*[a.ctypes.data for a in arrs])
    return st, arrs


def project(s, cam: Camera, dtype=np.float32) -> np.ndarray:
    st, keep = _splats(s, dtype)
    n = s.pos.shape[0]
    out = np.zeros((n, 12), dtype)
    fn = lib().orc_project_f32 if dtype == np.float32 else lib().orc_project_f64
    fn(C.byref(st), C.c_uint32(n), C.byref(cam), _p(out))
    return out


def tile_count(projected: np.ndarray, t: Tiling):
    n = projected.shape[0]
    counts = np.zeros(max(n, 1), np.uint32)
    pj = np.ascontiguousarray(projected, np.float32)
    total = lib().orc_tile_count(_p(pj), C.c_uint32(n), C.byref(t), _p(counts))
    return counts[:n], int(total)


def exclusive_scan(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, np.uint32)
    out = np.empty_like(a)
    lib().orc_exclusive_scan_u32(_p(a), C.c_uint64(a.size), _p(out))
    return out


def sort_pairs(keys: np.ndarray, vals: np.ndarray):
    k = np.array(keys, copy=True)
    v = np.array(vals, dtype=np.uint32, copy=True)
    if k.dtype == np.uint32:
        lib().orc_sort_pairs_u32(_p(k), _p(v), C.c_uint64(k.size))
    else:
        k = k.astype(np.uint64)
        lib().orc_sort_pairs_u64(_p(k), _p(v), C.c_uint64(k.size))
    return k, v


def tile_ranges(sorted_keys: np.ndarray, num_tiles: int) -> np.ndarray:
    """tile_ranges.wgsl over sorted u32 (tile << 16 | depth16) keys: 2 u32
    per tile, [begin, end), (0, 0) for a tile with no key; keys whose tile is
    num_tiles or more lie in no range."""
    k = np.ascontiguousarray(sorted_keys, np.uint32)
    ranges = np.zeros(2 * max(num_tiles, 1), np.uint32)
    lib().orc_tile_ranges(_p(k), C.c_uint64(k.size), C.c_uint32(num_tiles), _p(ranges))
    return ranges[: 2 * num_tiles]


def bin(projected: np.ndarray, t: Tiling, cull: bool = False):
    """Full TileBinner::bin: (sorted u32 keys, payloads, ranges[2*tiles]).
    cull: only the tiles each record's alpha >= 1/255 ellipse can reach, as
    the device's batched and serving paths bin."""
    projected = np.ascontiguousarray(projected, np.float32)
    n = projected.shape[0]
    f = lib().orc_bin_ex
    total = f(_p(projected), C.c_uint32(n), C.byref(t), C.c_int(int(cull)), None, None, None)
    ntiles = (-(-t.width // t.tile_size)) * (-(-t.height // t.tile_size))
    keys = np.zeros(max(total, 1), np.uint32)
    pay = np.zeros(max(total, 1), np.uint32)
    ranges = np.zeros(2 * ntiles, np.uint32)
    f(_p(projected), C.c_uint32(n), C.byref(t), C.c_int(int(cull)), _p(keys), _p(pay), _p(ranges))
    return keys[:total], pay[:total], ranges


def composite(projected: np.ndarray, payloads: np.ndarray, ranges: np.ndarray, t: Tiling, with_state=False):
    projected = np.ascontiguousarray(projected, np.float32)
    payloads = np.ascontiguousarray(payloads, np.uint32)
    if payloads.size == 0:
        payloads = np.zeros(1, np.uint32)
    ranges = np.ascontiguousarray(ranges, np.uint32)
    img = np.zeros((t.height, t.width, 4), np.float32)
    ft = np.zeros((t.height, t.width), np.float32)
    li = np.zeros((t.height, t.width), np.uint32)
    lib().orc_composite(_p(projected), _p(payloads), _p(ranges), C.byref(t), _p(img), _p(ft), _p(li))
    return (img, ft, li) if with_state else img


def composite_used(projected: np.ndarray, payloads: np.ndarray, ranges: np.ndarray, t: Tiling) -> np.ndarray:
    """Per pixel, the number of splats the forward blend blends into it."""
    projected = np.ascontiguousarray(projected, np.float32)
    payloads = np.ascontiguousarray(payloads, np.uint32)
    if payloads.size == 0:
        payloads = np.zeros(1, np.uint32)
    ranges = np.ascontiguousarray(ranges, np.uint32)
    img = np.zeros((t.height, t.width, 4), np.float32)
    used = np.zeros((t.height, t.width), np.uint32)
    lib().orc_composite_used(_p(projected), _p(payloads), _p(ranges), C.byref(t), _p(img), _p(used))
    return used


def composite_f64(projected: np.ndarray, payloads: np.ndarray, ranges: np.ndarray, t: Tiling,
                  with_state: bool = False):
    """float64 blend; with_state also returns (last_idx, n_used) per pixel."""
    projected = np.ascontiguousarray(projected, np.float64)
    payloads = np.ascontiguousarray(payloads, np.uint32)
    if payloads.size == 0:
        payloads = np.zeros(1, np.uint32)
    img = np.zeros((t.height, t.width, 4), np.float64)
    li = np.zeros((t.height, t.width), np.uint32)
    nu = np.zeros((t.height, t.width), np.uint32)
    rg = np.ascontiguousarray(ranges, np.uint32)
    lib().orc_composite_f64(_p(projected), _p(payloads), _p(rg), C.byref(t), _p(img), _p(li), _p(nu))
    return (img, li, nu) if with_state else img


def composite_backward_ref(projected, payloads, ranges, t: Tiling, forward_out, dl_dc) -> np.ndarray:
    projected = np.ascontiguousarray(projected, np.float32)
    n = projected.shape[0]
    payloads = np.ascontiguousarray(payloads, np.uint32)
    if payloads.size == 0:
        payloads = np.zeros(1, np.uint32)
    out = np.zeros((n, 12), np.float32)
    rg = np.ascontiguousarray(ranges, np.uint32)
    fo = np.ascontiguousarray(forward_out, np.float32)
    dl = np.ascontiguousarray(dl_dc, np.float32)
    lib().orc_composite_backward_ref(_p(projected), _p(payloads), _p(rg), C.byref(t), C.c_uint32(n), _p(fo),
                                     _p(dl), _p(out))
    return out


def composite_backward_recompute(projected, payloads, ranges, t: Tiling, final_t, last_idx, dl_dc) -> np.ndarray:
    projected = np.ascontiguousarray(projected, np.float32)
    n = projected.shape[0]
    payloads = np.ascontiguousarray(payloads, np.uint32)
    if payloads.size == 0:
        payloads = np.zeros(1, np.uint32)
    out = np.zeros((n, 12), np.float32)
    rg = np.ascontiguousarray(ranges, np.uint32)
    ft = np.ascontiguousarray(final_t, np.float32)
    li = np.ascontiguousarray(last_idx, np.uint32)
    dl = np.ascontiguousarray(dl_dc, np.float32)
    lib().orc_composite_backward_recompute(_p(projected), _p(payloads), _p(rg), C.byref(t), C.c_uint32(n), _p(ft),
                                           _p(li), _p(dl), _p(out))
    return out


def project_backward(s, cam: Camera, projected_grad: np.ndarray, full: bool = False):
    """Returns gradients in the host SoA layout (HostSplats-like tuple of
    arrays).  full=True: the exact chain rule (recompute mode)."""
    from nanotracer_amd.synth import HostSplats  # data container only

    st, keep = _splats(s, np.float32)
    n = s.pos.shape[0]
    g = HostSplats.zeros(n)
    gs = _Splats32(*[a.ctypes.data for a in g.arrays()])
    pg = np.ascontiguousarray(projected_grad, np.float32)
    fn = lib().orc_project_backward_full if full else lib().orc_project_backward
    fn(C.byref(st), C.c_uint32(n), C.byref(cam), _p(pg), C.byref(gs))
    return g


def powi(a: float, b: int) -> float:
    return lib().orc_powi_f32(a, b)


def adam_step(params, grads, m, v, lr, beta1, beta2, eps, t):
    """In place on float32 arrays (adam.rs:70-90)."""
    gr = np.ascontiguousarray(grads, np.float32)
    lib().orc_adam_step(_p(params), _p(gr), _p(m), _p(v),
                        C.c_uint64(params.size), C.c_float(lr), C.c_float(beta1), C.c_float(beta2),
                        C.c_float(eps), C.c_uint64(t))


def apply_constraints(rot, log_scale, logit, smin=-10.0, smax=5.0, omax=10.0):
    lib().orc_apply_constraints(_p(rot), _p(log_scale), _p(logit), C.c_uint32(logit.size), C.c_float(smin),
                                C.c_float(smax), C.c_float(omax))


def accumulate_grad(dpos, acc):
    dp = np.ascontiguousarray(dpos, np.float32)
    lib().orc_accumulate_grad(_p(dp), _p(acc), C.c_uint32(acc.size))


def loss_grad(pred, target, kind=0):
    pred = np.ascontiguousarray(pred, np.float32)
    h, w = pred.shape[:2]
    dl = np.zeros_like(pred)
    tg = np.ascontiguousarray(target, np.float32)
    loss = lib().orc_loss_grad(_p(pred), _p(tg), C.c_uint32(w),
                               C.c_uint32(h), C.c_int(kind), _p(dl))
    return loss, dl


class _Mut(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("pos", "rot", "log_scale", "logit", "sh_dc", "sh_rest")]


def _mut(s, cap):
    arrs = []
    for a, w in zip(s.arrays(), (3, 4, 3, 1, 3, 45)):
        b = np.zeros(cap * w, np.float32)
        b[: a.size] = a.reshape(-1)
        arrs.append(b)
    return _Mut(*[a.ctypes.data for a in arrs]), arrs


def _unmut(arrs, n):
    from nanotracer_amd.synth import HostSplats

    ws = (3, 4, 3, 1, 3, 45)
    out = [arrs[k][: n * ws[k]].reshape(n, ws[k]) if ws[k] > 1 else arrs[k][:n].copy() for k in range(6)]
    return HostSplats(*[np.ascontiguousarray(a) for a in out])


def _moments(moments, cap):
    """moments: list of 12 flat arrays (m, v per attribute) or None."""
    if moments is None:
        return None, None
    ws = (3, 4, 3, 1, 3, 45)
    bufs = []
    for k in range(12):
        b = np.zeros(cap * ws[k // 2], np.float32)
        b[: moments[k].size] = moments[k]
        bufs.append(b)
    ptrs = (C.c_void_p * 12)(*[b.ctypes.data for b in bufs])
    return ptrs, bufs


def prune(s, threshold=-5.3, moments=None, grad_acc=None):
    n = s.n
    st, arrs = _mut(s, n)
    mp, mb = _moments(moments, n)
    ga = None if grad_acc is None else np.ascontiguousarray(grad_acc, np.float32).copy()
    n2 = lib().orc_prune(C.byref(st), C.c_uint32(n), C.c_float(threshold), mp, None if ga is None else _p(ga))
    ws = (3, 4, 3, 1, 3, 45)
    mo = None if mb is None else [mb[k][: n2 * ws[k // 2]] for k in range(12)]
    return _unmut(arrs, n2), mo, (None if ga is None else ga[:n2])


def densify(s, grad_count, grad_threshold, split_thr, max_splats, it, moments=None, grad_acc=None):
    n = s.n
    cap = max(n * 3, max_splats) + 4
    st, arrs = _mut(s, cap)
    mp, mb = _moments(moments, cap)
    ga = np.zeros(cap, np.float32)
    ga[:n] = grad_acc
    counts = (C.c_uint32 * 3)()
    n2 = lib().orc_densify(C.byref(st), C.c_uint32(n), C.c_uint32(grad_count), C.c_float(grad_threshold),
                           C.c_float(split_thr), C.c_uint32(max_splats), C.c_uint32(it), mp, _p(ga), counts)
    ws = (3, 4, 3, 1, 3, 45)
    mo = None if mb is None else [mb[k][: n2 * ws[k // 2]] for k in range(12)]
    return _unmut(arrs, n2), mo, ga[:n2], tuple(counts)


def ssim_kernel() -> np.ndarray:
    k = np.zeros(11, np.float32)
    lib().orc_ssim_kernel(_p(k))
    return k


def ssim_blur(plane: np.ndarray, adjoint: bool = False) -> np.ndarray:
    plane = np.ascontiguousarray(plane, np.float32)
    out = np.zeros_like(plane)
    h, w = plane.shape
    lib().orc_ssim_blur(_p(plane), C.c_int(w), C.c_int(h), _p(out), C.c_int(int(adjoint)))
    return out


def ssim_loss_grad(pred: np.ndarray, target: np.ndarray):
    """loss.rs:346-383: (1 - mean SSIM over channels, dL/dC float4)."""
    pred = np.ascontiguousarray(pred, np.float32)
    target = np.ascontiguousarray(target, np.float32)
    h, w = pred.shape[:2]
    grad = np.zeros((h, w, 4), np.float32)
    lib().orc_ssim_loss_grad.restype = C.c_float
    loss = lib().orc_ssim_loss_grad(_p(pred), _p(target), C.c_int(w), C.c_int(h), _p(grad))
    return float(loss), grad


def combined_loss_grad(pred: np.ndarray, target: np.ndarray, lam: float):
    """loss.rs:390-415: (total, mse, ssim, dL/dC) of (1 - lam) MSE + lam (1 - SSIM)."""
    lam = np.float32(min(max(np.float32(lam), np.float32(0.0)), np.float32(1.0)))
    mse, mg = loss_grad(pred, target, 1)
    if lam == 0.0:
        return mse, mse, 0.0, mg
    ssim, sg = ssim_loss_grad(pred, target)
    if lam == 1.0:
        return ssim, 0.0, ssim, sg
    a, b = np.float32(1.0) - lam, lam
    return float(a * np.float32(mse) + b * np.float32(ssim)), mse, ssim, mg * a + sg * b


ATTRS = ("pos", "rot", "log_scale", "logit", "sh_dc", "sh_rest")


class FdContext:
    """Float64 local central differences of the mean-square loss against a
    fixed float32 binning (orc_fd_local_mse): build once per scene and
    camera, then fd(idx, attr, comp, h) per parameter."""

    def __init__(self, s, cam: Camera, t: Tiling, target: np.ndarray, weights: np.ndarray | None = None):
        self.s64 = type(s)(*[np.ascontiguousarray(a, np.float64) for a in s.arrays()])
        self.st, self._keep = _splats(self.s64, np.float64)
        self.n = s.n
        self.cam, self.t = cam, t
        self.proj32 = project(s, cam)
        self.proj64 = np.ascontiguousarray(project(self.s64, cam, np.float64), np.float64)
        self.keys, self.pay, self.rg = bin(self.proj32, t)
        if self.pay.size == 0:
            self.pay = np.zeros(1, np.uint32)
        self.target = np.ascontiguousarray(target, np.float32)
        # weights: differentiate sum <w, C> (fixed dL/dC) instead of the MSE
        self.weights = None if weights is None else np.ascontiguousarray(weights, np.float32)

    def _w(self):
        return None if self.weights is None else _p(self.weights)

    def fd(self, idx: int, attr: int, comp: int, h: float):
        out = C.c_double()
        rc = lib().orc_fd_local_mse(C.byref(self.st), C.c_uint32(self.n), C.byref(self.cam), C.byref(self.t),
                                    _p(self.proj64), _p(self.proj32), _p(self.pay), _p(self.rg), _p(self.target),
                                    self._w(), C.c_uint32(idx), C.c_int(attr), C.c_int(comp), C.c_double(h),
                                    C.byref(out))
        if rc < 0:
            raise ValueError("bad parameter index")
        return out.value, bool(rc)

    def fd_many(self, idx, attr, comp, h):
        """Arrays in, (fd, flag) arrays out (flag 0 continuous, 1 discontinuous
        under +-h, 2 float32 / float64 patterns differ); threads over
        parameters."""
        idx = np.ascontiguousarray(idx, np.uint32)
        k = idx.size
        attr = np.ascontiguousarray(np.broadcast_to(attr, (k,)), np.int32)
        comp = np.ascontiguousarray(np.broadcast_to(comp, (k,)), np.int32)
        h = np.ascontiguousarray(np.broadcast_to(h, (k,)), np.float64)
        fd = np.zeros(k, np.float64)
        disc = np.zeros(k, np.int32)
        lib().orc_fd_local_mse_batch(C.byref(self.st), C.c_uint32(self.n), C.byref(self.cam), C.byref(self.t),
                                     _p(self.proj64), _p(self.proj32), _p(self.pay), _p(self.rg), _p(self.target),
                                     self._w(), C.c_uint32(k), _p(idx), _p(attr), _p(comp), _p(h), _p(fd), _p(disc))
        if (disc < 0).any():
            raise ValueError("bad parameter index")
        return fd, disc


def tonemap(image: np.ndarray) -> np.ndarray:
    image = np.ascontiguousarray(image, np.float32)
    h, w = image.shape[:2]
    out = np.zeros((h, w, 4), np.uint8)
    lib().orc_tonemap(_p(image), C.c_uint32(w), C.c_uint32(h), _p(out))
    return out


def time_step_views(s, cams, tilings, targets, warmup=1, reps=5, threads=0):
    """The CPU baseline's timed step (orc_time_step_views): per-view and Adam
    wall times in seconds, shape (reps, n_views) and (reps,), measured by
    steady_clock inside the restatement."""
    st, keep = _splats(s, np.float32)
    nv = len(cams)
    cam_arr = (Camera * nv)(*cams)
    til_arr = (Tiling * nv)(*tilings)
    tg = [np.ascontiguousarray(t, np.float32) for t in targets]
    ptrs = (C.c_void_p * nv)(*[t.ctypes.data for t in tg])
    vs = np.zeros((reps, nv), np.float64)
    ad = np.zeros(reps, np.float64)
    rc = lib().orc_time_step_views(C.byref(st), C.c_uint32(s.n), cam_arr, til_arr, ptrs, C.c_uint32(nv),
                                   C.c_int(warmup), C.c_int(reps), C.c_int(threads), _p(vs), _p(ad))
    if rc != 0:
        raise ValueError("orc_time_step_views: bad arguments")
    return vs, ad
