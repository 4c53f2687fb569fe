"""ctypes binding of include/nts.h (libnts.so, built in-tree for gfx950).

There is no fallback: if the library is missing or a call fails, an
exception says so.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libnts.so"

NTS_OK = 0
NTS_E_INVALID, NTS_E_HIP, NTS_E_NOMEM, NTS_E_CAPACITY, NTS_E_STATE = -1, -2, -3, -4, -5
ERRORS = {-1: "NTS_E_INVALID", -2: "NTS_E_HIP", -3: "NTS_E_NOMEM", -4: "NTS_E_CAPACITY", -5: "NTS_E_STATE"}
LOSS_L1, LOSS_MSE, LOSS_SSIM = 0, 1, 2
SCENE_TRAINABLE = 1
SCHED_GRID, SCHED_PERSISTENT = 0, 1
BIN_CULLED, BIN_REFERENCE = 0, 1
SORT_NOALLOC_MAX_PAIRS = 1 << 21
BWD_REF_PARITY, BWD_RECOMPUTE = 0, 1


class NtsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Camera(C.Structure):
    """CameraUniform (crates/nano-optimize/src/raster.rs:61-71), 96 bytes."""

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
    """TilingParams (crates/nano-optimize/src/tile_binner.rs:21-42)."""

    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("tile_size", C.c_uint32), ("depth_max", C.c_float)]


class SplatsHost(C.Structure):
    _fields_ = [
        ("n", C.c_uint32),
        ("pos", C.POINTER(C.c_float)),
        ("rot", C.POINTER(C.c_float)),
        ("log_scale", C.POINTER(C.c_float)),
        ("logit", C.POINTER(C.c_float)),
        ("sh_dc", C.POINTER(C.c_float)),
        ("sh_rest", C.POINTER(C.c_float)),
    ]


class AdamCfg(C.Structure):
    """AdamConfig (crates/nano-optimize/src/adam.rs:22-39)."""

    _fields_ = [("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float)]


class Constraints(C.Structure):
    _fields_ = [("log_scale_min", C.c_float), ("log_scale_max", C.c_float), ("logit_max", C.c_float)]


class ViewDesc(C.Structure):
    """nts_view_desc: camera, depth range, image size."""

    _fields_ = [("cam", Camera), ("depth_max", C.c_float), ("width", C.c_uint32), ("height", C.c_uint32),
                ("pad", C.c_uint32)]


STEP_GRAPH, STEP_WRITE_GRADS = 1, 2
OUT_RGBA8, OUT_F32 = 0, 1


class ServerCfg(C.Structure):
    """nts_server_cfg."""

    _fields_ = [("hbm_budget", C.c_uint64), ("slots", C.c_uint32), ("max_views", C.c_uint32),
                ("max_pixels", C.c_uint64), ("max_records", C.c_uint64), ("pair_capacity", C.c_uint64),
                ("tile_size", C.c_uint32), ("lds_batch", C.c_uint32), ("slot_bytes", C.c_uint64)]


class ServeStats(C.Structure):
    """nts_serve_stats."""

    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("batches", C.c_uint64), ("views", C.c_uint64), ("overflow_reruns", C.c_uint64),
                ("max_in_flight", C.c_uint32), ("scenes", C.c_uint32), ("pending", C.c_uint32),
                ("memory_limited", C.c_uint32), ("hbm_in_use", C.c_uint64)]


class RequestInfo(C.Structure):
    """nts_request_info."""

    _fields_ = [("latency_ms", C.c_double), ("pairs", C.c_uint64), ("batch_views", C.c_uint32),
                ("status", C.c_int)]

assert C.sizeof(Camera) == 96
assert C.sizeof(ViewDesc) == 112


_lib = None

_SIGS = {
    "nts_version": (C.c_int, []),
    "nts_last_error": (C.c_char_p, []),
    "nts_ctx_create": (C.c_int, [C.c_int, C.c_uint64, C.POINTER(C.c_void_p)]),
    "nts_ctx_destroy": (C.c_int, [C.c_void_p]),
    "nts_sync": (C.c_int, [C.c_void_p]),
    "nts_prof_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "nts_prof_reset": (C.c_int, [C.c_void_p]),
    "nts_prof_query": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "nts_prof_report": (C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t]),
    "nts_probe_peaks": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                  C.POINTER(C.c_double)]),
    "nts_probe_blend_mix": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_double)]),
    "nts_probe_valu": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "nts_buffer_create": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    "nts_buffer_destroy": (C.c_int, [C.c_void_p, C.c_void_p]),
    "nts_buffer_size": (C.c_uint64, [C.c_void_p]),
    "nts_buffer_upload": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64]),
    "nts_buffer_download": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64]),
    "nts_buffer_zero": (C.c_int, [C.c_void_p, C.c_void_p]),
    "nts_scene_create": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p)]),
    "nts_scene_sync_params": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(SplatsHost)]),
    "nts_scene_destroy": (C.c_int, [C.c_void_p, C.c_void_p]),
    "nts_scene_count": (C.c_uint32, [C.c_void_p]),
    "nts_scene_upload": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(SplatsHost)]),
    "nts_scene_download": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(SplatsHost)]),
    "nts_scene_download_grads": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(SplatsHost)]),
    "nts_scene_upload_grads": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(SplatsHost)]),
    "nts_scene_zero_grads": (C.c_int, [C.c_void_p, C.c_void_p]),
    "nts_scene_download_moments": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(SplatsHost), C.POINTER(SplatsHost)]),
    "nts_scene_upload_moments": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(SplatsHost), C.POINTER(SplatsHost)]),
    "nts_scene_download_grad_acc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "nts_scene_upload_grad_acc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "nts_camera_from_pose": (C.c_int, [C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_float,
                                        C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(Camera)]),
    "nts_project": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Camera), C.c_void_p]),
    "nts_bin": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(Tiling), C.POINTER(C.c_void_p)]),
    "nts_bins_destroy": (C.c_int, [C.c_void_p, C.c_void_p]),
    "nts_bins_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]),
    "nts_bins_download_keys_u32": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "nts_bins_download_keys_u64": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "nts_bins_download_payloads": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "nts_bins_download_ranges": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "nts_scan_u32": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "nts_sort_pairs_u32": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32]),
    "nts_sort_pairs_u64": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32]),
    "nts_sort_pairs_u32_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32]),
    "nts_sort_tile_pairs": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32,
                                      C.c_void_p, C.c_uint32]),
    "nts_composite": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(Tiling)]),
    "nts_composite_ex": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(Tiling), C.c_void_p,
                                   C.c_void_p]),
    "nts_composite_backward_ex": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_void_p, C.POINTER(Tiling), C.c_int]),
    "nts_project_backward_ex": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(Camera), C.c_int]),
    "nts_loss_grad_ssim": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                     C.c_float, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]),
    "nts_loss_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_int,
                                C.POINTER(C.c_double)]),
    "nts_composite_backward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.POINTER(Tiling)]),
    "nts_project_backward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(Camera)]),
    "nts_adam_step_flat": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                     C.POINTER(AdamCfg), C.c_uint64]),
    "nts_adam_step_all": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(AdamCfg), C.POINTER(AdamCfg), C.c_uint64]),
    "nts_apply_constraints": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Constraints)]),
    "nts_accumulate_grad": (C.c_int, [C.c_void_p, C.c_void_p]),
    "nts_adam_step": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(AdamCfg), C.POINTER(AdamCfg), C.c_uint64,
                                C.POINTER(Constraints)]),
    "nts_tonemap": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32]),
    "nts_batch_create": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint64,
                                   C.POINTER(C.c_void_p)]),
    "nts_batch_destroy": (C.c_int, [C.c_void_p, C.c_void_p]),
    "nts_train_step_views": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(ViewDesc), C.c_uint32,
                                       C.c_void_p, C.c_uint32, C.c_int, C.c_int, C.POINTER(AdamCfg),
                                       C.POINTER(AdamCfg), C.c_uint64, C.POINTER(Constraints), C.c_int]),
    "nts_render_views": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(ViewDesc), C.c_uint32,
                                   C.c_uint32]),
    "nts_device_memory": (C.c_int, [C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "nts_server_create": (C.c_int, [C.c_int, C.POINTER(ServerCfg), C.POINTER(C.c_void_p)]),
    "nts_server_destroy": (C.c_int, [C.c_void_p]),
    "nts_server_add_scene": (C.c_int, [C.c_void_p, C.POINTER(SplatsHost), C.POINTER(C.c_uint32)]),
    "nts_server_add_scene_ply": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_uint32)]),
    "nts_server_submit": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(ViewDesc), C.c_int, C.c_void_p,
                                    C.POINTER(C.c_uint64)]),
    "nts_server_poll": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_int), C.POINTER(RequestInfo)]),
    "nts_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(RequestInfo)]),
    "nts_server_stats": (C.c_int, [C.c_void_p, C.POINTER(ServeStats)]),
    "nts_ply_count": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint32)]),
    "nts_ply_read": (C.c_int, [C.c_char_p, C.POINTER(SplatsHost)]),
    "nts_ply_write": (C.c_int, [C.c_char_p, C.POINTER(SplatsHost)]),
    "nts_scene_zero_grad_acc": (C.c_int, [C.c_void_p, C.c_void_p]),
    "nts_prune": (C.c_int, [C.c_void_p, C.c_void_p, C.c_float, C.POINTER(C.c_uint32)]),
    "nts_densify": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_float, C.c_float, C.c_uint32, C.c_uint32,
                              C.POINTER(C.c_uint32)]),
    "nts_batch_set_lds_batch": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32]),
    "nts_batch_set_ssim_lambda": (C.c_int, [C.c_void_p, C.c_void_p, C.c_float]),
    "nts_batch_set_scheduler": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "nts_batch_set_eval_counting": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "nts_batch_set_binning": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "nts_batch_eval_counts": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "nts_batch_losses": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "nts_batch_stats": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32),
                                  C.POINTER(C.c_uint32)]),
    "nts_batch_download_image": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "nts_batch_download_state": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "nts_batch_download_projected_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]),
    "nts_batch_download_bins": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint64), C.c_void_p,
                                          C.c_void_p, C.c_void_p]),
}

# symbols of include/nts.h, checked by the CPU test suite
EXPORTED = sorted(_SIGS)


def lib():
    """Load libnts.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is not built; run `python -m nanotracer_amd.build`")
        h = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def check(rc: int) -> None:
    if rc != NTS_OK:
        msg = lib().nts_last_error()
        raise NtsError(rc, msg.decode() if msg else "")


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


def vptr(a: np.ndarray):
    assert a.flags.c_contiguous
    return C.c_void_p(a.ctypes.data)
