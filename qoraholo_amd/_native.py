"""ctypes bindings of libqoraholo_hip.so (the C-ABI in include/qoraholo_hip.h).

Status codes become the exceptions the reference's numpy bindings raise for the matching C++
exceptions: invalid arguments ``ValueError``, device and runtime failures ``RuntimeError``,
misuse ``LogicError`` (a ``RuntimeError``). The library is required: nothing here falls back
to another implementation when it is missing.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_float, c_int, c_size_t, c_uint, c_uint16, c_uint64, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libqoraholo_hip.so"


class LogicError(RuntimeError):
    """Misuse of a stateful object, e.g. Pipeline.run before set_background."""


class OutOfDeviceMemory(RuntimeError):
    """The device could not hold the buffers of a call."""


class Geometry_t(ctypes.Structure):
    _fields_ = [("wavelength", c_float), ("n_medium", c_float), ("na", c_float), ("pixel", c_float)]


class Illumination_t(ctypes.Structure):
    _fields_ = [("fy", c_float), ("fx", c_float)]


class CropPlan_t(ctypes.Structure):
    _fields_ = [("cy", c_int), ("cx", c_int), ("radius", c_float), ("out", c_int)]


class OdtGrid_t(ctypes.Structure):
    _fields_ = [("n", c_int), ("nz", c_int), ("df", c_float), ("dfz", c_float), ("fm2", c_float),
                ("fna2", c_float), ("gain", c_float)]


class OdtOptions_t(ctypes.Structure):
    _fields_ = [("nz", c_int), ("iterations", c_int), ("solver", c_int), ("gp_transform", c_int),
                ("tv_weight", c_float), ("tv_inner", c_int), ("mapping", c_int)]


class OdtTimings_t(ctypes.Structure):
    _fields_ = [("bytes", c_size_t)] + [(k, c_double) for k in (
        "setup", "upload", "fft2", "clear", "mapping", "normalize", "solve", "finish", "download",
        "fgp_ms")] + [("fgp_steps", c_int), ("fgp_launches", c_int)]


class PipelineTimings_t(ctypes.Structure):
    _fields_ = [(k, c_double) for k in (
        "upload", "fields", "rytov", "fft2", "clear", "mapping", "normalize", "solve", "finish",
        "download", "total", "fgp_ms")] + [("fgp_steps", c_int), ("fgp_launches", c_int)]


class ServerStats_t(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("completed", ctypes.c_int64),
                ("failed", ctypes.c_int64), ("queued", ctypes.c_int64), ("p50_ms", c_double),
                ("p99_ms", c_double), ("mean_ms", c_double), ("recon_per_s", c_double),
                ("device_bytes", c_size_t), ("budget_bytes", c_size_t), ("slots", c_int),
                ("busy_slots", c_int), ("peak_busy_slots", c_int), ("graph_slots", c_int),
                ("batched_requests", ctypes.c_int64), ("peak_batch", c_int)]


class BpmOptions_t(ctypes.Structure):
    _fields_ = [("obliquity", c_int), ("chunk", c_int), ("checkpoint", c_int)]


class BpmTableView_t(ctypes.Structure):
    _fields_ = [("n", c_int), ("nz", c_int), ("count", c_int), ("phase", c_float),
                ("step", c_void_p), ("last", c_void_p), ("q", c_void_p), ("lipschitz", c_float)]


class BpmSolveOptions_t(ctypes.Structure):
    _fields_ = [("iterations", c_int), ("batch", c_int), ("step", c_float),
                ("tv_weight", c_float), ("tv_inner", c_int)]


class BpmRequestOptions_t(ctypes.Structure):
    _fields_ = [("n", c_int), ("nz", c_int), ("count", c_int),
                ("illuminations", POINTER(Illumination_t)), ("geometry", Geometry_t),
                ("bpm", BpmOptions_t), ("start_iterations", c_int), ("solve", BpmSolveOptions_t)]


class PipelineOptions_t(ctypes.Structure):
    _fields_ = [("count", c_int), ("hologram_size", c_int), ("crop", CropPlan_t),
                ("illuminations", POINTER(Illumination_t)), ("geometry", Geometry_t),
                ("odt", OdtOptions_t), ("upload_chunks", c_int)]


_SIGNATURES = {
    "qh_hip_last_error": (ctypes.c_char_p, []),
    "qh_hip_device_present": (c_int, []),
    "qh_make_grid": (c_int, [c_int, c_int, POINTER(Geometry_t), POINTER(OdtGrid_t)]),
    "qh_resolve_crop": (c_int, [c_float, c_float, c_float, c_int, c_int, POINTER(CropPlan_t)]),
    "qh_circular_scan": (c_int, [c_int, c_float, POINTER(Geometry_t), c_void_p]),
    "qh_gaussian_noise": (c_int, [c_uint64, c_float, c_size_t, c_void_p]),
    "qh_host_copy": (c_int, [c_void_p, c_void_p, c_size_t]),
    "qh_field_ratio": (c_int, [c_void_p, c_int, c_void_p, c_int, c_int, c_void_p]),
    "qh_hip_extract_fields": (c_int, [c_void_p, c_int, c_int, POINTER(CropPlan_t), c_void_p]),
    "qh_hip_unwrap": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p]),
    "qh_hip_rytov": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p]),
    "qh_hip_reconstruct": (c_int, [c_void_p, c_int, c_int, c_void_p, POINTER(OdtGrid_t),
                                   POINTER(Geometry_t), POINTER(OdtOptions_t), c_void_p,
                                   POINTER(c_double), POINTER(OdtTimings_t)]),
    "qh_hip_map_kspace": (c_int, [c_void_p, c_int, c_int, c_void_p, POINTER(OdtGrid_t), c_void_p,
                                  c_void_p, POINTER(c_double)]),
    "qh_hip_find_sidebands": (c_int, [c_void_p, c_int, c_int, c_float, c_void_p]),
    "qh_hip_estimate_illuminations": (c_int, [c_void_p, c_int, c_int, c_float, c_float, c_void_p,
                                              c_void_p]),
    "qh_hip_pipeline_create": (c_int, [POINTER(PipelineOptions_t), POINTER(c_void_p)]),
    "qh_hip_pipeline_destroy": (None, [c_void_p]),
    "qh_hip_pipeline_set_background": (c_int, [c_void_p, c_void_p]),
    "qh_hip_pipeline_input": (c_void_p, [c_void_p]),
    "qh_hip_pipeline_output": (c_void_p, [c_void_p]),
    "qh_hip_pipeline_process": (c_int, [c_void_p]),
    "qh_hip_pipeline_run": (c_int, [c_void_p, c_void_p, c_void_p]),
    "qh_hip_pipeline_use_graph": (c_int, [c_void_p, c_int]),
    "qh_hip_pipeline_process_device": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "qh_hip_pipeline_filled": (c_double, [c_void_p]),
    "qh_hip_pipeline_timings": (c_int, [c_void_p, POINTER(PipelineTimings_t)]),
    "qh_hip_pipeline_setup_ms": (c_double, [c_void_p]),
    "qh_hip_pipeline_device_bytes": (c_size_t, [c_void_p]),
    "qh_hip_pipeline_output_shape": (c_int, [c_void_p, POINTER(c_int), POINTER(c_int), POINTER(c_int)]),
    "qh_hip_pipeline_time_fgp": (c_int, [c_void_p, c_int]),
    "qh_hip_server_create": (c_int, [c_size_t, c_int, POINTER(c_void_p)]),
    "qh_hip_server_destroy": (None, [c_void_p]),
    "qh_hip_server_add_geometry": (c_int, [c_void_p, POINTER(PipelineOptions_t), c_void_p,
                                           POINTER(c_int)]),
    "qh_hip_server_add_geometry_batched": (c_int, [c_void_p, POINTER(PipelineOptions_t), c_void_p,
                                                   c_int, POINTER(c_int)]),
    "qh_hip_server_submit": (c_int, [c_void_p, c_int, c_void_p, c_void_p, POINTER(ctypes.c_int64)]),
    "qh_hip_server_add_bpm": (c_int, [c_void_p, POINTER(BpmRequestOptions_t), POINTER(c_int)]),
    "qh_hip_server_submit_bpm": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                         POINTER(ctypes.c_int64)]),
    "qh_hip_server_wait": (c_int, [c_void_p, ctypes.c_int64, POINTER(c_double), POINTER(c_double)]),
    "qh_hip_server_poll": (c_int, [c_void_p, ctypes.c_int64, POINTER(c_int)]),
    "qh_hip_server_stats": (c_int, [c_void_p, POINTER(ServerStats_t)]),
    "qh_hip_server_reset_stats": (c_int, [c_void_p]),
    "qh_bpm_tables": (c_int, [c_int, c_int, c_void_p, c_int, POINTER(Geometry_t), c_int, c_void_p,
                              c_void_p, c_void_p, POINTER(c_float), POINTER(c_float)]),
    "qh_hip_bpm_forward": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int, POINTER(Geometry_t),
                                   POINTER(BpmOptions_t), c_void_p]),
    "qh_hip_bpm_vjp": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_int,
                               POINTER(Geometry_t), POINTER(BpmOptions_t), c_void_p]),
    "qh_hip_bpm_reconstruct": (c_int, [c_void_p, c_int, c_int, c_void_p, POINTER(Geometry_t),
                                       c_void_p, c_int, POINTER(BpmOptions_t),
                                       POINTER(BpmSolveOptions_t), c_void_p]),
    "qh_hip_bpm_jvp": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_int,
                               POINTER(Geometry_t), POINTER(BpmOptions_t), c_void_p]),
    "qh_hip_bpm_forward_tables": (c_int, [c_void_p, POINTER(BpmTableView_t), POINTER(BpmOptions_t),
                                          c_void_p]),
    "qh_hip_bpm_vjp_tables": (c_int, [c_void_p, c_void_p, POINTER(BpmTableView_t),
                                      POINTER(BpmOptions_t), c_void_p]),
    "qh_hip_bpm_jvp_tables": (c_int, [c_void_p, c_void_p, POINTER(BpmTableView_t),
                                      POINTER(BpmOptions_t), c_void_p]),
    "qh_hip_bpm_reconstruct_tables": (c_int, [c_void_p, c_void_p, POINTER(BpmTableView_t),
                                              POINTER(BpmOptions_t), POINTER(BpmSolveOptions_t),
                                              c_void_p]),
    "qh_hip_bpm_plan_create": (c_int, [c_int, c_int, c_void_p, c_int, POINTER(Geometry_t),
                                       POINTER(BpmOptions_t), POINTER(c_void_p)]),
    "qh_hip_bpm_plan_destroy": (None, [c_void_p]),
    "qh_hip_bpm_plan_info": (c_int, [c_void_p, POINTER(c_float), POINTER(c_float), POINTER(c_int),
                                     POINTER(c_size_t)]),
    "qh_hip_bpm_plan_checkpoint": (c_int, [c_void_p, POINTER(c_int)]),
    "qh_hip_bpm_forward_dev": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "qh_hip_bpm_vjp_dev": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "qh_hip_bpm_jvp_dev": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "qh_hip_bpm_gradient_dev": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p,
                                        POINTER(c_double), c_void_p]),
    "qh_hip_fft_c2c": (c_int, [c_void_p, c_int, c_int, c_int, c_uint, c_int, c_float]),
    "qh_hip_rfft3": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p]),
    "qh_hip_irfft3": (c_int, [c_void_p, c_int, c_int, c_int, c_float, c_void_p]),
    "qh_hip_tv_prox": (c_int, [c_void_p, c_int, c_int, c_int, c_float, c_int, c_int, c_void_p]),
    "qh_hip_tv_prox_batch": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_float, c_int, c_int,
                                     c_void_p]),
    "qh_hip_tv_prox_dev_noalloc": (c_int, [c_void_p, c_int, c_int, c_int, c_float, c_int, c_int,
                                           c_void_p, c_void_p]),
}

_lib = None


def lib() -> ctypes.CDLL:
    """The loaded library; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH.name} is not built: run `python -m qoraholo_amd.build`")
        handle = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def symbols() -> list[str]:
    return list(_SIGNATURES)


def check(rc: int) -> None:
    if rc == 0:
        return
    msg = lib().qh_hip_last_error().decode()
    if rc == 1:
        raise ValueError(msg)
    if rc == 3:
        raise OutOfDeviceMemory(msg)
    if rc == 4:
        raise LogicError(msg)
    raise RuntimeError(msg)


def ptr(a) -> c_void_p:
    """Address of a numpy array's data (or an integer address)."""
    if isinstance(a, int):
        return c_void_p(a)
    return c_void_p(a.ctypes.data)
