"""ctypes binding of libctlproj.so (the C ABI declared in include/ctlproj.h).

The library is loaded from ``ctl_amd/lib/libctlproj.so``; a missing library raises instead of
falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libctlproj.so"

OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_NO_DEVICE = 2
ERR_DEVICE = 3
ERR_OUT_OF_MEMORY = 4
ERR_NOT_CONFIGURED = 5
ERR_NO_DATA = 6
ERR_OVER_BUDGET = 7
ERR_INJECTED = 8
ERR_TIMEOUT = 9
ERR_IO = 10

F32, F16, BF16 = 0, 1, 2
CARM, TUBULAR = 0, 1
SHORT_SCAN, WOBBLE, CIRCLE_PLUS_LINE, HELICAL, AXIAL = 0, 1, 2, 3, 4


class Settings(C.Structure):
    _fields_ = [
        ("rays_per_pixel", C.c_uint32 * 2),
        ("ray_sampling", C.c_float),
        ("volume_upsampling", C.c_uint32),
        ("interpolate", C.c_int32),
        ("tile_w", C.c_uint32),
        ("tile_h", C.c_uint32),
        ("unroll", C.c_uint32),
        ("blocks_per_cu", C.c_uint32),
        ("lds_chunk", C.c_int32),
        ("lds_texels", C.c_uint32),
    ]


class NrrdInfo(C.Structure):
    _fields_ = [
        ("version", C.c_uint32),
        ("ndim", C.c_uint32),
        ("nsizes", C.c_uint32),
        ("elem", C.c_int32),
        ("elem_bytes", C.c_uint32),
        ("sizes", C.c_uint64 * 4),
        ("count", C.c_uint64),
        ("header_bytes", C.c_uint64),
        ("has_spacing", C.c_int32),
        ("spacing", C.c_double * 3),
        ("has_origin", C.c_int32),
        ("origin", C.c_double * 3),
        ("data_complete", C.c_int32),
    ]


class Detector(C.Structure):
    _fields_ = [
        ("pixels", C.c_uint32 * 2),
        ("pixel_mm", C.c_double * 2),
        ("modules", C.c_uint32),
        ("cylindrical", C.c_int32),
        ("angulation", C.c_double),
        ("module_spacing", C.c_double),
        ("skew", C.c_double),
    ]


class Gantry(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("carm_span", C.c_double),
        ("sdd", C.c_double),
        ("sid", C.c_double),
        ("rotation", C.c_double),
        ("pitch", C.c_double),
        ("tilt", C.c_double),
    ]


class Trajectory(C.Structure):
    _fields_ = [("type", C.c_int32), ("p", C.c_double * 6)]


class ServerStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("submitted", "completed", "failed", "rejected", "views",
                                          "launches")] + \
               [(n, C.c_double) for n in ("p50_ms", "p99_ms", "mean_ms", "max_ms", "p50_single_ms",
                                          "p99_single_ms", "p50_multi_ms", "p99_multi_ms", "span_s")] + \
               [(n, C.c_uint32) for n in ("slots", "cu_masked_slots", "max_scan_launches_in_flight",
                                          "reserved")] + \
               [("drr_launches_beside_scan", C.c_uint64)]


PROGRESS_FN = C.CFUNCTYPE(None, C.c_int, C.c_void_p)

_vp = C.c_void_p
_pp = C.POINTER(C.c_void_p)
_u32p = C.POINTER(C.c_uint32)
_f32p = C.POINTER(C.c_float)
_f64p = C.POINTER(C.c_double)
_u64p = C.POINTER(C.c_uint64)
_i = C.c_int
_u32 = C.c_uint32
_sz = C.c_size_t

# name -> (restype, argtypes); every symbol of include/ctlproj.h (and the batch/server headers).
SIGNATURES: dict[str, tuple] = {
    "ctlproj_settings_default": (None, [C.POINTER(Settings)]),
    "ctlproj_last_error": (C.c_char_p, []),
    "ctlproj_device_count": (_i, [C.POINTER(_i)]),
    "ctlproj_context_create": (_i, [_i, _pp]),
    "ctlproj_context_destroy": (_i, [_vp]),
    "ctlproj_volume_create": (_i, [_vp, _f32p, _u32p, _f32p, _f32p, _i, _pp]),
    "ctlproj_volume_create_from_device": (_i, [_vp, _vp, _u32p, _f32p, _f32p, _i, _pp]),
    "ctlproj_volume_destroy": (_i, [_vp]),
    "ctlproj_volume_bytes": (_i, [_vp, C.POINTER(_sz)]),
    "ctlproj_projector_create": (_i, [_vp, C.POINTER(Settings), _pp]),
    "ctlproj_projector_destroy": (_i, [_vp]),
    "ctlproj_projector_set_settings": (_i, [_vp, C.POINTER(Settings)]),
    "ctlproj_projector_get_settings": (_i, [_vp, C.POINTER(Settings)]),
    "ctlproj_projector_configure": (_i, [_vp, _f64p, _u32, _u32, _u32, _u32]),
    "ctlproj_project": (_i, [_vp, _vp, _f32p, _sz, PROGRESS_FN, _vp]),
    "ctlproj_project_host_volume": (_i, [_vp, _f32p, _u32p, _f32p, _f32p, _f32p, _sz]),
    "ctlproj_project_device": (_i, [_vp, _vp, _vp, _sz]),
    "ctlproj_project_device_noalloc": (_i, [_vp, _vp, _vp, _sz]),
    "ctlproj_projector_synchronize": (_i, [_vp]),
    "ctlproj_projector_last_kernel_ms": (_i, [_vp, _f32p]),
    "ctlproj_projector_count_samples": (_i, [_vp, _vp, _u64p]),
    "ctlproj_projector_count_taken": (_i, [_vp, _vp, _vp, _sz, _u64p]),
    "ctlproj_encode_trajectory": (_i, [C.POINTER(Gantry), C.POINTER(Detector), C.POINTER(Trajectory),
                                       _u32, _f64p, _sz]),
    "ctlproj_encode_carm_poses": (_i, [C.c_double, C.POINTER(Detector), _f64p, _f64p, _u32, _f64p, _sz]),
    "ctlproj_phantom_ellipsoids": (_i, [_u32p, _f32p, _u32, C.c_uint64, C.c_float, C.c_float,
                                        C.c_float, C.c_float, C.c_float, _f32p]),
    "ctlproj_debug_inject_failure": (_i, [_vp, _i]),
    "ctlproj_settings_optimized_for": (_i, [_f32p, C.c_double, C.c_double, C.POINTER(Settings)]),
    "ctlproj_encode_trajectory_shifted": (_i, [C.POINTER(Gantry), C.POINTER(Detector), C.POINTER(Trajectory),
                                               _u32, _f64p, _f64p, _sz]),
    "ctlproj_encode_carm_poses_shifted": (_i, [C.c_double, C.POINTER(Detector), _f64p, _f64p, _u32, _f64p, _f64p,
                                               _sz]),
    "ctlproj_project_composite": (_i, [_vp, C.POINTER(_vp), _u32, _f32p, _sz]),
    "ctlproj_project_composite_device": (_i, [_vp, C.POINTER(_vp), _u32, _vp, _sz]),
    "ctlproj_average_projections": (_i, [_vp, _vp, _u32, _sz, _i, _vp]),
    "ctlproj_decode_geometry": (_i, [_f64p, _u32, _u32, _u32, _u32, _i, C.c_double, C.c_double, _f64p, _f64p,
                                     _f64p, _f64p, _f64p, _f64p]),
    "ctlproj_encode_generic": (_i, [_f64p, _f64p, _f64p, _f64p, _f64p, _u32, _u32, _u32, _u32, _f64p, _sz]),
    "ctlproj_den_info": (_i, [C.c_char_p, _u32p, _u32p, _u32p, _u32p]),
    "ctlproj_den_read": (_i, [C.c_char_p, _vp, _sz]),
    "ctlproj_den_write": (_i, [C.c_char_p, _vp, _u32, _u32, _u32, _u32]),
    "ctlproj_nrrd_read_header": (_i, [C.c_char_p, C.POINTER(NrrdInfo), C.c_char_p, _sz]),
    "ctlproj_nrrd_read": (_i, [C.c_char_p, _vp, _sz]),
    "ctlproj_nrrd_write": (_i, [C.c_char_p, _vp, _i, _u32, C.POINTER(C.c_uint64), C.c_char_p, _f64p, _f64p,
                                C.c_char_p]),
    "ctlproj_diag_gather_probe": (_i, [_vp, _vp, _i, _vp, _sz, _f32p, _f64p]),
    "ctlproj_spectral_extinction": (_i, [_vp, _vp, _u32, _u32, _sz, _f32p, _f64p, _f32p, _u32, _f64p, _vp]),
    "ctlproj_combine_views": (_i, [_vp, _vp, _f32p, _vp, _f32p, _u32, _sz, _vp]),
    "ctlproj_project_spectral": (_i, [_vp, C.POINTER(_vp), _u32, _f32p, _f64p, _f32p, _u32, _f64p, _f32p, _sz]),
    "ctlproj_project_linear_dynamic": (_i, [_vp, _vp, _vp, _f64p, _f32p, _sz]),
    "ctlproj_slicer_default_side": (_i, [_u32p, _u32p]),
    "ctlproj_slice": (_i, [_vp, _f64p, _f64p, _u32, _u32, _u32, C.c_float, _f32p, _sz]),
    "ctlproj_slice_device": (_i, [_vp, _f64p, _f64p, _u32, _u32, _u32, C.c_float, _vp, _sz]),
    "ctlproj_resampler_default_ranges": (_i, [_vp, _f32p]),
    "ctlproj_resample": (_i, [_vp, _f32p, _f32p, _u32, _f32p, _u32, _f32p, _u32, _f32p, _sz]),
    "ctlproj_sample": (_i, [_vp, _f32p, _f32p, _sz, _f32p]),
    "ctlproj_sample_device": (_i, [_vp, _f32p, _vp, _sz, _vp]),
    "ctlproj_radon3d_default_slice": (_i, [_u32p, _f32p, _u32p, _f32p]),
    "ctlproj_radon3d_side_for_resolution": (_i, [_u32, C.c_float, C.c_float, _u32p]),
    "ctlproj_radon3d_plane_integrals": (_i, [_vp, _f64p, _u32, _f32p, _u32, _u32, C.c_float, _f32p, _sz]),
    "ctlproj_radon3d_plane_integrals_device": (_i, [_vp, _f64p, _u32, _f32p, _u32, _u32, C.c_float, _vp, _sz]),
    "ctlproj_radon3d_sample_transform": (_i, [_vp, _f32p, _u32, _f32p, _u32, _f32p, _u32, _u32, C.c_float,
                                              _f32p, _sz]),
    "ctlproj_projector_autotune": (_i, [_vp, _vp, _f32p]),
    "ctlproj_autotune_table": (_i, [C.c_char_p, _sz, C.POINTER(_sz)]),
    "ctlproj_batch_create": (_i, [_vp, _pp]),
    "ctlproj_batch_destroy": (_i, [_vp]),
    "ctlproj_batch_add": (_i, [_vp, _vp, _f64p, _u32, _u32, _u32, _u32, C.POINTER(Settings), _vp, _sz,
                               C.POINTER(_u32)]),
    "ctlproj_batch_finalize": (_i, [_vp, _i]),
    "ctlproj_batch_launch": (_i, [_vp]),
    "ctlproj_batch_synchronize": (_i, [_vp]),
    "ctlproj_batch_last_ms": (_i, [_vp, _f32p]),
    "ctlproj_batch_info": (_i, [_vp, _u32p, _u32p, _u32p, _u32p, _u64p]),
    "ctlproj_batch_count_samples": (_i, [_vp, _u64p]),
    "ctlproj_batch_shape": (_i, [_vp, _vp]),
    "ctlproj_server_create": (_i, [_vp, _sz, C.c_uint64, _pp]),
    "ctlproj_server_destroy": (_i, [_vp]),
    "ctlproj_server_submit": (_i, [_vp, _vp, _f64p, _u32, _u32, _u32, _u32, C.POINTER(Settings), _vp, _sz,
                                   _u64p]),
    "ctlproj_server_wait": (_i, [_vp, C.c_uint64, C.c_double, C.POINTER(C.c_double)]),
    "ctlproj_server_drain": (_i, [_vp]),
    "ctlproj_server_get_stats": (_i, [_vp, C.POINTER(ServerStats)]),
    "ctlproj_server_reset_stats": (_i, [_vp]),
}

_lib: C.CDLL | None = None


class CtlProjError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, status: int, message: str):
        super().__init__(f"[status {status}] {message}")
        self.status = status
        self.message = message


def lib() -> C.CDLL:
    """Loads libctlproj.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        # torch's wheel carries its own libamdhip64.so.7 (same soname as ROCm's): load torch
        # first so that both use one HIP runtime in this process.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: run `python -m ctl_amd.build` first")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    msg = lib().ctlproj_last_error()
    return msg.decode() if msg else ""


def check(status: int) -> None:
    if status != OK:
        raise CtlProjError(status, last_error())


def fptr(a) -> C.POINTER(C.c_float):
    return a.ctypes.data_as(_f32p)


def dptr(a) -> C.POINTER(C.c_double):
    return a.ctypes.data_as(_f64p)


def u32arr(*v) -> C.Array:
    return (C.c_uint32 * len(v))(*v)


def f32arr(*v) -> C.Array:
    return (C.c_float * len(v))(*v)
