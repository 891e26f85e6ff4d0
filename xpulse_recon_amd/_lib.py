"""ctypes binding of the C-ABI in include/xpr_hip.h.

Loads ``xpulse_recon_amd/lib/libxpr_hip.so`` (built in-tree by
``python -m xpulse_recon_amd.build``).  There is no fallback: when the library
is missing every entry point raises, so no result can silently come from
anywhere but the HIP kernels.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libxpr_hip.so")

XPR_OK = 0
XPR_F32, XPR_F16, XPR_BF16 = 0, 1, 2
DTYPES = {"f32": XPR_F32, "f16": XPR_F16, "bf16": XPR_BF16}


class XprError(RuntimeError):
    """A negative XPR_E_* status returned by the library."""


class ProjArgs(C.Structure):
    """xpr_proj_args (include/xpr_hip.h): the Forward/Backward kernel
    arguments (KernelForward.cl:10-21) passed by value."""

    _fields_ = [
        ("det_width", C.c_uint32),
        ("n_views", C.c_uint32),
        ("modules", C.c_uint32),
        ("vmin", C.c_int32 * 3),
        ("vmax", C.c_int32 * 3),
        ("line_offset", C.c_int32),
        ("line_rows", C.c_int32),
        ("origin", C.c_float * 3),
        ("voxel", C.c_float * 3),
        ("interp", C.c_int32),
        ("view_index", C.c_void_p),
    ]


XPR_INTERP_NEAREST, XPR_INTERP_BILINEAR = 0, 1
XPR_ACC_U32, XPR_ACC_U64, XPR_ACC_F32 = 32, 64, -32


def proj_args(det_width, n_views, modules, vmin, vmax, line_offset, line_rows, origin, voxel,
              view_index=None, interp=XPR_INTERP_NEAREST):
    a = ProjArgs()
    a.det_width, a.n_views, a.modules = int(det_width), int(n_views), int(modules)
    for i in range(3):
        a.vmin[i], a.vmax[i] = int(vmin[i]), int(vmax[i])
        a.origin[i], a.voxel[i] = float(origin[i]), float(voxel[i])
    a.line_offset, a.line_rows = int(line_offset), int(line_rows)
    a.view_index = None if view_index is None else int(view_index)
    a.interp = int(interp)
    return a


class GeometryParams(C.Structure):
    """xpr_geometry_params: JSON geometry/detector sections (global_parameters.hpp:23-54)."""

    _fields_ = [
        ("so", C.c_double), ("sd", C.c_double), ("angle", C.c_double),
        ("heli_offset", C.c_double),
        ("px", C.c_double), ("rx", C.c_double), ("sx", C.c_double), ("sy", C.c_double),
        ("module_angle_offset", C.c_double), ("module_center_offset_z", C.c_double),
        ("roll", C.c_double), ("pitch", C.c_double), ("yaw", C.c_double),
        ("module_number", C.c_int64),
        ("dwidth", C.c_int64), ("dheight", C.c_int64), ("projections", C.c_int64),
        ("has_module_angle", C.c_int32), ("fp64", C.c_int32),
        ("angle_list", C.POINTER(C.c_double)),
    ]


class Grid(C.Structure):
    _fields_ = [("vx", C.c_double), ("vwidth", C.c_int64), ("vheight", C.c_int64),
                ("orig", C.c_double * 3)]


class Ellipsoid(C.Structure):
    _fields_ = [("c", C.c_double * 3), ("a", C.c_double * 3), ("rot", C.c_double * 9),
                ("mu", C.c_double)]


class Job(C.Structure):
    """xpr_job: one reconstruction submitted to a server (host buffers)."""

    _fields_ = [
        ("geometry", GeometryParams),
        ("it", C.c_int64), ("sit", C.c_int64),
        ("weight", C.c_float), ("weight_factor", C.c_float),
        ("mlog", C.c_int32), ("normalize", C.c_int32),
        ("storage", C.c_int32), ("final_fp32", C.c_int32), ("acc_bits", C.c_int32),
        ("jitter", C.c_int32), ("jitter_seed", C.c_int64), ("slab_height", C.c_int64),
        ("projections", C.c_void_p), ("volume_out", C.c_void_p),
    ]


class JobStats(C.Structure):
    _fields_ = [
        ("submit_ms", C.c_double), ("start_ms", C.c_double), ("end_ms", C.c_double),
        ("device_ms", C.c_double), ("device_bytes", C.c_uint64), ("voxel_views", C.c_uint64),
        ("streamed", C.c_int32), ("status", C.c_int32),
    ]


# name -> (restype, argtypes)
_P = C.c_void_p
_SIGS = {
    "xpr_abi_version": (C.c_int, []),
    "xpr_last_error": (C.c_char_p, []),
    "xpr_set_device": (C.c_int, [C.c_int]),
    "xpr_mem_info": (C.c_int, [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "xpr_alloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
    "xpr_free": (C.c_int, [_P]),
    "xpr_host_alloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
    "xpr_host_free": (C.c_int, [_P]),
    "xpr_alloc_async": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t, _P]),
    "xpr_free_async": (C.c_int, [_P, _P]),
    "xpr_upload": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "xpr_download": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "xpr_stream_create": (C.c_int, [C.POINTER(C.c_void_p)]),
    "xpr_stream_destroy": (C.c_int, [_P]),
    "xpr_stream_sync": (C.c_int, [_P]),
    "xpr_event_create": (C.c_int, [C.POINTER(C.c_void_p)]),
    "xpr_event_destroy": (C.c_int, [_P]),
    "xpr_event_record": (C.c_int, [_P, _P]),
    "xpr_stream_wait_event": (C.c_int, [_P, _P]),
    "xpr_event_sync": (C.c_int, [_P]),
    "xpr_event_elapsed_ms": (C.c_int, [_P, _P, C.POINTER(C.c_float)]),
    "xpr_fill_u32": (C.c_int, [_P, _P, C.c_uint32, C.c_uint64]),
    "xpr_fill_u64": (C.c_int, [_P, _P, C.c_uint64, C.c_uint64]),
    "xpr_fill": (C.c_int, [_P, _P, C.c_int, C.c_float, C.c_uint64]),
    "xpr_convert": (C.c_int, [_P, _P, C.c_int, _P, C.c_int, C.c_uint64]),
    "xpr_forward": (C.c_int, [_P, _P, C.c_int, _P, C.c_int, _P, C.POINTER(ProjArgs)]),
    "xpr_error": (C.c_int, [_P, _P, C.c_int, _P, C.c_int, _P, C.c_int, C.c_uint64, C.c_float]),
    "xpr_backward": (C.c_int, [_P, _P, C.c_int, _P, C.c_int, _P, C.POINTER(ProjArgs), _P]),
    "xpr_backward_max": (C.c_int, [_P, _P, C.c_int, _P, C.c_int, _P, C.POINTER(ProjArgs), _P, _P]),
    "xpr_subiteration": (C.c_int, [_P, _P, C.c_int, _P, C.c_int, _P, C.POINTER(ProjArgs),
                                   C.c_float, C.c_int, _P, C.c_int, _P, C.c_int, _P]),
    "xpr_subiteration_static": (C.c_int, [_P, _P, C.c_int, _P, C.c_int, _P,
                                          C.POINTER(ProjArgs), C.c_float, C.c_int, C.c_int,
                                          C.c_int, _P]),
    "xpr_tune_forward": (C.c_int, [_P, _P, C.c_int, _P, C.c_int, _P, C.POINTER(ProjArgs),
                                   C.POINTER(C.c_int)]),
    "xpr_tune_backward": (C.c_int, [_P, _P, C.c_int, _P, C.c_int, _P, C.POINTER(ProjArgs),
                                    C.POINTER(C.c_int)]),
    "xpr_tune_reset": (C.c_int, []),
    "xpr_check_fast_paths": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "xpr_q22_to_float": (C.c_int, [_P, _P, C.c_int, _P, C.c_uint64]),
    "xpr_valu_ceiling": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "xpr_valu_rates": (C.c_int, [_P, _P, C.c_int]),
    "xpr_scale": (C.c_int, [_P, _P, C.c_uint64, C.c_float]),
    "xpr_minmax_init": (C.c_int, [_P, _P]),
    "xpr_minmax_normal": (C.c_int, [_P, _P, C.c_uint64, _P]),
    "xpr_minmax_read": (C.c_int, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "xpr_preprocess": (C.c_int, [_P, _P, _P, C.c_int, C.c_uint64, C.c_int, C.c_int,
                                 C.c_float, C.c_float]),
    "xpr_preprocess_views": (C.c_int, [_P, _P, C.c_uint64, _P, C.c_int, C.c_uint64, C.c_uint32,
                                       C.c_int, C.c_int, C.c_float, C.c_float]),
    "xpr_geometry": (C.c_int, [C.POINTER(GeometryParams), _P, C.POINTER(Grid)]),
    "xpr_geometry_frames": (C.c_int, [C.POINTER(GeometryParams), _P]),
    "xpr_hit_bound": (C.c_int, [_P, C.POINTER(ProjArgs), C.POINTER(C.c_double)]),
    "xpr_reconstruct": (C.c_int, [C.c_char_p, C.c_char_p, C.c_size_t]),
    "xpr_template_json": (C.c_int, [C.c_char_p, C.c_size_t]),
    "xpr_reconstruct_dev": (C.c_int, [C.c_char_p, _P, C.c_int64, C.c_int64, C.c_int64, _P,
                                      C.c_char_p, C.c_size_t]),
    "xpr_plan_chunks": (C.c_int, [C.c_char_p, C.c_int64, C.c_int64, C.c_int64, _P, C.c_int64,
                                  C.POINTER(C.c_int64)]),
    "xpr_reconstruct_mem": (C.c_int, [C.c_char_p, _P, C.c_int64, C.c_int64, C.c_int64, _P,
                                      C.c_char_p, C.c_size_t]),
    "xpr_server_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint64, C.c_int, C.c_int]),
    "xpr_server_submit": (C.c_int, [_P, C.POINTER(Job), C.POINTER(C.c_uint64)]),
    "xpr_server_wait": (C.c_int, [_P, C.c_uint64, C.POINTER(JobStats)]),
    "xpr_job_device_bytes": (C.c_int, [C.POINTER(Job), C.POINTER(C.c_uint64)]),
    "xpr_server_destroy": (C.c_int, [_P]),
    "xpr_synth_phantom": (C.c_int, [_P, C.c_double, C.c_int64, C.c_int64,
                                    C.POINTER(C.c_double), _P, C.c_int, _P]),
    "xpr_synth_line_integrals": (C.c_int, [_P, _P, C.c_uint32, C.c_uint32, C.c_uint32,
                                           C.c_uint32, _P, C.c_int, _P]),
}

_lib = None


def load():
    """Load the HIP library; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"HIP library not built: {LIB_PATH} is missing "
                "(run `python -m xpulse_recon_amd.build`)")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(lib, name)
            f.restype, f.argtypes = res, args
        _lib = lib
    return _lib


def exported_symbols():
    return list(_SIGS)


def check(rc: int, what: str = "") -> None:
    if rc != XPR_OK:
        msg = load().xpr_last_error().decode(errors="replace")
        raise XprError(f"{what or 'xpr call'} failed ({rc}): {msg}")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)
