"""ctypes front end of the CPU oracle (oracle/xpr_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker.  Only tests/, ``__graft_entry__.smoke()``
and bench.py's ``cpu_baseline`` leg import this module; the product package
never does.

The C library restates the reference's OpenCL kernels (KernelForward.cl,
KernelError.cl, KernelBackward.cl), its preprocessing (data_loader.hpp:48-101)
and its geometry (geometry.hpp:35-328) under one canonical arithmetic; see the
header comment of xpr_oracle.c.  Parity pins live in tests/test_oracle_*.py.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libxpr_oracle.so")
_lib = None


def build() -> None:
    """Compile the oracle library with its Makefile (in-tree, oracle/build/)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


class Geom(C.Structure):
    """orc_geom: the Forward/Backward kernel arguments (KernelForward.cl:10-21)."""

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
    ]


class Geom64(C.Structure):
    _fields_ = [
        ("det_width", C.c_uint32),
        ("n_views", C.c_uint32),
        ("modules", C.c_uint32),
        ("vmin", C.c_int32 * 3),
        ("vmax", C.c_int32 * 3),
        ("line_offset", C.c_int32),
        ("line_rows", C.c_int32),
        ("origin", C.c_double * 3),
        ("voxel", C.c_double * 3),
    ]


class GeomParams(C.Structure):
    _fields_ = [
        ("so", C.c_double), ("sd", C.c_double), ("angle", C.c_double),
        ("heli_offset", C.c_double),
        ("px", C.c_double), ("rx_px", C.c_double), ("sx", C.c_double), ("sy", C.c_double),
        ("module_number", C.c_int64),
        ("has_module_angle", C.c_int32),
        ("module_angle_offset", C.c_double),
        ("module_center_offset_z", C.c_double),
        ("dwidth", C.c_int64), ("dheight", C.c_int64), ("projections", C.c_int64),
        ("angle_list", C.POINTER(C.c_double)),
    ]


class Grid(C.Structure):
    _fields_ = [("vx", C.c_double), ("vwidth", C.c_int64), ("vheight", C.c_int64),
                ("orig", C.c_double * 3)]


class Ellipsoid(C.Structure):
    _fields_ = [("c", C.c_double * 3), ("a", C.c_double * 3), ("rot", C.c_double * 9),
                ("mu", C.c_double)]


_P = C.c_void_p


def _declare(L):
    L.orc_forward_f32.argtypes = [_P, _P, C.c_int, _P, C.POINTER(Geom), C.c_int]
    L.orc_forward_f32.restype = C.c_uint64
    L.orc_error_f32.argtypes = [_P, _P, C.c_int, _P, C.c_uint64, C.c_float]
    L.orc_backward_f32.argtypes = [_P, _P, _P, C.POINTER(Geom), C.c_int]
    L.orc_backward_f32.restype = C.c_uint64
    L.orc_hits_f32.argtypes = [_P, _P, C.POINTER(Geom)]
    L.orc_forward_f64.argtypes = [_P, _P, _P, C.POINTER(Geom64)]
    L.orc_error_f64.argtypes = [_P, _P, _P, C.c_uint64, C.c_double]
    L.orc_backward_f64.argtypes = [_P, _P, _P, C.POINTER(Geom64)]
    L.orc_backward_f64.restype = C.c_uint64
    L.orc_minmax_normal.argtypes = [_P, C.c_uint64, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.orc_preprocess.argtypes = [_P, C.c_uint64, C.c_int, C.c_int, C.c_float, C.c_float]
    L.orc_geometry.argtypes = [C.POINTER(GeomParams), C.c_int, _P, _P]
    L.orc_geometry.restype = Grid
    L.orc_project_physical.argtypes = [C.POINTER(GeomParams), C.c_int64, C.c_int64,
                                       C.c_double, C.c_double, C.c_double,
                                       C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.orc_line_integrals.argtypes = [C.POINTER(GeomParams), C.c_int64, C.POINTER(Ellipsoid),
                                     C.c_int, _P]
    L.orc_phantom_voxels.argtypes = [C.POINTER(Grid), C.POINTER(Ellipsoid), C.c_int, _P]


def _ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


# --------------------------------------------------------------------------
# geometry
# --------------------------------------------------------------------------
@dataclass
class GeometrySpec:
    """JSON ``geometry`` + ``detector`` sections (global_parameters.hpp:23-54)."""

    dwidth: int
    dheight: int
    projections: int
    so: float = 100.0
    sd: float = 1000.0
    angle: float = 360.0
    heli_offset: float = 0.0
    px: float = 0.1
    rx: float = 0.0            # pixels (multiplied by px at load, parameters.hpp:39)
    sx: float = 0.0            # mm (geometry.hpp:61,252)
    sy: float = 0.0
    module_number: int = 1
    module_angle_offset: float | None = None
    module_center_offset_z: float = 0.0
    angle_list: list | None = None

    def _params(self):
        p = GeomParams()
        p.so, p.sd, p.angle, p.heli_offset = self.so, self.sd, self.angle, self.heli_offset
        p.px, p.rx_px, p.sx, p.sy = self.px, self.rx, self.sx, self.sy
        p.module_number = self.module_number
        p.has_module_angle = 0 if self.module_angle_offset is None else 1
        p.module_angle_offset = 0.0 if self.module_angle_offset is None else self.module_angle_offset
        p.module_center_offset_z = self.module_center_offset_z
        p.dwidth, p.dheight, p.projections = self.dwidth, self.dheight, self.projections
        keep = None
        if self.angle_list is not None:
            keep = np.ascontiguousarray(self.angle_list, dtype=np.float64)
            p.angle_list = keep.ctypes.data_as(C.POINTER(C.c_double))
        return p, keep


def geometry(spec: GeometrySpec, fp64: bool):
    """ProjData for every (angle, module), shape [projections*modules, 20], and the grid."""
    p, keep = spec._params()
    n = spec.projections * spec.module_number
    if fp64:
        out = np.zeros((n, 20), np.float64)
        g = lib().orc_geometry(C.byref(p), 1, _ptr(out), None)
    else:
        out = np.zeros((n, 20), np.float32)
        g = lib().orc_geometry(C.byref(p), 0, None, _ptr(out))
    del keep
    grid = {"vx": g.vx, "vwidth": g.vwidth, "vheight": g.vheight, "orig": tuple(g.orig)}
    return out, grid


def project_physical(spec: GeometrySpec, a: int, m: int, x: float, y: float, z: float):
    p, keep = spec._params()
    u, v = C.c_double(), C.c_double()
    lib().orc_project_physical(C.byref(p), a, m, x, y, z, C.byref(u), C.byref(v))
    del keep
    return u.value, v.value


def line_integrals(spec: GeometrySpec, a: int, ellipsoids) -> np.ndarray:
    p, keep = spec._params()
    arr = (Ellipsoid * len(ellipsoids))(*ellipsoids)
    out = np.zeros((spec.dheight, spec.dwidth), np.float64)
    lib().orc_line_integrals(C.byref(p), a, arr, len(ellipsoids), _ptr(out))
    del keep
    return out


def phantom_voxels(grid: dict, ellipsoids) -> np.ndarray:
    g = Grid()
    g.vx, g.vwidth, g.vheight = grid["vx"], grid["vwidth"], grid["vheight"]
    for i in range(3):
        g.orig[i] = grid["orig"][i]
    arr = (Ellipsoid * len(ellipsoids))(*ellipsoids)
    out = np.zeros((grid["vheight"], grid["vwidth"], grid["vwidth"]), np.float64)
    lib().orc_phantom_voxels(C.byref(g), arr, len(ellipsoids), _ptr(out))
    return out


def make_ellipsoid(c, a, rot, mu) -> Ellipsoid:
    e = Ellipsoid()
    for i in range(3):
        e.c[i], e.a[i] = c[i], a[i]
    r = np.asarray(rot, np.float64).reshape(9)
    for i in range(9):
        e.rot[i] = r[i]
    e.mu = mu
    return e


# --------------------------------------------------------------------------
# kernels
# --------------------------------------------------------------------------
def make_geom(det_width, n_views, modules, vmin, vmax, line_offset, line_rows, origin, voxel,
              fp64=False):
    g = Geom64() if fp64 else Geom()
    g.det_width, g.n_views, g.modules = det_width, n_views, modules
    for i in range(3):
        g.vmin[i], g.vmax[i] = int(vmin[i]), int(vmax[i])
        g.origin[i] = origin[i]
        g.voxel[i] = voxel[i]
    g.line_offset, g.line_rows = line_offset, line_rows
    return g


def forward_f32(vol: np.ndarray, sums: np.ndarray, projdata: np.ndarray, g: Geom,
                threads: int = 1) -> int:
    assert vol.dtype == np.float32 and projdata.dtype == np.float32
    bits = {np.dtype(np.uint32): 32, np.dtype(np.uint64): 64}[sums.dtype]
    return lib().orc_forward_f32(_ptr(vol), _ptr(sums), bits, _ptr(projdata), C.byref(g), threads)


def error_f32(images: np.ndarray, sums: np.ndarray, weight: float, out: np.ndarray | None = None):
    bits = {np.dtype(np.uint32): 32, np.dtype(np.uint64): 64}[sums.dtype]
    if out is None:
        out = np.empty(sums.shape, np.float32)
    assert images.dtype == np.float32 and images.size == sums.size == out.size
    lib().orc_error_f32(_ptr(images), _ptr(sums), bits, _ptr(out), images.size, weight)
    return out


def backward_f32(vol: np.ndarray, err: np.ndarray, projdata: np.ndarray, g: Geom,
                 threads: int = 1) -> int:
    assert vol.dtype == np.float32 and err.dtype == np.float32
    return lib().orc_backward_f32(_ptr(vol), _ptr(err), _ptr(projdata), C.byref(g), threads)


def hits_f32(projdata: np.ndarray, g: Geom) -> np.ndarray:
    nx, nl, nk = (g.vmax[i] - g.vmin[i] for i in range(3))
    out = np.zeros((nl, nk, nx), np.uint32)
    lib().orc_hits_f32(_ptr(out), _ptr(projdata), C.byref(g))
    return out


def forward_f64(vol, sums, projdata, g: Geom64):
    lib().orc_forward_f64(_ptr(vol), _ptr(sums), _ptr(projdata), C.byref(g))


def error_f64(images, sums, weight, out=None):
    if out is None:
        out = np.empty(sums.shape, np.float64)
    lib().orc_error_f64(_ptr(images), _ptr(sums), _ptr(out), images.size, weight)
    return out


def backward_f64(vol, err, projdata, g: Geom64) -> int:
    return lib().orc_backward_f64(_ptr(vol), _ptr(err), _ptr(projdata), C.byref(g))


def minmax_normal(data: np.ndarray, mn=np.finfo(np.float32).max, mx=np.finfo(np.float32).tiny):
    a, b = C.c_float(mn), C.c_float(mx)
    lib().orc_minmax_normal(_ptr(data), data.size, C.byref(a), C.byref(b))
    return a.value, b.value


def preprocess(data: np.ndarray, normalize: bool, mlog: bool, mn: float = 0.0, mx: float = 1.0):
    assert data.dtype == np.float32
    lib().orc_preprocess(_ptr(data), data.size, int(normalize), int(mlog), mn, mx)
    return data


# --------------------------------------------------------------------------
# sub-iteration schedule (reconstruction_n.hpp:51-78)
# --------------------------------------------------------------------------
def sit_views(n_proj: int, sit: int, s: int) -> list:
    """Views of sub-iteration s: {s, s+S, ...} truncated to floor(N/S)
    (dataset.hpp:133-140, :225; geometry.hpp:306-307)."""
    a_s = n_proj // sit
    return list(range(s, n_proj, sit))[:a_s]


def reconstruct_naive(images: np.ndarray, projdata: np.ndarray, grid: dict, it: int, sit: int,
                      weight: float, weight_factor: float, jitter, modules: int = 1,
                      acc_bits: int = 64, threads: int = 1):
    """FP32 canonical naive schedule (reconstruction_n.hpp:51-78).

    images: [N, H, W] preprocessed projections; projdata: [N*modules, 20] f32;
    jitter: array [it*sit, 3] of origin offsets, one row per sub-iteration
    (src/reconstruction.hpp:268-272 with explicit draws)."""
    n, h, w = images.shape
    W, Hv = grid["vwidth"], grid["vheight"]
    a_s = n // sit
    vol = np.ones((Hv, W, W), np.float32)
    sums = np.zeros((a_s, h, w), np.uint64 if acc_bits == 64 else np.uint32)
    sums[...] = 1 << 22
    wgt = np.float32(weight)
    vx = grid["vx"]
    for mit in range(it):
        for s in range(sit):
            views = sit_views(n, sit, s)
            pd = np.ascontiguousarray(
                np.concatenate([projdata[v * modules:(v + 1) * modules] for v in views]))
            img = np.ascontiguousarray(images[views])
            o = [np.float32(np.float32(grid["orig"][i]) + np.float32(jitter[mit * sit + s][i]))
                 for i in range(3)]
            g = make_geom(w, a_s, modules, (0, 0, 0), (W, Hv, W), 0, h, o, (vx, vx, vx))
            if s > 0 or mit > 0:
                forward_f32(vol, sums, pd, g, threads)
            err = error_f32(img, sums, float(wgt))
            backward_f32(vol, err, pd, g, threads)
            sums[...] = 0
        wgt = np.float32(wgt * np.float32(weight_factor))
    return vol


def reconstruct_naive_f64(images: np.ndarray, projdata64: np.ndarray, grid: dict, it: int,
                          sit: int, weight: float, weight_factor: float, jitter,
                          modules: int = 1):
    """FP64 model of the naive schedule (reconstruction_n.hpp:51-78): the same
    voxel-driven NN operators in double precision with real-valued ray sums
    (no Q10.22 quantisation), FP64 ProjData and the same jitter offsets.
    The accuracy reference for storage-precision studies."""
    n, h, w = images.shape
    W, Hv = grid["vwidth"], grid["vheight"]
    a_s = n // sit
    vol = np.ones((Hv, W, W), np.float64)
    sums = np.ones((a_s, h, w), np.float64)
    wgt = float(weight)
    vx = grid["vx"]
    for mit in range(it):
        for s in range(sit):
            views = sit_views(n, sit, s)
            pd = np.ascontiguousarray(
                np.concatenate([projdata64[v * modules:(v + 1) * modules] for v in views]))
            img = np.ascontiguousarray(images[views], np.float64)
            o = [grid["orig"][i] + float(jitter[mit * sit + s][i]) for i in range(3)]
            g = make_geom(w, a_s, modules, (0, 0, 0), (W, Hv, W), 0, h, o, (vx, vx, vx), fp64=True)
            if s > 0 or mit > 0:
                forward_f64(vol, sums, pd, g)
            err = error_f64(img, sums, wgt)
            backward_f64(vol, err, pd, g)
            sums[...] = 0.0
        wgt = wgt * weight_factor
    return vol
