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
        ("roll", C.c_double), ("pitch", C.c_double), ("yaw", C.c_double),
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
    L.orc_forward_f64nn.argtypes = [_P, _P, _P, C.POINTER(Geom), C.c_int, C.c_int]
    L.orc_backward_f64nn.argtypes = [_P, _P, _P, C.POINTER(Geom), C.c_int]
    L.orc_backward_f64nn.restype = C.c_uint64
    L.orc_backward_bilinear_f32.argtypes = [_P, _P, _P, C.POINTER(Geom), C.c_int]
    L.orc_backward_bilinear_f32.restype = C.c_uint64
    L.orc_backward_bilinear_f64nn.argtypes = [_P, _P, _P, C.POINTER(Geom), C.c_int]
    L.orc_backward_bilinear_f64nn.restype = C.c_uint64
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
    roll: float = 0.0          # deg, module rotation extension (FP64 geometry only)
    pitch: float = 0.0
    yaw: float = 0.0

    def _params(self):
        p = GeomParams()
        p.so, p.sd, p.angle, p.heli_offset = self.so, self.sd, self.angle, self.heli_offset
        p.px, p.rx_px, p.sx, p.sy = self.px, self.rx, self.sx, self.sy
        p.module_number = self.module_number
        p.has_module_angle = 0 if self.module_angle_offset is None else 1
        p.module_angle_offset = 0.0 if self.module_angle_offset is None else self.module_angle_offset
        p.module_center_offset_z = self.module_center_offset_z
        p.dwidth, p.dheight, p.projections = self.dwidth, self.dheight, self.projections
        p.roll, p.pitch, p.yaw = self.roll, self.pitch, self.yaw
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
        if g.vwidth < 0:
            raise ValueError("roll/pitch/yaw: the oracle builds these geometries in FP64 only")
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


def forward_f64nn(vol, sums, projdata32, g: Geom, threads: int = 1, q22: bool = False):
    """q22: truncate each contribution to a multiple of 2^-22 first
    (FLOAT_TO_FIXED, constants.cl:22), modelling the Q.22 sums' quantisation."""
    assert vol.dtype == np.float64 and sums.dtype == np.float64 and projdata32.dtype == np.float32
    lib().orc_forward_f64nn(_ptr(vol), _ptr(sums), _ptr(projdata32), C.byref(g), threads,
                            int(q22))


def backward_f64nn(vol, err, projdata32, g: Geom, threads: int = 1) -> int:
    assert vol.dtype == np.float64 and err.dtype == np.float64 and projdata32.dtype == np.float32
    return lib().orc_backward_f64nn(_ptr(vol), _ptr(err), _ptr(projdata32), C.byref(g), threads)


def backward_bilinear_f32(vol, err, projdata, g: Geom, threads: int = 1) -> int:
    """Opt-in bilinear backprojector, FP32 canonical form (the intent of the
    unused INTERP_BILIN branch, KernelForward.cl:38-53): taps (floor u, floor
    v)..(+1, +1) clamped to the pixels the nearest-neighbour test accepts, over
    the voxel-views it accepts; three fmaf lerps, update
    v * expf(sum / n).  Returns the count of voxels with no view."""
    assert vol.dtype == np.float32 and err.dtype == np.float32 and projdata.dtype == np.float32
    return lib().orc_backward_bilinear_f32(_ptr(vol), _ptr(err), _ptr(projdata), C.byref(g),
                                           threads)


def backward_bilinear_f64nn(vol, err, projdata32, g: Geom, threads: int = 1) -> int:
    """FP64 model of the bilinear backprojector on the float32 accept set and
    taps (double weights, sample, sum and update)."""
    assert vol.dtype == np.float64 and err.dtype == np.float64 and projdata32.dtype == np.float32
    return lib().orc_backward_bilinear_f64nn(_ptr(vol), _ptr(err), _ptr(projdata32), C.byref(g),
                                             threads)


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
                      acc_bits: int = 64, threads: int = 1, interp: str = "nearest"):
    """FP32 canonical naive schedule (reconstruction_n.hpp:51-78).

    images: [N, H, W] preprocessed projections; projdata: [N*modules, 20] f32;
    jitter: array [it*sit, 3] of origin offsets, one row per sub-iteration
    (src/reconstruction.hpp:268-272 with explicit draws).  interp "bilinear":
    the opt-in bilinear backprojector (backward_bilinear_f32)."""
    bwd = backward_bilinear_f32 if interp == "bilinear" else backward_f32
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
            bwd(vol, err, pd, g, threads)
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


# --------------------------------------------------------------------------
# method_2: chunk planner and chunked schedule (geometry_chunk.hpp:92-197,
# reconstruction_chunk.hpp:36-131)
# --------------------------------------------------------------------------
def layer_row_bounds(projdata: np.ndarray, grid: dict):
    """computeLayerYBounds (geometry_chunk.hpp:120-145): per layer, the lowest
    and highest detector row coordinate of the four vertical edges of the
    volume (x = +-orig.x, z = +-orig.z) at the layer's top and bottom, over
    every ProjData record.  float32, evaluated as the sums
    ((m4*x + m5*y) + m6*z) + m7 over ((m12*x + m13*y) + m14*z) + m15."""
    f = np.float32
    H, vx = int(grid["vheight"]), f(grid["vx"])
    o = [f(grid["orig"][i]) for i in range(3)]
    cx = np.array([o[0], -o[0], o[0], -o[0]], f)
    cz = np.array([-o[2], -o[2], o[2], o[2]], f)
    m = projdata.reshape(-1, 20).astype(f)
    lmin = np.empty(H, f)
    lmax = np.empty(H, f)
    for l0 in range(0, H, 256):
        l1 = min(H, l0 + 256)
        # y of the top and bottom edge of every layer of the block: [L, 2]
        y = (o[1] + (np.arange(l0, l1)[:, None] + np.arange(2)[None, :]).astype(f) * vx).astype(f)
        y = y[:, :, None, None]                                     # [L, 2, 1, 1]
        r = lambda k: m[:, k][None, None, :, None]                  # [1, 1, R, 1]
        Y = ((r(4) * cx + r(5) * y) + r(6) * cz) + r(7)
        Wh = ((r(12) * cx + r(13) * y) + r(14) * cz) + r(15)
        v = (Y / Wh).reshape(l1 - l0, -1)
        lmin[l0:l1] = v.min(axis=1)
        lmax[l0:l1] = v.max(axis=1)
    return lmin, lmax


def plan_chunks(projdata: np.ndarray, grid: dict, det_width: int, det_height: int,
                concurrent: int, storage_bytes: int, budget: float, overlap_fct: float = 1.0):
    """GeometryChunk::computeChunks and its helpers (geometry_chunk.hpp:92-197)
    under the build's intended semantics, as [vOffset, vSize, iOffset, iSize,
    vPadTop, vPadBot] per chunk:

    * every layer 0..H-1 belongs to one chunk (the reference's loop runs over
      its first..last complete layers and drops the layer that does not fit,
      :94-111); a layer that does not fit starts the next chunk;
    * a chunk's detector rows are [floor(min)-1, ceil(max - iOffset)+2) of its
      first and last layers' bounds, clipped to the detector (:164-168);
    * padding: the layers above (below) whose row bounds reach the rows
      (:175-188), the bottom one up to layer H-1 (the reference stops at
      H-2), each count scaled by overlap_fct and rounded up;
    * a chunk fits when its device buffers (padded volume in the storage
      type; images, u64 sums and error over the chunk's rows) plus the
      ProjData are below the budget (the reference's sizeInMemory, :62-72,
      with this build's buffers)."""
    f = np.float32
    lmin, lmax = layer_row_bounds(projdata, grid)
    H, W = int(grid["vheight"]), int(grid["vwidth"])
    pd_bytes = projdata.size * 4

    def rows_of(off, size):
        mn = min(lmin[off], lmin[off + size - 1])
        mx = max(lmax[off], lmax[off + size - 1])
        i_off = int(max(np.floor(mn) - f(1), f(0)))
        i_size = int(max(min(np.ceil(f(mx) - f(i_off)) + f(2), f(det_height - i_off)), f(0)))
        return i_off, i_size

    def pads(off, size, i_off, i_size):
        top = 0
        while off - top > 0 and lmax[off - top - 1] > f(i_off - 1):
            top += 1
        bot = 0
        while off + size + bot < H and lmin[off + size + bot] < f(i_off + i_size + 1):
            bot += 1
        fct = f(overlap_fct)
        return int(np.ceil(f(top) * fct)), int(np.ceil(f(bot) * fct))

    def make(off, size):
        i_off, i_size = rows_of(off, size)
        top, bot = pads(off, size, i_off, i_size)
        return [off, size, i_off, i_size, top, bot]

    def nbytes(c):
        vol = (c[1] + c[4] + c[5]) * W * W * storage_bytes
        img = c[3] * det_width * concurrent
        return vol + img * (storage_bytes + 8 + storage_bytes)

    chunks = []
    off, size = 0, 0
    cur = None
    for _ in range(H):
        trial = make(off, size + 1)
        if nbytes(trial) + pd_bytes < budget:
            cur, size = trial, size + 1
            continue
        if size == 0:
            raise RuntimeError("Not enough GPU memory for the chunk division")
        chunks.append(cur)
        off, size = off + size, 1
        cur = make(off, 1)
        if not nbytes(cur) + pd_bytes < budget:
            raise RuntimeError("Not enough GPU memory for the chunk division")
    chunks.append(cur)
    return chunks


def reconstruct_chunked(images: np.ndarray, projdata: np.ndarray, grid: dict, it: int, sit: int,
                        weight: float, weight_factor: float, jitter, chunks, modules: int = 1,
                        threads: int = 1):
    """FP32 canonical method_2 schedule (reconstruction_chunk.hpp:36-131): each
    chunk [vOffset, vSize, iOffset, iSize, vPadTop, vPadBot] is an independent
    naive reconstruction of its padded layers against the detector rows
    [iOffset, iOffset+iSize) (kernels bounded by those rows, the Error over
    W*iSize*A_s pixels), with the next it*sit rows of jitter; only the chunk's
    own layers are kept.  images: [N, H, W] preprocessed."""
    n, h, w = images.shape
    W = grid["vwidth"]
    a_s = n // sit
    vx = grid["vx"]
    out = np.empty((grid["vheight"], W, W), np.float32)
    k = 0
    for (v_off, v_size, i_off, i_size, top, bot) in chunks:
        l0, l1 = v_off - top, v_off + v_size + bot
        vol = np.ones((l1 - l0, W, W), np.float32)
        sums = np.full((a_s, i_size, w), 1 << 22, np.uint64)
        wgt = np.float32(weight)
        for mit in range(it):
            for s in range(sit):
                views = sit_views(n, sit, s)
                pd = np.ascontiguousarray(
                    np.concatenate([projdata[v * modules:(v + 1) * modules] for v in views]))
                img = np.ascontiguousarray(images[views][:, i_off:i_off + i_size])
                o = [np.float32(np.float32(grid["orig"][i]) + np.float32(jitter[k][i]))
                     for i in range(3)]
                k += 1
                g = make_geom(w, a_s, modules, (0, l0, 0), (W, l1, W), i_off, i_size, o,
                              (vx, vx, vx))
                if s > 0 or mit > 0:
                    forward_f32(vol, sums, pd, g, threads)
                err = error_f32(img, sums, float(wgt))
                backward_f32(vol, err, pd, g, threads)
                sums[...] = 0
            wgt = np.float32(wgt * np.float32(weight_factor))
        out[v_off:v_off + v_size] = vol[top:top + v_size]
    return out


def reproject(volume: np.ndarray, projdata: np.ndarray, grid: dict, det_width: int,
              det_height: int, n_proj: int, sit: int, jitter, modules: int = 1,
              samples: int = 32, threads: int = 1):
    """forward_exec (src/reconstruction.hpp:66-158): the volume divided by
    `samples` and zeroed outside the inscribed circle of each layer
    (sqrt((i-W/2)^2 + (j-W/2)^2) >= W/2 - 1, float32), forward projected
    `samples` times per sub-iteration with the jitter rows in order
    (sub-iteration-major) into Q.22 sums; returns {view: float32 image},
    image = float(sum) / 2^22 (FIXED_TO_FLOAT)."""
    f = np.float32
    W, H = int(grid["vwidth"]), int(grid["vheight"])
    vol = (volume.astype(f) / f(samples)).astype(f)
    i = np.arange(W, dtype=f)
    di = (i - f(W) * f(0.5)).astype(f)
    r = np.sqrt((di[None, :] * di[None, :] + di[:, None] * di[:, None]).astype(f)).astype(f)
    vol[:, r >= f(W) * f(0.5) - f(1.0)] = f(0.0)
    vol = np.ascontiguousarray(vol)
    a_s = n_proj // sit
    vx = grid["vx"]
    h, w = int(det_height), int(det_width)
    out = {}
    k = 0
    for s in range(sit):
        views = sit_views(n_proj, sit, s)
        pd = np.ascontiguousarray(
            np.concatenate([projdata[v * modules:(v + 1) * modules] for v in views]))
        sums = np.zeros((a_s, h, w), np.uint64)
        for _ in range(samples):
            o = [f(f(grid["orig"][c]) + f(jitter[k][c])) for c in range(3)]
            k += 1
            g = make_geom(w, a_s, modules, (0, 0, 0), (W, H, W), 0, h, o, (vx, vx, vx))
            forward_f32(vol, sums, pd, g, threads)
        img = (sums.astype(np.float64).astype(f) / f(1 << 22)).astype(f)
        for j, v in enumerate(views):
            out[v] = img[j]
    return out


def reconstruct_naive_f64nn(images: np.ndarray, projdata: np.ndarray, grid: dict, it: int,
                            sit: int, weight: float, weight_factor: float, jitter,
                            modules: int = 1, threads: int = 1, q22: bool = False,
                            interp: str = "nearest"):
    """The naive schedule (reconstruction_n.hpp:51-78) with FP64 values on the
    canonical float32 pixel choice: the same float32 ProjData, float32 origin
    (orig + jitter) and voxel positions as the FP32 path decide which pixel
    each voxel-view meets; sums, error, weights and the update are double
    (first-pass sums 1.0).  q22 truncates each forward contribution to a
    multiple of 2^-22 as the Q.22 sums and the float sums (whose tiles bin
    the same Q.22 integers) do; without it there is no quantisation.  interp
    "bilinear": the
    bilinear backprojector's FP64 model (backward_bilinear_f64nn)."""
    bwd = backward_bilinear_f64nn if interp == "bilinear" else backward_f64nn
    n, h, w = images.shape
    W, Hv = grid["vwidth"], grid["vheight"]
    a_s = n // sit
    vol = np.ones((Hv, W, W), np.float64)
    sums = np.ones((a_s, h, w), np.float64)
    wgt = float(np.float32(weight))
    vx = grid["vx"]
    for mit in range(it):
        for s in range(sit):
            views = sit_views(n, sit, s)
            pd = np.ascontiguousarray(
                np.concatenate([projdata[v * modules:(v + 1) * modules] for v in views]))
            img = np.ascontiguousarray(images[views], np.float64)
            o = [np.float32(np.float32(grid["orig"][i]) + np.float32(jitter[mit * sit + s][i]))
                 for i in range(3)]
            g = make_geom(w, a_s, modules, (0, 0, 0), (W, Hv, W), 0, h, o, (vx, vx, vx))
            if s > 0 or mit > 0:
                forward_f64nn(vol, sums, pd, g, threads, q22)
            err = error_f64(img, sums, wgt)
            bwd(vol, err, pd, g, threads)
            sums[...] = 0.0
        wgt = float(np.float32(np.float32(wgt) * np.float32(weight_factor)))
    return vol
