"""Cone-beam geometry (Geometry, dataset/geometry.hpp:35-328) through the C-ABI.

``projdata(...)`` returns the 20-float ProjData record of every (angle, module)
and the reconstruction grid, computed by the library's C++ geometry in float32
(the reference's arithmetic) or FP64.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib


@dataclass
class Geometry:
    """The JSON ``geometry`` + ``detector`` keys (global_parameters.hpp:23-54)
    plus the detector size and projection count read from the input TIFFs."""

    dwidth: int
    dheight: int
    projections: int
    so: float = 100.0
    sd: float = 1000.0
    angle: float = 360.0
    heli_offset: float = 0.0
    px: float = 0.1
    rx: float = 0.0                     # pixels
    sx: float = 0.0                     # mm
    sy: float = 0.0                     # mm
    module_number: int = 1
    module_angle_offset: float | None = None
    module_center_offset_z: float = 0.0
    roll: float = 0.0                   # deg, extension (ignored by the reference)
    pitch: float = 0.0
    yaw: float = 0.0
    angle_list: list = field(default_factory=list)

    def _params(self, fp64: bool):
        p = _lib.GeometryParams()
        for k in ("so", "sd", "angle", "heli_offset", "px", "rx", "sx", "sy",
                  "module_center_offset_z", "roll", "pitch", "yaw"):
            setattr(p, k, float(getattr(self, k)))
        p.module_number, p.dwidth, p.dheight = self.module_number, self.dwidth, self.dheight
        p.projections = self.projections
        p.has_module_angle = int(self.module_angle_offset is not None)
        p.module_angle_offset = float(self.module_angle_offset or 0.0)
        p.fp64 = int(fp64)
        keep = None
        if self.angle_list:
            if len(self.angle_list) != self.projections:
                raise ValueError("The size of the angle list is different from the number "
                                 "of projections")
            keep = np.ascontiguousarray(self.angle_list, np.float64)
            p.angle_list = keep.ctypes.data_as(C.POINTER(C.c_double))
        return p, keep

    def projdata(self, fp64: bool = False):
        """(ProjData float32 [projections*module_number, 20], grid dict)."""
        p, keep = self._params(fp64)
        out = np.zeros((self.projections * self.module_number, 20), np.float32)
        g = _lib.Grid()
        _lib.call("xpr_geometry", C.byref(p), out.ctypes.data_as(C.c_void_p), C.byref(g))
        del keep
        grid = {"vx": g.vx, "vwidth": g.vwidth, "vheight": g.vheight, "orig": tuple(g.orig)}
        return out, grid

    def frames(self) -> np.ndarray:
        """FP64 detector frames [projections*module_number, 14] (source,
        upper-left corner, column step, row step, vpl, vpb)."""
        p, keep = self._params(True)
        out = np.zeros((self.projections * self.module_number, 14), np.float64)
        _lib.call("xpr_geometry_frames", C.byref(p), out.ctypes.data_as(C.c_void_p))
        del keep
        return out

    def sit_projdata(self, pd: np.ndarray, sit: int, s: int) -> np.ndarray:
        """ProjData of sub-iteration s: views {s, s+S, ...} truncated to
        floor(N/S) (geometry.hpp:306-323, dataset.hpp:225)."""
        a_s = self.projections // sit
        m = self.module_number
        views = list(range(s, self.projections, sit))[:a_s]
        return np.ascontiguousarray(np.concatenate([pd[v * m:(v + 1) * m] for v in views]))
