"""Concurrent reconstructions under a device-memory budget (the C-ABI's
``xpr_server_*``).  The reference runs one reconstruction per process; a
server admits jobs in arrival order while their resident footprint fits the
budget and runs each on its own streams, overlapping the others.

Each job carries raw projections (as read from the TIFFs, float32
``[projections, dheight, dwidth]``) and receives the float32 volume
``[vheight, vwidth, vwidth]`` (layer 0 at the top), with the same
preprocessing, schedule arithmetic and seeded jitter as ``XPRecon``.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .geometry import Geometry

_STORAGE = {"f32": _lib.XPR_F32, "f16": _lib.XPR_F16, "bf16": _lib.XPR_BF16}


@dataclass
class JobResult:
    volume: np.ndarray
    submit_ms: float
    start_ms: float
    end_ms: float
    device_ms: float
    device_bytes: int
    voxel_views: int
    streamed: bool

    @property
    def latency_ms(self) -> float:
        return self.end_ms - self.submit_ms


class _Pending:
    def __init__(self, job, geo_keep, projections, volume):
        self.job, self.geo_keep, self.projections, self.volume = job, geo_keep, projections, volume


def _job(geometry: Geometry, projections: np.ndarray, *, it: int, sit: int, weight: float,
         weight_factor: float, mlog: bool, normalize: bool, storage: str, final_fp32: bool,
         acc_bits: int, jitter: bool, jitter_seed: int, slab_height: int, fp64: bool):
    if storage not in _STORAGE:
        raise _lib.XprError(f"unknown storage {storage!r}")
    proj = np.ascontiguousarray(projections, np.float32)
    if proj.shape != (geometry.projections, geometry.dheight, geometry.dwidth):
        raise _lib.XprError(f"projections shape {proj.shape} does not match the geometry")
    gp, keep = geometry._params(fp64)
    _, grid = geometry.projdata()
    vol = np.empty((grid["vheight"], grid["vwidth"], grid["vwidth"]), np.float32)
    j = _lib.Job()
    j.geometry = gp
    j.it, j.sit = it, sit
    j.weight, j.weight_factor = weight, weight_factor
    j.mlog, j.normalize = int(mlog), int(normalize)
    j.storage, j.final_fp32, j.acc_bits = _STORAGE[storage], int(final_fp32), acc_bits
    j.jitter, j.jitter_seed, j.slab_height = int(jitter), jitter_seed, slab_height
    j.projections = proj.ctypes.data_as(C.c_void_p)
    j.volume_out = vol.ctypes.data_as(C.c_void_p)
    return j, keep, proj, vol


def job_device_bytes(geometry: Geometry, *, sit: int = 10, storage: str = "f32",
                     final_fp32: bool = True, acc_bits: int = 64) -> int:
    """Resident device memory of a job (its admission figure)."""
    dummy = np.zeros((geometry.projections, geometry.dheight, geometry.dwidth), np.float32)
    j, keep, _, _ = _job(geometry, dummy, it=1, sit=sit, weight=1.0, weight_factor=1.0,
                         mlog=True, normalize=False, storage=storage, final_fp32=final_fp32,
                         acc_bits=acc_bits, jitter=False, jitter_seed=0, slab_height=64,
                         fp64=False)
    out = C.c_uint64()
    _lib.call("xpr_job_device_bytes", C.byref(j), C.byref(out))
    del keep
    return int(out.value)


class Server:
    """Admission-controlled job server on one device."""

    def __init__(self, budget_bytes: int, max_concurrent: int = 8, device: int = 0):
        h = C.c_void_p()
        _lib.call("xpr_server_create", C.byref(h), int(budget_bytes), int(max_concurrent),
                  int(device))
        self._h = h
        self._jobs: dict[int, _Pending] = {}

    def submit(self, geometry: Geometry, projections: np.ndarray, *, it: int = 1, sit: int = 10,
               weight: float = 1.0, weight_factor: float = 0.9, mlog: bool = True,
               normalize: bool = False, storage: str = "f32", final_fp32: bool = True,
               acc_bits: int = 64, jitter: bool = True, jitter_seed: int = 0,
               slab_height: int = 64, fp64: bool = False) -> int:
        j, keep, proj, vol = _job(geometry, projections, it=it, sit=sit, weight=weight,
                                  weight_factor=weight_factor, mlog=mlog, normalize=normalize,
                                  storage=storage, final_fp32=final_fp32, acc_bits=acc_bits,
                                  jitter=jitter, jitter_seed=jitter_seed,
                                  slab_height=slab_height, fp64=fp64)
        jid = C.c_uint64()
        _lib.call("xpr_server_submit", self._h, C.byref(j), C.byref(jid))
        self._jobs[jid.value] = _Pending(j, keep, proj, vol)
        return jid.value

    def wait(self, job_id: int) -> JobResult:
        st = _lib.JobStats()
        rc = _lib.load().xpr_server_wait(self._h, C.c_uint64(job_id), C.byref(st))
        pend = self._jobs.pop(job_id, None)
        _lib.check(rc, "xpr_server_wait")
        return JobResult(pend.volume, st.submit_ms, st.start_ms, st.end_ms, st.device_ms,
                         int(st.device_bytes), int(st.voxel_views), bool(st.streamed))

    def close(self) -> None:
        if self._h is not None:
            _lib.call("xpr_server_destroy", self._h)
            self._h = None
            self._jobs.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
