"""MI355X-native iterative cone-beam CT hot path (XPRecon's Forward / Error /
Backward operators) behind a C-ABI (include/xpr_hip.h).

This module is the Python face of the C-ABI for tests and benchmarks: device
memory comes from torch tensors (plumbing only), every computation runs in the
HIP library ``lib/libxpr_hip.so``.  The operator names follow the reference's
``Reconstruction::forward/error/backward`` (src/reconstruction.hpp:294-312).
"""
from __future__ import annotations

import ctypes as C

from . import _lib
from ._lib import ProjArgs, XprError, proj_args  # noqa: F401

__all__ = ["forward", "error", "backward", "subiteration", "fill", "fill_u32", "fill_u64", "convert",
           "preprocess", "preprocess_views", "minmax_normal", "proj_args", "ProjArgs", "XprError",
           "dtype_code", "acc_bits", "tune_forward", "tune_backward", "tune_reset",
           "check_fast_paths", "hit_bound"]


def _stream(stream):
    if stream is None:
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return C.c_void_p(stream)
    return C.c_void_p(stream.cuda_stream)


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def dtype_code(t) -> int:
    import torch
    m = {torch.float32: _lib.XPR_F32, torch.float16: _lib.XPR_F16,
         torch.bfloat16: _lib.XPR_BF16}
    if t.dtype not in m:
        raise XprError(f"unsupported storage dtype {t.dtype}")
    return m[t.dtype]


def acc_bits(t) -> int:
    """The accumulator a sums tensor stands for: 32- or 64-bit integers are
    Q.22 sums (XPR_ACC_U32 / XPR_ACC_U64), float32 the opt-in float sums
    (XPR_ACC_F32)."""
    import torch
    m = {torch.int32: _lib.XPR_ACC_U32, torch.uint32: _lib.XPR_ACC_U32,
         torch.int64: _lib.XPR_ACC_U64, torch.uint64: _lib.XPR_ACC_U64,
         torch.float32: _lib.XPR_ACC_F32}
    if t.dtype not in m:
        raise XprError(f"sums must be a 32- or 64-bit integer or a float32 tensor, got {t.dtype}")
    return m[t.dtype]


def _need_cuda(*ts):
    for t in ts:
        if not t.is_cuda:
            raise XprError("tensor is not on the GPU")


def _check_sizes(volume, det, projdata, a: ProjArgs):
    """Host-side shape check before any launch: the buffers must hold the slab,
    the detector views and the ProjData records the arguments describe."""
    n = [a.vmax[i] - a.vmin[i] for i in range(3)]
    if min(n) < 0:
        raise XprError("volumeMax < volumeMin")
    if volume.numel() < n[0] * n[1] * n[2] or not volume.is_contiguous():
        raise XprError(f"volume holds {volume.numel()} voxels, slab needs {n[0] * n[1] * n[2]}")
    need = a.n_views * a.line_rows * a.det_width
    if a.view_index:
        need = 0        # the caller's view list addresses the detector buffer
    if det.numel() < need or not det.is_contiguous():
        raise XprError(f"detector buffer holds {det.numel()} pixels, needs {need}")
    if projdata.numel() < 20 * a.n_views * a.modules or projdata.dtype.itemsize != 4:
        raise XprError("projdata must hold 20 float32 per (view, module)")


def forward(volume, sums, projdata, args: ProjArgs, stream=None) -> None:
    """Forward projection: sums[view,row,col] += trunc(v*2^22) (KernelForward.cl:10-55);
    float32 sums (the opt-in float accumulator) add the same Q.22 values
    times 2^-22 with float additions."""
    _need_cuda(volume, sums, projdata)
    _check_sizes(volume, sums, projdata, args)
    _lib.call("xpr_forward", _stream(stream), _ptr(volume), dtype_code(volume), _ptr(sums),
              acc_bits(sums), _ptr(projdata), C.byref(args))


def error(images, sums, err, weight: float, n: int | None = None, stream=None) -> None:
    """err = log(max(img / max(sum*2^-22, eps), eps)) * weight (KernelError.cl:8-24)."""
    _need_cuda(images, sums, err)
    n = images.numel() if n is None else n
    if n > min(images.numel(), sums.numel(), err.numel()):
        raise XprError("error: n exceeds a buffer")
    _lib.call("xpr_error", _stream(stream), _ptr(images), dtype_code(images), _ptr(sums),
              acc_bits(sums), _ptr(err), dtype_code(err), n, float(weight))


MAX_SLOTS = 1024          # XPR_MAX_SLOTS


def backward(volume, err, projdata, args: ProjArgs, empty=None, stream=None, max_bits=None) -> None:
    """Backward projection + multiplicative update (KernelBackward.cl:8-41).
    max_bits (optional, a device tensor of MAX_SLOTS 32-bit elements, cleared
    by the caller): its slots are raised to the bits of the largest voxel
    values after the update; the largest slot is the volume's
    (xpr_backward_max)."""
    _need_cuda(volume, err, projdata)
    _check_sizes(volume, err, projdata, args)
    if max_bits is not None:
        if not max_bits.is_cuda or max_bits.numel() < MAX_SLOTS or max_bits.element_size() != 4:
            raise XprError("max_bits must be a device tensor of MAX_SLOTS 32-bit elements")
        _lib.call("xpr_backward_max", _stream(stream), _ptr(volume), dtype_code(volume), _ptr(err),
                  dtype_code(err), _ptr(projdata), C.byref(args),
                  None if empty is None else _ptr(empty), _ptr(max_bits))
        return
    _lib.call("xpr_backward", _stream(stream), _ptr(volume), dtype_code(volume), _ptr(err),
              dtype_code(err), _ptr(projdata), C.byref(args),
              None if empty is None else _ptr(empty))


def hit_bound(projdata, args: ProjArgs) -> float:
    """Bound on the voxels of a forward launch of args whose nearest pixel is
    one pixel of one view, for any jitter of args.origin up to one voxel
    (xpr_hit_bound; projdata: host float32 records).  Below 2^32 / the largest
    Q.22 voxel value, 32-bit sums equal 64-bit sums."""
    import numpy as np
    pd = np.ascontiguousarray(projdata, dtype=np.float32)
    if pd.size < 20 * args.n_views * max(args.modules, 1):
        raise XprError("projdata holds fewer records than args.n_views * args.modules")
    out = C.c_double()
    _lib.call("xpr_hit_bound", pd.ctypes.data_as(C.c_void_p), C.byref(args), C.byref(out))
    return out.value


def subiteration(volume, images, projdata, args: ProjArgs, weight: float, first_pass: bool,
                 sums, err, empty=None, stream=None) -> None:
    """One sub-iteration of the slab (reconstruction_n.hpp:66-76): sums <- 0
    and Forward (none on the first pass: sums = 1.0), Error into err,
    Backward + update.  sums and err are the workspace (their contents on
    entry do not matter)."""
    _need_cuda(volume, images, projdata, sums, err)
    _check_sizes(volume, sums, projdata, args)
    _check_sizes(volume, err, projdata, args)
    _check_sizes(volume, images, projdata, args)
    _lib.call("xpr_subiteration", _stream(stream), _ptr(volume), dtype_code(volume),
              _ptr(images), dtype_code(images), _ptr(projdata), C.byref(args), float(weight),
              int(bool(first_pass)), _ptr(sums), acc_bits(sums), _ptr(err), dtype_code(err),
              None if empty is None else _ptr(empty))


def subiteration_static(volume, images, projdata, args: ProjArgs, weight: float,
                        first_pass: bool, acc: int = 64, err_dtype: str = "f32", empty=None,
                        stream=None) -> None:
    """subiteration() with the sums and error workspace in the library's
    static device scratch (no device allocation; calls that overlap on one
    GPU race)."""
    _need_cuda(volume, images, projdata)
    _check_sizes(volume, images, projdata, args)
    _lib.call("xpr_subiteration_static", _stream(stream), _ptr(volume), dtype_code(volume),
              _ptr(images), dtype_code(images), _ptr(projdata), C.byref(args), float(weight),
              int(bool(first_pass)), int(acc), _lib.DTYPES[err_dtype],
              None if empty is None else _ptr(empty))


def fill(t, value: float, stream=None) -> None:
    _need_cuda(t)
    _lib.call("xpr_fill", _stream(stream), _ptr(t), dtype_code(t), float(value), t.numel())


def fill_u32(t, value: int, stream=None) -> None:
    _lib.call("xpr_fill_u32", _stream(stream), _ptr(t), value & 0xFFFFFFFF, t.numel())


def fill_u64(t, value: int, stream=None) -> None:
    _lib.call("xpr_fill_u64", _stream(stream), _ptr(t), value & 0xFFFFFFFFFFFFFFFF, t.numel())


def convert(dst, src, stream=None) -> None:
    n = min(dst.numel(), src.numel())
    _lib.call("xpr_convert", _stream(stream), _ptr(dst), dtype_code(dst), _ptr(src),
              dtype_code(src), n)


def minmax_normal(data, stream=None):
    """Global min/max over isnormal() samples (data_loader.hpp:55-74)."""
    import torch
    mm = torch.empty(2, dtype=torch.int32, device=data.device)
    s = _stream(stream)
    _lib.call("xpr_minmax_init", s, _ptr(mm))
    _lib.call("xpr_minmax_normal", s, _ptr(data), data.numel(), _ptr(mm))
    host = mm.cpu()
    a, b = C.c_float(), C.c_float()
    _lib.call("xpr_minmax_read", C.c_void_p(host.data_ptr()), C.byref(a), C.byref(b))
    return a.value, b.value


def preprocess(data, out, normalize: bool, mlog: bool, mn: float = 0.0, mx: float = 1.0,
               stream=None) -> None:
    """Normalise and -log clamp (data_loader.hpp:80-86) into out's dtype."""
    if out.numel() < data.numel():
        raise XprError("preprocess: output smaller than input")
    _lib.call("xpr_preprocess", _stream(stream), _ptr(data), _ptr(out), dtype_code(out),
              data.numel(), int(normalize), int(mlog), float(mn), float(mx))


def preprocess_views(data, out, step: int, normalize: bool, mlog: bool, mn: float = 0.0,
                     mx: float = 1.0, stream=None) -> None:
    """preprocess of the views 0, step, 2 step, ... of data ([N][rows][cols]),
    as many as out ([n][rows][cols]) holds, in one launch."""
    n_views, per = out.shape[0], out[0].numel()
    if data[0].numel() != per or (n_views - 1) * step >= data.shape[0]:
        raise XprError("preprocess_views: shapes do not match")
    _lib.call("xpr_preprocess_views", _stream(stream), _ptr(data), step * per, _ptr(out),
              dtype_code(out), per, n_views, int(normalize), int(mlog), float(mn), float(mx))


def tune_forward(volume, sums_scratch, projdata, args: ProjArgs, stream=None) -> int:
    """Time the forward's tile capacities on these buffers (sums_scratch gets
    several passes) and keep the fastest for this launch shape; returns it."""
    _need_cuda(volume, sums_scratch, projdata)
    _check_sizes(volume, sums_scratch, projdata, args)
    out = C.c_int()
    _lib.call("xpr_tune_forward", _stream(stream), _ptr(volume), dtype_code(volume),
              _ptr(sums_scratch), acc_bits(sums_scratch), _ptr(projdata), C.byref(args),
              C.byref(out))
    return out.value


def tune_backward(volume_scratch, err, projdata, args: ProjArgs, stream=None) -> int:
    """Time the backward's voxels-per-lane shapes on these buffers
    (volume_scratch gets several updates) and keep the fastest; returns it."""
    _need_cuda(volume_scratch, err, projdata)
    _check_sizes(volume_scratch, err, projdata, args)
    out = C.c_int()
    _lib.call("xpr_tune_backward", _stream(stream), _ptr(volume_scratch),
              dtype_code(volume_scratch), _ptr(err), dtype_code(err), _ptr(projdata),
              C.byref(args), C.byref(out))
    return out.value


def tune_reset() -> None:
    _lib.call("xpr_tune_reset")


def check_fast_paths(stream=None):
    """Counts of float32 inputs where the exact fast forms differ from the
    canonical ones: (reciprocal in [1e-6, 1e6], v_cvt_rpi vs roundf on
    (-0.5, 2^23), reciprocal outside the range).  The first two are 0 on a
    correct device."""
    out = (C.c_uint64 * 3)()
    _lib.call("xpr_check_fast_paths", _stream(stream), out)
    return int(out[0]), int(out[1]), int(out[2])
