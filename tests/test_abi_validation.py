"""Argument checking of the C-ABI, on the CPU: every call below is rejected
before any device work, with the documented XPR_E_* code and a message in
xpr_last_error() (the reference throws std::runtime_error from the same
places, gpu/error.hpp:20-26; no exception crosses this ABI)."""
import ctypes as C

import pytest

from xpulse_recon_amd import _lib
from xpulse_recon_amd._lib import ProjArgs

E_INVALID, E_UNSUPPORTED = -1, -4


def _lib_or_skip():
    try:
        return _lib.load()
    except RuntimeError:
        pytest.skip("library not built")


def _err(L):
    return L.xpr_last_error().decode()


def _args(vmin=(0, 0, 0), vmax=(8, 8, 8), modules=1):
    a = ProjArgs()
    a.det_width, a.n_views, a.modules = 8, 2, modules
    for i in range(3):
        a.vmin[i], a.vmax[i] = vmin[i], vmax[i]
        a.voxel[i] = 0.01
    a.line_offset, a.line_rows = 0, 8
    return a


DUMMY = C.c_void_p(0x1000)    # never dereferenced: every call fails its checks first


def test_projector_argument_checks():
    L = _lib_or_skip()
    assert L.xpr_forward(None, DUMMY, 0, DUMMY, 64, DUMMY, None) == E_INVALID
    assert "xpr_proj_args" in _err(L)
    a = _args(vmin=(0, 5, 0), vmax=(8, 4, 8))
    assert L.xpr_forward(None, DUMMY, 0, DUMMY, 64, DUMMY, C.byref(a)) == E_INVALID
    assert "volumeMax" in _err(L)
    a = _args()
    assert L.xpr_forward(None, DUMMY, 0, DUMMY, 16, DUMMY, C.byref(a)) == E_INVALID
    assert "acc_bits" in _err(L)
    assert L.xpr_forward(None, DUMMY, 7, DUMMY, 64, DUMMY, C.byref(a)) == E_UNSUPPORTED
    assert L.xpr_forward(None, None, 0, DUMMY, 64, DUMMY, C.byref(a)) == E_INVALID
    assert L.xpr_backward(None, DUMMY, 0, DUMMY, 9, DUMMY, C.byref(a), None) == E_UNSUPPORTED
    a0 = _args(modules=0)
    assert L.xpr_backward(None, DUMMY, 0, DUMMY, 0, DUMMY, C.byref(a0), None) == E_INVALID
    assert "modules" in _err(L)
    big = _args(vmax=(8, 8, (1 << 24) + 1))
    assert L.xpr_backward(None, DUMMY, 0, DUMMY, 0, DUMMY, C.byref(big), None) == E_INVALID
    out = C.c_int()
    assert L.xpr_tune_forward(None, None, 0, DUMMY, 64, DUMMY, C.byref(a), C.byref(out)) == E_INVALID
    assert L.xpr_tune_backward(None, DUMMY, 0, DUMMY, 0, DUMMY, None, C.byref(out)) == E_INVALID


def test_error_argument_checks():
    L = _lib_or_skip()
    # in-place error (E written over the sums) only for u32 sums and f32 error
    assert L.xpr_error(None, DUMMY, 0, DUMMY, 64, DUMMY, 0, 10, C.c_float(1.0)) == E_INVALID
    assert "in-place" in _err(L)
    assert L.xpr_error(None, DUMMY, 0, DUMMY, 48, C.c_void_p(0x2000), 0, 10,
                       C.c_float(1.0)) == E_INVALID
    assert L.xpr_error(None, None, 0, DUMMY, 64, C.c_void_p(0x2000), 0, 10,
                       C.c_float(1.0)) == E_INVALID


def test_server_submit_checks():
    L = _lib_or_skip()
    h = C.c_void_p()
    assert L.xpr_server_create(C.byref(h), 0, 1, 0) == E_INVALID
    assert L.xpr_server_create(C.byref(h), 1 << 30, 0, 0) == E_INVALID
    assert L.xpr_server_create(C.byref(h), 1 << 30, 2, 0) == 0
    try:
        job = _lib.Job()
        job.geometry.dwidth, job.geometry.dheight, job.geometry.projections = 8, 8, 4
        job.it, job.sit, job.acc_bits = 1, 2, 64
        jid = C.c_uint64()
        assert L.xpr_server_submit(h, C.byref(job), C.byref(jid)) == E_INVALID   # no buffers
        assert "projections" in _err(L)
        job.projections, job.volume_out = DUMMY, DUMMY
        job.sit = 5                                    # 4 projections / 5 sub-iterations
        assert L.xpr_server_submit(h, C.byref(job), C.byref(jid)) == E_INVALID
        job.sit, job.acc_bits = 2, 16
        assert L.xpr_server_submit(h, C.byref(job), C.byref(jid)) == E_INVALID
        job.acc_bits, job.storage = 64, 5
        assert L.xpr_server_submit(h, C.byref(job), C.byref(jid)) == E_UNSUPPORTED
        assert L.xpr_server_wait(h, 12345, None) == E_INVALID                   # unknown id
    finally:
        assert L.xpr_server_destroy(h) == 0


def test_job_device_bytes_orders_storage():
    """The admission figure grows with the storage width and with the FP32
    copies of a 16-bit job's last sub-iteration, by the volume and sample
    sizes those imply."""
    L = _lib_or_skip()

    def need(storage, final):
        job = _lib.Job()
        g = job.geometry
        g.dwidth, g.dheight, g.projections, g.module_number = 64, 32, 40, 1
        g.so, g.sd, g.px = 100.0, 1000.0, 0.1
        job.it, job.sit, job.acc_bits, job.storage, job.final_fp32 = 1, 4, 64, storage, final
        out = C.c_uint64()
        assert L.xpr_job_device_bytes(C.byref(job), C.byref(out)) == 0
        return out.value

    V, P, As = 64 * 64 * 32, 64 * 32, 10
    f32, f16, f16f = need(0, 1), need(1, 0), need(1, 1)
    assert f32 - f16 == V * 2 + 4 * As * P * 2 + As * P * 2      # volume, images, error at 2 B less
    assert f16f - f16 == V * 4 + As * P * 8                      # float32 volume, image and error
    assert need(2, 0) == f16


def test_reconstruct_mem_rejects_missing_buffers():
    L = _lib_or_skip()
    assert L.xpr_reconstruct_mem(b"{}", None, 8, 8, 4, DUMMY, None, 0) == -1
    assert L.xpr_reconstruct_mem(b"{}", DUMMY, 0, 8, 4, DUMMY, None, 0) == -1
    assert L.xpr_reconstruct_mem(b"{not json", DUMMY, 8, 8, 4, DUMMY, None, 0) == -2
