"""The method_2 chunk planner (GeometryChunk, geometry_chunk.hpp:92-197) on the
CPU: the library's planner (xpr_plan_chunks, the engine's plan_chunks)
against the oracle's independent restatement, and against what the plan must
satisfy whatever the implementation:

* the chunks cover the layers 0..H-1 in order, without gaps or overlap;
* each chunk's detector rows hold the projection of every one of its layers
  (rows [floor(min)-1, ceil(max)+2) of its first and last layers' bounds);
* with overlap_fct 1, every layer outside a chunk whose projection reaches the
  chunk's rows is in its padding (brute force over all layers);
* each chunk fits the budget, and none could have taken its next layer;

at small sizes, several budgets, two modules, overlap_fct < 1, and at the
out-of-core workload's full size: 2048^2 detector -> 2048^3 grid, 2048 views,
sit 16, FP32, 16 GB device budget."""
import ctypes as C
import json
import math

import numpy as np
import pytest

from oracle import oracle as O
from xpulse_recon_amd import _lib


def _lib_or_skip():
    try:
        return _lib.load()
    except RuntimeError:
        pytest.skip("library not built")


def _params(spec: O.GeometrySpec, sit, budget_gb, storage="f32", overlap=1.0):
    det = {"px": spec.px, "rx": spec.rx, "sx": spec.sx, "sy": spec.sy,
           "module_number": spec.module_number}
    if spec.module_angle_offset is not None:
        det["module_angle_offset"] = spec.module_angle_offset
    return {"recon": {"method": "method_2", "it": 1, "sit": sit},
            "geometry": {"so": spec.so, "sd": spec.sd, "angle": spec.angle,
                         "heli_offset": spec.heli_offset},
            "detector": det, "method2": {"overlap_fct": overlap},
            "xpr": {"device_budget_gb": budget_gb, "storage": storage}}


def engine_plan(spec, sit, budget_gb, storage="f32", overlap=1.0):
    L = _lib_or_skip()
    out = (C.c_int64 * (6 * 4096))()
    n = C.c_int64()
    rc = L.xpr_plan_chunks(json.dumps(_params(spec, sit, budget_gb, storage, overlap)).encode(),
                           spec.dwidth, spec.dheight, spec.projections, out, 4096, C.byref(n))
    if rc != 0:
        raise RuntimeError(f"xpr_plan_chunks {rc}: {L.xpr_last_error().decode()}")
    return [list(out[6 * i:6 * i + 6]) for i in range(n.value)]


def oracle_plan(spec, sit, budget_gb, storage="f32", overlap=1.0):
    pd, grid = O.geometry(spec, False)
    sb = 4 if storage == "f32" else 2
    return O.plan_chunks(pd, grid, spec.dwidth, spec.dheight, spec.projections // sit, sb,
                         budget_gb * 1e9, overlap), pd, grid


def check_invariants(chunks, pd, grid, spec, sit, budget_gb, storage_bytes, overlap):
    H, W = grid["vheight"], grid["vwidth"]
    lmin, lmax = O.layer_row_bounds(pd, grid)
    a_s = spec.projections // sit
    dh = spec.dheight

    def rows(off, size):
        lo = min(lmin[off:off + size])
        hi = max(lmax[off:off + size])
        i_off = max(math.floor(lo) - 1, 0)
        return i_off, min(math.ceil(hi - i_off) + 2, dh - i_off)

    def pads(off, size, i_off, i_size):
        # brute force: every layer outside the chunk whose bounds reach its rows
        top = [l for l in range(off) if lmax[l] > i_off - 1]
        bot = [l for l in range(off + size, H) if lmin[l] < i_off + i_size + 1]
        return top, bot

    def nbytes(size, i_size, top, bot):
        return ((size + top + bot) * W * W * storage_bytes +
                i_size * spec.dwidth * a_s * (2 * storage_bytes + 8) + pd.size * 4)

    budget = budget_gb * 1e9
    assert chunks[0][0] == 0
    for a, b in zip(chunks[:-1], chunks[1:]):
        assert b[0] == a[0] + a[1]
    assert chunks[-1][0] + chunks[-1][1] == H
    for k, (off, size, i_off, i_size, top, bot) in enumerate(chunks):
        assert size >= 1 and 0 <= i_off and i_off + i_size <= dh
        # the rows hold the projection of every layer of the chunk (the
        # planner looks only at its first and last layers)
        assert (i_off, i_size) == rows(off, size)
        assert i_off == 0 or min(lmin[off:off + size]) > i_off
        assert i_off + i_size == dh or max(lmax[off:off + size]) < i_off + i_size
        assert nbytes(size, i_size, top, bot) < budget
        above, below = pads(off, size, i_off, i_size)
        if overlap == 1.0:
            assert above == list(range(off - top, off))
            assert below == list(range(off + size, off + size + bot))
        else:
            assert top == math.ceil(np.float32(len(above)) * np.float32(overlap))
            assert bot == math.ceil(np.float32(len(below)) * np.float32(overlap))
        if k + 1 < len(chunks):
            # none could have taken its next layer
            r_off, r_size = rows(off, size + 1)
            ab, be = pads(off, size + 1, r_off, r_size)
            fct = np.float32(overlap)
            t = math.ceil(np.float32(len(ab)) * fct)
            b = math.ceil(np.float32(len(be)) * fct)
            assert not nbytes(size + 1, r_size, t, b) < budget


SMALL = [
    # spec kwargs, sit, budget (GB)
    (dict(dwidth=40, dheight=32, projections=24, rx=1.5), 4, 0.30 * (40 * 40 * 32 * 4 + 6 * 40 * 32 * 16) / 1e9),
    (dict(dwidth=40, dheight=32, projections=24, rx=1.5), 4, 0.55 * (40 * 40 * 32 * 4 + 6 * 40 * 32 * 16) / 1e9),
    (dict(dwidth=64, dheight=96, projections=40, sx=0.2, sy=0.1), 5, 1.2e-3),
    (dict(dwidth=48, dheight=64, projections=16, module_number=2, module_angle_offset=4.0), 4, 6e-4),
    (dict(dwidth=96, dheight=128, projections=36, heli_offset=0.05), 6, 2.5e-3),
]


@pytest.mark.parametrize("overlap", [1.0, 0.5])
@pytest.mark.parametrize("case", range(len(SMALL)))
def test_engine_plan_equals_oracle_restatement(case, overlap):
    kw, sit, budget = SMALL[case]
    spec = O.GeometrySpec(**kw)
    got = engine_plan(spec, sit, budget, overlap=overlap)
    ref, pd, grid = oracle_plan(spec, sit, budget, overlap=overlap)
    assert got == ref
    assert len(got) >= 2
    check_invariants(got, pd, grid, spec, sit, budget, 4, overlap)


def test_engine_plan_f16_budget():
    kw, sit, budget = SMALL[2]
    spec = O.GeometrySpec(**kw)
    got = engine_plan(spec, sit, budget, storage="f16")
    ref, pd, grid = oracle_plan(spec, sit, budget, storage="f16")
    assert got == ref
    assert len(got) < len(engine_plan(spec, sit, budget))      # halved volume bytes
    check_invariants(got, pd, grid, spec, sit, budget, 2, 1.0)


def test_single_layer_over_budget_is_refused():
    L = _lib_or_skip()
    spec = O.GeometrySpec(dwidth=40, dheight=32, projections=24)
    out = (C.c_int64 * 60)()
    n = C.c_int64()
    rc = L.xpr_plan_chunks(json.dumps(_params(spec, 4, 1e-6)).encode(), 40, 32, 24, out, 10,
                           C.byref(n))
    assert rc == -3 and "Not enough GPU memory" in L.xpr_last_error().decode()
    with pytest.raises(RuntimeError, match="Not enough GPU memory"):
        oracle_plan(spec, 4, 1e-6)
    rc = L.xpr_plan_chunks(json.dumps(_params(spec, 4, 0.0)).encode(), 40, 32, 24, out, 10,
                           C.byref(n))
    assert rc == -1 and "device_budget_gb" in L.xpr_last_error().decode()


def test_out_of_core_plan_full_size():
    """2048^3 grid, 2048 views of 2048^2, sit 16 (128 views per sub-iteration),
    FP32, 16 GB: four chunks."""
    spec = O.GeometrySpec(dwidth=2048, dheight=2048, projections=2048)
    got = engine_plan(spec, 16, 16.0)
    ref, pd, grid = oracle_plan(spec, 16, 16.0)
    assert got == ref
    assert [(c[0], c[0] + c[1]) for c in got] == [(0, 681), (681, 1260), (1260, 1743),
                                                  (1743, 2048)]
    check_invariants(got, pd, grid, spec, 16, 16.0, 4, 1.0)
