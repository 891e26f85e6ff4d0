"""Serving: jobs submitted to an admission-controlled server (xpr_server_*).

Bars: a job's volume is bit-identical to XPRecon's naive schedule on the same
projections, geometry and jitter seed; bit-identical alone and in a mix of
concurrent jobs of other sizes; a job over the budget runs with slab
streaming and is still bit-identical (method_1a equals naive exactly); with a
budget that holds one job at a time the jobs run one after another."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from tests.tiff_util import read_f32, write_f32

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "xpulse_recon_amd", "bin", "XPRecon")


def _projections(w, h, n, seed, so=100.0, sd=1000.0):
    spec = O.GeometrySpec(dwidth=w, dheight=h, projections=n, so=so, sd=sd, rx=1.5)
    _, grid = O.geometry(spec, False)
    from xpulse_recon_amd import synth
    ell = synth.random_ellipsoids(seed, 5, w * grid["vx"], 64, mu_lo=0.02, mu_hi=0.06)
    oe = [O.make_ellipsoid(e["c"], e["a"], e["rot"], e["mu"]) for e in ell]
    rng = np.random.default_rng(seed + 1)
    out = np.empty((n, h, w), np.float32)
    for a in range(n):
        p = O.line_integrals(spec, a, oe)
        out[a] = (rng.poisson(2e4 * np.exp(-p)) / 2e4).astype(np.float32)
    return out


def _geometry(w, h, n, so=100.0, sd=1000.0):
    from xpulse_recon_amd.geometry import Geometry
    return Geometry(dwidth=w, dheight=h, projections=n, so=so, sd=sd, px=0.1, rx=1.5)


def _cli_naive(tmp_path, name, proj, it, sit, seed):
    d = tmp_path / f"{name}_in"
    d.mkdir()
    for a in range(proj.shape[0]):
        write_f32(d / f"proj_{a:04d}.tif", proj[a])
    out = tmp_path / name
    params = {"recon": {"input": str(d), "output": str(out), "it": it, "sit": sit,
                        "method": "naive"},
              "geometry": {"so": 100.0, "sd": 1000.0}, "detector": {"px": 0.1, "rx": 1.5},
              "xpr": {"jitter_seed": seed}}
    pf = tmp_path / f"{name}.json"
    pf.write_text(json.dumps(params))
    print(f"XPRecon {name}", flush=True)
    r = subprocess.run([CLI, str(pf)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return np.stack([read_f32(f) for f in sorted(out.glob("layer_*.tif"))])


@pytest.fixture(scope="module")
def jobs(cuda):
    return {"a": (_geometry(40, 32, 24), _projections(40, 32, 24, 3)),
            "b": (_geometry(48, 20, 30), _projections(48, 20, 30, 7)),
            "c": (_geometry(32, 28, 20, so=90.0, sd=1100.0),
                  _projections(32, 28, 20, 9, so=90.0, sd=1100.0))}


def test_server_job_equals_cli_naive(jobs, tmp_path):
    from xpulse_recon_amd.serve import Server
    geo, proj = jobs["a"]
    ref = _cli_naive(tmp_path, "a", proj, it=2, sit=4, seed=11)
    with Server(budget_bytes=1 << 30, max_concurrent=2) as sv:
        r = sv.wait(sv.submit(geo, proj, it=2, sit=4, jitter_seed=11))
    assert not r.streamed and r.device_ms > 0
    np.testing.assert_array_equal(r.volume, ref)


def test_alone_equals_in_mix(jobs):
    """Three jobs of different sizes and precisions, each run alone, then all
    submitted at once with room for all: results are bit-identical."""
    from xpulse_recon_amd.serve import Server
    cfg = {"a": dict(it=2, sit=4, jitter_seed=1), "b": dict(it=1, sit=5, jitter_seed=2,
                                                             storage="f16"),
           "c": dict(it=2, sit=2, jitter_seed=3, storage="bf16", acc_bits=32)}
    alone = {}
    with Server(budget_bytes=1 << 30, max_concurrent=1) as sv:
        for k, kw in cfg.items():
            alone[k] = sv.wait(sv.submit(*jobs[k], **kw)).volume
    print("mix", flush=True)
    with Server(budget_bytes=1 << 30, max_concurrent=3) as sv:
        ids = {k: sv.submit(*jobs[k], **kw) for k, kw in cfg.items()}
        mixed = {k: sv.wait(i) for k, i in ids.items()}
    overlap = max(r.start_ms for r in mixed.values()) < min(r.end_ms for r in mixed.values())
    assert overlap                                  # all three were admitted together
    for k in cfg:
        assert np.isfinite(mixed[k].volume).all()
        np.testing.assert_array_equal(mixed[k].volume, alone[k])


def test_over_budget_job_streams(jobs):
    """A budget below the job's resident footprint: the server runs it with
    slab streaming (volume on the host), which equals the naive schedule."""
    from xpulse_recon_amd.serve import Server, job_device_bytes
    geo, proj = jobs["a"]
    need = job_device_bytes(geo, sit=4)
    with Server(budget_bytes=need * 2, max_concurrent=1) as sv:
        ref = sv.wait(sv.submit(geo, proj, it=2, sit=4, jitter_seed=5))
    with Server(budget_bytes=need - 1, max_concurrent=1) as sv:
        r = sv.wait(sv.submit(geo, proj, it=2, sit=4, jitter_seed=5, slab_height=7))
    assert r.streamed and not ref.streamed
    assert r.device_bytes < need
    np.testing.assert_array_equal(r.volume, ref.volume)


def test_budget_serialises_admission(jobs):
    """Budget for one job at a time: the second starts after the first ends,
    and both give the same result as a roomy server."""
    from xpulse_recon_amd.serve import Server, job_device_bytes
    geo, proj = jobs["a"]
    need = job_device_bytes(geo, sit=4)
    with Server(budget_bytes=need + need // 2, max_concurrent=4) as sv:
        i1 = sv.submit(geo, proj, it=1, sit=4, jitter_seed=8)
        i2 = sv.submit(geo, proj, it=1, sit=4, jitter_seed=8)
        r1, r2 = sv.wait(i1), sv.wait(i2)
    assert r2.start_ms >= r1.end_ms
    np.testing.assert_array_equal(r1.volume, r2.volume)


def test_failed_job_reports_error(jobs):
    """A job whose projection stack is inconsistent with sit raises on submit;
    the server keeps serving afterwards."""
    from xpulse_recon_amd._lib import XprError
    from xpulse_recon_amd.serve import Server
    geo, proj = jobs["a"]
    with Server(budget_bytes=1 << 30) as sv:
        with pytest.raises(XprError):
            sv.submit(geo, proj, it=1, sit=1000)
        r = sv.wait(sv.submit(geo, proj, it=1, sit=4, jitter_seed=1))
        assert np.isfinite(r.volume).all()


def _four_jobs(jobs):
    """Jobs covering the batch engine's item kinds: FP32, FP16 and BF16 with
    the FP32 switch of the last sub-iteration, normalisation, two modules."""
    from xpulse_recon_amd.geometry import Geometry
    g2 = Geometry(dwidth=48, dheight=20, projections=30, px=0.1, rx=1.5, module_number=2,
                  module_angle_offset=4.0)
    return [
        (jobs["a"], dict(it=2, sit=4, jitter_seed=1)),
        (jobs["b"], dict(it=2, sit=5, jitter_seed=2, storage="f16")),
        (jobs["c"], dict(it=1, sit=4, jitter_seed=3, storage="bf16", normalize=True)),
        ((g2, jobs["b"][1]), dict(it=1, sit=3, jitter_seed=4, weight=0.8, weight_factor=0.7)),
    ]


def _run_all(cases, **server_kw):
    from xpulse_recon_amd.serve import Server
    with Server(budget_bytes=1 << 30, **server_kw) as sv:
        ids = [sv.submit(*job, **kw) for job, kw in cases]
        return [sv.wait(i) for i in ids]


@pytest.mark.parametrize("step", ["5e5", "2e5"])
def test_batched_steps_equal_per_job_streams(jobs, monkeypatch, step):
    """The continuous-batching engine (all jobs in shared steps; the step
    targets are below every job's total work, so all four are batched, and cut
    into layer ranges over many steps, each replaying the captured graph)
    gives every job's volume bit for bit as the job run alone on its own
    streams."""
    cases = _four_jobs(jobs)
    monkeypatch.setenv("XPR_SERVE_MODE", "streams")
    ref = [r.volume for r in _run_all(cases, max_concurrent=1)]
    monkeypatch.setenv("XPR_SERVE_MODE", "batch")
    monkeypatch.setenv("XPR_SERVE_STEP", step)
    print(f"batched, step {step}", flush=True)
    got = _run_all(cases, max_concurrent=4)
    for r, v in zip(got, ref):
        assert np.isfinite(r.volume).all() and r.device_ms > 0
        np.testing.assert_array_equal(r.volume, v)
