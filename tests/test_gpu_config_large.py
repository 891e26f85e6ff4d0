"""The out-of-core and serving configurations at their workloads, on the GPU.

Out-of-core: 2048^2 detector -> 2048^3 grid (32 GiB FP32), 2048 views of
2048^2 (32 GiB), 40 random ellipsoids plus 200 spheres of radius 2-6 voxels
(mu 0.05 per voxel), Poisson I0 = 5e4, it=2, sit=16 (128 views per
sub-iteration), method_2 under a 16 GB device budget, projections and volume
in host memory (xpr_reconstruct_mem).  Bars: the plan the engine logs equals
the CPU planner's (xpr_plan_chunks) and the oracle's restatement
(geometry_chunk.hpp:92-197); the volume is finite; its RMSE against the
phantom mu*vx at the voxel centres is at most 5.2e-3 (the reconstruction
measured 4.7e-3; the bar leaves 10 %).

Serving: a Poisson stream (4 jobs/s, arrivals seed 5005) of reconstructions
drawn from {128^3/180 views, 384^3/360, 768^3/720}, so U[80,150], sd
U[800,1200] mm, 12-ellipsoid phantoms, FP16 storage with a final FP32
sub-iteration, admitted by a 200 GB budget, it=1, sit=10.  Bar: every job's
volume in the mix equals the same job run alone on an idle server, bit for
bit (each job owns its buffers, streams and seeded jitter); p50/p99
time-to-reconstruction printed."""
import ctypes as C
import json
import time

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_out_of_core_2048(cuda):
    import torch
    from xpulse_recon_amd import _lib, synth
    from xpulse_recon_amd.geometry import Geometry
    D, N, IT, S = 2048, 2048, 2, 16
    geom = Geometry(dwidth=D, dheight=D, projections=N, so=100.0, sd=1000.0, px=0.1)
    _, grid = geom.projdata()
    W, H, vx = grid["vwidth"], grid["vheight"], grid["vx"]
    fov = W * vx
    ell = synth.random_ellipsoids(1004, 40, fov, D)
    ell += synth.spheres(1005, 200, fov, 2 * vx, 6 * vx, 0.05 / vx)
    proj = np.empty((N, D, D), np.float32)
    t0 = time.perf_counter()
    for v0 in range(0, N, 32):
        views = list(range(v0, min(N, v0 + 32)))
        p = synth.line_integrals(geom, views, ell, cuda)
        proj[v0:v0 + len(views)] = synth.transmission(p, 5e4, 6000 + v0).cpu().numpy()
        del p
    print(f"2048 projections of 2048^2 in {time.perf_counter() - t0:.1f} s", flush=True)
    params = {"recon": {"method": "method_2", "it": IT, "sit": S, "mlog": True, "weight": 1.0,
                        "weight_factor": 0.9},
              "geometry": {"so": 100.0, "sd": 1000.0}, "detector": {"px": 0.1},
              "xpr": {"device_budget_gb": 16.0, "jitter_seed": 7, "quiet": True}}
    text = json.dumps(params).encode()
    L = _lib.load()
    plan = (C.c_int64 * (6 * 64))()
    nplan = C.c_int64()
    assert L.xpr_plan_chunks(text, D, D, N, plan, 64, C.byref(nplan)) == 0
    cpu_plan = [list(plan[6 * i:6 * i + 6]) for i in range(nplan.value)]
    spec = O.GeometrySpec(dwidth=D, dheight=D, projections=N)
    pd, ogrid = O.geometry(spec, False)
    assert cpu_plan == O.plan_chunks(pd, ogrid, D, D, N // S, 4, 16e9)
    vol = np.empty((H, W, W), np.float32)
    rec = C.create_string_buffer(1 << 20)
    print("xpr_reconstruct_mem method_2 2048^3, 16 GB budget", flush=True)
    t0 = time.perf_counter()
    rc = L.xpr_reconstruct_mem(text, proj.ctypes.data_as(C.c_void_p), D, D, N,
                               vol.ctypes.data_as(C.c_void_p), rec, len(rec))
    print(f"reconstructed in {time.perf_counter() - t0:.1f} s", flush=True)
    assert rc == 0, L.xpr_last_error().decode()
    del proj
    record = json.loads(rec.value.decode())
    assert [[int(x) for x in c] for c in record["chunks"]] == cpu_plan
    assert len(cpu_plan) == 4
    assert len(record["subiteration_ms"]) == len(cpu_plan) * IT * S
    assert np.isfinite(vol).all()
    sq, n = 0.0, 0
    for l0 in range(0, H, 128):
        l1 = min(H, l0 + 128)
        g = dict(grid, vheight=l1 - l0,
                 orig=(grid["orig"][0], grid["orig"][1] + l0 * vx, grid["orig"][2]))
        truth = synth.phantom(g, ell, cuda).double() * vx
        d = torch.from_numpy(vol[l0:l1]).to(cuda).double() - truth
        sq += float((d * d).sum())
        n += d.numel()
        del truth, d
    rmse = (sq / n) ** 0.5
    print(f"RMSE vs phantom {rmse:.4e}", flush=True)
    assert rmse <= 5.2e-3


SERVE_CLASSES = [(128, 180, 128), (384, 360, 256), (768, 720, 512)]


def test_serving_mix_equals_alone(cuda):
    from xpulse_recon_amd import synth
    from xpulse_recon_amd.geometry import Geometry
    from xpulse_recon_amd.serve import Server
    rng = np.random.default_rng(5005)
    jobs = []
    for j in range(9):
        D, N, n_ph = SERVE_CLASSES[j % 3] if j < 3 else SERVE_CLASSES[int(rng.integers(0, 3))]
        so, sd = float(rng.uniform(80, 150)), float(rng.uniform(800, 1200))
        geom = Geometry(dwidth=D, dheight=D, projections=N, so=so, sd=sd, px=0.1)
        _, grid = geom.projdata()
        ell = synth.random_ellipsoids(3000 + j, 12, grid["vwidth"] * grid["vx"], n_ph)
        proj = np.empty((N, D, D), np.float32)
        for v0 in range(0, N, 90):
            views = list(range(v0, min(N, v0 + 90)))
            p = synth.line_integrals(geom, views, ell, cuda)
            proj[v0:v0 + len(views)] = synth.transmission(p, 1e4, 5000 + j * 8 + v0).cpu().numpy()
            del p
        jobs.append((geom, proj, D))
    gaps = rng.exponential(1.0 / 4.0, len(jobs))
    arrivals = np.cumsum(gaps) - gaps[0]
    kw = dict(it=1, sit=10, storage="f16", final_fp32=True)
    print(f"serving mix of {len(jobs)} jobs, sizes {[d for _, _, d in jobs]}", flush=True)
    with Server(int(200e9), 8) as sv:
        ids = []
        start = time.perf_counter()
        for j, ((geom, proj, _), t) in enumerate(zip(jobs, arrivals)):
            dt = start + t - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            ids.append(sv.submit(geom, proj, jitter_seed=100 + j, **kw))
        mixed = [sv.wait(i) for i in ids]
    lat = np.array([r.latency_ms for r in mixed])
    print(f"p50 {np.percentile(lat, 50):.1f} ms, p99 {np.percentile(lat, 99):.1f} ms", flush=True)
    assert sum(r.voxel_views for r in mixed) > 0
    with Server(int(200e9), 1) as sv:
        for j, ((geom, proj, D), r) in enumerate(zip(jobs, mixed)):
            alone = sv.wait(sv.submit(geom, proj, jitter_seed=100 + j, **kw))
            assert np.isfinite(alone.volume).all()
            np.testing.assert_array_equal(r.volume, alone.volume, err_msg=f"job {j} ({D}^3)")
