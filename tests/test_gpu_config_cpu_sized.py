"""The reference's CPU-runnable configuration at its full size: 128^2 detector
-> 128^3 grid, 180 views over 360 deg, so=100, sd=1000, px=0.1 mm, 12 random
ellipsoids, Poisson I0 = 1e4, mlog, it=1, sit=10 (18 views per sub-iteration),
naive schedule.  The complete reconstruction through XPRecon against the
oracle's naive schedule on the same preprocessed images and logged jitter.

Bar: rel-L2 1e-5 (the oracle and the kernels share the canonical arithmetic;
device expf/logf differ in the last ulp) and bit-identical Forward sums on the
final volume."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import rel_l2
from tests.tiff_util import read_f32, write_f32

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "xpulse_recon_amd", "bin", "XPRecon")
D, N, SIT = 128, 180, 10


def test_full_size_naive_matches_oracle(cuda, tmp_path):
    from xpulse_recon_amd import synth
    spec = O.GeometrySpec(dwidth=D, dheight=D, projections=N)
    pd, grid = O.geometry(spec, False)
    ell = synth.random_ellipsoids(1001, 12, D * grid["vx"], 128)
    oe = [O.make_ellipsoid(e["c"], e["a"], e["rot"], e["mu"]) for e in ell]
    rng = np.random.default_rng(1002)
    d = tmp_path / "in"
    d.mkdir()
    raw = np.empty((N, D, D), np.float32)
    for a in range(N):
        p = O.line_integrals(spec, a, oe)
        raw[a] = (rng.poisson(1e4 * np.exp(-p)) / 1e4).astype(np.float32)
        write_f32(d / f"proj_{a:04d}.tif", raw[a])
    out = tmp_path / "out"
    params = {"recon": {"input": str(d), "output": str(out), "it": 1, "sit": SIT,
                        "method": "naive"},
              "xpr": {"jitter_seed": 1003, "run_log": str(tmp_path / "run.json")}}
    pf = tmp_path / "params.json"
    pf.write_text(json.dumps(params))
    print("XPRecon 128^3 / 180 views", flush=True)
    r = subprocess.run([CLI, str(pf)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    vol = np.stack([read_f32(f) for f in sorted(out.glob("layer_*.tif"))])
    log = json.loads((tmp_path / "run.json").read_text())
    assert vol.shape == (D, D, D) and len(log["subiteration_ms"]) == SIT
    imgs = raw.copy()
    O.preprocess(imgs.reshape(-1), False, True)
    print("oracle", flush=True)
    ref = O.reconstruct_naive(imgs, pd, grid, 1, SIT, 1.0, 0.9,
                              np.array(log["jitter"], np.float32), threads=16)
    assert rel_l2(vol, ref) < 1e-5
    # the final volumes' forward projections of one sub-iteration agree bit for bit
    import torch
    import xpulse_recon_amd as X
    views = O.sit_views(N, SIT, 0)
    pds = np.ascontiguousarray(np.concatenate([pd[v:v + 1] for v in views]))
    o = [np.float32(grid["orig"][i]) for i in range(3)]
    g = O.make_geom(D, len(views), 1, (0, 0, 0), (D, D, D), 0, D, o, (grid["vx"],) * 3)
    s_ref = np.zeros((len(views), D, D), np.uint64)
    O.forward_f32(ref, s_ref, pds, g, 16)
    s_gpu = torch.zeros((len(views), D, D), dtype=torch.int64, device=cuda)
    a = X.proj_args(D, len(views), 1, (0, 0, 0), (D, D, D), 0, D, o, (grid["vx"],) * 3)
    X.forward(torch.from_numpy(ref).to(cuda), s_gpu, torch.from_numpy(pds).to(cuda), a)
    np.testing.assert_array_equal(s_gpu.cpu().numpy().view(np.uint64), s_ref)
