"""Randomised geometries for the projectors' test-free paths.

The forward's exact/inside views and the backward's inside segments skip the
per-voxel bounds test when a rounding bound proves it redundant.  Over seeded
random scans (source and detector distances, pixel size, detector offsets,
detector roll, one or two modules, slabs, cropped row bands, origin jitter) the sums and the
updated volume must be bit-identical with those paths switched off, and the
forward sums must equal the oracle's."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import geom_args, oracle_geom, projdata_for, small_spec

pytestmark = pytest.mark.gpu


def _roll(pd, deg, w, h):
    """The matrices of a detector rolled by deg about its centre pixel: the u
    and v rows turn in the detector plane (float64, rounded to float32), so
    the u row gains a layer coefficient and the w row keeps its own."""
    t = np.deg2rad(deg)
    c, s = np.cos(t), np.sin(t)
    cu, cv = w / 2.0, h / 2.0
    m = pd.astype(np.float64)
    u, v, wr = m[:, 0:4], m[:, 4:8], m[:, 12:16]
    out = pd.copy()
    out[:, 0:4] = c * u - s * v + (cu - c * cu + s * cv) * wr
    out[:, 4:8] = s * u + c * v + (cv - s * cu - c * cv) * wr
    return out


def _case(seed):
    rng = np.random.default_rng(seed)
    w = int(rng.integers(64, 193))
    h = int(rng.integers(16, 97))
    n = int(rng.integers(8, 61))
    so = float(rng.uniform(50.0, 200.0))
    sd = float(rng.uniform(so + 100.0, 1500.0))
    mods = int(rng.integers(1, 3))
    kw = dict(so=so, sd=sd, px=float(rng.uniform(0.05, 0.2)), rx=float(rng.uniform(-5.0, 5.0)),
              sx=float(rng.uniform(-0.3, 0.3)), module_number=mods)
    if mods == 2:
        kw["module_angle_offset"] = float(rng.uniform(2.0, 8.0))
    spec = small_spec(w, h, n, **kw)
    nv = int(rng.integers(4, min(n, 24) + 1))
    views = sorted(rng.choice(n, nv, replace=False).tolist())
    pd, grid = projdata_for(spec, views)
    if rng.random() < 0.5:
        pd = _roll(pd, float(rng.uniform(-2.0, 2.0)), w, h)
    W, H = grid["vwidth"], grid["vheight"]
    l0 = int(rng.integers(0, H // 2))
    l1 = int(rng.integers(l0 + 1, H + 1))
    if rng.random() < 0.5:
        lo = int(rng.integers(0, h // 2))
        crop = (lo, int(rng.integers(1, h - lo + 1)))
    else:
        crop = (0, None)
    jit = tuple(float(x) for x in rng.uniform(0.0, grid["vx"], 3))
    d = geom_args(spec, grid, len(views), vmin=(0, l0, 0), vmax=(W, l1, W), jitter=jit,
                  line_offset=crop[0], line_rows=crop[1])
    vol = rng.uniform(0.0, 2.0, (l1 - l0, W, W)).astype(np.float32)
    err = rng.normal(0.0, 0.3, (d["n_views"], d["line_rows"], d["det_width"])).astype(np.float32)
    return spec, pd, d, vol, err


def _args(d):
    import xpulse_recon_amd as X
    return X.proj_args(d["det_width"], d["n_views"], d["modules"], d["vmin"], d["vmax"],
                       d["line_offset"], d["line_rows"], d["origin"], d["voxel"])


@pytest.mark.parametrize("seed", range(16))
def test_fast_paths_equal_checked_random(cuda, monkeypatch, seed):
    import torch
    import xpulse_recon_amd as X
    spec, pd, d, vol, err = _case(5000 + seed)
    print(f"seed {5000 + seed}: grid {vol.shape}, views {d['n_views']}x{d['modules']}, "
          f"rows {d['line_offset']}+{d['line_rows']}", flush=True)
    tp = torch.from_numpy(pd).to(cuda)
    tv = torch.from_numpy(vol).to(cuda)
    te = torch.from_numpy(err).to(cuda)
    shape = (d["n_views"], d["line_rows"], d["det_width"])
    sums, vols = {}, {}
    for on in ("1", "0"):
        monkeypatch.setenv("XPR_FORWARD_EXACT", on)
        monkeypatch.setenv("XPR_BACKWARD_INSIDE", on)
        s = torch.zeros(shape, dtype=torch.int64, device=cuda)
        X.forward(tv, s, tp, _args(d))
        sums[on] = s.cpu().numpy()
        v = tv.clone()
        X.backward(v, te, tp, _args(d))
        vols[on] = v.cpu().numpy()
    np.testing.assert_array_equal(sums["1"], sums["0"])
    np.testing.assert_array_equal(vols["1"], vols["0"])
    ref = np.zeros(shape, np.uint64)
    O.forward_f32(vol, ref, pd, oracle_geom(d), 8)
    np.testing.assert_array_equal(sums["1"].view(np.uint64), ref)
