"""xpr_hit_bound: the bound on voxels per detector pixel that lets the engine
take 32-bit detector sums, against brute-force hit counts.

The counts come from the oracle's Forward over a volume whose every Q.22
value is 1 (v = 2^-22), so each 64-bit sum is the number of voxels that round
to that pixel, at several origin jitters (the bound is stated for any jitter
up to one voxel).  A bound below some count would let a 32-bit sum wrap; a
bound far above every count would make the 32-bit path useless, so both
directions are checked.  Host-only: no GPU.
"""
import numpy as np
import pytest

import xpulse_recon_amd as X
from oracle import oracle as O
from xpulse_recon_amd.geometry import Geometry

CASES = [
    dict(dwidth=48, dheight=40, projections=24),
    dict(dwidth=48, dheight=40, projections=24, rx=2.5, roll=3.0),
    dict(dwidth=40, dheight=32, projections=16, so=60.0, sd=600.0, pitch=6.0, yaw=4.0),
    dict(dwidth=64, dheight=16, projections=12, module_number=2, module_angle_offset=1.5,
         module_center_offset_z=300.0),
    dict(dwidth=32, dheight=32, projections=10, heli_offset=0.05, angle=720.0),
    dict(dwidth=40, dheight=24, projections=12, so=30.0, sd=200.0, px=0.2),
]


def _counts(kw, pd, grid, jitter):
    W, H = grid["vwidth"], grid["vheight"]
    origin = [np.float32(np.float32(grid["orig"][i]) + np.float32(jitter[i])) for i in range(3)]
    vx = np.float32(grid["vx"])
    m = kw.get("module_number", 1)
    n = kw["projections"]
    g = O.make_geom(kw["dwidth"], n, m, (0, 0, 0), (W, H, W), 0, kw["dheight"], origin, (vx, vx, vx))
    vol = np.full((H, W, W), np.float32(2.0 ** -22), np.float32)
    sums = np.zeros((n, kw["dheight"], kw["dwidth"]), np.uint64)
    O.forward_f32(vol, sums, pd, g, threads=8)
    return int(sums.max())


def _bound(kw, pd, grid):
    W, H = grid["vwidth"], grid["vheight"]
    vx = np.float32(grid["vx"])
    orig = [np.float32(grid["orig"][i]) for i in range(3)]
    a = X.proj_args(kw["dwidth"], kw["projections"], kw.get("module_number", 1), (0, 0, 0),
                    (W, H, W), 0, kw["dheight"], orig, (vx, vx, vx))
    return X.hit_bound(pd, a)


@pytest.mark.parametrize("kw", CASES)
def test_bound_holds_for_any_jitter_and_is_useful(kw):
    fp64 = any(kw.get(k, 0.0) for k in ("roll", "pitch", "yaw"))
    pd, grid = Geometry(**kw).projdata(fp64=fp64)
    bound = _bound(kw, pd, grid)
    vx = float(np.float32(grid["vx"]))
    rng = np.random.default_rng(11)
    jitters = [(0.0, 0.0, 0.0), (0.999 * vx,) * 3] + [tuple(rng.uniform(0, vx, 3)) for _ in range(3)]
    worst = max(_counts(kw, pd, grid, j) for j in jitters)
    print(f"{kw}: bound {bound:.0f}, largest count {worst}", flush=True)
    assert worst <= bound
    assert bound <= 8 * worst


def test_source_inside_the_box_has_no_bound():
    kw = dict(dwidth=32, dheight=32, projections=4)
    pd, grid = Geometry(**kw).projdata()
    vx = np.float32(grid["vx"])
    orig = [np.float32(grid["orig"][i]) for i in range(3)]
    # a box reaching past the source (so = 100 mm: 10000 voxels of 10 um)
    n = 45000
    a = X.proj_args(kw["dwidth"], kw["projections"], 1, (-n, -n, -n), (n, n, n), 0,
                    kw["dheight"], orig, (vx, vx, vx))
    assert X.hit_bound(pd, a) == float("inf")


def test_bound_scales_with_the_planes_and_sums_the_modules():
    kw = dict(dwidth=48, dheight=40, projections=6)
    pd, grid = Geometry(**kw).projdata()
    W, H = grid["vwidth"], grid["vheight"]
    vx = np.float32(grid["vx"])
    orig = [np.float32(grid["orig"][i]) for i in range(3)]

    def bound(vmax, mods=1, records=pd):
        a = X.proj_args(kw["dwidth"], kw["projections"] // mods, mods, (0, 0, 0), vmax, 0,
                        kw["dheight"], orig, (vx, vx, vx))
        return X.hit_bound(records, a)

    full = bound((W, H, W))
    half = bound((W // 2, H, W // 2))
    assert 0 < half < full
    # the same records read as 2 modules of 3 views: a view's bound is the sum
    # of its modules' (here each pair of records)
    assert bound((W, H, W), mods=2) > full
    assert bound((0, H, W)) == 0.0
