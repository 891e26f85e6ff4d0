"""Writes tests/golden/subiteration_small.npz: one canonical sub-iteration
(Forward, Error, Backward) of three small scans, computed by the oracle.

The reference ships no tests or fixtures and cannot run here (no OpenCL
device), so these vectors come from the oracle, whose functions are pinned by
closed forms and brute force in tests/test_oracle_*.py.  They freeze the
canonical results: tests/test_golden.py checks that the oracle still
reproduces them bit for bit, and the GPU tests check the HIP projectors
against them.  Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402
from tests.helpers import geom_args, oracle_geom, projdata_for, small_spec  # noqa: E402

# (spec kwargs, views, slab, crop, jitter, seed)
CASES = [
    (dict(w=16, h=12, n=12), [0, 3, 6, 9], None, None, (0.0, 0.0, 0.0), 11),
    (dict(w=20, h=14, n=10, rx=1.5, so=90.0, sd=1100.0, module_number=2, module_angle_offset=4.0),
     [1, 4, 7], ((0, 2, 0), (20, 12, 20)), None, (0.002, 0.001, 0.003), 12),
    (dict(w=18, h=16, n=8, sx=0.1), [0, 2, 5], ((0, 4, 0), (18, 13, 18)), (3, 9),
     (0.001, 0.0, 0.002), 13),
]


def case_arrays(i):
    kw, views, slab, crop, jit, seed = CASES[i]
    kw = dict(kw)
    spec = small_spec(kw.pop("w"), kw.pop("h"), kw.pop("n"), **kw)
    pd, grid = projdata_for(spec, views)
    vmin, vmax = slab if slab else (None, None)
    lo, rows = crop if crop else (0, None)
    d = geom_args(spec, grid, len(views), vmin=vmin, vmax=vmax, jitter=jit, line_offset=lo,
                  line_rows=rows)
    nx, nl, nk = (d["vmax"][k] - d["vmin"][k] for k in range(3))
    rng = np.random.default_rng(seed)
    vol = rng.uniform(0.2, 2.0, (nl, nk, nx)).astype(np.float32)
    shape = (d["n_views"], d["line_rows"], d["det_width"])
    images = rng.uniform(0.05, 4.0, shape).astype(np.float32)
    return pd, d, vol, images


def compute(pd, d, vol, images, weight=0.9):
    g = oracle_geom(d)
    shape = (d["n_views"], d["line_rows"], d["det_width"])
    s64 = np.zeros(shape, np.uint64)
    O.forward_f32(vol, s64, pd, g)
    s32 = np.zeros(shape, np.uint32)
    O.forward_f32(vol, s32, pd, g)
    err = O.error_f32(images, s64, weight)
    upd = vol.copy()
    empty = O.backward_f32(upd, err, pd, g)
    return s64, s32, err, upd, empty


def main():
    out = {}
    for i in range(len(CASES)):
        pd, d, vol, images = case_arrays(i)
        s64, s32, err, upd, empty = compute(pd, d, vol, images)
        out[f"c{i}_projdata"] = pd
        out[f"c{i}_args"] = np.array([d["det_width"], d["n_views"], d["modules"], *d["vmin"],
                                      *d["vmax"], d["line_offset"], d["line_rows"]], np.int64)
        out[f"c{i}_origin_voxel"] = np.array([*d["origin"], *d["voxel"]], np.float32)
        out[f"c{i}_volume"] = vol
        out[f"c{i}_images"] = images
        out[f"c{i}_sums_u64"] = s64
        out[f"c{i}_sums_u32"] = s32
        out[f"c{i}_error"] = err
        out[f"c{i}_updated"] = upd
        out[f"c{i}_empty"] = np.array([empty], np.int64)
    path = os.path.join(HERE, "subiteration_small.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
