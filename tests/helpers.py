"""Shared test inputs: small cone-beam geometries and volumes."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O


def small_spec(w=32, h=24, n=12, **kw):
    return O.GeometrySpec(dwidth=w, dheight=h, projections=n, **kw)


def projdata_for(spec: O.GeometrySpec, views, fp64=False):
    pd, grid = O.geometry(spec, fp64)
    m = spec.module_number
    sel = np.concatenate([pd[v * m:(v + 1) * m] for v in views])
    return np.ascontiguousarray(sel), grid


def geom_args(spec, grid, n_views, vmin=None, vmax=None, jitter=(0.0, 0.0, 0.0),
              line_offset=0, line_rows=None, fp64=False):
    W, H = grid["vwidth"], grid["vheight"]
    vmin = (0, 0, 0) if vmin is None else vmin
    vmax = (W, H, W) if vmax is None else vmax
    rows = spec.dheight if line_rows is None else line_rows
    if fp64:
        origin = [grid["orig"][i] + jitter[i] for i in range(3)]
        vx = grid["vx"]
    else:
        origin = [np.float32(np.float32(grid["orig"][i]) + np.float32(jitter[i])) for i in range(3)]
        vx = np.float32(grid["vx"])
    return dict(det_width=spec.dwidth, n_views=n_views, modules=spec.module_number,
                vmin=vmin, vmax=vmax, line_offset=line_offset, line_rows=rows,
                origin=origin, voxel=(vx, vx, vx))


def oracle_geom(d, fp64=False):
    return O.make_geom(d["det_width"], d["n_views"], d["modules"], d["vmin"], d["vmax"],
                       d["line_offset"], d["line_rows"], d["origin"], d["voxel"], fp64=fp64)


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))
