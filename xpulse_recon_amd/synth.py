"""Synthetic inputs: seeded random-ellipsoid phantoms and their FP64 analytic
projections (transmission with Poisson noise), generated on the GPU.

Phantom recipe: FOV = W*vx, R = FOV/2; centres uniform in the ball of radius
0.4 R; semi-axes U[0.05, 0.3]*R; ZYX Euler angles U[0, 2pi); attenuation
U[mu_lo, mu_hi] per phantom-voxel length FOV/n_phantom, additive.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .geometry import Geometry


def _euler_zyx(a, b, c):
    ca, sa, cb, sb, cc, sc = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(c), np.sin(c)
    rz = np.array([[ca, -sa, 0], [sa, ca, 0], [0, 0, 1]])
    ry = np.array([[cb, 0, sb], [0, 1, 0], [-sb, 0, cb]])
    rx = np.array([[1, 0, 0], [0, cc, -sc], [0, sc, cc]])
    return rz @ ry @ rx


def random_ellipsoids(seed: int, n: int, fov: float, n_phantom: int, mu_lo=0.005, mu_hi=0.02):
    """List of dicts (c, a, rot, mu) with mu in 1/mm."""
    rng = np.random.default_rng(seed)
    R = 0.5 * fov
    out = []
    for _ in range(n):
        while True:
            c = rng.uniform(-0.4 * R, 0.4 * R, 3)
            if np.linalg.norm(c) <= 0.4 * R:
                break
        a = rng.uniform(0.05, 0.3, 3) * R
        rot = _euler_zyx(*rng.uniform(0.0, 2 * np.pi, 3)).T      # world -> body
        mu = rng.uniform(mu_lo, mu_hi) / (fov / n_phantom)
        out.append(dict(c=c, a=a, rot=rot, mu=mu))
    return out


def spheres(seed: int, n: int, fov: float, r_lo: float, r_hi: float, mu: float):
    rng = np.random.default_rng(seed)
    R = 0.5 * fov
    out = []
    for _ in range(n):
        while True:
            c = rng.uniform(-0.8 * R, 0.8 * R, 3)
            if np.linalg.norm(c[[0, 2]]) <= 0.8 * R:
                break
        r = rng.uniform(r_lo, r_hi)
        out.append(dict(c=c, a=np.array([r, r, r]), rot=np.eye(3), mu=mu))
    return out


def _pack(ellipsoids) -> np.ndarray:
    arr = np.zeros((len(ellipsoids), 16), np.float64)
    for i, e in enumerate(ellipsoids):
        arr[i, 0:3], arr[i, 3:6] = e["c"], e["a"]
        arr[i, 6:15], arr[i, 15] = np.asarray(e["rot"]).reshape(9), e["mu"]
    return arr


def line_integrals(geom: Geometry, views, ellipsoids, device="cuda", stream=None):
    """FP64 line integrals [len(views), H, W] (torch, on device)."""
    import torch
    frames = geom.frames()
    m = geom.module_number
    sel = np.ascontiguousarray(np.concatenate([frames[v * m:(v + 1) * m] for v in views]))
    tf = torch.from_numpy(sel).to(device)
    te = torch.from_numpy(_pack(ellipsoids)).to(device)
    out = torch.empty((len(views), geom.dheight, geom.dwidth), dtype=torch.float64, device=device)
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream if stream is None else stream)
    _lib.call("xpr_synth_line_integrals", s, C.c_void_p(tf.data_ptr()), len(views), m,
              geom.dwidth, geom.dheight, C.c_void_p(te.data_ptr()), len(ellipsoids),
              C.c_void_p(out.data_ptr()))
    return out


def phantom(grid: dict, ellipsoids, device="cuda", stream=None):
    """Phantom attenuation (1/mm) sampled at the voxel centres of the grid,
    float32 [vheight, vwidth, vwidth] (torch, on device)."""
    import torch
    W, H = grid["vwidth"], grid["vheight"]
    te = torch.from_numpy(_pack(ellipsoids)).to(device)
    out = torch.empty((H, W, W), dtype=torch.float32, device=device)
    orig = (C.c_double * 3)(*grid["orig"])
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream if stream is None else stream)
    _lib.call("xpr_synth_phantom", s, float(grid["vx"]), W, H, orig, C.c_void_p(te.data_ptr()),
              len(ellipsoids), C.c_void_p(out.data_ptr()))
    return out


def transmission(p, i0: float, seed: int):
    """Poisson counts N ~ P(I0 exp(-p)), returned as transmission N/I0 (float32)."""
    import torch
    g = torch.Generator(device=p.device).manual_seed(seed)
    lam = torch.exp(-p) * i0
    return (torch.poisson(lam, generator=g) / i0).to(torch.float32)
