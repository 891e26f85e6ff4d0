"""Pins of the oracle's Forward / Error / Backward / preprocessing restatement.

Expected values come from first principles, not from the oracle's own code
path: pixels from the physical ray/plane projection (FP64, no matrix), sums
from integer arithmetic, updates from the closed form v*exp(c), and the
elementwise kernels from numpy float64.  A swapped axis in to3D, a wrong view
or row stride in IMAGE_OFFSET_2D, a wrong rounding mode, a dropped weight or a
sign error in the update fails one of these.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import geom_args, oracle_geom, projdata_for, small_spec

Q = 1 << 22


def _voxel_pos(d, i, l, k):
    return (float(d["origin"][0]) + (i + d["vmin"][0]) * float(d["voxel"][0]),
            float(d["origin"][1]) + (l + d["vmin"][1]) * float(d["voxel"][1]),
            float(d["origin"][2]) + (k + d["vmin"][2]) * float(d["voxel"][2]))


def _physical_pixel(spec, a, pos, margin=2e-3):
    """(col, row) the voxel at pos rounds to in view a, or None if out of
    bounds; 'amb' when within margin of a rounding boundary."""
    u, v = O.project_physical(spec, a, 0, *pos)
    for c in (u, v):
        if abs((c - math.floor(c)) - 0.5) < margin:
            return "amb"
    cu, cv = math.floor(u + 0.5), math.floor(v + 0.5)
    if 0 <= cu < spec.dwidth and 0 <= cv < spec.dheight:
        return cu, cv
    return None


@pytest.mark.parametrize("jit", [(0.0, 0.0, 0.0), (0.0031, 0.0072, 0.0055)])
def test_forward_single_voxel_lands_on_physical_pixel(jit):
    spec = small_spec(20, 14, 12, rx=1.5)
    views = list(range(0, 12, 2))
    pd, grid = projdata_for(spec, views, fp64=True)
    pd = pd.astype(np.float32)
    d = geom_args(spec, grid, len(views), jitter=jit)
    g = oracle_geom(d)
    W, H = grid["vwidth"], grid["vheight"]
    rng = np.random.default_rng(5)
    checked = 0
    for _ in range(40):
        i, l, k = int(rng.integers(W)), int(rng.integers(H)), int(rng.integers(W))
        vol = np.zeros((H, W, W), np.float32)
        vol[l, k, i] = 1.7
        sums = np.zeros((len(views), spec.dheight, spec.dwidth), np.uint64)
        O.forward_f32(vol, sums, pd, g)
        q = int(np.float32(1.7) * np.float32(Q))
        for j, a in enumerate(views):
            px = _physical_pixel(spec, a, _voxel_pos(d, i, l, k))
            if px == "amb":
                continue
            expect = np.zeros((spec.dheight, spec.dwidth), np.uint64)
            if px is not None:
                expect[px[1], px[0]] = q
            np.testing.assert_array_equal(sums[j], expect)
            checked += 1
    assert checked > 150


def test_forward_slab_additivity_and_linearity_bit_exact():
    spec = small_spec(24, 16, 10)
    views = [1, 3, 5, 7, 9]
    pd, grid = projdata_for(spec, views)
    W, H = grid["vwidth"], grid["vheight"]
    rng = np.random.default_rng(6)
    # multiples of 2^-22: FLOAT_TO_FIXED is exact, so F is linear in integers
    a = (rng.integers(0, 1 << 23, (H, W, W)) / Q).astype(np.float32)
    b = (rng.integers(0, 1 << 23, (H, W, W)) / Q).astype(np.float32)
    d = geom_args(spec, grid, len(views), jitter=(0.001, 0.002, 0.003))
    g = oracle_geom(d)
    shape = (len(views), spec.dheight, spec.dwidth)
    fa, fb, fab = (np.zeros(shape, np.uint64) for _ in range(3))
    O.forward_f32(a, fa, pd, g)
    O.forward_f32(b, fb, pd, g)
    O.forward_f32(np.ascontiguousarray(a + b), fab, pd, g)
    np.testing.assert_array_equal(fab, fa + fb)
    # slab split along layers, same origin
    parts = np.zeros(shape, np.uint64)
    for l0, l1 in [(0, 5), (5, 6), (6, H)]:
        ds = dict(d, vmin=(0, l0, 0), vmax=(W, l1, W))
        O.forward_f32(np.ascontiguousarray(a[l0:l1]), parts, pd, oracle_geom(ds))
    np.testing.assert_array_equal(parts, fa)
    # total = sum over in-bounds voxel-views of trunc(v*2^22)
    hits = O.hits_f32(pd, g)
    assert int(fa.sum()) == int((np.round(a.astype(np.float64) * Q).astype(np.uint64)
                                 * hits.astype(np.uint64)).sum())


def test_forward_u32_is_u64_mod_2_32():
    spec = small_spec(16, 12, 6)
    pd, grid = projdata_for(spec, range(6))
    W, H = grid["vwidth"], grid["vheight"]
    vol = np.full((H, W, W), 900.0, np.float32)     # sums far above 1024
    g = oracle_geom(geom_args(spec, grid, 6))
    s32 = np.zeros((6, 12, 16), np.uint32)
    s64 = np.zeros((6, 12, 16), np.uint64)
    O.forward_f32(vol, s32, pd, g)
    O.forward_f32(vol, s64, pd, g)
    assert s64.max() >= (1 << 32)
    np.testing.assert_array_equal(s32, (s64 & 0xFFFFFFFF).astype(np.uint32))


def test_forward_hit_counts_match_physical_brute_force():
    spec = small_spec(18, 12, 8, rx=-1.0)
    views = list(range(8))
    pd, grid = projdata_for(spec, views, fp64=True)
    pd = pd.astype(np.float32)
    d = geom_args(spec, grid, len(views))
    W, H = grid["vwidth"], grid["vheight"]
    vol = np.full((H, W, W), 1.0, np.float32)
    sums = np.zeros((len(views), spec.dheight, spec.dwidth), np.uint64)
    O.forward_f32(vol, sums, pd, oracle_geom(d))
    counts = np.zeros(sums.shape, np.int64)
    amb = 0
    for l in range(H):
        for k in range(W):
            for i in range(W):
                for j, a in enumerate(views):
                    px = _physical_pixel(spec, a, _voxel_pos(d, i, l, k), margin=1e-3)
                    if px == "amb":
                        amb += 1
                    elif px is not None:
                        counts[j, px[1], px[0]] += 1
    diff = np.abs(sums // Q - counts).sum()
    assert diff <= 2 * amb
    assert amb < 0.01 * H * W * W * len(views)


def test_error_closed_form():
    rng = np.random.default_rng(7)
    n = 5000
    ref = rng.uniform(0.0, 3.0, n).astype(np.float32)
    ref[:5] = [0.0, 1e-9, 2.0, 5.0, 0.5]
    sums32 = rng.integers(0, 1 << 31, n).astype(np.uint32)
    sums32[:5] = [0, 1, 1 << 22, 3 << 22, 7]
    w = 0.81
    out = O.error_f32(ref, sums32, w)
    got = sums32.astype(np.float64) / Q
    eps = 2.0 ** -22
    exp = np.log(np.maximum(ref.astype(np.float64) / np.maximum(got, eps), eps)) * np.float64(np.float32(w))
    np.testing.assert_allclose(out, exp, rtol=2e-6, atol=2e-6)
    assert out[0] == pytest.approx(math.log(eps) * np.float32(w), rel=1e-6)   # ref 0 -> log(eps)
    # first pass: sums filled with 1<<22 means got = 1.0
    one = O.error_f32(ref, np.full(n, 1 << 22, np.uint64), 1.0)
    np.testing.assert_allclose(one, np.log(np.maximum(ref.astype(np.float64), eps)), rtol=2e-6,
                               atol=2e-6)
    # in place, like the reference (u32 sums buffer reused for the error)
    buf = sums32.copy()
    O.error_f32(ref, buf, w, out=buf.view(np.float32))
    np.testing.assert_array_equal(buf.view(np.float32), out)


def test_backward_constant_error_gives_exp_update():
    spec = small_spec(20, 16, 10)
    views = [0, 2, 4, 6, 8]
    pd, grid = projdata_for(spec, views)
    W, H = grid["vwidth"], grid["vheight"]
    d = geom_args(spec, grid, len(views), jitter=(0.002, 0.001, 0.004))
    g = oracle_geom(d)
    rng = np.random.default_rng(8)
    vol = rng.uniform(0.5, 2.0, (H, W, W)).astype(np.float32)
    err = np.full((len(views), spec.dheight, spec.dwidth), 0.25, np.float32)
    out = vol.copy()
    empty = O.backward_f32(out, err, pd, g)
    hits = O.hits_f32(pd, g)
    seen = hits > 0
    np.testing.assert_allclose(out[seen], vol[seen] * np.exp(0.25), rtol=3e-7)
    np.testing.assert_array_equal(out[~seen], vol[~seen])
    assert empty == int((~seen).sum())


def test_backward_per_voxel_mean_matches_physical():
    spec = small_spec(16, 12, 12, rx=0.5)
    views = list(range(1, 12, 3))
    pd, grid = projdata_for(spec, views, fp64=True)
    pd = pd.astype(np.float32)
    W, H = grid["vwidth"], grid["vheight"]
    d = geom_args(spec, grid, len(views), jitter=(0.004, 0.003, 0.001))
    rng = np.random.default_rng(9)
    err = rng.normal(0.0, 0.3, (len(views), spec.dheight, spec.dwidth)).astype(np.float32)
    vol = np.ones((H, W, W), np.float32)
    O.backward_f32(vol, err, pd, oracle_geom(d))
    checked = 0
    for l in range(H):
        for k in range(W):
            for i in range(W):
                pix = [_physical_pixel(spec, a, _voxel_pos(d, i, l, k)) for a in views]
                if "amb" in pix:
                    continue
                vals = [err[j, p[1], p[0]] for j, p in enumerate(pix) if p is not None]
                if not vals:
                    assert vol[l, k, i] == 1.0
                    continue
                assert vol[l, k, i] == pytest.approx(math.exp(np.mean(vals)), rel=2e-6)
                checked += 1
    assert checked > 0.9 * H * W * W * 0.5


def test_preprocess_closed_form():
    rng = np.random.default_rng(10)
    x = rng.uniform(0.0, 1.2, 4000).astype(np.float32)
    x[:6] = [0.0, 1.0, 1e-40, np.inf, 1.0 - 2 ** -23, 2 ** -30]
    ref = x.astype(np.float64)
    y = O.preprocess(x.copy(), normalize=False, mlog=True)
    eps = 2.0 ** -22
    exp = -np.log(np.clip(ref, eps, 1.0 - eps))
    np.testing.assert_allclose(y, exp, rtol=2e-6, atol=1e-7)
    mn, mx = O.minmax_normal(x)
    normal = x[np.isfinite(x) & (np.abs(x) >= np.finfo(np.float32).tiny)]
    assert (mn, mx) == (float(normal.min()), float(normal.max()))
    z = O.preprocess(x.copy(), normalize=True, mlog=False, mn=mn, mx=mx)
    isn = np.isfinite(x) & (np.abs(x) >= np.finfo(np.float32).tiny)
    np.testing.assert_allclose(z[isn], (ref[isn] - mn) / (mx - mn), rtol=1e-6, atol=1e-7)
    assert np.all(z[~isn] == np.float32((mx + mn) / 2))


def test_fp64_model_agrees_with_fp32_canonical():
    spec = small_spec(20, 16, 8)
    views = [0, 2, 4, 6]
    pd64, grid = projdata_for(spec, views, fp64=True)
    pd32 = pd64.astype(np.float32)
    W, H = grid["vwidth"], grid["vheight"]
    d32 = geom_args(spec, grid, len(views))
    d64 = geom_args(spec, grid, len(views), fp64=True)
    rng = np.random.default_rng(11)
    vol = rng.uniform(0.5, 1.5, (H, W, W)).astype(np.float32)
    s32 = np.zeros((len(views), 16, 20), np.uint64)
    O.forward_f32(vol, s32, pd32, oracle_geom(d32))
    s64 = np.zeros((len(views), 16, 20), np.float64)
    O.forward_f64(vol.astype(np.float64), s64, pd64, oracle_geom(d64, fp64=True))
    rel = np.linalg.norm(s32 / Q - s64) / np.linalg.norm(s64)
    assert rel < 2e-2          # only NN flips at rounding boundaries differ


def test_fp64_schedule_single_pass_closed_form():
    """it=1, sit=1: one Backward with got = 1, so every voxel seen by at least
    one view becomes exp(mean log(c)) = c for constant images c; voxels no view
    reaches keep 1."""
    spec = small_spec(20, 16, 8)
    pd64, grid = O.geometry(spec, True)
    c = 0.37
    images = np.full((8, 16, 20), c, np.float64)
    vol = O.reconstruct_naive_f64(images, pd64, grid, 1, 1, 1.0, 0.9, np.zeros((1, 3)))
    seen = vol != 1.0
    assert seen.mean() > 0.5
    np.testing.assert_allclose(vol[seen], c, rtol=1e-12)


def test_fp64_schedule_tracks_fp32_canonical():
    """The FP64 model of the naive schedule against the FP32 canonical one on
    the same (float32-valued) matrices and jitter: they differ by float
    rounding, Q10.22 truncation and rare NN flips only."""
    spec = small_spec(24, 16, 12)
    pd32, grid = O.geometry(spec, False)
    rng = np.random.default_rng(5)
    images = rng.uniform(0.05, 0.6, (12, 16, 24)).astype(np.float32)
    jit = rng.uniform(0.0, grid["vx"], (6, 3)).astype(np.float32)
    v32 = O.reconstruct_naive(images, pd32, grid, 2, 3, 1.0, 0.9, jit)
    v64 = O.reconstruct_naive_f64(images.astype(np.float64), pd32.astype(np.float64), grid, 2, 3,
                                  1.0, 0.9, jit.astype(np.float64))
    rel = np.linalg.norm(v32 - v64) / np.linalg.norm(v64)
    assert rel < 1e-4
    assert np.max(np.abs(v32 / v64 - 1.0)) < 1e-4


def test_phantom_voxels_quadratic_form():
    """Phantom sampling against the quadratic form x^T R^T diag(a^-2) R x <= 1."""
    spec = small_spec(24, 20, 4)
    _, grid = O.geometry(spec, False)
    rng = np.random.default_rng(9)
    ells = []
    for mu in (0.3, 0.7):
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        c = rng.uniform(-0.05, 0.05, 3)
        a = rng.uniform(0.04, 0.1, 3)
        ells.append((c, a, q, mu))
    got = O.phantom_voxels(grid, [O.make_ellipsoid(c, a, q, mu) for c, a, q, mu in ells])
    W, Hv, vx = grid["vwidth"], grid["vheight"], grid["vx"]
    l, k, i = np.meshgrid(np.arange(Hv), np.arange(W), np.arange(W), indexing="ij")
    pts = np.stack([grid["orig"][0] + (i + 0.5) * vx, grid["orig"][1] + (l + 0.5) * vx,
                    grid["orig"][2] + (k + 0.5) * vx], -1)
    want = np.zeros_like(got)
    for c, a, q, mu in ells:
        M = q.T @ np.diag(1.0 / a ** 2) @ q
        x = pts - c
        r2 = np.einsum("...i,ij,...j->...", x, M, x)
        want += np.where(r2 <= 1.0, mu, 0.0)
    assert (got > 0).mean() > 0.01
    # points within rounding of the surface may fall either side
    assert np.mean(got != want) < 1e-3


def test_fp64_on_float32_pixels_equals_q22_sums_for_exact_values():
    """The FP64 model on the canonical float32 pixel choice meets exactly the
    voxel-views of the FP32 Forward: for volumes whose values are whole
    multiples of 2^-22 (truncation exact, double sums exact) its sums are the
    FP32 Q.22 sums times 2^-22, pixel for pixel; its hit counts are the
    brute-force count of the FP32 path."""
    spec = small_spec(40, 36, 10, rx=1.5, module_number=2, module_angle_offset=4.0)
    pd, grid = projdata_for(spec, [1, 4, 7])
    d = geom_args(spec, grid, 3, vmin=(0, 10, 0), vmax=(40, 22, 40), jitter=(0.002, 0.003, 0.001),
                  line_offset=8, line_rows=20)
    g = oracle_geom(d)
    rng = np.random.default_rng(61)
    q = rng.integers(0, 1 << 21, (12, 40, 40))
    vol32 = (q / Q).astype(np.float32)
    ref = np.zeros((3, 20, 40), np.uint64)
    O.forward_f32(vol32, ref, pd, g)
    s64 = np.zeros((3, 20, 40), np.float64)
    O.forward_f64nn(vol32.astype(np.float64), s64, pd, g, 4)
    np.testing.assert_array_equal(s64 * Q, ref.astype(np.float64))
    ones = np.zeros((3, 20, 40), np.float64)
    O.forward_f64nn(np.ones((12, 40, 40)), ones, pd, g)
    assert ones.sum() == O.hits_f32(pd, g).sum()


def test_fp64_q22_model_equals_q22_sums_for_any_float32_volume():
    """With q22 the FP64 model truncates each contribution to a multiple of
    2^-22; for any float32 volume (v * 2^22 is exact in float, so the FP32
    truncation is the same floor) its sums are the FP32 Q.22 sums times
    2^-22 exactly, while the unquantised model is strictly above them
    wherever a voxel with a fractional part contributes."""
    spec = small_spec(40, 36, 10, rx=1.5)
    pd, grid = projdata_for(spec, [0, 5, 8])
    d = geom_args(spec, grid, 3, jitter=(0.001, -0.002, 0.003))
    g = oracle_geom(d)
    rng = np.random.default_rng(63)
    vol32 = rng.uniform(0.0, 0.01, (36, 40, 40)).astype(np.float32)
    ref = np.zeros((3, 36, 40), np.uint64)
    O.forward_f32(vol32, ref, pd, g)
    s_q = np.zeros((3, 36, 40), np.float64)
    O.forward_f64nn(vol32.astype(np.float64), s_q, pd, g, 4, q22=True)
    np.testing.assert_array_equal(s_q * Q, ref.astype(np.float64))
    s_raw = np.zeros((3, 36, 40), np.float64)
    O.forward_f64nn(vol32.astype(np.float64), s_raw, pd, g, 4)
    hit = ref > 0
    assert hit.sum() > 1000
    assert (s_raw[hit] > s_q[hit]).all()
    # the truncation loses less than 2^-22 per contribution
    cnt = np.zeros((3, 36, 40), np.float64)
    O.forward_f64nn(np.ones((36, 40, 40)), cnt, pd, g)
    assert ((s_raw - s_q)[hit] < cnt[hit] / Q).all()
    assert (s_raw[~hit] == 0).all()


def test_fp64_on_float32_pixels_backward_constant_error():
    """Constant error c: v * exp(c) in double wherever the FP32 path sees a
    view, untouched elsewhere (the count of untouched voxels returned)."""
    spec = small_spec(32, 24, 12)
    pd, grid = projdata_for(spec, [0, 3, 6, 9])
    d = geom_args(spec, grid, 4, jitter=(0.001, 0.002, 0.0))
    g = oracle_geom(d)
    rng = np.random.default_rng(62)
    vol = rng.uniform(0.5, 2.0, (24, 32, 32))
    out = vol.copy()
    c = 0.3
    empty = O.backward_f64nn(out, np.full((4, 24, 32), c), pd, g, 2)
    hits = O.hits_f32(pd, g)
    seen = hits > 0
    np.testing.assert_allclose(out[seen], vol[seen] * math.exp(c), rtol=1e-15)
    np.testing.assert_array_equal(out[~seen], vol[~seen])
    assert empty == int((~seen).sum())
