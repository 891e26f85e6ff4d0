"""Pins of the oracle's geometry (geometry.hpp:35-328) by closed forms.

The matrices are checked against an independent ray/plane intersection with the
detector corners (no matrix), and that projection is checked against the
textbook cone-beam formulas (magnification sd/so, axis on column W/2 + rx,
rotation about +Y).  A transposed matrix, a wrong composition order, a sign
error in the frustum or a column/row mix-up fails here.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import small_spec


def _proj(pd, p):
    M = np.asarray(pd[:16], np.float64).reshape(4, 4)    # M[col][row] (glm storage)
    ph = np.array([p[0], p[1], p[2], 1.0])
    return (M[0] @ ph) / (M[3] @ ph), (M[1] @ ph) / (M[3] @ ph)


SPECS = [
    small_spec(),
    small_spec(64, 48, 12, rx=2.5, sx=0.3, sy=-0.2),
    small_spec(60, 40, 8, so=80.0, sd=900.0, px=0.15, module_number=3),
    small_spec(48, 32, 10, heli_offset=0.05, angle=720.0),
    small_spec(64, 32, 6, module_number=2, module_angle_offset=1.5, module_center_offset_z=300.0),
]


@pytest.mark.parametrize("spec", SPECS)
def test_matrix_equals_physical_projection(spec):
    pd, grid = O.geometry(spec, fp64=True)
    rng = np.random.default_rng(1)
    half = 0.5 * grid["vwidth"] * grid["vx"]
    worst = 0.0
    for a in range(spec.projections):
        for m in range(spec.module_number):
            for _ in range(20):
                p = rng.uniform(-half, half, 3)
                u, v = _proj(pd[a * spec.module_number + m], p)
                up, vp = O.project_physical(spec, a, m, *p)
                worst = max(worst, abs(u - up), abs(v - vp))
    assert worst < 1e-8


@pytest.mark.parametrize("rx,sx,sy", [(0.0, 0.0, 0.0), (2.5, 0.3, -0.2), (-4.0, 0.0, 0.1)])
def test_rotation_centre_lands_on_axis_column(rx, sx, sy):
    spec = small_spec(64, 48, 9, rx=rx, sx=sx, sy=sy)
    for a in range(spec.projections):
        u, v = O.project_physical(spec, a, 0, 0.0, 0.0, 0.0)
        assert u == pytest.approx(spec.dwidth / 2 + rx, abs=1e-9)
        assert v == pytest.approx(spec.dheight / 2, abs=1e-9)


def test_magnification_at_angle_zero():
    spec = small_spec(64, 48, 4, so=120.0, sd=960.0, px=0.2)
    for (x, y, z) in [(0.3, 0.0, 0.0), (0.0, -0.25, 0.0), (0.2, 0.1, 5.0), (-0.4, 0.3, -7.5)]:
        u, v = O.project_physical(spec, 0, 0, x, y, z)
        mag = spec.sd / (spec.so + z)
        assert u == pytest.approx(spec.dwidth / 2 + x * mag / spec.px, abs=1e-9)
        assert v == pytest.approx(spec.dheight / 2 + y * mag / spec.px, abs=1e-9)


def test_rotation_is_about_plus_y():
    spec = small_spec(64, 48, 8)   # angles 0, 45, ..., 315
    rng = np.random.default_rng(2)
    for a in range(spec.projections):
        th = math.radians(a * 360.0 / spec.projections)
        for _ in range(5):
            x, y, z = rng.uniform(-0.3, 0.3, 3)
            # the source turns by +th about +Y == the object turns by -th
            xr = x * math.cos(-th) + z * math.sin(-th)
            zr = -x * math.sin(-th) + z * math.cos(-th)
            assert O.project_physical(spec, a, 0, x, y, z) == pytest.approx(
                O.project_physical(spec, 0, 0, xr, y, zr), abs=1e-9)


def test_helical_shift():
    spec = small_spec(32, 24, 10, heli_offset=0.04)
    _, grid = O.geometry(spec, fp64=True)
    assert grid["vheight"] == spec.dheight + int(abs(0.04 / grid["vx"]))
    mag = spec.sd / spec.so
    for a in range(spec.projections):
        h = (0.04 / spec.projections) * a - 0.02
        _, v = O.project_physical(spec, a, 0, 0.0, 0.0, 0.0)
        assert v == pytest.approx(spec.dheight / 2 - h * mag / spec.px, abs=1e-9)


def test_grid():
    spec = small_spec(40, 30, 6, so=90.0, sd=1100.0, px=0.12)
    _, grid = O.geometry(spec, fp64=True)
    vx = 0.12 * 90.0 / 1100.0
    assert grid["vx"] == pytest.approx(vx, rel=1e-15)
    assert (grid["vwidth"], grid["vheight"]) == (40, 30)
    assert grid["orig"] == pytest.approx((-20 * vx, -15 * vx, -20 * vx), rel=1e-15)


def test_central_module_matches_single_module():
    one = small_spec(60, 40, 6)
    three = small_spec(60, 40, 6, module_number=3)
    rng = np.random.default_rng(3)
    for a in range(6):
        for _ in range(5):
            p = rng.uniform(-0.05, 0.05, 3)
            assert O.project_physical(three, a, 1, *p) == pytest.approx(
                O.project_physical(one, a, 0, *p), abs=1e-9)


@pytest.mark.parametrize("spec", SPECS[:3])
def test_float32_compat_close_to_fp64(spec):
    pd64, _ = O.geometry(spec, fp64=True)
    pd32, g32 = O.geometry(spec, fp64=False)
    assert g32["vx"] == pytest.approx(spec.px * spec.so / spec.sd, rel=1e-6)
    rng = np.random.default_rng(4)
    half = 0.5 * spec.dwidth * g32["vx"]
    worst = 0.0
    for k in range(pd64.shape[0]):
        for _ in range(10):
            p = rng.uniform(-half, half, 3)
            a = np.subtract(_proj(pd64[k], p), _proj(pd32[k], p))
            worst = max(worst, float(np.abs(a).max()))
    assert worst < 5e-3
    np.testing.assert_array_equal(pd32[:, 16:], pd64[:, 16:].astype(np.float32))


@pytest.mark.parametrize("roll", [0.3, -7.5, 30.0])
def test_roll_is_a_rotation_of_the_detector_image(roll):
    """A detector rolled by phi about the beam axis through its centre lies in
    the same plane; a point's pixel coordinates are the unrolled ones turned
    about the detector centre (W/2, H/2) by -phi in the (column, row) frame:
    (u' - W/2, v' - H/2) = [[cos, sin], [-sin, cos]] (u - W/2, v - H/2)."""
    kw = dict(dwidth=64, dheight=48, projections=7, rx=2.5)
    flat = O.GeometrySpec(**kw)
    rolled = O.GeometrySpec(**kw, roll=roll)
    c, s = math.cos(math.radians(roll)), math.sin(math.radians(roll))
    rng = np.random.default_rng(3)
    pd, grid = O.geometry(rolled, fp64=True)
    half = 0.5 * grid["vwidth"] * grid["vx"]
    for a in range(7):
        for _ in range(20):
            p = rng.uniform(-half, half, 3)
            u0, v0 = O.project_physical(flat, a, 0, *p)
            u1, v1 = O.project_physical(rolled, a, 0, *p)
            du, dv = u0 - 32.0, v0 - 24.0
            assert u1 - 32.0 == pytest.approx(c * du + s * dv, abs=1e-9)
            assert v1 - 24.0 == pytest.approx(-s * du + c * dv, abs=1e-9)
            # and the matrices project the same way
            mu, mv = _proj(pd[a], p)
            assert mu == pytest.approx(u1, abs=1e-8) and mv == pytest.approx(v1, abs=1e-8)


def test_pitch_and_yaw_matrices_equal_physical_projection():
    """Pitch and yaw tilt the detector plane: the matrices still agree with the
    ray/plane intersection through the turned corners, and a point on the
    beam axis through the detector centre stays at the centre."""
    spec = small_spec(64, 48, 6, pitch=4.0, yaw=-6.0, roll=2.0)
    pd, grid = O.geometry(spec, fp64=True)
    rng = np.random.default_rng(4)
    half = 0.5 * grid["vwidth"] * grid["vx"]
    for a in range(6):
        for _ in range(10):
            p = rng.uniform(-half, half, 3)
            u, v = O.project_physical(spec, a, 0, *p)
            mu, mv = _proj(pd[a], p)
            assert mu == pytest.approx(u, abs=1e-8) and mv == pytest.approx(v, abs=1e-8)
        # the source-detector axis passes through the rotation centre (rx = 0)
        cu, cv = O.project_physical(spec, a, 0, 0.0, 0.0, 0.0)
        assert cu == pytest.approx(32.0, abs=1e-9) and cv == pytest.approx(24.0, abs=1e-9)
