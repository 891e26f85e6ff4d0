"""The library's C++ geometry (xpr_geometry) against the oracle's restatement
(itself pinned by closed forms in test_oracle_geometry.py)."""
import numpy as np
import pytest

from oracle import oracle as O
from xpulse_recon_amd.geometry import Geometry

CASES = [
    dict(dwidth=32, dheight=24, projections=12),
    dict(dwidth=64, dheight=48, projections=30, rx=2.5, sx=0.3, sy=-0.2),
    dict(dwidth=60, dheight=40, projections=8, so=80.0, sd=900.0, px=0.15, module_number=3),
    dict(dwidth=48, dheight=32, projections=10, heli_offset=0.05, angle=720.0),
    dict(dwidth=64, dheight=32, projections=6, module_number=2, module_angle_offset=1.5,
         module_center_offset_z=300.0),
    dict(dwidth=16, dheight=16, projections=5, angle_list=[0.0, 10.0, 33.3, 90.0, 181.0]),
]


@pytest.mark.parametrize("kw", CASES)
def test_float32_geometry_bit_exact(kw):
    pd, grid = Geometry(**kw).projdata(fp64=False)
    ref, rgrid = O.geometry(O.GeometrySpec(**kw), fp64=False)
    np.testing.assert_array_equal(pd, ref)
    assert grid == rgrid


@pytest.mark.parametrize("kw", CASES)
def test_fp64_geometry(kw):
    pd, grid = Geometry(**kw).projdata(fp64=True)
    ref, rgrid = O.geometry(O.GeometrySpec(**kw), fp64=True)
    np.testing.assert_allclose(pd, ref.astype(np.float32), rtol=2e-7, atol=1e-12)
    assert grid["vwidth"] == rgrid["vwidth"] and grid["vheight"] == rgrid["vheight"]
    assert grid["vx"] == pytest.approx(rgrid["vx"], rel=1e-15)


def test_module_rotation_extension():
    """roll about the beam axis turns the detector image: the centre stays put,
    a point on the central row moves off that row by the roll angle."""
    base = Geometry(dwidth=64, dheight=64, projections=4)
    rolled = Geometry(dwidth=64, dheight=64, projections=4, roll=3.0)
    pd0, _ = base.projdata(fp64=True)
    pd1, _ = rolled.projdata(fp64=True)

    def proj(pd, p):
        M = pd[:16].astype(np.float64).reshape(4, 4)
        ph = np.array([*p, 1.0])
        return np.array([M[0] @ ph, M[1] @ ph]) / (M[3] @ ph)

    np.testing.assert_allclose(proj(pd1[0], (0, 0, 0)), proj(pd0[0], (0, 0, 0)), atol=1e-4)
    du, dv = proj(pd1[0], (0.2, 0, 0)) - proj(pd1[0], (0, 0, 0))
    assert abs(np.degrees(np.arctan2(dv, du))) == pytest.approx(3.0, abs=0.05)


def test_bad_geometry_raises():
    from xpulse_recon_amd import XprError
    with pytest.raises(XprError, match="positive"):
        Geometry(dwidth=8, dheight=8, projections=2, so=0.0).projdata()
    with pytest.raises(ValueError):
        Geometry(dwidth=8, dheight=8, projections=2, angle_list=[1.0]).projdata()


@pytest.mark.parametrize("rot", [dict(roll=0.3), dict(roll=-7.5, rx=2.5), dict(pitch=4.0, yaw=-6.0),
                                 dict(roll=2.0, pitch=-3.0, yaw=5.0, module_number=2)])
def test_module_rotation_matches_oracle(rot):
    """The library's roll/pitch/yaw extension against the oracle's own FP64
    construction (rotation matrices, itself pinned by the rolled-image closed
    form): FP64 ProjData to float32 rounding, float32 ProjData projecting every
    test point within 5e-3 px of the FP64 oracle."""
    kw = dict(dwidth=64, dheight=48, projections=9, **rot)
    pd64, _ = Geometry(**kw).projdata(fp64=True)
    pd32, grid = Geometry(**kw).projdata(fp64=False)
    ref, _ = O.geometry(O.GeometrySpec(**kw), fp64=True)
    np.testing.assert_allclose(pd64, ref.astype(np.float32), rtol=2e-7, atol=1e-12)
    rng = np.random.default_rng(5)
    half = 0.5 * grid["vwidth"] * grid["vx"]

    def proj(pd, p):
        M = np.asarray(pd[:16], np.float64).reshape(4, 4)
        ph = np.array([*p, 1.0])
        return np.array([M[0] @ ph, M[1] @ ph]) / (M[3] @ ph)

    for r in range(pd32.shape[0]):
        for _ in range(10):
            p = rng.uniform(-half, half, 3)
            np.testing.assert_allclose(proj(pd32[r], p), proj(ref[r], p), atol=5e-3)
