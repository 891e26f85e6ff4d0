"""Pins of the oracle's FP64 line integrals (orc_line_integrals), on which
every reconstruction-error figure against a phantom rests, by closed forms.

At angle 0 with no detector shift or axis offset the geometry is explicit:
the source sits at (0, 0, -so) and pixel (col, row) is centred at
(px (col - W/2), px (row - H/2), sd - so) (geometry.hpp:58-68, 244-265; pixel
k centred at coordinate k).  A sphere of radius r whose centre lies at
distance d from a ray adds mu * 2 sqrt(r^2 - d^2); an axis-aligned ellipsoid
centred at (cx, cy, cz) adds mu * 2 a_z sqrt(1 - (cx/a_x)^2 - (cy/a_y)^2) to
the central ray (which runs along z through x = y = 0); disjoint bodies add."""
import math

import numpy as np
import pytest

from oracle import oracle as O

W, H = 48, 40
SPEC = O.GeometrySpec(dwidth=W, dheight=H, projections=4)


def _ray(col, row):
    src = np.array([0.0, 0.0, -SPEC.so])
    pix = np.array([SPEC.px * (col - W / 2), SPEC.px * (row - H / 2), SPEC.sd - SPEC.so])
    d = pix - src
    return src, d / np.linalg.norm(d)


def _sphere_chord(c, r, src, d):
    q = np.asarray(c) - src
    x = np.cross(q, d)                 # |q x d| = distance, without cancellation
    dist2 = float(x @ x)
    return 2.0 * math.sqrt(r * r - dist2) if dist2 < r * r else 0.0


# the detector (48 x 40 px of 0.1 mm, magnification 10) sees +-0.24 mm by
# +-0.2 mm of the object
@pytest.mark.parametrize("centre,r", [((0.0, 0.0, 0.0), 0.2), ((0.1, -0.05, 0.08), 0.12),
                                      ((-0.12, 0.1, -0.05), 0.08), ((0.0, 0.0, 0.0), 0.9)])
def test_sphere_chords_every_pixel(centre, r):
    mu = 0.7
    e = O.make_ellipsoid(centre, (r, r, r), np.eye(3), mu)
    got = O.line_integrals(SPEC, 0, [e])
    ref = np.zeros((H, W))
    for row in range(H):
        for col in range(W):
            ref[row, col] = mu * _sphere_chord(centre, r, *_ray(col, row))
    # rays grazing the sphere lose digits to cancellation in the square root:
    # an absolute floor of 1e-9 (the chords reach 1 mm)
    np.testing.assert_allclose(got, ref, rtol=1e-9, atol=1e-9)
    assert (ref > 0).sum() > 20


@pytest.mark.parametrize("c,a", [((0.0, 0.0, 0.1), (0.5, 0.3, 0.8)),
                                 ((0.2, -0.1, -0.3), (0.6, 0.4, 0.25))])
def test_axis_aligned_ellipsoid_central_ray(c, a):
    mu = 1.3
    e = O.make_ellipsoid(c, a, np.eye(3), mu)
    got = O.line_integrals(SPEC, 0, [e])
    expect = mu * 2.0 * a[2] * math.sqrt(1.0 - (c[0] / a[0]) ** 2 - (c[1] / a[1]) ** 2)
    assert got[H // 2, W // 2] == pytest.approx(expect, rel=1e-10)


def test_rotation_of_a_sphere_changes_nothing_and_bodies_add():
    r1 = O.make_ellipsoid((0.2, 0.1, 0.0), (0.4, 0.4, 0.4), np.eye(3), 0.5)
    a = 0.7
    rot = np.array([[math.cos(a), -math.sin(a), 0], [math.sin(a), math.cos(a), 0], [0, 0, 1]])
    r2 = O.make_ellipsoid((0.2, 0.1, 0.0), (0.4, 0.4, 0.4), rot, 0.5)
    np.testing.assert_allclose(O.line_integrals(SPEC, 0, [r1]), O.line_integrals(SPEC, 0, [r2]),
                               rtol=1e-12, atol=1e-14)
    s = O.make_ellipsoid((-0.6, -0.5, 0.2), (0.2, 0.2, 0.2), np.eye(3), 2.0)
    np.testing.assert_allclose(O.line_integrals(SPEC, 0, [r1, s]),
                               O.line_integrals(SPEC, 0, [r1]) + O.line_integrals(SPEC, 0, [s]),
                               rtol=1e-12, atol=1e-14)


def test_views_rotate_the_object_about_y():
    """View a sees the object turned by -a about +Y (geometry.hpp:271-272): a
    sphere at (x, 0, 0) seen at 90 deg equals, at 0 deg, the sphere whose
    centre is (x, 0, 0) turned by -90 deg about +Y."""
    spec = O.GeometrySpec(dwidth=W, dheight=H, projections=4)   # views at 0, 90, 180, 270
    x = 0.35
    e_at_x = O.make_ellipsoid((x, 0.0, 0.0), (0.3, 0.3, 0.3), np.eye(3), 1.0)
    v90 = O.line_integrals(spec, 1, [e_at_x])
    # the equivalent body for the view at 0: the centre rotated by -90 deg about +Y
    rad = -math.pi / 2
    c = (x * math.cos(rad), 0.0, -x * math.sin(rad))
    e_rot = O.make_ellipsoid(c, (0.3, 0.3, 0.3), np.eye(3), 1.0)
    v0 = O.line_integrals(spec, 0, [e_rot])
    np.testing.assert_allclose(v90, v0, rtol=1e-9, atol=1e-9)
    assert v0.max() > 0.1
