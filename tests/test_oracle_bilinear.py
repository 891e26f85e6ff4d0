"""Pins of the oracle's opt-in bilinear backprojector (backward_bilinear_f32,
backward_bilinear_f64nn), by closed forms computed here from first
principles, not through the oracle's code path.

The sampler counts the voxel-views the nearest-neighbour bounds test accepts
(KernelForward.cl:34-35) and clamps its taps to the accepted pixels, so on
an affine image E[row][col] = a col + b row + c a voxel that projects to
(u, v) gathers a clamp(u, cl, cr-1) + b clamp(v, rl, rr-1) + c: one view
updates it to the exponential of that (u, v from the float32 voxel position
and ProjData evaluated here in double).  With a != b a swapped axis, a
transposed tap or a weight taken from the wrong side fails; a constant image
(weights summing to 1) gives v e^c; its accept set is the nearest-neighbour
hit map (hits_f32) exactly; an image that is NaN outside the accepted pixels
shows that no tap leaves them."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import geom_args, oracle_geom, projdata_for, small_spec


def _positions(d, W, H):
    """float32 voxel positions fmaf(idx, vx, origin) as float64, [H, W, W]
    (layer, depth, x); the float64 sum of the exact product and the origin,
    rounded once to float32, is the fused result."""
    def axis(n, a):
        idx = np.arange(n, dtype=np.float64) + d["vmin"][a]
        return (idx * np.float64(np.float32(d["voxel"][a])) +
                np.float64(np.float32(d["origin"][a]))).astype(np.float32).astype(np.float64)
    px, py, pz = axis(W, 0), axis(H, 1), axis(W, 2)
    return np.broadcast_arrays(px[None, None, :], py[:, None, None], pz[None, :, None])


def _project(pd_row, px, py, pz, line_offset):
    p = pd_row.astype(np.float64)
    X = p[0] * px + p[1] * py + p[2] * pz + p[3]
    Y = p[4] * px + p[5] * py + p[6] * pz + p[7]
    Wd = p[12] * px + p[13] * py + p[14] * pz + p[15]
    return X / Wd, Y / Wd - line_offset


def _linear_image(n, rows, cols, a, b, c):
    r, k = np.meshgrid(np.arange(rows), np.arange(cols), indexing="ij")
    return np.broadcast_to(a * k + b * r + c, (n, rows, cols)).astype(np.float64).copy()


@pytest.mark.parametrize("tilt", [False, True])
def test_affine_image_is_reproduced(tilt):
    kw = dict(rx=1.5, roll=0.3, pitch=0.2) if tilt else dict(rx=1.5)
    spec = small_spec(40, 36, 8, **kw)
    # (a rolled and pitched detector: FP64 matrices rounded to float32)
    pd, grid = projdata_for(spec, [3], fp64=tilt)
    pd = pd.astype(np.float32)
    d = geom_args(spec, grid, 1, jitter=(0.001, 0.002, -0.001))
    g = oracle_geom(d)
    W, H = grid["vwidth"], grid["vheight"]
    a, b, c = 0.013, -0.021, 0.1
    E = _linear_image(1, spec.dheight, spec.dwidth, a, b, c)
    vol = np.ones((H, W, W))
    empty = O.backward_bilinear_f64nn(vol, E, pd, g)
    px, py, pz = _positions(d, W, H)
    u, v = _project(pd[0], px, py, pz, 0)
    cu = np.clip(u, 0, spec.dwidth - 1)
    cv = np.clip(v, 0, spec.dheight - 1)
    expect = np.exp(a * cu + b * cv + c)
    hit = vol != 1.0
    assert hit.sum() > 0.3 * vol.size and empty == vol.size - hit.sum()
    # the accept set is the nearest-neighbour hit map, voxel for voxel
    np.testing.assert_array_equal(hit, O.hits_f32(pd, g).reshape(vol.shape) > 0)
    # a voxel within 1e-6 px of a detector edge may clamp on the other side
    # of it in float32 than in double (a difference of at most a 1e-6 px)
    np.testing.assert_allclose(vol[hit], expect[hit], rtol=1e-7)
    interior = hit & (u > 0.5) & (u < spec.dwidth - 1.5) & (v > 0.5) & (v < spec.dheight - 1.5)
    np.testing.assert_allclose(vol[interior], expect[interior], rtol=1e-12)
    assert (hit & ~interior).sum() > 0            # some voxels sample clamped taps
    # swapped axes are told apart
    assert np.abs(np.exp(a * cv + b * cu + c)[hit] - vol[hit]).max() > 1e-2
    # the FP32 canonical form, against the same closed form
    vol32 = np.ones((H, W, W), np.float32)
    empty32 = O.backward_bilinear_f32(vol32, E.astype(np.float32), pd, g)
    assert empty32 == empty
    np.testing.assert_allclose(vol32[hit], expect[hit], rtol=2e-6)


def test_constant_image_multiplies_by_exp_c():
    """Weights sum to 1: a constant error c updates every voxel with a view
    to v e^c (several views, two modules, a cropped row band), and leaves the
    others unchanged."""
    spec = small_spec(40, 36, 10, rx=1.5, module_number=2, module_angle_offset=4.0)
    pd, grid = projdata_for(spec, [1, 4, 7])
    d = geom_args(spec, grid, 3, jitter=(0.002, 0.003, 0.001), line_offset=8, line_rows=20)
    g = oracle_geom(d)
    W, H = grid["vwidth"], grid["vheight"]
    rng = np.random.default_rng(71)
    v0 = rng.uniform(0.5, 2.0, (H, W, W))
    c = 0.375
    E = np.full((3, 20, 40), c)
    vol = v0.copy()
    empty = O.backward_bilinear_f64nn(vol, E, pd, g)
    hit = vol != v0
    assert 0 < empty < vol.size and empty == vol.size - hit.sum()
    np.testing.assert_allclose(vol[hit], v0[hit] * np.exp(c), rtol=1e-14)
    vol32 = v0.astype(np.float32)
    assert O.backward_bilinear_f32(vol32, E.astype(np.float32), pd, g) == empty
    ref32 = (v0.astype(np.float32) * np.float32(np.exp(np.float32(c)))).astype(np.float32)
    np.testing.assert_allclose(vol32[hit], ref32[hit], rtol=2.5e-7)
    np.testing.assert_array_equal(vol32[~hit], v0.astype(np.float32)[~hit])


def test_taps_stay_inside_the_viewport():
    """A viewport narrower than the detector on every side, the error image
    NaN outside its accepted pixels (columns ceil(vpl)..ceil(vpr)-1, rows
    likewise): no counted voxel-view reads a NaN tap, the voxels counted are
    the nearest-neighbour hit map's, and on an affine image inside the
    viewport each gathers the closed form at the clamped coordinate."""
    spec = small_spec(40, 36, 8, rx=1.5)
    pd, grid = projdata_for(spec, [2])
    pd = pd.copy()
    pd[:, 16], pd[:, 17], pd[:, 18], pd[:, 19] = 3.5, 30.2, 2.0, 33.7
    d = geom_args(spec, grid, 1, jitter=(0.001, 0.0, 0.002))
    g = oracle_geom(d)
    W, H = grid["vwidth"], grid["vheight"]
    a, b, c = 0.011, 0.017, -0.2
    E = np.full((1, spec.dheight, spec.dwidth), np.nan)
    E[:, 2:34, 4:31] = _linear_image(1, spec.dheight, spec.dwidth, a, b, c)[:, 2:34, 4:31]
    vol = np.ones((H, W, W))
    O.backward_bilinear_f64nn(vol, E, pd, g)
    assert np.isfinite(vol).all()
    hit = vol != 1.0
    assert hit.sum() > 100
    np.testing.assert_array_equal(hit, O.hits_f32(pd, g).reshape(vol.shape) > 0)
    px, py, pz = _positions(d, W, H)
    u, v = _project(pd[0], px, py, pz, 0)
    expect = np.exp(a * np.clip(u, 4, 30) + b * np.clip(v, 2, 33) + c)
    np.testing.assert_allclose(vol[hit], expect[hit], rtol=1e-7)
    vol32 = np.ones((H, W, W), np.float32)
    O.backward_bilinear_f32(vol32, E.astype(np.float32), pd, g)
    assert np.isfinite(vol32).all()
    np.testing.assert_array_equal(vol32 != 1.0, hit)


def test_fp32_form_agrees_with_fp64_model():
    """Random error images, several views and modules: the FP32 canonical
    form within float rounding of the FP64 model on the same accept set.  The
    float u carries a few ulp of rounding (4e-6 px at 48 px), which moves the
    weight fx by as much; with taps differing by up to 0.6 a sample moves by
    up to ~3e-6, hence rtol 1e-5 per voxel and 1e-6 in relative L2."""
    spec = small_spec(48, 40, 12, rx=2.0, module_number=2, module_angle_offset=3.0)
    pd, grid = projdata_for(spec, [0, 3, 5, 9])
    d = geom_args(spec, grid, 4, jitter=(0.003, 0.001, 0.002))
    g = oracle_geom(d)
    W, H = grid["vwidth"], grid["vheight"]
    rng = np.random.default_rng(72)
    E = rng.uniform(-0.3, 0.3, (4, spec.dheight, spec.dwidth)).astype(np.float32)
    v0 = rng.uniform(0.5, 2.0, (H, W, W)).astype(np.float32)
    vol32 = v0.copy()
    vol64 = v0.astype(np.float64)
    e32 = O.backward_bilinear_f32(vol32, E, pd, g, 4)
    e64 = O.backward_bilinear_f64nn(vol64, E.astype(np.float64), pd, g, 4)
    assert e32 == e64
    np.testing.assert_allclose(vol32, vol64, rtol=1e-5)
    assert np.linalg.norm(vol32 - vol64) / np.linalg.norm(vol64) < 1e-6
