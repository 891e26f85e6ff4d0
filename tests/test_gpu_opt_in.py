"""The opt-in modes outside the reference's arithmetic, on the GPU through the
C-ABI, against the oracle: bilinear sampling in the backprojector
(xpr_proj_args.interp = XPR_INTERP_BILINEAR) and float detector sums
(acc_bits XPR_ACC_F32, a float32 sums tensor here).

Bars:
  * bilinear Backward, FP32 storage: the oracle's FP32 canonical form
    (backward_bilinear_f32) within the expf tolerance, rtol 4e-7 (the sums
    are the same floats added in the same order); inside segments equal the
    checked ones bit for bit; the FP64 model (backward_bilinear_f64nn)
    within rtol 5e-5 per voxel and 5e-6 in relative L2 (u carries about two
    ulp of float rounding, ~1e-5 px at 130 px, which moves the weights; taps
    of these N(0, 0.4) error images differ by 0.57 rms and up to ~3; see
    tests/test_oracle_bilinear.py).
  * float Forward: the voxel values are quantised to Q.22 as in the integer
    sums (FLOAT_TO_FIXED, constants.cl:22) and binned exactly in integer LDS
    tiles; only the additions into the detector pixels are float.  Each
    pixel is within the summation bound of its exact Q.22 sum S (the
    oracle's u64 sums x 2^-22): every term (a tile entry or a direct
    voxel-view) is rounded once to float and at most n - 1 float additions
    follow, n the pixel's voxel-views, so |s - S| <= n 2^-24 S (1 + 2 n 2^-24);
    relative L2 below 1e-6.  Volumes of multiples of 2^-8 below 1 (every
    term and partial sum exact in float) give sums equal to the Q.22 sums x
    2^-22 bit for bit, whatever the order of the atomics.
  * complete reconstructions through XPRecon with "accumulator": "f32" and
    with "interp": "bilinear", against the oracle's FP64 naive schedule on
    the float32 pixel choice with the same Q.22 quantisation and sampling: relative RMSE
    below 1e-4 of the volume norm.  The bilinear gather is not the adjoint
    of the nearest-pixel forward; its reconstruction error against the
    phantom stays within 1.6x that of the nearest-pixel reconstruction of
    the same scan (1.2-1.6x measured over jitter draws on the oracle)."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import geom_args, oracle_geom, projdata_for, rel_l2, small_spec
from tests.test_gpu_kernels import CASES, INSIDE_CASES, _case
from tests.tiff_util import read_f32, write_f32

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "xpulse_recon_amd", "bin", "XPRecon")
Q = 1 << 22


def _args(d, bilinear=False):
    import xpulse_recon_amd as X
    from xpulse_recon_amd import _lib
    return X.proj_args(d["det_width"], d["n_views"], d["modules"], d["vmin"], d["vmax"],
                       d["line_offset"], d["line_rows"], d["origin"], d["voxel"],
                       interp=_lib.XPR_INTERP_BILINEAR if bilinear else _lib.XPR_INTERP_NEAREST)


def _bilinear_gpu(cuda, vol, err, pd, d, dtype=None):
    import torch
    import xpulse_recon_amd as X
    dtype = dtype or torch.float32
    tv = torch.from_numpy(vol).to(cuda).to(dtype)
    empty = torch.zeros(1, dtype=torch.int64, device=cuda)
    X.backward(tv, torch.from_numpy(err).to(cuda).to(dtype), torch.from_numpy(pd).to(cuda),
               _args(d, bilinear=True), empty=empty)
    return tv.float().cpu().numpy(), int(empty.item())


@pytest.mark.parametrize("tilt", ["none", "roll", "pitch", "residue"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_bilinear_backward_parity(cuda, case, tilt):
    spec, pd, d = _case(case, tilt)
    nx, nl, nk = (d["vmax"][i] - d["vmin"][i] for i in range(3))
    rng = np.random.default_rng(300 + case)
    vol = rng.uniform(0.5, 2.0, (nl, nk, nx)).astype(np.float32)
    err = rng.normal(0.0, 0.4, (d["n_views"], d["line_rows"], d["det_width"])).astype(np.float32)
    print(f"bilinear backward case {case} tilt {tilt}", flush=True)
    got, empty = _bilinear_gpu(cuda, vol, err, pd, d)
    ref = vol.copy()
    empty_ref = O.backward_bilinear_f32(ref, err, pd, oracle_geom(d), 8)
    assert empty == empty_ref
    np.testing.assert_allclose(got, ref, rtol=4e-7, atol=0)
    ref64 = vol.astype(np.float64)
    O.backward_bilinear_f64nn(ref64, err.astype(np.float64), pd, oracle_geom(d), 8)
    np.testing.assert_allclose(got, ref64, rtol=5e-5, atol=0)
    print(f"vs FP64 model: relative L2 {rel_l2(got, ref64):.3e}", flush=True)
    assert rel_l2(got, ref64) < 5e-6


@pytest.mark.parametrize("case", range(len(INSIDE_CASES)))
def test_bilinear_inside_equal_checked(cuda, monkeypatch, case):
    """Bilinear row segments proven inside skip the per-voxel test: the volume
    equals the checked one bit for bit, FP32 and FP16 storage."""
    import torch
    kw, views, crop = INSIDE_CASES[case]
    kw = dict(kw)
    spec = small_spec(kw.pop("w"), kw.pop("h"), kw.pop("n"), **kw)
    pd, grid = projdata_for(spec, views)
    lo, rows = crop if crop else (0, None)
    d = geom_args(spec, grid, len(views), jitter=(0.002, 0.004, 0.001), line_offset=lo,
                  line_rows=rows)
    W, H = grid["vwidth"], grid["vheight"]
    rng = np.random.default_rng(600 + case)
    vol = rng.uniform(0.5, 2.0, (H, W, W)).astype(np.float32)
    err = rng.normal(0.0, 0.4, (d["n_views"], d["line_rows"], d["det_width"])).astype(np.float32)
    print(f"bilinear inside case {case}", flush=True)
    out = {}
    for inside in ("1", "0"):
        monkeypatch.setenv("XPR_BACKWARD_INSIDE", inside)
        out[inside] = _bilinear_gpu(cuda, vol, err, pd, d)
        out[inside + "h"] = _bilinear_gpu(cuda, vol, err, pd, d, torch.float16)
    np.testing.assert_array_equal(out["1"][0], out["0"][0])
    assert out["1"][1] == out["0"][1]
    np.testing.assert_array_equal(out["1h"][0], out["0h"][0])
    ref = vol.copy()
    assert O.backward_bilinear_f32(ref, err, pd, oracle_geom(d), 8) == out["1"][1]
    np.testing.assert_allclose(out["1"][0], ref, rtol=4e-7, atol=0)


def test_bilinear_16bit_storage(cuda):
    """FP16 volume and error: taps widened to float, FP32 weights and sums;
    against the oracle fed the same FP16 values, within one FP16 ulp of the
    stored result."""
    import torch
    spec, pd, d = _case(1)
    nx, nl, nk = (d["vmax"][i] - d["vmin"][i] for i in range(3))
    rng = np.random.default_rng(310)
    vol = rng.uniform(0.5, 2.0, (nl, nk, nx)).astype(np.float16).astype(np.float32)
    err = rng.normal(0.0, 0.4, (d["n_views"], d["line_rows"], d["det_width"]))
    err = err.astype(np.float16).astype(np.float32)
    got, empty = _bilinear_gpu(cuda, vol, err, pd, d, torch.float16)
    ref = vol.copy()
    assert O.backward_bilinear_f32(ref, err, pd, oracle_geom(d), 8) == empty
    np.testing.assert_allclose(got, ref, rtol=2.0 ** -10 * 1.01, atol=0)


def _float_forward(cuda, vol, pd, d):
    import torch
    import xpulse_recon_amd as X
    sums = torch.zeros((d["n_views"], d["line_rows"], d["det_width"]), dtype=torch.float32,
                       device=cuda)
    X.forward(torch.from_numpy(vol).to(cuda), sums, torch.from_numpy(pd).to(cuda), _args(d))
    return sums.cpu().numpy()


@pytest.mark.parametrize("tilt", ["none", "roll", "pitch"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_float_forward_dyadic_values_exact(cuda, case, tilt):
    """Multiples of 2^-8 below 1: every partial sum of a pixel is exact in
    float, so the float sums equal the Q.22 sums x 2^-22 bit for bit."""
    spec, pd, d = _case(case, tilt)
    nx, nl, nk = (d["vmax"][i] - d["vmin"][i] for i in range(3))
    rng = np.random.default_rng(400 + case)
    vol = (rng.integers(0, 256, (nl, nk, nx)) / 256.0).astype(np.float32)
    ref = np.zeros((d["n_views"], d["line_rows"], d["det_width"]), np.uint64)
    O.forward_f32(vol, ref, pd, oracle_geom(d))
    assert ref.max() < (1 << 37)                  # sums below 2^15: exact in float
    print(f"float forward dyadic case {case} tilt {tilt}", flush=True)
    got = _float_forward(cuda, vol, pd, d)
    np.testing.assert_array_equal(got, (ref.astype(np.float64) / Q).astype(np.float32))


@pytest.mark.parametrize("tilt", ["none", "roll", "pitch"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_float_forward_within_summation_bound(cuda, case, tilt):
    spec, pd, d = _case(case, tilt)
    nx, nl, nk = (d["vmax"][i] - d["vmin"][i] for i in range(3))
    rng = np.random.default_rng(450 + case)
    vol = rng.uniform(0.0, 2.0, (nl, nk, nx)).astype(np.float32)
    g = oracle_geom(d)
    shape = (d["n_views"], d["line_rows"], d["det_width"])
    q22 = np.zeros(shape, np.uint64)
    O.forward_f32(vol, q22, pd, g)
    exact = q22.astype(np.float64) / Q
    cnt = np.zeros(shape)
    O.forward_f64nn(np.ones(vol.shape), cnt, pd, g, 8)
    print(f"float forward case {case} tilt {tilt}", flush=True)
    got = _float_forward(cuda, vol, pd, d).astype(np.float64)
    u = 2.0 ** -24
    bound = cnt * u * exact * (1.0 + 2.0 * cnt * u)
    assert (np.abs(got - exact) <= bound).all()
    assert rel_l2(got, exact) < 1e-6


def test_float_sums_subiteration(cuda):
    """xpr_subiteration with float sums: sums from 0, forward, error on the
    float sums, backward; against the oracle's pieces (exact Q.22 sums rounded
    to float, error in float64 from them, FP32 backward), rtol 1e-5."""
    import torch
    import xpulse_recon_amd as X
    spec, pd, d = _case(1)
    nx, nl, nk = (d["vmax"][i] - d["vmin"][i] for i in range(3))
    rng = np.random.default_rng(470)
    vol = rng.uniform(0.5, 2.0, (nl, nk, nx)).astype(np.float32)
    shape = (d["n_views"], d["line_rows"], d["det_width"])
    img = rng.uniform(1.0, 40.0, shape).astype(np.float32)
    tv = torch.from_numpy(vol).to(cuda)
    sums = torch.empty(shape, dtype=torch.float32, device=cuda)
    err = torch.empty(shape, dtype=torch.float32, device=cuda)
    X.subiteration(tv, torch.from_numpy(img).to(cuda), torch.from_numpy(pd).to(cuda), _args(d),
                   0.8, False, sums, err)
    g = oracle_geom(d)
    q22 = np.zeros(shape, np.uint64)
    O.forward_f32(vol, q22, pd, g)
    s32 = (q22.astype(np.float64) / Q).astype(np.float32).astype(np.float64)
    eps = 2.0 ** -22
    e = (np.log(np.maximum(img / np.maximum(s32, eps), eps)) * np.float32(0.8)).astype(np.float32)
    np.testing.assert_allclose(err.cpu().numpy(), e, rtol=1e-5, atol=1e-6)
    ref = vol.copy()
    O.backward_f32(ref, e, pd, g, 8)
    np.testing.assert_allclose(tv.cpu().numpy(), ref, rtol=1e-5)


W_REC = H_REC = 64
N_REC, SIT_REC, IT_REC = 90, 10, 2


@pytest.fixture(scope="module")
def scan(tmp_path_factory):
    from xpulse_recon_amd import synth
    t = tmp_path_factory.mktemp("optin")
    spec = O.GeometrySpec(dwidth=W_REC, dheight=H_REC, projections=N_REC)
    pd32, grid = O.geometry(spec, False)
    ell = synth.random_ellipsoids(1101, 12, W_REC * grid["vx"], 128)
    oe = [O.make_ellipsoid(e["c"], e["a"], e["rot"], e["mu"]) for e in ell]
    truth = O.phantom_voxels(grid, oe) * grid["vx"]
    rng = np.random.default_rng(1102)
    d = t / "in"
    d.mkdir()
    raw = np.empty((N_REC, H_REC, W_REC), np.float32)
    for a in range(N_REC):
        raw[a] = (rng.poisson(1e4 * np.exp(-O.line_integrals(spec, a, oe))) / 1e4).astype(np.float32)
        write_f32(d / f"proj_{a:04d}.tif", raw[a])
    imgs = raw.copy()
    O.preprocess(imgs.reshape(-1), False, True)
    return t, d, pd32, grid, imgs, truth


def _cli(t, d, name, xpr):
    out, log = t / name, t / f"{name}.json"
    params = {"recon": {"input": str(d), "output": str(out), "it": IT_REC, "sit": SIT_REC,
                        "method": "naive"},
              "xpr": dict(xpr, jitter_seed=31, run_log=str(log))}
    pf = t / f"{name}_params.json"
    pf.write_text(json.dumps(params))
    print(f"XPRecon {name} {xpr}", flush=True)
    r = subprocess.run([CLI, str(pf)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    vol = np.stack([read_f32(f) for f in sorted(out.glob("layer_*.tif"))])
    return vol, json.loads(log.read_text())


_RMSE = {}


@pytest.mark.parametrize("mode", ["nearest", "float_sums", "bilinear", "both"])
def test_reconstruction_with_opt_in_modes(cuda, scan, mode):
    t, d, pd32, grid, imgs, truth = scan
    xpr = {"nearest": {}, "float_sums": {"accumulator": "f32"},
           "bilinear": {"interp": "bilinear"},
           "both": {"accumulator": "f32", "interp": "bilinear"}}[mode]
    vol, log = _cli(t, d, mode, xpr)
    assert log["interp"] == ("bilinear" if "interp" in xpr else "nearest")
    assert log["accumulator"] == ("f32" if "accumulator" in xpr else 64)
    jit = np.array(log["jitter"], np.float32)
    ref = O.reconstruct_naive_f64nn(imgs, pd32, grid, IT_REC, SIT_REC, 1.0, 0.9, jit,
                                    threads=16, q22=True,
                                    interp=xpr.get("interp", "nearest"))
    rel = rel_l2(vol, ref)
    rmse = float(np.sqrt(np.mean((vol.astype(np.float64) - truth) ** 2)))
    print(json.dumps({"mode": mode, "vs_fp64_model_rel": rel, "rmse_vs_truth": rmse}), flush=True)
    assert np.isfinite(vol).all()
    assert rel < 1e-4
    _RMSE[mode] = rmse
    if "interp" in xpr and "nearest" in _RMSE:
        assert rmse <= 1.6 * _RMSE["nearest"]
