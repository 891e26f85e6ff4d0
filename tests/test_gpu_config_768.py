"""The reference's at-scale configurations at their full size, through the
C-ABI: 768^2 detector -> 768^3 grid (vx = px*so/sd), 720 views over 360 deg,
sit 20 (36 views per sub-iteration), so=100, sd=1000, px=0.1 mm, rx=2.5 px,
one module rolled 0.3 deg, slab height 64 (12 slabs); FP32 storage (the
plain-FP32 at-scale configuration) and FP16 / BF16 storage (the mixed
precision configuration).

At this size the projectors' exact-tile and inside fast paths decide most
bounds tests from rounding bounds whose magnitudes grow with the grid, so
every property below runs with them on and is checked against them off.

Bars:
* 12 slab forwards sum to the whole-volume forward bit for bit; the Q.22
  mass of an all-ones volume is 2^22 times a whole number of in-bounds
  voxel-views (KernelForward.cl:22-37, integer atomics);
* slab backwards equal the whole-volume backward bit for bit (per-voxel
  update, KernelBackward.cl:20-40);
* XPR_FORWARD_EXACT=0/1 and XPR_BACKWARD_INSIDE=0/1 bit-identical;
* a central 16-layer slab against the oracle: forward bit-exact (on the
  widened values for 16-bit storage), backward rtol 4e-7 (FP32) or one
  storage ulp (16-bit);
* the engine's method_1a (slab 64, streamed through the host layer store)
  equals its naive schedule bit for bit after one iteration of 20
  sub-iterations (reconstruction_lpln.hpp:54-92 vs reconstruction_n.hpp:51-78).
"""
import ctypes as C
import json

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

D, NV, SIT, SLAB = 768, 720, 20, 64
A_S = NV // SIT
Q = 1 << 22
SUB = 7                      # the sub-iteration whose views the kernel tests use
CENTRAL = (376, 392)         # the central 16 layers for the oracle


def _geometry():
    from xpulse_recon_amd.geometry import Geometry
    return Geometry(dwidth=D, dheight=D, projections=NV, so=100.0, sd=1000.0, px=0.1, rx=2.5,
                    roll=0.3)


@pytest.fixture(scope="module")
def scan(cuda):
    import torch
    import xpulse_recon_amd as X
    from xpulse_recon_amd import synth
    geom = _geometry()
    pd_all, grid = geom.projdata()
    W, H, vx = grid["vwidth"], grid["vheight"], np.float32(grid["vx"])
    assert (W, H) == (D, D)
    pd = geom.sit_projdata(pd_all, SIT, SUB)
    views = list(range(SUB, NV, SIT))[:A_S]
    ell = synth.random_ellipsoids(1002, 12, W * grid["vx"], 512)
    print("synthesising 36 views of 768^2", flush=True)
    tr = synth.transmission(synth.line_integrals(geom, views, ell, cuda), 5e4, 1003)
    img = torch.empty_like(tr)
    X.preprocess(tr, img, normalize=False, mlog=True)
    phantom = synth.phantom(grid, ell, cuda) * float(vx)
    jit = (0.31 * vx, 0.77 * vx, 0.12 * vx)
    origin = [np.float32(np.float32(grid["orig"][i]) + np.float32(jit[i])) for i in range(3)]
    g = torch.Generator().manual_seed(2002)
    rnd = (torch.rand((H, W, W), generator=g) * 1.5).to(cuda)
    torch.cuda.synchronize()
    return dict(geom=geom, pd=pd, tpd=torch.from_numpy(pd).to(cuda), img=img, phantom=phantom,
                random=rnd, origin=origin, vx=vx, W=W, H=H, ell=ell)


def _args(sc, l0=0, l1=D):
    import xpulse_recon_amd as X
    vx = sc["vx"]
    return X.proj_args(D, A_S, 1, (0, l0, 0), (D, l1, D), 0, D, sc["origin"], (vx, vx, vx))


def _volume(sc, kind, dtype, cuda):
    import torch
    tdt = {"f32": torch.float32, "f16": torch.float16, "bf16": torch.bfloat16}[dtype]
    if kind == "ones":
        return torch.ones((D, D, D), dtype=tdt, device=cuda)
    return sc[kind].to(tdt).contiguous()


def _sums(cuda):
    import torch
    return torch.zeros((A_S, D, D), dtype=torch.int64, device=cuda)


def _slabs():
    return [(l0, min(D, l0 + SLAB)) for l0 in range(0, D, SLAB)]


@pytest.mark.parametrize("dtype,kind", [("f32", "phantom"), ("f32", "random"), ("f32", "ones"),
                                        ("f16", "phantom"), ("f16", "random"),
                                        ("bf16", "phantom"), ("bf16", "random")])
def test_slab_forwards_equal_whole_forward(scan, cuda, dtype, kind):
    import torch
    import xpulse_recon_amd as X
    vol = _volume(scan, kind, dtype, cuda)
    print(f"768^3 forward {dtype} {kind}: whole and 12 slabs", flush=True)
    whole = _sums(cuda)
    X.forward(vol, whole, scan["tpd"], _args(scan))
    parts = _sums(cuda)
    for l0, l1 in _slabs():
        X.forward(vol[l0:l1], parts, scan["tpd"], _args(scan, l0, l1))
    assert torch.equal(parts, whole)
    # the ellipsoids cover about an eighth of the detector, random and ones
    # volumes all of the field of view
    assert int(whole.count_nonzero()) > (0.05 if kind == "phantom" else 0.5) * whole.numel()
    if kind == "ones":
        mass = int(whole.sum())
        assert mass % Q == 0
        hits = mass // Q
        assert 0.6 * D ** 3 * A_S < hits <= D ** 3 * A_S


@pytest.mark.parametrize("dtype", ["f32", "f16", "bf16"])
def test_fast_paths_equal_canonical_tests(scan, cuda, monkeypatch, dtype):
    import torch
    import xpulse_recon_amd as X
    vol = _volume(scan, "phantom", dtype, cuda)
    out = {}
    for ex in ("1", "0"):
        monkeypatch.setenv("XPR_FORWARD_EXACT", ex)
        s = _sums(cuda)
        X.forward(vol, s, scan["tpd"], _args(scan))
        out[ex] = s
    assert torch.equal(out["1"], out["0"])
    err = torch.empty((A_S, D, D), dtype=vol.dtype, device=cuda)
    X.error(scan["img"], out["1"], err, 0.9)
    res = {}
    for ins in ("1", "0"):
        monkeypatch.setenv("XPR_BACKWARD_INSIDE", ins)
        v = vol.clone()
        empty = torch.zeros(1, dtype=torch.int64, device=cuda)
        X.backward(v, err, scan["tpd"], _args(scan), empty=empty)
        res[ins] = (v, int(empty.item()))
    assert torch.equal(res["1"][0], res["0"][0])
    assert res["1"][1] == res["0"][1]
    assert not torch.equal(res["1"][0], vol)


@pytest.mark.parametrize("dtype", ["f32", "f16", "bf16"])
def test_slab_backwards_equal_whole_backward(scan, cuda, dtype):
    import torch
    import xpulse_recon_amd as X
    vol = _volume(scan, "phantom", dtype, cuda)
    s = _sums(cuda)
    X.forward(vol, s, scan["tpd"], _args(scan))
    err = torch.empty((A_S, D, D), dtype=vol.dtype, device=cuda)
    X.error(scan["img"], s, err, 1.0)
    a = vol.clone()
    ea = torch.zeros(1, dtype=torch.int64, device=cuda)
    X.backward(a, err, scan["tpd"], _args(scan), empty=ea)
    b = vol.clone()
    eb = torch.zeros(1, dtype=torch.int64, device=cuda)
    for l0, l1 in _slabs():
        piece = b[l0:l1].contiguous()
        X.backward(piece, err, scan["tpd"], _args(scan, l0, l1), empty=eb)
        b[l0:l1] = piece
    assert torch.equal(a, b)
    assert int(ea.item()) == int(eb.item())


@pytest.mark.parametrize("dtype,kind", [("f32", "phantom"), ("f32", "random"), ("f32", "ones"),
                                        ("f16", "phantom"), ("bf16", "random")])
def test_central_slab_matches_oracle(scan, cuda, dtype, kind):
    import torch
    import xpulse_recon_amd as X
    l0, l1 = CENTRAL
    vol = _volume(scan, kind, dtype, cuda)[l0:l1].clone()
    a = _args(scan, l0, l1)
    g = O.make_geom(D, A_S, 1, (0, l0, 0), (D, l1, D), 0, D, scan["origin"], (scan["vx"],) * 3)
    wide = vol.float().cpu().numpy()
    print(f"oracle: central slab {dtype} {kind}", flush=True)
    ref = np.zeros((A_S, D, D), np.uint64)
    O.forward_f32(wide, ref, scan["pd"], g, 16)
    s = _sums(cuda)
    X.forward(vol, s, scan["tpd"], a)
    np.testing.assert_array_equal(s.cpu().numpy().view(np.uint64), ref)
    # backward fed the same error images (of the slab's own sums)
    err = torch.empty((A_S, D, D), dtype=vol.dtype, device=cuda)
    X.error(scan["img"], s, err, 0.9)
    wide_err = err.float().cpu().numpy()
    ref_v = wide.copy()
    empty_ref = O.backward_f32(ref_v, wide_err, scan["pd"], g, 16)
    empty = torch.zeros(1, dtype=torch.int64, device=cuda)
    X.backward(vol, err, scan["tpd"], a, empty=empty)
    assert int(empty.item()) == empty_ref
    got = vol.float().cpu().numpy()
    if dtype == "f32":
        np.testing.assert_allclose(got, ref_v, rtol=4e-7, atol=0)
    else:
        ref16 = torch.from_numpy(ref_v).to(vol.dtype).float().numpy()
        ulp = 2.0 ** (-10 if dtype == "f16" else -7)
        np.testing.assert_allclose(got, ref16, rtol=ulp * 1.01, atol=0)
        assert np.mean(got != ref16) < 0.01


@pytest.fixture(scope="module")
def raw_projections(scan, cuda):
    """All 720 transmission projections of the scan, host float32."""
    import torch
    from xpulse_recon_amd import synth
    proj = np.empty((NV, D, D), np.float32)
    print("synthesising 720 views of 768^2", flush=True)
    for v0 in range(0, NV, 60):
        views = list(range(v0, min(NV, v0 + 60)))
        p = synth.line_integrals(scan["geom"], views, scan["ell"], cuda)
        proj[v0:v0 + len(views)] = synth.transmission(p, 5e4, 1003 + v0).cpu().numpy()
        del p
    torch.cuda.synchronize()
    return proj


def _reconstruct(proj, method, storage, store="ram", **xpr):
    from xpulse_recon_amd import _lib
    params = {"recon": {"method": method, "it": 1, "sit": SIT, "mlog": True, "weight": 1.0,
                        "weight_factor": 0.9},
              "geometry": {"so": 100.0, "sd": 1000.0},
              "detector": {"px": 0.1, "rx": 2.5, "roll": 0.3},
              "xpr": {"jitter_seed": 7, "layer_store": store, "slab_height": SLAB,
                      "storage": storage, "final_fp32": False, "quiet": True, **xpr}}
    vol = np.empty((D, D, D), np.float32)
    rec = C.create_string_buffer(1 << 16)
    print(f"xpr_reconstruct_mem {method} {storage} {store} 768^3", flush=True)
    rc = _lib.load().xpr_reconstruct_mem(json.dumps(params).encode(),
                                         proj.ctypes.data_as(C.c_void_p), D, D, NV,
                                         vol.ctypes.data_as(C.c_void_p), rec, len(rec))
    assert rc == 0, _lib.load().xpr_last_error()
    return vol, json.loads(rec.value.decode())


@pytest.mark.parametrize("storage,store", [("f32", "ram"), ("f32", "device"), ("f16", "ram")])
def test_engine_method_1a_equals_naive(scan, raw_projections, storage, store):
    """The engine's method_1a with the layer store in pinned host memory (12
    slabs streamed in and out per pass) or in HBM (one launch per pass over the
    resident volume) equals its naive schedule bit for bit."""
    naive, rn = _reconstruct(raw_projections, "naive", storage)
    slab, rs = _reconstruct(raw_projections, "method_1a", storage, store)
    assert rn["jitter"] == rs["jitter"] and len(rs["subiteration_ms"]) == SIT
    assert rs["layer_store"] == store
    assert rs["launches_per_pass"] == (1 if store == "device" else D // SLAB)
    assert np.isfinite(naive).all()
    np.testing.assert_array_equal(slab, naive)


def test_narrow_sums_equal_wide_sums(scan, raw_projections):
    """At full size: method_1a over the resident volume with 32-bit detector
    sums wherever the hit bound allows them (the default) equals the naive
    schedule with 64-bit sums throughout, bit for bit."""
    wide, rw = _reconstruct(raw_projections, "naive", "f32", narrow_sums=False)
    narrow, rn = _reconstruct(raw_projections, "method_1a", "f32", "device")
    bits = rn["forward_sum_bits"]
    print(f"hit bound {rn['hit_bound']:.0f}; 32-bit forwards {bits.count(32)} of {len(bits)}", flush=True)
    assert set(rw["forward_sum_bits"]) == {64}
    assert bits.count(32) >= len(bits) // 2
    np.testing.assert_array_equal(narrow, wide)
