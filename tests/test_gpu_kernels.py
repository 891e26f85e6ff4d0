"""GPU parity of the HIP kernels against the oracle, through the C-ABI.

Bars: Forward bit-exact (integer Q.22 sums); Backward and Error within a few
float32 ulp (the only difference is the device expf/logf), stated per test as
rtol; storage conversions bit-exact.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import geom_args, oracle_geom, projdata_for, rel_l2, small_spec

pytestmark = pytest.mark.gpu

Q = 1 << 22


def _args(d):
    import xpulse_recon_amd as X
    return X.proj_args(d["det_width"], d["n_views"], d["modules"], d["vmin"], d["vmax"],
                       d["line_offset"], d["line_rows"], d["origin"], d["voxel"])


def _check_shapes(vol, det, d):
    nx, nl, nk = (d["vmax"][i] - d["vmin"][i] for i in range(3))
    assert vol.size == nx * nl * nk
    assert det.size == d["n_views"] * d["line_rows"] * d["det_width"]


CASES = [
    # (spec kwargs, views, slab (vmin, vmax) or None, line_offset/rows or None, jitter)
    (dict(w=32, h=24, n=12), [0, 3, 6, 9], None, None, (0.0, 0.0, 0.0)),
    (dict(w=100, h=40, n=20, rx=2.5), list(range(1, 20, 2)), None, None, (0.004, 0.001, 0.002)),
    (dict(w=130, h=8, n=6, sx=0.2), [0, 1, 2, 3, 4, 5], ((0, 2, 0), (130, 7, 130)), None,
     (0.001, 0.0, 0.003)),
    (dict(w=48, h=32, n=8, module_number=3), [0, 2, 4, 6], None, None, (0.0, 0.002, 0.0)),
    (dict(w=40, h=36, n=10), [1, 4, 7], ((0, 10, 0), (40, 22, 40)), (8, 20), (0.002, 0.003, 0.001)),
    (dict(w=1, h=3, n=4), [0, 1, 2, 3], None, None, (0.0, 0.0, 0.0)),
]


def _tilt(pd, tilt):
    """Odd records of pd with non-zero layer coefficients in the u and w rows.
    "roll": the u row takes 2 % of the v row's (a detector rolled about the
    beam axis) and the w row a residue of 3e-7, as a float32 roll leaves it:
    the w row terms stay equal across layers, the u row terms do not.
    "pitch": a w row coefficient large enough that every layer has its own W:
    the projectors take the canonical per-voxel projection there.
    ("residue": see _residue.)"""
    pd = pd.copy()
    if tilt == "roll":
        pd[1::2, 1] = 0.02 * pd[1::2, 5]
        pd[1::2, 13] = 3e-7
    elif tilt == "pitch":
        pd[1::2, 1] = 0.02 * pd[1::2, 5]
        pd[1::2, 13] = 2e-3 * np.abs(pd[1::2, 12]) + 1e-3
    return pd


def _residue(pd, d):
    """Odd records with w row coefficients around half an ulp of the w row
    term over the volume's layer positions: some bricks and row segments keep
    equal w row terms in all their layers, some in none, some in part of their
    rows (these must take the per-voxel projection)."""
    pd = pd.copy()
    rng = np.random.default_rng(77)
    n = pd[1::2].shape[0]
    y0 = d["origin"][1] + d["vmin"][1] * d["voxel"][1]
    y1 = d["origin"][1] + (d["vmax"][1] - 1) * d["voxel"][1]
    ymax = max(abs(y0), abs(y1), 1e-6)
    ulp = np.spacing(np.abs(pd[1::2, 15]).astype(np.float32)).astype(np.float64)
    pd[1::2, 13] = (rng.uniform(0.05, 3.0, n) * ulp / ymax * rng.choice([-1.0, 1.0], n)).astype(np.float32)
    return pd


def _case(i, tilt="none"):
    kw, views, slab, crop, jit = CASES[i]
    kw = dict(kw)
    w, h, n = kw.pop("w"), kw.pop("h"), kw.pop("n")
    spec = small_spec(w, h, n, **kw)
    pd, grid = projdata_for(spec, views)
    if tilt in ("roll", "pitch"):
        pd = _tilt(pd, tilt)
    vmin, vmax = slab if slab else (None, None)
    lo, rows = crop if crop else (0, None)
    d = geom_args(spec, grid, len(views), vmin=vmin, vmax=vmax, jitter=jit, line_offset=lo,
                  line_rows=rows)
    if tilt == "residue":
        pd = _residue(pd, d)
    return spec, pd, d


@pytest.mark.parametrize("scale", [3.0, 0.2])
@pytest.mark.parametrize("tilt", ["none", "roll", "pitch", "residue"])
@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("acc", [32, 64])
def test_forward_bit_exact(cuda, case, acc, tilt, scale):
    """Sums equal the oracle's bit for bit.  Volumes up to 0.2 keep every Q.22
    value below 2^20, so the bricks accumulate in 32-bit tile entries; up to
    3.0 they take 64-bit entries (with 64-bit sums)."""
    import torch
    import xpulse_recon_amd as X
    spec, pd, d = _case(case, tilt)
    nx, nl, nk = (d["vmax"][i] - d["vmin"][i] for i in range(3))
    rng = np.random.default_rng(100 + case)
    vol = rng.uniform(0.0, scale, (nl, nk, nx)).astype(np.float32)
    vol[0, 0, :] = 0.0
    shape = (d["n_views"], d["line_rows"], d["det_width"])
    ref = np.zeros(shape, np.uint32 if acc == 32 else np.uint64)
    _check_shapes(vol, ref, d)
    O.forward_f32(vol, ref, pd, oracle_geom(d))
    tdt = torch.int32 if acc == 32 else torch.int64
    sums = torch.zeros(shape, dtype=tdt, device=cuda)
    X.forward(torch.from_numpy(vol).to(cuda), sums, torch.from_numpy(pd).to(cuda), _args(d))
    got = sums.cpu().numpy().view(ref.dtype)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("tilt", ["none", "roll", "pitch", "residue"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_backward_parity(cuda, case, tilt):
    import torch
    import xpulse_recon_amd as X
    spec, pd, d = _case(case, tilt)
    nx, nl, nk = (d["vmax"][i] - d["vmin"][i] for i in range(3))
    rng = np.random.default_rng(200 + case)
    vol = rng.uniform(0.5, 2.0, (nl, nk, nx)).astype(np.float32)
    err = rng.normal(0.0, 0.4, (d["n_views"], d["line_rows"], d["det_width"])).astype(np.float32)
    _check_shapes(vol, err, d)
    ref = vol.copy()
    empty_ref = O.backward_f32(ref, err, pd, oracle_geom(d))
    tv = torch.from_numpy(vol).to(cuda)
    empty = torch.zeros(1, dtype=torch.int64, device=cuda)
    X.backward(tv, torch.from_numpy(err).to(cuda), torch.from_numpy(pd).to(cuda), _args(d),
               empty=empty)
    got = tv.cpu().numpy()
    assert int(empty.item()) == empty_ref
    np.testing.assert_allclose(got, ref, rtol=4e-7, atol=0)
    assert rel_l2(got, ref) < 1e-6


def _band_case():
    """All 36 layers against 8 held detector rows (14-21): the outer layers
    have no in-bounds view."""
    spec = small_spec(40, 36, 10)
    pd, grid = projdata_for(spec, [1, 4, 7])
    d = geom_args(spec, grid, 3, line_offset=14, line_rows=8, jitter=(0.002, 0.003, 0.001))
    return spec, pd, d


@pytest.mark.parametrize("case", list(range(len(CASES))) + ["band"])
@pytest.mark.parametrize("dtype", ["f32", "f16"])
def test_backward_max_reports_the_largest_value(cuda, case, dtype):
    """xpr_backward_max updates the volume exactly as xpr_backward and raises
    max_bits to the largest value after the update, kept voxels (no in-bounds
    view) included: in the band case one of them holds the maximum."""
    import torch
    import xpulse_recon_amd as X
    spec, pd, d = _band_case() if case == "band" else _case(case)
    nx, nl, nk = (d["vmax"][i] - d["vmin"][i] for i in range(3))
    rng = np.random.default_rng(400 + (9 if case == "band" else case))
    vol = rng.uniform(0.5, 2.0, (nl, nk, nx)).astype(np.float32)
    err = rng.normal(0.0, 0.4, (d["n_views"], d["line_rows"], d["det_width"])).astype(np.float32)
    # voxels with no in-bounds view keep their value even under a NaN error
    probe = vol.copy()
    O.backward_f32(probe, np.full_like(err, np.nan), pd, oracle_geom(d))
    kept = np.argwhere(np.isfinite(probe))
    assert (len(kept) > 0) == (case == "band")
    if len(kept):
        vol[tuple(kept[len(kept) // 2])] = 50.0
    tdt = {"f32": torch.float32, "f16": torch.float16}[dtype]
    te, tp = torch.from_numpy(err).to(cuda).to(tdt), torch.from_numpy(pd).to(cuda)
    plain = torch.from_numpy(vol).to(cuda).to(tdt)
    X.backward(plain, te, tp, _args(d))
    tv = torch.from_numpy(vol).to(cuda).to(tdt)
    top = torch.zeros(X.MAX_SLOTS, dtype=torch.int32, device=cuda)
    X.backward(tv, te, tp, _args(d), max_bits=top)
    assert torch.equal(tv, plain)
    got = float(np.array([int(top.max())], np.int32).view(np.float32)[0])
    stored = float(tv.float().max())
    if dtype == "f32":
        assert got == stored
    else:                                     # the value before the 16-bit store rounds it
        assert abs(got - stored) <= stored * 2.0 ** -11
    if len(kept):
        assert got == 50.0 == stored


def test_valu_rates_and_ceiling(cuda):
    """xpr_valu_rates measures every calibration kind; xpr_valu_ceiling is the
    best of them (separate launches: within 10 %).  The integer mixes issue at
    least as fast as FMAs alone."""
    import ctypes as C
    from xpulse_recon_amd import _lib
    v = (C.c_double * 6)()
    _lib.call("xpr_valu_rates", None, v, 6)
    best = C.c_double()
    _lib.call("xpr_valu_ceiling", None, C.byref(best))
    rates = list(v)
    print("VALU lane-instructions/s per kind", [f"{r:.3e}" for r in rates], flush=True)
    assert all(r > 1e12 for r in rates)
    assert 0.9 * max(rates) <= best.value <= 1.1 * max(rates)
    assert max(rates[3:]) >= 0.9 * rates[0]


def test_error_parity(cuda):
    import torch
    import xpulse_recon_amd as X
    rng = np.random.default_rng(7)
    n = 100003
    img = rng.uniform(0.0, 4.0, n).astype(np.float32)
    img[:4] = [0.0, 1e-30, 1.0, 3.0]
    for acc in (32, 64):
        hi = (1 << 31) if acc == 32 else (1 << 40)
        sums = rng.integers(0, hi, n).astype(np.uint32 if acc == 32 else np.uint64)
        sums[:4] = [0, 1, 1 << 22, 5 << 22]
        ref = O.error_f32(img, sums, 0.9)
        tdt = torch.int32 if acc == 32 else torch.int64
        ts = torch.from_numpy(sums.view(np.int32 if acc == 32 else np.int64)).to(cuda)
        out = torch.empty(n, dtype=torch.float32, device=cuda)
        X.error(torch.from_numpy(img).to(cuda), ts, out, 0.9)
        np.testing.assert_allclose(out.cpu().numpy(), ref, rtol=4e-7, atol=1e-7)
        if acc == 32:                       # in place in the sums buffer
            X.error(torch.from_numpy(img).to(cuda), ts, ts.view(torch.float32), 0.9)
            np.testing.assert_array_equal(ts.view(torch.float32).cpu().numpy(), out.cpu().numpy())
        assert ts.dtype == tdt


def test_preprocess_and_minmax_parity(cuda):
    import torch
    import xpulse_recon_amd as X
    rng = np.random.default_rng(12)
    x = rng.uniform(-0.2, 1.3, 70001).astype(np.float32)
    x[:6] = [0.0, np.inf, np.nan, 1e-40, 1.0, -3.0]
    mn, mx = O.minmax_normal(x)
    tx = torch.from_numpy(x).to(cuda)
    assert X.minmax_normal(tx) == (mn, mx)
    # 70001 pixels take the one-pixel kernel, 70000 (aligned) four per access
    for n in (70001, 70000):
        for normalize in (False, True):
            ref = O.preprocess(x[:n].copy(), normalize, True, mn, mx)
            out = torch.empty(n, dtype=torch.float32, device=cuda)
            X.preprocess(tx[:n], out, normalize, True, mn, mx)
            got = out.cpu().numpy()
            fin = np.isfinite(ref)
            np.testing.assert_array_equal(np.isnan(got), np.isnan(ref))
            np.testing.assert_allclose(got[fin], ref[fin], rtol=4e-7, atol=1e-7)


@pytest.mark.parametrize("dtype", ["f32", "f16", "bf16"])
def test_preprocess_views_equals_per_view(cuda, dtype):
    """The views s, s+S, ... of a projection stack preprocessed in one launch
    (xpr_preprocess_views) equal the same views preprocessed one call each,
    bit for bit: 36 views of 768^2, step 20, normalised or not; and the
    four-pixel kernel (aligned rows) equals the one-pixel kernel (a view read
    4 bytes off alignment)."""
    import torch
    import xpulse_recon_amd as X
    tdt = {"f32": torch.float32, "f16": torch.float16, "bf16": torch.bfloat16}[dtype]
    g = torch.Generator(device=cuda).manual_seed(14)
    raw = torch.rand((720, 768, 768), generator=g, device=cuda) * 1.4 - 0.1
    for normalize in (False, True):
        one = torch.empty((36, 768, 768), dtype=tdt, device=cuda)
        X.preprocess_views(raw[7:], one, 20, normalize, True, -0.05, 1.2)
        ref = torch.empty_like(one)
        for j in range(36):
            X.preprocess(raw[7 + 20 * j], ref[j], normalize, True, -0.05, 1.2)
        # one view through the one-pixel kernel: its input 4 bytes off alignment
        shifted = torch.empty(768 * 768 + 1, device=cuda)
        shifted[1:] = raw[7 + 20 * 5].reshape(-1)
        scalar = torch.empty((768, 768), dtype=tdt, device=cuda)
        X.preprocess(shifted[1:], scalar, normalize, True, -0.05, 1.2)
        torch.cuda.synchronize()
        bits = torch.int16 if dtype != "f32" else torch.int32
        assert torch.equal(one.view(bits), ref.view(bits))
        assert torch.equal(scalar.view(bits), one[5].view(bits))


def _bf16_bits(a):
    b = a.astype(np.float32).view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16).astype(np.uint16)
    nan = np.isnan(a)
    r[nan] = ((b[nan] >> 16) | 0x40).astype(np.uint16)
    return r


def test_convert_and_fill(cuda):
    import torch
    import xpulse_recon_amd as X
    rng = np.random.default_rng(13)
    x = (rng.standard_normal(5000) * 10.0 ** rng.uniform(-8, 4, 5000)).astype(np.float32)
    tx = torch.from_numpy(x).to(cuda)
    h = torch.empty(5000, dtype=torch.float16, device=cuda)
    X.convert(h, tx)
    np.testing.assert_array_equal(h.cpu().numpy().view(np.uint16), x.astype(np.float16).view(np.uint16))
    b = torch.empty(5000, dtype=torch.bfloat16, device=cuda)
    X.convert(b, tx)
    np.testing.assert_array_equal(b.cpu().view(torch.int16).numpy().view(np.uint16), _bf16_bits(x))
    back = torch.empty_like(tx)
    X.convert(back, b)
    np.testing.assert_array_equal(back.cpu().numpy(),
                                  (_bf16_bits(x).astype(np.uint32) << 16).view(np.float32))
    s = torch.empty(777, dtype=torch.int64, device=cuda)
    X.fill_u64(s, (1 << 40) + 3)
    assert (s.cpu() == (1 << 40) + 3).all()
    X.fill(h, 1.0)
    assert (h.cpu() == 1.0).all()


@pytest.mark.parametrize("hi", [2.0, 0.2])
@pytest.mark.parametrize("tilt", ["none", "roll"])
@pytest.mark.parametrize("vdt", ["f16", "bf16"])
def test_mixed_precision_storage(cuda, vdt, tilt, hi):
    """16-bit volume and error storage, FP32 arithmetic: forward is bit-exact
    on the widened values; backward equals the oracle's FP32 update rounded to
    the storage type (one storage ulp where device expf differs).  Upright and
    rolled matrices; volumes up to 2.0 (64-bit tile entries) and up to 0.2
    (32-bit entries)."""
    import torch
    import xpulse_recon_amd as X
    spec, pd, d = _case(1, tilt)
    nx, nl, nk = (d["vmax"][i] - d["vmin"][i] for i in range(3))
    tdt = {"f16": torch.float16, "bf16": torch.bfloat16}[vdt]
    rng = np.random.default_rng(14)
    vol32 = torch.from_numpy(rng.uniform(hi / 4.0, hi, (nl, nk, nx)).astype(np.float32))
    vol16 = vol32.to(tdt)
    err32 = torch.from_numpy(rng.normal(0, 0.3, (d["n_views"], d["line_rows"], d["det_width"]))
                             .astype(np.float32))
    err16 = err32.to(tdt)
    wide_vol = vol16.float().numpy()
    wide_err = err16.float().numpy()
    ref_sums = np.zeros(err32.shape, np.uint64)
    O.forward_f32(wide_vol, ref_sums, pd, oracle_geom(d))
    sums = torch.zeros(err32.shape, dtype=torch.int64, device=cuda)
    tpd = torch.from_numpy(pd).to(cuda)
    X.forward(vol16.to(cuda), sums, tpd, _args(d))
    np.testing.assert_array_equal(sums.cpu().numpy().view(np.uint64), ref_sums)
    ref = wide_vol.copy()
    O.backward_f32(ref, wide_err, pd, oracle_geom(d))
    ref16 = torch.from_numpy(ref).to(tdt).float().numpy()
    tv = vol16.to(cuda)
    X.backward(tv, err16.to(cuda), tpd, _args(d))
    got = tv.float().cpu().numpy()
    ulp = 2.0 ** (-10 if vdt == "f16" else -7)
    np.testing.assert_allclose(got, ref16, rtol=ulp * 1.01, atol=0)
    assert np.mean(got != ref16) < 0.01


def test_slab_split_is_exact_at_scale(cuda):
    """256^3 grid, 24 views: forward over 5 uneven slabs sums to the whole-volume
    forward bit for bit; backward over slabs equals whole-volume backward."""
    import torch
    import xpulse_recon_amd as X
    spec = small_spec(256, 256, 240, rx=1.5)
    views = list(range(3, 240, 10))
    pd, grid = projdata_for(spec, views)
    d = geom_args(spec, grid, len(views), jitter=(0.0031, 0.0012, 0.0077))
    W = H = 256
    g = torch.Generator(device="cpu").manual_seed(15)
    vol = (torch.rand((H, W, W), generator=g) * 2.0 + 0.1).to(cuda)
    tpd = torch.from_numpy(pd).to(cuda)
    shape = (len(views), 256, 256)
    whole = torch.zeros(shape, dtype=torch.int64, device=cuda)
    X.forward(vol, whole, tpd, _args(d))
    parts = torch.zeros_like(whole)
    cuts = [0, 37, 38, 100, 201, 256]
    for l0, l1 in zip(cuts[:-1], cuts[1:]):
        ds = dict(d, vmin=(0, l0, 0), vmax=(W, l1, W))
        X.forward(vol[l0:l1].contiguous(), parts, tpd, _args(ds))
    assert torch.equal(parts, whole)
    err = torch.randn(shape, generator=g).mul_(0.2).to(cuda)
    a = vol.clone()
    X.backward(a, err, tpd, _args(d))
    b = vol.clone()
    for l0, l1 in zip(cuts[:-1], cuts[1:]):
        ds = dict(d, vmin=(0, l0, 0), vmax=(W, l1, W))
        piece = b[l0:l1].contiguous()
        X.backward(piece, err, tpd, _args(ds))
        b[l0:l1] = piece
    assert torch.equal(a, b)
    # uniform volume of 1.0: total Q.22 mass = 2^22 * in-bounds voxel-views,
    # and every in-bounds voxel sees err == 0 -> stays exactly 1
    ones = torch.ones((H, W, W), device=cuda)
    s = torch.zeros(shape, dtype=torch.int64, device=cuda)
    X.forward(ones, s, tpd, _args(d))
    hits = int(s.sum().item()) // Q
    assert int(s.sum().item()) % Q == 0
    assert 0.5 * H * W * W * len(views) < hits <= H * W * W * len(views)


def test_synthetic_line_integrals_match_oracle(cuda):
    from xpulse_recon_amd import synth
    from xpulse_recon_amd.geometry import Geometry
    kw = dict(dwidth=48, dheight=40, projections=8, rx=2.0, sx=0.1, module_number=2)
    geom = Geometry(**kw)
    _, grid = geom.projdata()
    ell = synth.random_ellipsoids(5, 6, grid["vwidth"] * grid["vx"], 128)
    got = synth.line_integrals(geom, [1, 5], ell, cuda).cpu().numpy()
    oe = [O.make_ellipsoid(e["c"], e["a"], e["rot"], e["mu"]) for e in ell]
    spec = O.GeometrySpec(**kw)
    for j, a in enumerate([1, 5]):
        ref = O.line_integrals(spec, a, oe)
        np.testing.assert_allclose(got[j], ref, rtol=1e-9, atol=1e-12)
    assert got.max() > 0.1


def test_synthetic_phantom_matches_oracle(cuda):
    from xpulse_recon_amd import synth
    from xpulse_recon_amd.geometry import Geometry
    geom = Geometry(dwidth=40, dheight=24, projections=4, heli_offset=0.05)
    _, grid = geom.projdata()
    ell = synth.random_ellipsoids(8, 9, grid["vwidth"] * grid["vx"], 40)
    got = synth.phantom(grid, ell, cuda).cpu().numpy()
    oe = [O.make_ellipsoid(e["c"], e["a"], e["rot"], e["mu"]) for e in ell]
    ref = O.phantom_voxels(grid, oe).astype(np.float32)
    assert got.shape == ref.shape and (ref > 0).mean() > 0.02
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("cap", ["0", "40", "2048"])
@pytest.mark.parametrize("acc", [32, 64])
def test_forward_paths_agree_with_large_values(cuda, monkeypatch, cap, acc):
    """Direct atomics (cap 0), a tile too small for most footprints (cap 40) and
    the full tile give the oracle's sums bit for bit, including voxels whose
    Q.22 value does not fit 32 bits (u64 exact, u32 saturated then wrapped)."""
    import torch
    import xpulse_recon_amd as X
    monkeypatch.setenv("XPR_FORWARD_TILE_CAP", cap)
    spec, pd, d = _case(1)
    nx, nl, nk = (d["vmax"][i] - d["vmin"][i] for i in range(3))
    rng = np.random.default_rng(300)
    vol = rng.uniform(0.0, 3.0, (nl, nk, nx)).astype(np.float32)
    vol[5, 7, 3:9] = 2000.0
    vol[20, 1, 50] = 1023.9999
    shape = (d["n_views"], d["line_rows"], d["det_width"])
    ref = np.zeros(shape, np.uint32 if acc == 32 else np.uint64)
    O.forward_f32(vol, ref, pd, oracle_geom(d))
    sums = torch.zeros(shape, dtype=torch.int32 if acc == 32 else torch.int64, device=cuda)
    X.forward(torch.from_numpy(vol).to(cuda), sums, torch.from_numpy(pd).to(cuda), _args(d))
    np.testing.assert_array_equal(sums.cpu().numpy().view(ref.dtype), ref)


@pytest.mark.parametrize("scale", [1e-3, 0.05, 3.0])
def test_forward_bit_exact_across_tile_widths(cuda, scale):
    """Bricks whose total Q.22 mass is below 2^32 accumulate in 32-bit LDS
    entries, heavier ones in 64-bit entries; both give the oracle's u64 sums."""
    import torch
    import xpulse_recon_amd as X
    spec, pd, d = _case(1)
    nx, nl, nk = (d["vmax"][i] - d["vmin"][i] for i in range(3))
    rng = np.random.default_rng(400)
    vol = (rng.uniform(0.0, 1.0, (nl, nk, nx)) * scale).astype(np.float32)
    vol[::7, 3, :] *= 200.0           # a few heavy rows: mixed bricks in one launch
    shape = (d["n_views"], d["line_rows"], d["det_width"])
    ref = np.zeros(shape, np.uint64)
    O.forward_f32(vol, ref, pd, oracle_geom(d))
    sums = torch.zeros(shape, dtype=torch.int64, device=cuda)
    X.forward(torch.from_numpy(vol).to(cuda), sums, torch.from_numpy(pd).to(cuda), _args(d))
    np.testing.assert_array_equal(sums.cpu().numpy().view(np.uint64), ref)


@pytest.mark.parametrize("tilt", ["none", "roll", "pitch", "residue"])
@pytest.mark.parametrize("case", [1, 2, 4])
def test_backward_lane_shapes_bit_identical(cuda, monkeypatch, case, tilt):
    """Every compiled voxels-per-lane shape of the backward (4, 8, 16: 1, 2 or
    4 x positions in 4 layers) gives the same volume bit for bit (widths 100
    and 130 leave partial segments, odd layer counts a short layer group),
    including the cropped band of case 4, and matches the oracle."""
    import torch
    import xpulse_recon_amd as X
    spec, pd, d = _case(case, tilt)
    nx, nl, nk = (d["vmax"][i] - d["vmin"][i] for i in range(3))
    rng = np.random.default_rng(300 + case)
    vol = rng.uniform(0.5, 2.0, (nl, nk, nx)).astype(np.float32)
    err = rng.normal(0.0, 0.4, (d["n_views"], d["line_rows"], d["det_width"])).astype(np.float32)
    ref = vol.copy()
    O.backward_f32(ref, err, pd, oracle_geom(d))
    te, tp = torch.from_numpy(err).to(cuda), torch.from_numpy(pd).to(cuda)
    outs = []
    for j in ("4", "8", "16"):
        monkeypatch.setenv("XPR_BACKWARD_J", j)
        tv = torch.from_numpy(vol).to(cuda)
        X.backward(tv, te, tp, _args(d))
        outs.append(tv.cpu().numpy())
    for o in outs[1:]:
        np.testing.assert_array_equal(o, outs[0])
    np.testing.assert_allclose(outs[0], ref, rtol=4e-7, atol=0)


def test_tuned_shapes_keep_results(cuda):
    """xpr_tune_* time the shapes on scratch buffers and keep one per launch
    shape; launches after tuning give the untuned results bit for bit."""
    import torch
    import xpulse_recon_amd as X
    spec, pd, d = _case(1)
    nx, nl, nk = (d["vmax"][i] - d["vmin"][i] for i in range(3))
    rng = np.random.default_rng(9)
    vol = rng.uniform(0.5, 2.0, (nl, nk, nx)).astype(np.float32)
    err = rng.normal(0.0, 0.4, (d["n_views"], d["line_rows"], d["det_width"])).astype(np.float32)
    tp, te = torch.from_numpy(pd).to(cuda), torch.from_numpy(err).to(cuda)
    shape = (d["n_views"], d["line_rows"], d["det_width"])
    X.tune_reset()

    def run():
        s = torch.zeros(shape, dtype=torch.int64, device=cuda)
        X.forward(torch.from_numpy(vol).to(cuda), s, tp, _args(d))
        tv = torch.from_numpy(vol).to(cuda)
        X.backward(tv, te, tp, _args(d))
        return s.cpu().numpy(), tv.cpu().numpy()

    s0, v0 = run()
    cap = X.tune_forward(torch.from_numpy(vol).to(cuda),
                         torch.zeros(shape, dtype=torch.int64, device=cuda), tp, _args(d))
    j = X.tune_backward(torch.from_numpy(vol).to(cuda), te, tp, _args(d))
    assert cap in (384, 768, 1536) and j in (4, 8, 16)
    s1, v1 = run()
    np.testing.assert_array_equal(s1, s0)
    np.testing.assert_array_equal(v1, v0)
    X.tune_reset()


@pytest.mark.parametrize("acc", [32, 64])
def test_forward_exact_views_equal_checked(cuda, monkeypatch, acc):
    """Views whose rounding bound lets the tile test decide the bounds (the
    exact-tile splat), and those whose tile the viewport does not clip in a
    full brick (the inside splat, no test), give the sums of the canonical
    bounds test bit for bit, on a 160^2 x 96 slab with 24 views (partial
    bricks along x, edge bricks clipped), a cropped row band and the oracle."""
    import torch
    import xpulse_recon_amd as X
    spec = small_spec(160, 96, 48, rx=1.5)
    views = list(range(0, 48, 2))
    pd, grid = projdata_for(spec, views)
    d = geom_args(spec, grid, len(views), jitter=(0.003, 0.001, 0.006))
    rng = np.random.default_rng(42)
    W, H = grid["vwidth"], grid["vheight"]
    vol = rng.uniform(0.0, 1.5, (H, W, W)).astype(np.float32)
    shape = (d["n_views"], d["line_rows"], d["det_width"])
    tdt = torch.int32 if acc == 32 else torch.int64
    tv, tp = torch.from_numpy(vol).to(cuda), torch.from_numpy(pd).to(cuda)
    out = {}
    for ex in ("1", "0"):
        monkeypatch.setenv("XPR_FORWARD_EXACT", ex)
        s = torch.zeros(shape, dtype=tdt, device=cuda)
        X.forward(tv, s, tp, _args(d))
        out[ex] = s.cpu().numpy()
    np.testing.assert_array_equal(out["1"], out["0"])
    ref = np.zeros(shape, np.uint32 if acc == 32 else np.uint64)
    O.forward_f32(vol, ref, pd, oracle_geom(d), 8)
    np.testing.assert_array_equal(out["1"].view(ref.dtype), ref)
    # a cropped band of detector rows
    dc = geom_args(spec, grid, len(views), jitter=(0.003, 0.001, 0.006), line_offset=30,
                   line_rows=40)
    shape_c = (dc["n_views"], dc["line_rows"], dc["det_width"])
    out = {}
    for ex in ("1", "0"):
        monkeypatch.setenv("XPR_FORWARD_EXACT", ex)
        s = torch.zeros(shape_c, dtype=tdt, device=cuda)
        X.forward(tv, s, tp, _args(dc))
        out[ex] = s.cpu().numpy()
    np.testing.assert_array_equal(out["1"], out["0"])


INSIDE_CASES = [
    # (spec kwargs, views, line_offset/rows or None): 100 view records (two
    # set-up chunks of 64 lanes), 2 modules x 40 views, a cropped row band
    (dict(w=160, h=96, n=100, rx=1.5), list(range(100)), None),
    (dict(w=96, h=40, n=40, module_number=2, module_angle_offset=6.0), list(range(40)), None),
    (dict(w=160, h=96, n=48, rx=-2.0), list(range(0, 48, 2)), (30, 40)),
]


@pytest.mark.parametrize("case", range(len(INSIDE_CASES)))
def test_backward_inside_views_equal_checked(cuda, monkeypatch, case):
    """Row segments whose rounding bound puts every voxel in bounds skip the
    per-voxel test: the volume equals the canonical test's bit for bit (FP32
    at J = 4 and 16, and FP16 storage), and the oracle within the expf
    tolerance."""
    import torch
    import xpulse_recon_amd as X
    kw, views, crop = INSIDE_CASES[case]
    kw = dict(kw)
    spec = small_spec(kw.pop("w"), kw.pop("h"), kw.pop("n"), **kw)
    pd, grid = projdata_for(spec, views)
    lo, rows = crop if crop else (0, None)
    d = geom_args(spec, grid, len(views), jitter=(0.002, 0.004, 0.001), line_offset=lo,
                  line_rows=rows)
    W, H = grid["vwidth"], grid["vheight"]
    rng = np.random.default_rng(500 + case)
    vol = rng.uniform(0.5, 2.0, (H, W, W)).astype(np.float32)
    err = rng.normal(0.0, 0.4, (d["n_views"], d["line_rows"], d["det_width"])).astype(np.float32)
    te, tp = torch.from_numpy(err).to(cuda), torch.from_numpy(pd).to(cuda)

    def run(dtype, inside, j):
        monkeypatch.setenv("XPR_BACKWARD_INSIDE", inside)
        monkeypatch.setenv("XPR_BACKWARD_J", j)
        tv = torch.from_numpy(vol).to(cuda).to(dtype)
        empty = torch.zeros(1, dtype=torch.int64, device=cuda)
        X.backward(tv, te.to(dtype), tp, _args(d), empty=empty)
        return tv.float().cpu().numpy(), int(empty.item())

    for j in ("4", "16"):
        a, ea = run(torch.float32, "1", j)
        b, eb = run(torch.float32, "0", j)
        np.testing.assert_array_equal(a, b)
        assert ea == eb
    h1, _ = run(torch.float16, "1", "4")
    h0, _ = run(torch.float16, "0", "4")
    np.testing.assert_array_equal(h1, h0)
    ref = vol.copy()
    empty_ref = O.backward_f32(ref, err, pd, oracle_geom(d), 8)
    assert ea == empty_ref
    np.testing.assert_allclose(a, ref, rtol=4e-7, atol=0)


@pytest.mark.parametrize("fast", ["1", "0"])
def test_indexed_views_two_modules(cuda, monkeypatch, fast):
    """Indexed launches (a subset of the detector views through view_index,
    the reference's ForwardH/BackwardH) with two modules per view on a 128^2 x
    64 grid (full bricks, so the inside splat and inside segments run): the
    forward writes exactly the indexed views with the oracle's sums, the
    backward equals the oracle fed the indexed error images; with the exact /
    inside fast paths on and off (XPR_FORWARD_EXACT, XPR_BACKWARD_INSIDE)."""
    import torch
    import xpulse_recon_amd as X
    monkeypatch.setenv("XPR_FORWARD_EXACT", fast)
    monkeypatch.setenv("XPR_BACKWARD_INSIDE", fast)
    spec = small_spec(128, 64, 40, module_number=2, module_angle_offset=5.0, rx=1.0)
    idx = [3, 7, 8, 20, 33, 39]
    pd, grid = projdata_for(spec, idx)
    W, H = grid["vwidth"], grid["vheight"]
    d = geom_args(spec, grid, len(idx), jitter=(0.002, 0.003, 0.001))
    rng = np.random.default_rng(77)
    vol = rng.uniform(0.2, 1.8, (H, W, W)).astype(np.float32)
    rows, width = d["line_rows"], d["det_width"]
    tidx = torch.tensor(idx, dtype=torch.int32, device=cuda)
    a = X.proj_args(width, len(idx), 2, d["vmin"], d["vmax"], 0, rows, d["origin"], d["voxel"],
                    view_index=tidx.data_ptr())
    tp = torch.from_numpy(pd).to(cuda)
    sums = torch.zeros((40, rows, width), dtype=torch.int64, device=cuda)
    X.forward(torch.from_numpy(vol).to(cuda), sums, tp, a)
    ref = np.zeros((len(idx), rows, width), np.uint64)
    O.forward_f32(vol, ref, pd, oracle_geom(d), 8)
    got = sums.cpu().numpy().view(np.uint64)
    np.testing.assert_array_equal(got[idx], ref)
    others = [v for v in range(40) if v not in idx]
    assert not got[others].any()
    err = rng.normal(0.0, 0.3, (40, rows, width)).astype(np.float32)
    tv = torch.from_numpy(vol).to(cuda)
    X.backward(tv, torch.from_numpy(err).to(cuda), tp, a)
    ref_v = vol.copy()
    O.backward_f32(ref_v, np.ascontiguousarray(err[idx]), pd, oracle_geom(d), 8)
    np.testing.assert_allclose(tv.cpu().numpy(), ref_v, rtol=4e-7, atol=0)
