"""End-to-end: the XPRecon drop-in on small synthetic scans, every schedule,
against the oracle's naive schedule (reconstruction_n.hpp:51-78) fed the same
preprocessed images, geometry and logged jitter.

Bars: naive within rel-L2 1e-5 of the oracle (only device expf/logf differ);
method_1a/1b/1c bit-identical to naive (integer sums, per-voxel updates);
method_2 within rel-L2 1e-5 of the oracle's chunked schedule on the same plan
(independent chunks, reconstruction_chunk.hpp:36-131); the reprojected views
(forward_exec) equal the oracle's bit for bit."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import rel_l2
from tests.tiff_util import read_f32, write_f32

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "xpulse_recon_amd", "bin", "XPRecon")

W, H, N = 40, 32, 24


def _scan(tmp_path, seed=3, i0=2e4):
    spec = O.GeometrySpec(dwidth=W, dheight=H, projections=N, rx=1.5)
    _, grid = O.geometry(spec, False)
    from xpulse_recon_amd import synth
    ell = synth.random_ellipsoids(seed, 5, W * grid["vx"], 64, mu_lo=0.02, mu_hi=0.06)
    oe = [O.make_ellipsoid(e["c"], e["a"], e["rot"], e["mu"]) for e in ell]
    rng = np.random.default_rng(seed + 1)
    d = tmp_path / "in"
    d.mkdir()
    for a in range(N):
        p = O.line_integrals(spec, a, oe)
        t = rng.poisson(i0 * np.exp(-p)) / i0
        write_f32(d / f"proj_{a:04d}.tif", t.astype(np.float32))
    return spec, d


def _run(tmp_path, name, indir, **recon):
    out = tmp_path / name
    xpr = recon.pop("xpr", {})
    method2 = recon.pop("method2", {})
    xpr.setdefault("jitter_seed", 11)
    xpr["run_log"] = str(tmp_path / f"{name}.json")
    params = {"recon": dict(input=str(indir), output=str(out), it=2, sit=4, **recon),
              "geometry": {"so": 100.0, "sd": 1000.0}, "detector": {"px": 0.1, "rx": 1.5},
              "method2": method2, "xpr": xpr}
    pf = tmp_path / f"{name}_params.json"
    pf.write_text(json.dumps(params))
    print(f"XPRecon {name}", flush=True)
    r = subprocess.run([CLI, str(pf)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    log = json.loads((tmp_path / f"{name}.json").read_text())
    layers = sorted(out.glob("layer_*.tif"))
    vol = np.stack([read_f32(f) for f in layers]) if layers else None
    return vol, log


def _oracle(spec, indir, log, normalize=False, it=2, sit=4):
    imgs = np.stack([read_f32(indir / f"proj_{a:04d}.tif") for a in range(N)])
    if normalize:
        mn, mx = O.minmax_normal(imgs)
        O.preprocess(imgs.reshape(-1), True, True, mn, mx)
    else:
        O.preprocess(imgs.reshape(-1), False, True)
    pd, grid = O.geometry(spec, False)
    jit = np.array(log["jitter"], np.float32)
    vol = O.reconstruct_naive(imgs, pd, grid, it, sit, 1.0, 0.9, jit, threads=8)
    if normalize:
        vol = vol * np.float32(mx - mn)
    return vol


@pytest.fixture(scope="module")
def scan(tmp_path_factory, cuda):
    t = tmp_path_factory.mktemp("scan")
    spec, d = _scan(t)
    return t, spec, d


def test_naive_matches_oracle(scan):
    t, spec, d = scan
    vol, log = _run(t, "naive", d, method="naive")
    assert vol.shape == (H, W, W)
    assert len(log["jitter"]) == 8 and len(log["subiteration_ms"]) == 8
    ref = _oracle(spec, d, log)
    assert rel_l2(vol, ref) < 1e-5


@pytest.mark.parametrize("method,xpr", [
    ("method_1a", {"slab_height": 7, "layer_store": "zstd"}),
    ("method_1a", {"slab_height": 7, "layer_store": "ram"}),
    ("method_1a", {"slab_height": 7, "layer_store": "device"}),
    ("method_1a", {"slab_height": 7, "layer_store": "ram", "storage": "f16"}),
    ("method_1b", {"slab_height": 16, "layer_store": "ram", "threads": 4}),
    ("method_1b", {"slab_height": 6, "layer_store": "zstd", "threads": 4, "storage": "bf16"}),
    ("method_1c", {"slab_height": 5, "layer_store": "zstd"}),
    ("method_1c", {"slab_height": 5, "layer_store": "device"}),
    # slab height from the budget: room for three 13-layer slab buffers (the
    # pipeline's ring) beside the sums, error and images of one sub-iteration,
    # ProjData and the two preprocessing upload slots
    ("method_1a", {"layer_store": "ram",
                   "device_budget_gb": (3 * 13 * W * W * 4 + 6 * W * H * 16 + 24 * 80 +
                                        2 * 6 * W * H * 4 + 100) / 1e9}),
])
def test_streaming_schedules_equal_naive(scan, method, xpr):
    """method_1a/1b/1c over every layer store (zstd spill files, pinned RAM,
    HBM) and storage type give the naive schedule's volume bit for bit."""
    t, spec, d = scan
    storage = xpr.get("storage", "f32")
    naive, _ = _run(t, f"naive_ref_{storage}", d, method="naive",
                    xpr={"storage": storage, "final_fp32": False})
    name = f"{method}_{xpr['layer_store']}_{storage}_{xpr.get('slab_height', 0)}"
    vol, log = _run(t, name, d, method=method, xpr=dict(xpr, final_fp32=False))
    np.testing.assert_array_equal(vol, naive)
    assert log["layer_store"] == xpr["layer_store"]
    if xpr["layer_store"] == "zstd":
        assert log["spill_compression"] == "zstd"          # libzstd.so.1 present
    if "slab_height" not in xpr:
        assert log["slab_height"] == 13
    # an HBM store is one launch per pass over the resident volume, except
    # method_1c's per-slab view lists
    whole = xpr["layer_store"] == "device" and method != "method_1c"
    assert log["launches_per_pass"] == (1 if whole else -(-H // log["slab_height"]))
    assert len(log["forward_pass_ms"]) == 7 and len(log["backward_pass_ms"]) == 8
    assert not list((t / name).glob("*.zstd"))          # spill removed at the end


@pytest.mark.parametrize("method,xpr", [
    ("naive", {}),
    ("naive", {"storage": "f16"}),
    ("method_1a", {"layer_store": "device", "slab_height": 7}),
    ("method_2", {"device_budget_gb": 0.30 * (W * W * H * 4 + 6 * W * H * 16) / 1e9}),
])
def test_narrow_sums_equal_wide_sums(scan, method, xpr):
    """32-bit detector sums wherever the hit bound times the volume's largest
    Q.22 value stays below 2^32 (xpr.narrow_sums, the default) give the volume
    of 64-bit sums bit for bit; the run record says which forwards took them
    (method_2: several chunks, each its own bound on its layers and rows)."""
    t, spec, d = scan
    tag = f"{method}_{xpr.get('storage', 'f32')}"
    wide, lw = _run(t, f"wide_{tag}", d, method=method, xpr=dict(xpr, narrow_sums=False))
    narrow, ln = _run(t, f"narrow_{tag}", d, method=method, xpr=dict(xpr))
    np.testing.assert_array_equal(narrow, wide)
    assert set(lw["forward_sum_bits"]) == {64} and "hit_bound" not in lw
    assert 32 in ln["forward_sum_bits"]
    assert len(ln["forward_sum_bits"]) == len(ln["forward_pass_ms"])
    assert 0 < ln["hit_bound"] < W * W * H


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_slab_final_fp32_equals_naive(scan, storage):
    """16-bit storage with the slab schedule's store in HBM: its final
    sub-iteration in float32 storage equals the naive schedule's bit for bit."""
    t, spec, d = scan
    naive, ln = _run(t, f"naive_f32last_{storage}", d, method="naive",
                     xpr={"storage": storage, "final_fp32": True})
    vol, log = _run(t, f"slab_f32last_{storage}", d, method="method_1a",
                    xpr={"storage": storage, "final_fp32": True, "layer_store": "device",
                         "slab_height": 7})
    assert log["final_fp32"] is True
    np.testing.assert_array_equal(vol, naive)
    plain, _ = _run(t, f"naive_no32_{storage}", d, method="naive",
                    xpr={"storage": storage, "final_fp32": False})
    assert not np.array_equal(vol, plain)


def test_auto_layer_store(scan):
    """layer_store auto: HBM when the whole volume fits the budget (one launch
    per pass when no slab height is given), else zstd spill files."""
    t, spec, d = scan
    _, log = _run(t, "auto_dev", d, method="method_1a", xpr={})
    assert log["layer_store"] == "device" and log["slab_height"] == H
    assert log["images_resident"]
    _, log = _run(t, "auto_zstd", d, method="method_1a", xpr={"device_budget_gb": 0.0003})
    assert log["layer_store"] == "zstd" and 1 <= log["slab_height"] < H


def test_method_2_single_chunk_equals_naive(scan):
    """With a budget that holds the whole problem the planner makes one chunk
    without padding or cropping, so method_2 is the naive schedule exactly."""
    t, spec, d = scan
    naive, _ = _run(t, "naive_ref2", d, method="naive")
    vol, log = _run(t, "m2one", d, method="method_2", xpr={"device_budget_gb": 1.0})
    assert log["chunks"] == [[0, H, 0, H, 0, 0]]
    np.testing.assert_array_equal(vol, naive)


@pytest.mark.parametrize("overlap", [1.0, 0.5])
def test_method_2_chunks_match_oracle(scan, overlap):
    """Several independent chunks (padding + cropped detector rows) against the
    oracle's chunked schedule on the plan the engine logged, with the same
    jitter (reconstruction_chunk.hpp:36-131): rel-L2 1e-5 like the naive
    schedule.  The plan equals the oracle's restatement of the planner
    (geometry_chunk.hpp:92-197)."""
    t, spec, d = scan
    budget = 0.30 * (W * W * H * 4 + 6 * W * H * 16) / 1e9
    vol, log = _run(t, f"m2_{overlap}", d, method="method_2",
                    xpr={"device_budget_gb": budget}, method2={"overlap_fct": overlap})
    chunks = log["chunks"]
    assert len(chunks) >= 2 and sum(c[1] for c in chunks) == H
    assert all(c[4] > 0 or c[0] == 0 for c in chunks)          # top padding inside
    imgs = np.stack([read_f32(d / f"proj_{a:04d}.tif") for a in range(N)])
    O.preprocess(imgs.reshape(-1), False, True)
    pd, grid = O.geometry(spec, False)
    assert chunks == O.plan_chunks(pd, grid, W, H, N // 4, 4, budget * 1e9, overlap)
    jit = np.array(log["jitter"], np.float32)
    assert len(jit) == len(chunks) * 2 * 4
    ref = O.reconstruct_chunked(imgs, pd, grid, 2, 4, 1.0, 0.9, jit, chunks, threads=8)
    assert rel_l2(vol, ref) < 1e-5


def test_normalize_and_reprojection(scan):
    t, spec, d = scan
    pout = t / "reproj"
    vol, log = _run(t, "norm", d, method="naive", normalize=True, proj_output=str(pout))
    ref = _oracle(spec, d, log, normalize=True)
    assert rel_l2(vol, ref) < 1e-5
    projs = sorted(pout.glob("p_*.tif"))
    assert len(projs) == N
    # forward_exec (src/reconstruction.hpp:66-158) against the oracle: 32
    # jittered forwards per sub-iteration of the written volume, masked and
    # divided by 32, with the jitter rows after the reconstruction's 8
    pd, grid = O.geometry(spec, False)
    jit = np.array(log["jitter"], np.float32)
    assert len(jit) == 8 + 4 * 32
    ref = O.reproject(vol, pd, grid, W, H, N, 4, jit[8:], threads=8)
    for a in range(N):
        got = read_f32(pout / f"p_{a:010d}.tif")
        assert got.shape == (H, W)
        np.testing.assert_array_equal(got, ref[a])
    assert max(float(ref[a].max()) for a in range(N)) > 0


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_mixed_precision_storage_close_to_fp32(scan, storage):
    t, spec, d = scan
    ref, _ = _run(t, "naive_ref3", d, method="naive")
    vol, _ = _run(t, f"mp_{storage}", d, method="naive", xpr={"storage": storage})
    # f16 flushes background voxels below 6e-8 to 0 (the multiplicative update
    # cannot revive them); the bulk of the volume agrees to the storage precision
    assert rel_l2(vol, ref) < (1.5e-2 if storage == "f16" else 3e-2)


def test_bad_tiff_is_a_reconstruction_error(tmp_path, cuda):
    d = tmp_path / "bad"
    d.mkdir()
    import struct
    # a 16-bit TIFF header: the reference throws "incompatble format"
    write_f32(d / "a.tif", np.zeros((4, 4), np.float32))
    b = bytearray((d / "a.tif").read_bytes())
    # BitsPerSample entry (3rd tag) -> 16
    struct.pack_into("<H", b, 8 + 2 + 12 * 2 + 8, 16)
    (d / "a.tif").write_bytes(bytes(b))
    r = subprocess.run([CLI, json.dumps({"recon": {"input": str(d), "output": str(tmp_path / "o")}})],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 256 - 3
    assert "incompatble format" in r.stderr


@pytest.mark.parametrize("method,xpr", [("naive", {}),
                                        ("method_2", {"device_budget_gb": 0.30 * (W * W * H * 4 + 6 * W * H * 16) / 1e9}),
                                        ("method_1a", {"slab_height": 9, "layer_store": "ram"})])
def test_reconstruct_mem_equals_cli(scan, method, xpr):
    """xpr_reconstruct_mem (projections and volume in host memory) gives the
    CLI's layers bit for bit, for the resident, chunked and slab schedules."""
    import ctypes as C
    from xpulse_recon_amd import _lib
    t, spec, d = scan
    ref, log = _run(t, f"cli_{method}", d, method=method, xpr=dict(xpr))
    raw = np.stack([read_f32(d / f"proj_{a:04d}.tif") for a in range(N)])
    params = {"recon": dict(it=2, sit=4, method=method),
              "geometry": {"so": 100.0, "sd": 1000.0}, "detector": {"px": 0.1, "rx": 1.5},
              "xpr": dict(xpr, jitter_seed=11)}
    vol = np.empty_like(ref)
    rec = C.create_string_buffer(1 << 16)
    print(f"xpr_reconstruct_mem {method}", flush=True)
    rc = _lib.load().xpr_reconstruct_mem(json.dumps(params).encode(),
                                         raw.ctypes.data_as(C.c_void_p), W, H, N,
                                         vol.ctypes.data_as(C.c_void_p), rec, len(rec))
    assert rc == 0
    record = json.loads(rec.value.decode())
    assert record["jitter"] == log["jitter"]
    if method == "method_2":
        assert len(record["chunks"]) >= 2
    np.testing.assert_array_equal(vol, ref)


@pytest.mark.parametrize("method,xpr,normalize", [
    ("naive", {}, False),
    ("naive", {"storage": "f16"}, True),
    ("method_2", {"device_budget_gb": 0.30 * (W * W * H * 4 + 6 * W * H * 16) / 1e9}, True),
    ("method_1a", {"slab_height": 9, "layer_store": "device"}, False),
    ("method_1a", {"slab_height": 9, "layer_store": "ram"}, True),
    ("method_1c", {"slab_height": 5, "layer_store": "zstd"}, False),
])
def test_reconstruct_dev_equals_cli(scan, method, xpr, normalize):
    """xpr_reconstruct_dev (projections read from and the volume written to
    device memory, no host copy) gives the CLI's layers bit for bit, with and
    without normalisation, for every kind of schedule; the reprojection of a
    device job reads the device volume."""
    import ctypes as C
    import torch
    from xpulse_recon_amd import _lib
    t, spec, d = scan
    pout = t / f"reproj_dev_{method}"
    ref, log = _run(t, f"clid_{method}_{normalize}", d, method=method, xpr=dict(xpr),
                    normalize=normalize, proj_output=str(t / f"reproj_cli_{method}"))
    raw = np.stack([read_f32(d / f"proj_{a:04d}.tif") for a in range(N)])
    params = {"recon": dict(it=2, sit=4, method=method, normalize=normalize,
                            output=str(t / f"devout_{method}"), proj_output=str(pout)),
              "geometry": {"so": 100.0, "sd": 1000.0}, "detector": {"px": 0.1, "rx": 1.5},
              "xpr": dict(xpr, jitter_seed=11)}
    dproj = torch.from_numpy(raw).cuda()
    dvol = torch.full(ref.shape, -1.0, device="cuda")
    torch.cuda.synchronize()
    rec = C.create_string_buffer(1 << 16)
    print(f"xpr_reconstruct_dev {method}", flush=True)
    rc = _lib.load().xpr_reconstruct_dev(json.dumps(params).encode(),
                                         C.c_void_p(dproj.data_ptr()), W, H, N,
                                         C.c_void_p(dvol.data_ptr()), rec, len(rec))
    assert rc == 0
    record = json.loads(rec.value.decode())
    assert record["jitter"] == log["jitter"]
    np.testing.assert_array_equal(dvol.cpu().numpy(), ref)
    for a in (0, 7, N - 1):
        np.testing.assert_array_equal(read_f32(pout / f"p_{a:010d}.tif"),
                                      read_f32(t / f"reproj_cli_{method}" / f"p_{a:010d}.tif"))
