"""xpr_subiteration_static (the sub-iteration with its workspace in the
library's static device scratch, no allocation) against xpr_subiteration with
a caller workspace, through the C-ABI: two sizes in a row, the second smaller
(the scratch keeps the first call's contents, so the sums must be cleared
inside the call), each over a first pass and a forward pass.

Bar: bit-identical volumes and empty-voxel counts (the same kernels on the
same inputs)."""
import numpy as np
import pytest

from tests.helpers import geom_args, projdata_for, small_spec

pytestmark = pytest.mark.gpu

MAX_BYTES = 64 << 20


def _case(w, h, n, views, slab=None, crop=None, rx=1.5):
    spec = small_spec(w, h, n, rx=rx)
    pd, grid = projdata_for(spec, views)
    vmin, vmax = slab if slab else (None, None)
    lo, rows = crop if crop else (0, None)
    d = geom_args(spec, grid, len(views), vmin=vmin, vmax=vmax, jitter=(0.003, 0.002, 0.001),
                  line_offset=lo, line_rows=rows)
    return pd, d


def _args(d):
    import xpulse_recon_amd as X
    return X.proj_args(d["det_width"], d["n_views"], d["modules"], d["vmin"], d["vmax"],
                       d["line_offset"], d["line_rows"], d["origin"], d["voxel"])


def _run_pair(cuda, pd, d, acc, err_dt, seed):
    import torch
    import xpulse_recon_amd as X
    nx, nl, nk = (d["vmax"][i] - d["vmin"][i] for i in range(3))
    shape = (d["n_views"], d["line_rows"], d["det_width"])
    g = torch.Generator().manual_seed(seed)
    vol = (torch.rand((nl, nk, nx), generator=g) + 0.5).to(cuda)
    img = (torch.rand(shape, generator=g) * 3.0).to(cuda)
    tpd = torch.from_numpy(pd).to(cuda)
    a, b = vol.clone(), vol.clone()
    edt = {"f32": torch.float32, "f16": torch.float16, "bf16": torch.bfloat16}[err_dt]
    sums = torch.empty(shape, dtype=torch.int64 if acc == 64 else torch.int32, device=cuda)
    err = torch.empty(shape, dtype=edt, device=cuda)
    for k, w in enumerate((1.0, 0.9)):
        ea = torch.zeros(1, dtype=torch.int64, device=cuda)
        eb = torch.zeros(1, dtype=torch.int64, device=cuda)
        print(f"static sub-iteration {k}: views {shape[0]}, detector {shape[2]}x{shape[1]}, "
              f"slab {nx}x{nl}x{nk}, acc {acc}, err {err_dt}", flush=True)
        X.subiteration_static(a, img, tpd, _args(d), w, k == 0, acc=acc, err_dtype=err_dt,
                              empty=ea)
        X.subiteration(b, img, tpd, _args(d), w, k == 0, sums, err, empty=eb)
        torch.cuda.synchronize()
        assert torch.equal(a, b)
        assert int(ea.item()) == int(eb.item())
    assert not torch.equal(a, vol)          # the update did something


@pytest.mark.parametrize("acc,err_dt", [(64, "f32"), (32, "f16")])
def test_static_scratch_equals_workspace_two_sizes(cuda, acc, err_dt):
    # 36 views of 384 x 384: 63.7 MB of 64-bit sums and f32 error, near the maximum
    pd, d = _case(384, 384, 720, list(range(0, 720, 20)), slab=((0, 176, 0), (384, 208, 384)))
    assert d["n_views"] * d["line_rows"] * d["det_width"] * 12 + 256 <= MAX_BYTES
    _run_pair(cuda, pd, d, acc, err_dt, 1)
    # then a smaller scan with a cropped band of detector rows
    pd, d = _case(100, 40, 20, list(range(1, 20, 2)), crop=(8, 20))
    _run_pair(cuda, pd, d, acc, err_dt, 2)


def test_static_scratch_rejects_oversize(cuda):
    import torch
    import xpulse_recon_amd as X
    # 36 views of 768 x 768 need 255 MB: refused before any launch
    pd, d = _case(768, 8, 720, list(range(0, 720, 20)))
    nx, nl, nk = (d["vmax"][i] - d["vmin"][i] for i in range(3))
    vol = torch.ones((nl, nk, nx), device=cuda)
    big = dict(d, line_rows=768)
    img = torch.ones((big["n_views"], 768, big["det_width"]), device=cuda)
    with pytest.raises(X.XprError, match="static scratch"):
        X.subiteration_static(vol, img, torch.from_numpy(pd).to(cuda), _args(big), 1.0, False)
    assert torch.equal(vol, torch.ones_like(vol))
