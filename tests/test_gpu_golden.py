"""The HIP sub-iteration against the committed golden vectors
(tests/golden/subiteration_small.npz): Forward sums bit for bit (u64 and the
reference's wrapping u32), Error within the device logf tolerance, Backward
within the device expf tolerance, and the same empty-voxel count."""
import numpy as np
import pytest

from tests.test_golden import load_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("i", range(3))
def test_hip_subiteration_matches_golden(cuda, i):
    import torch
    import xpulse_recon_amd as X
    d, f = load_case(i)
    a = X.proj_args(d["det_width"], d["n_views"], d["modules"], d["vmin"], d["vmax"],
                    d["line_offset"], d["line_rows"], d["origin"], d["voxel"])
    tp = torch.from_numpy(f["projdata"]).to(cuda)
    tv = torch.from_numpy(f["volume"]).to(cuda)
    shape = f["sums_u64"].shape
    s64 = torch.zeros(shape, dtype=torch.int64, device=cuda)
    X.forward(tv, s64, tp, a)
    np.testing.assert_array_equal(s64.cpu().numpy().view(np.uint64), f["sums_u64"])
    s32 = torch.zeros(shape, dtype=torch.int32, device=cuda)
    X.forward(tv, s32, tp, a)
    np.testing.assert_array_equal(s32.cpu().numpy().view(np.uint32), f["sums_u32"])
    err = torch.empty(shape, dtype=torch.float32, device=cuda)
    X.error(torch.from_numpy(f["images"]).to(cuda), s64, err, 0.9)
    np.testing.assert_allclose(err.cpu().numpy(), f["error"], rtol=4e-7, atol=1e-7)
    empty = torch.zeros(1, dtype=torch.int64, device=cuda)
    X.backward(tv, torch.from_numpy(f["error"]).to(cuda), tp, a, empty=empty)
    np.testing.assert_allclose(tv.cpu().numpy(), f["updated"], rtol=4e-7, atol=0)
    assert int(empty.item()) == int(f["empty"][0])
