"""The oracle against the committed golden sub-iterations
(tests/golden/subiteration_small.npz, written by tests/golden/make_golden.py):
Forward sums (u64 and u32), Error, Backward update and empty-voxel count,
bit for bit.  A change of the canonical arithmetic in the oracle shows here."""
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "subiteration_small.npz")


def load_case(i):
    z = np.load(GOLDEN)
    a = z[f"c{i}_args"]
    ov = z[f"c{i}_origin_voxel"]
    d = dict(det_width=int(a[0]), n_views=int(a[1]), modules=int(a[2]),
             vmin=tuple(int(x) for x in a[3:6]), vmax=tuple(int(x) for x in a[6:9]),
             line_offset=int(a[9]), line_rows=int(a[10]),
             origin=tuple(np.float32(x) for x in ov[:3]), voxel=tuple(np.float32(x) for x in ov[3:]))
    return d, {k[len(f"c{i}_"):]: z[k] for k in z.files if k.startswith(f"c{i}_")}


def n_cases():
    return len([k for k in np.load(GOLDEN).files if k.endswith("_args")])


@pytest.mark.parametrize("i", range(3))
def test_oracle_reproduces_golden(i):
    assert n_cases() == 3
    d, f = load_case(i)
    g = O.make_geom(d["det_width"], d["n_views"], d["modules"], d["vmin"], d["vmax"],
                    d["line_offset"], d["line_rows"], d["origin"], d["voxel"])
    shape = f["sums_u64"].shape
    s64 = np.zeros(shape, np.uint64)
    O.forward_f32(f["volume"], s64, f["projdata"], g)
    np.testing.assert_array_equal(s64, f["sums_u64"])
    s32 = np.zeros(shape, np.uint32)
    O.forward_f32(f["volume"], s32, f["projdata"], g)
    np.testing.assert_array_equal(s32, f["sums_u32"])
    err = O.error_f32(f["images"], s64, 0.9)
    np.testing.assert_array_equal(err, f["error"])
    upd = f["volume"].copy()
    empty = O.backward_f32(upd, err, f["projdata"], g)
    np.testing.assert_array_equal(upd, f["updated"])
    assert empty == int(f["empty"][0])
