"""Skeletons that share a stream share its workspace.  Its status region holds
only tagged words; a skeleton's untagged per-tile data (foldSeg's carry
records, segment ranges and starts) lives in the stream's scratch instead.

Regression: with foldSeg's records in the status region, a later scan on the
same stream could take a record whose upper half equalled the current tag for
a published prefix (a wrong carry from one tile on, about once in a hundred
alternations on random data).  Here the foldSeg input is all ones in short
segments, so its records are full of small integers -- the very values the
first epochs' tags take on a fresh stream -- and every scan is checked."""
import numpy as np
import pytest

import acc_hip as A
import oracle_py as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dt", [np.int32, np.int64])
def test_scan_after_fold_seg_on_a_fresh_stream(dt):
    s = A.Stream(0)  # fresh: the epoch starts at 0, so the tags are small
    n_seg = 1 << 14
    lens = np.random.default_rng(3).integers(1, 40, size=n_seg)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    ones = np.ones(int(off[-1]), dt)
    d_ones, d_off = A.DeviceArray.from_numpy(s, ones), A.DeviceArray.from_numpy(s, off)
    seg_out = A.DeviceArray(n_seg, dt)
    n = (1 << 22) + 4096
    x = O.gen_uniform_i32(9, -1000, 1000, n).astype(dt)
    d_x = A.DeviceArray.from_numpy(s, x)
    want = np.cumsum(x, dtype=dt)
    out = A.DeviceArray(n, dt)
    for rep in range(40):
        A.fold_seg(s, A.ADD, None, d_ones, 1, ones.size, d_off, n_seg, seg_out)
        A.scan(s, A.ADD, A.LEFT_TO_RIGHT, None, d_x, 1, n, out)
        got = out.to_numpy(s)
        bad = np.flatnonzero(got != want)
        assert bad.size == 0, f"rep {rep}: first wrong output {bad[0]} (delta {int(got[bad[0]]) - int(want[bad[0]])})"
        assert np.array_equal(seg_out.to_numpy(s), lens.astype(dt))
    s.destroy()
