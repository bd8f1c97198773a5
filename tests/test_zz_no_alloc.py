"""The allocation-free foldSeg entry point (acc_hip_fold_seg_noalloc) against
acc_hip_fold_seg, through the C ABI: identical results (the same kernels, so
bit-exact also for Float32) on two calls in a row, the second smaller; no
device allocation by the call (checked under an allocation limit that refuses
any allocation); a call beyond the fixed scratch refused before launching."""
import ctypes

import numpy as np
import pytest

import acc_hip as A
import oracle_py as O

pytestmark = pytest.mark.gpu

ADD, MAX = 0, 3


def case(stream, dt, n, m, seed):
    lens = O.gen_segment_lengths(seed, m, 16384, n)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    x = O.gen_normal_f32(seed, n).astype(dt) if np.dtype(dt).kind == "f" else \
        O.gen_uniform_i32(seed, -1000, 1000, n).astype(dt)
    return x, off, A.DeviceArray.from_numpy(stream, x), A.DeviceArray.from_numpy(stream, off)


@pytest.mark.parametrize("dt", [np.float32, np.int32, np.int64])
@pytest.mark.parametrize("op", [ADD, MAX])
def test_noalloc_matches_fold_seg_twice_second_smaller(stream, dt, op):
    s = A.Stream(0)
    for n, m, seed in (((1 << 24) + 77, 1 << 14, 3), (300001, 1000, 4)):
        x, off, dx, doff = case(s, dt, n, m, seed)
        want, got = A.DeviceArray(m, dt), A.DeviceArray(m, dt)
        A.fold_seg(s, op, None, dx, 1, n, doff, m, want)
        A.fold_seg_noalloc(s, op, None, dx, 1, n, doff, m, got)
        w, g = want.to_numpy(s), got.to_numpy(s)
        assert np.array_equal(g.view(np.uint8), w.view(np.uint8)), (np.dtype(dt).name, op, n)
        if np.dtype(dt).kind == "i":
            assert np.array_equal(g, O.fold_seg(op, None, x, off))
    s.destroy()


def test_noalloc_allocates_nothing(stream):
    s = A.Stream(0)  # a fresh stream: its own workspace has never been allocated
    x, off, dx, doff = case(s, np.float32, 1 << 22, 4096, 5)
    out = A.DeviceArray(4096, np.float32)
    ref = A.DeviceArray(4096, np.float32)
    used = ctypes.c_size_t()
    A.check(A.lib().acc_hip_allocated_bytes(0, ctypes.byref(used)))
    A.check(A.lib().acc_hip_set_alloc_limit(0, used.value))  # any further allocation fails
    try:
        for seed in (np.float32(1.5), None):
            A.fold_seg_noalloc(s, ADD, seed, dx, 1, x.size, doff, 4096, out)
            s.synchronize()
        after = ctypes.c_size_t()
        A.check(A.lib().acc_hip_allocated_bytes(0, ctypes.byref(after)))
        assert after.value == used.value
    finally:
        A.check(A.lib().acc_hip_set_alloc_limit(0, 0))
    A.fold_seg(s, ADD, None, dx, 1, x.size, doff, 4096, ref)
    assert np.array_equal(out.to_numpy(s), ref.to_numpy(s))
    s.destroy()


def test_noalloc_refuses_beyond_its_scratch(stream):
    # 2^30 Float32 elements in one row need more than the 16 MiB scratch; the
    # call is refused on the host (no launch), whatever the buffers hold
    n = 1 << 30
    dx = A.DeviceArray(n, np.float32)
    doff = A.DeviceArray.from_numpy(stream, np.array([0, n], np.int64))
    out = A.DeviceArray(1, np.float32)
    with pytest.raises(A.AccError, match="scratch"):
        A.fold_seg_noalloc(stream, ADD, None, dx, 1, n, doff, 1, out)
    # the entry point still works normally afterwards
    x = np.arange(1 << 20, dtype=np.float32) % 5
    d_small = A.DeviceArray.from_numpy(stream, x)
    d_off = A.DeviceArray.from_numpy(stream, np.array([0, 1 << 20], np.int64))
    A.fold_seg_noalloc(stream, ADD, None, d_small, 1, 1 << 20, d_off, 1, out)
    assert out.to_numpy(stream)[0] == np.float32(x.astype(np.float64).sum())
