"""The remote memory table on the GPU: used arrays evicted and uploaded again
under memory pressure, dirty results copied back before eviction, lazy
copy-back, the nursery, and out of memory when nothing can be evicted
(PTX/Array/Table.hs:35-49, PTX/Array/Remote.hs:63-74, PTX/Array/Data.hs:91-152).
Memory pressure comes from the library's allocation limit."""
import ctypes
from contextlib import contextmanager

import numpy as np
import pytest

import acc_hip as A

pytestmark = pytest.mark.gpu
MiB = 1 << 20


@contextmanager
def limit(extra):
    used = ctypes.c_size_t()
    A.check(A.lib().acc_hip_allocated_bytes(0, ctypes.byref(used)))
    A.check(A.lib().acc_hip_set_alloc_limit(0, used.value + extra))
    try:
        yield
    finally:
        A.check(A.lib().acc_hip_set_alloc_limit(0, 0))


def test_used_arrays_are_evicted_and_uploaded_again(stream):
    t = A.Table(0)
    n = 4 * MiB // 4  # 4 MiB of Int32 each
    arrays = {k: np.random.default_rng(k).integers(-1000, 1000, size=n).astype(np.int32) for k in range(1, 11)}
    out = A.DeviceArray(1, np.int32)
    with limit(20 * MiB):  # room for about four of the ten arrays
        for k, a in arrays.items():
            d = t.use(stream, k, a)
            A.fold(stream, A.ADD, np.int32(0), d, 1, n, out)
            t.mark(stream, k)
            assert out.to_numpy(stream)[0] == np.sum(a, dtype=np.int32)
        st = t.stats()
        assert st.evictions >= 5 and st.resident_bytes <= 20 * MiB
        # every array is still there: the evicted ones come back from their host copies
        for k in (1, 10, 2, 9, 3):
            d = t.get(stream, k, n, np.int32)
            A.fold(stream, A.ADD, np.int32(0), d, 1, n, out)
            t.mark(stream, k)
            assert out.to_numpy(stream)[0] == np.sum(arrays[k], dtype=np.int32), k
        assert t.stats().uploads > 10
    t.destroy()


def test_dirty_results_are_copied_back_before_eviction(stream):
    t = A.Table(0)
    n = 4 * MiB // 4
    o = A.DeviceArray(1, np.int64)
    x = A.DeviceArray(n, np.int64)  # outside the table, before the limit
    with limit(14 * MiB):
        r = t.alloc(100, n, np.int32)
        A.generate(stream, A.GEN_IOTA, n, r, 7, 3)  # result = 7 + 3i, on the device only
        t.mark(stream, 100, writes=True)
        # fill the limit with used arrays until the result has to go
        for k in range(1, 6):
            t.use(stream, k, np.full(n, k, np.int32))
            t.mark(stream, k)
        st = t.stats()
        assert st.downloads >= 1 and st.evictions >= 1
        host, copied = t.peek(100)  # already clean: nothing to copy
        assert not copied
        assert np.array_equal(host, 7 + 3 * np.arange(n, dtype=np.int32))
        d = t.get(stream, 100, n, np.int32)  # back on the device from the copy
        A.map_(stream, A.UNOP_CAST, d, n, x)
        A.fold(stream, A.ADD, np.int64(0), x, 1, n, o)
        assert o.to_numpy(stream)[0] == int((7 + 3 * np.arange(n, dtype=np.int64)).sum())
    t.destroy()


def test_lazy_copy_back_only_when_newer(stream):
    t = A.Table(0)
    r = t.alloc(7, 1000, np.float32)
    A.generate(stream, A.GEN_FILL, 1000, r, 2.5)
    t.mark(stream, 7, writes=True)
    host, copied = t.peek(7)
    assert copied and np.all(host == 2.5)
    host, copied = t.peek(7)
    assert not copied
    t.destroy()


def test_nursery_reuses_a_freed_block(stream):
    t = A.Table(0)
    a = t.alloc(1, 123457, np.int32)
    t.free(1)
    b = t.alloc(2, 123000, np.int32)  # same 4096-byte multiple
    assert b.ptr == a.ptr and t.stats().nursery_hits == 1
    t.destroy()


def test_out_of_memory_when_nothing_can_be_evicted(stream):
    t = A.Table(0)
    with limit(10 * MiB):
        t.alloc(1, 2 * MiB, np.int32, host_backing=False)  # 8 MiB, device-only
        with pytest.raises(A.OutOfMemory):
            t.alloc(2, MiB, np.int32, host_backing=False)
        t.free(1)  # into the nursery
        t.alloc(3, MiB, np.int32, host_backing=False)  # the nursery is released and the allocation fits
        assert t.stats().oom_retries >= 2
    t.destroy()


def test_used_array_is_read_only(stream):
    t = A.Table(0)
    t.use(stream, 5, np.arange(10, dtype=np.int32))
    with pytest.raises(A.AccError, match="read-only"):
        t.mark(stream, 5, writes=True)
    t.destroy()
