"""The C-ABI library loads and exports every function declared in
include/accelerate_hip.h; error paths that never reach the GPU behave as
documented.  Runs without a GPU (no compute calls)."""
import ctypes
import os
import re

import pytest

import acc_hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", f) for f in sorted(os.listdir(os.path.join(ROOT, "include")))
           if f.endswith(".h")]


def declared_functions():
    names = []
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(acc_hip_[a-z_0-9]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(acc_hip.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_functions()) <= set(acc_hip.SIGNATURES)


def test_abi_version():
    assert acc_hip.lib().acc_hip_abi_version() == 1


def test_null_stream_is_invalid_value_with_message():
    L = acc_hip.lib()
    rc = L.acc_hip_fold(None, acc_hip.INT32, acc_hip.ADD, None, None, 1, 10, None)
    assert rc == acc_hip.INVALID_VALUE
    assert b"null stream" in L.acc_hip_last_error()


def test_error_codes_map_to_exceptions():
    with pytest.raises(acc_hip.AccError):
        acc_hip.check(acc_hip.INVALID_VALUE)
    with pytest.raises(acc_hip.OutOfMemory):
        acc_hip.check(acc_hip.OUT_OF_MEMORY)
    with pytest.raises(acc_hip.EmptyArray):
        acc_hip.check(acc_hip.EMPTY_ARRAY)


def test_bad_device_index_is_rejected():
    L = acc_hip.lib()
    assert L.acc_hip_set_alloc_limit(-1, 0) == acc_hip.INVALID_VALUE
    p = ctypes.c_void_p()
    assert L.acc_hip_malloc(999, 16, ctypes.byref(p)) != acc_hip.SUCCESS


def test_workload_generators():
    from acc_hip import workloads as W
    import numpy as np
    u = W.uniform_i32(42, -1000, 1000, 1 << 16)
    assert u.dtype == np.int32 and u.min() == -1000 and u.max() == 1000
    assert np.array_equal(u, W.uniform_i32(42, -1000, 1000, 1 << 16))
    lens = W.segment_lengths(42, 1 << 12, 16384, 1 << 22)
    assert lens.sum() == 1 << 22 and lens.min() >= 1
    off = W.offsets(lens)
    assert off[0] == 0 and off[-1] == 1 << 22 and np.all(np.diff(off) >= 1)
    z = W.zipf_i32(42, 1.1, 4096, 1 << 20)
    p1 = 1.0 / sum(k ** -1.1 for k in range(1, 4097))
    assert z.min() >= 0 and z.max() < 4096 and abs((z == 0).mean() - p1) < 0.01
    f = W.uniform_f32(1, -1.0, 1.0, 1000)
    assert f.dtype == np.float32 and f.min() >= -1 and f.max() < 1
    assert W.normal_f32(3, 10).dtype == np.float32


def test_segment_lengths_dense_stay_positive():
    # nearly as many segments as elements: the rounding excess must not drive
    # a length below 1 (offsets stay monotonic)
    from acc_hip import workloads as W
    import numpy as np
    for nseg, maxlen, total in ((3145745, 3, 3145745), (1000, 5, 1200), (7, 1000, 7)):
        lens = W.segment_lengths(21, nseg, maxlen, total)
        assert lens.min() >= 1 and int(lens.sum()) == total
    with pytest.raises(ValueError):
        W.segment_lengths(1, 10, 3, 9)
