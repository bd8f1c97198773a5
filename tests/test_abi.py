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
