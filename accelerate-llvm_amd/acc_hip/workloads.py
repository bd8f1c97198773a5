"""Seeded synthetic inputs of the benchmark configurations.

Every array is one xoshiro256** stream whose state is the first four outputs
of splitmix64 started at the seed (42 unless stated), generated by the host
library lib/libacc_synth.so (csrc/synth.c), so the benchmark, the tests and
the CPU baseline see the same arrays.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import c_double, c_float, c_int, c_int32, c_int64, c_uint64, c_void_p

import numpy as np

SYNTH_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lib", "libacc_synth.so")
_lib = None


def _synth():
    global _lib
    if _lib is None:
        if not os.path.exists(SYNTH_PATH):
            raise RuntimeError(f"{SYNTH_PATH} is not built (make -C accelerate-llvm_amd)")
        L = ctypes.CDLL(SYNTH_PATH)
        sig = {
            "acc_synth_splitmix64": (None, [c_uint64, c_void_p, c_int64]),
            "acc_synth_xoshiro": (None, [c_void_p, c_void_p, c_int64]),
            "acc_synth_uniform_i32": (None, [c_uint64, c_int32, c_int32, c_void_p, c_int64]),
            "acc_synth_uniform_f32": (None, [c_uint64, c_float, c_float, c_void_p, c_int64]),
            "acc_synth_normal_f32": (None, [c_uint64, c_void_p, c_int64]),
            "acc_synth_zipf_i32": (c_int, [c_uint64, c_double, c_int32, c_void_p, c_int64]),
            "acc_synth_segment_lengths": (c_int, [c_uint64, c_int64, c_int64, c_int64, c_void_p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _lib = L
    return _lib


def _p(a: np.ndarray) -> c_void_p:
    return a.ctypes.data_as(c_void_p)


def splitmix64(x: int, n: int) -> np.ndarray:
    """n outputs of splitmix64 from state x."""
    out = np.empty(n, np.uint64)
    _synth().acc_synth_splitmix64(c_uint64(x), _p(out), n)
    return out


def xoshiro256ss(state, n: int) -> np.ndarray:
    """n outputs of xoshiro256** from the four given state words."""
    st = np.asarray(state, dtype=np.uint64)
    assert st.shape == (4,)
    out = np.empty(n, np.uint64)
    _synth().acc_synth_xoshiro(_p(st), _p(out), n)
    return out


def uniform_i32(seed: int, lo: int, hi: int, n: int) -> np.ndarray:
    """n Int32 values uniform on the closed range [lo, hi]."""
    out = np.empty(n, np.int32)
    _synth().acc_synth_uniform_i32(c_uint64(seed), lo, hi, _p(out), n)
    return out


def uniform_f32(seed: int, lo: float, hi: float, n: int) -> np.ndarray:
    """n Float32 values uniform on [lo, hi)."""
    out = np.empty(n, np.float32)
    _synth().acc_synth_uniform_f32(c_uint64(seed), lo, hi, _p(out), n)
    return out


def normal_f32(seed: int, n: int) -> np.ndarray:
    """n Float32 values ~ N(0, 1) (Box-Muller from the same stream)."""
    out = np.empty(n, np.float32)
    _synth().acc_synth_normal_f32(c_uint64(seed), _p(out), n)
    return out


def zipf_i32(seed: int, s: float, nbins: int, n: int) -> np.ndarray:
    """n bin indices in [0, nbins) with P(k) proportional to (k+1)^-s (inverse CDF)."""
    out = np.empty(n, np.int32)
    if _synth().acc_synth_zipf_i32(c_uint64(seed), s, nbins, _p(out), n):
        raise ValueError("zipf: nbins must be positive")
    return out


def segment_lengths(seed: int, nseg: int, maxlen: int, total: int) -> np.ndarray:
    """nseg Int64 lengths, log-uniform in [1, maxlen], scaled to sum to total
    (each at least 1; a rounding shortfall goes to the longest, an excess is
    taken off the segments in index order, never below 1)."""
    if nseg > total:
        raise ValueError("more segments than elements")
    out = np.empty(nseg, np.int64)
    if _synth().acc_synth_segment_lengths(c_uint64(seed), nseg, maxlen, total, _p(out)):
        raise ValueError("segment lengths: need 1 <= nseg <= total and maxlen >= 1")
    return out


def offsets(lengths: np.ndarray) -> np.ndarray:
    """`scanl (+) 0` of the segment lengths, as the frontend hands them to foldSeg."""
    return np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
