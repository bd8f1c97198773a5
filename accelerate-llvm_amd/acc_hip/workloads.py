"""Seeded synthetic inputs of the benchmark configurations.

Every generator is a pure function of its seed (numpy's PCG64), so the
benchmark, the tests and the CPU baseline see the same arrays.
"""
from __future__ import annotations

import numpy as np


def uniform_i32(seed: int, lo: int, hi: int, n: int) -> np.ndarray:
    """n Int32 values uniform on the closed range [lo, hi]."""
    return np.random.default_rng(seed).integers(lo, hi, size=n, dtype=np.int32, endpoint=True)


def uniform_f32(seed: int, lo: float, hi: float, n: int) -> np.ndarray:
    """n Float32 values uniform on [lo, hi)."""
    return np.random.default_rng(seed).uniform(lo, hi, size=n).astype(np.float32)


def normal_f32(seed: int, n: int) -> np.ndarray:
    """n Float32 values ~ N(0, 1)."""
    return np.random.default_rng(seed).standard_normal(n, dtype=np.float32)


def zipf_i32(seed: int, s: float, nbins: int, n: int) -> np.ndarray:
    """n bin indices in [0, nbins) with P(k) proportional to (k+1)^-s (inverse CDF)."""
    cdf = np.cumsum(np.arange(1, nbins + 1, dtype=np.float64) ** -s)
    cdf /= cdf[-1]
    u = np.random.default_rng(seed).random(n)
    return np.minimum(np.searchsorted(cdf, u, side="right"), nbins - 1).astype(np.int32)


def segment_lengths(seed: int, nseg: int, maxlen: int, total: int) -> np.ndarray:
    """nseg Int64 lengths, log-uniform in [1, maxlen], scaled to sum to total
    (each at least 1; a rounding shortfall goes to the longest, an excess is
    taken off the longest ones, never below 1)."""
    if nseg > total:
        raise ValueError("more segments than elements")
    rng = np.random.default_rng(seed)
    l = np.exp(rng.uniform(0.0, np.log(maxlen), size=nseg))
    out = np.maximum(1, np.rint(l * (total / l.sum()))).astype(np.int64)
    diff = total - int(out.sum())
    if diff >= 0:
        out[np.argmax(out)] += diff
    else:
        order = np.argsort(-out, kind="stable")
        room = np.cumsum(out[order] - 1)  # what the longest k can give up
        k = int(np.searchsorted(room, -diff)) + 1
        take = out[order[:k]] - 1
        take[-1] -= int(room[k - 1]) + diff  # the k-th gives only the rest
        out[order[:k]] -= take
    return out


def offsets(lengths: np.ndarray) -> np.ndarray:
    """`scanl (+) 0` of the segment lengths, as the frontend hands them to foldSeg."""
    return np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
