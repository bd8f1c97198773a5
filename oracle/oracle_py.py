"""numpy front end of liboracle.so (the C restatement in oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import c_double, c_float, c_int, c_int32, c_int64, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_SUFFIX = {
    np.dtype(np.int32): "i32",
    np.dtype(np.int64): "i64",
    np.dtype(np.uint32): "u32",
    np.dtype(np.uint64): "u64",
    np.dtype(np.float32): "f32",
    np.dtype(np.float64): "f64",
}

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
        for name in ("or_native_fold_add_i32",):
            getattr(_lib, name).restype = c_int32
        _lib.or_divide_work1.restype = c_int64
        _lib.or_build_info.restype = ctypes.c_char_p
    return _lib


def build_info() -> str:
    """Compiler and -march of the loaded oracle library."""
    return lib().or_build_info().decode()


def _ptr(a):
    return None if a is None else a.ctypes.data_as(c_void_p)


def _seed(dtype, seed):
    """A scalar, or a sequence (the function boundary's c0, cy, cx), as a C array."""
    if seed is None:
        return None, None
    a = np.ascontiguousarray(np.array(seed, dtype=dtype).reshape(-1))
    return a, _ptr(a)


def _check(rc, what):
    if rc == -2:
        raise ValueError(f"{what}: empty array")
    if rc != 0:
        raise ValueError(f"{what}: unsupported combine function")


def fold(op, seed, x, outer=1):
    x = np.ascontiguousarray(x)
    inner = x.size // outer if outer else 0
    out = np.empty(outer, x.dtype)
    keep, sp = _seed(x.dtype, seed)
    f = getattr(lib(), "or_fold_" + _SUFFIX[x.dtype])
    _check(f(c_int(op), sp, _ptr(x), c_int64(outer), c_int64(inner), _ptr(out)), "fold")
    return out


def fold_seg(op, seed, x, offsets, outer=1):
    x = np.ascontiguousarray(x)
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    nseg = off.size - 1
    inner = x.size // outer if outer else 0
    out = np.zeros(outer * nseg, x.dtype)
    keep, sp = _seed(x.dtype, seed)
    f = getattr(lib(), "or_fold_seg_" + _SUFFIX[x.dtype])
    _check(f(c_int(op), sp, _ptr(x), c_int64(outer), c_int64(inner), _ptr(off), c_int64(nseg), _ptr(out)), "foldSeg")
    return out


def scan(op, direction, seed, x, outer=1):
    x = np.ascontiguousarray(x)
    inner = x.size // outer if outer else 0
    on = inner + 1 if seed is not None else inner
    out = np.empty(outer * on, x.dtype)
    keep, sp = _seed(x.dtype, seed)
    f = getattr(lib(), "or_scan_" + _SUFFIX[x.dtype])
    _check(f(c_int(op), c_int(direction), sp, _ptr(x), c_int64(outer), c_int64(inner), _ptr(out)), "scan")
    return out


def scan_prime(op, direction, seed, x, outer=1):
    x = np.ascontiguousarray(x)
    inner = x.size // outer if outer else 0
    out = np.empty(outer * inner, x.dtype)
    sums = np.empty(outer, x.dtype)
    keep, sp = _seed(x.dtype, seed)
    f = getattr(lib(), "or_scan_prime_" + _SUFFIX[x.dtype])
    _check(f(c_int(op), c_int(direction), sp, _ptr(x), c_int64(outer), c_int64(inner), _ptr(out), _ptr(sums)),
           "scan'")
    return out, sums


def permute(op, defaults, src, fill, idx):
    out = np.array(defaults, copy=True)
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    s = None if src is None else np.ascontiguousarray(src, dtype=out.dtype)
    keep, fp = _seed(out.dtype, fill)
    f = getattr(lib(), "or_permute_" + _SUFFIX[out.dtype])
    _check(f(c_int(op), _ptr(out), c_int64(out.size), _ptr(s), fp, _ptr(idx), c_int64(idx.size)), "permute")
    return out


def stencil2d(ry, rx, weights, boundary, constant, img):
    img = np.ascontiguousarray(img)
    h, w = img.shape
    wt = np.ascontiguousarray(np.asarray(weights, dtype=img.dtype).reshape(-1))
    out = np.empty_like(img)
    keep, cp = _seed(img.dtype, constant if constant is not None else 0)
    f = getattr(lib(), "or_stencil2d_" + _SUFFIX[img.dtype])
    _check(f(c_int(ry), c_int(rx), _ptr(wt), c_int(boundary), cp, _ptr(img), c_int64(h), c_int64(w), _ptr(out)),
           "stencil")
    return out


# ---- producers (numpy restatements; numpy float32/float64 arithmetic rounds
# after every operation, integer arrays wrap, astype is the C conversion)
UNOP_ID, UNOP_NEG, UNOP_ABS, UNOP_AFFINE, UNOP_CAST = range(5)
GEN_FILL, GEN_IOTA = range(2)
IX_GATHER, IX_REVERSE, IX_TRANSPOSE = range(3)


def _wrap_int(fn, x, *args):
    """fn over the unsigned type of the same width (wrapping), back to x's type."""
    u = np.dtype(f"u{x.dtype.itemsize}")
    return fn(x.view(u), *[np.asarray(a, dtype=x.dtype).astype(u) for a in args]).astype(u).view(x.dtype)


def map_unop(f, x, out_dtype=None, a=None, b=None):
    """map f (PTX/CodeGen/Map.hs:37-63) with the element functions of
    LLVM/CodeGen/Arithmetic.hs:69-86: negate x = x * (-1), abs = fabs /
    llvm.abs (abs minBound == minBound) / identity for unsigned types."""
    x = np.ascontiguousarray(x)
    if f == UNOP_CAST:
        return x.astype(out_dtype)
    if f == UNOP_ID:
        return x.copy()
    integral = x.dtype.kind in "iu"
    if f == UNOP_NEG:
        return _wrap_int(lambda v: np.uint64(0).astype(v.dtype) - v, x) if integral else -x
    if f == UNOP_ABS:
        if x.dtype.kind == "u":
            return x.copy()
        if integral:
            return np.where(x < 0, _wrap_int(lambda v: np.uint64(0).astype(v.dtype) - v, x), x)
        return np.abs(x)
    if f == UNOP_AFFINE:
        if integral:
            return _wrap_int(lambda v, aa, bb: v * aa + bb, x, a, b)
        t = x.dtype.type
        return (t(a) * x) + t(b)
    raise ValueError(f"element function {f}")


def zipwith(op, a, b):
    """zipWith f a b for the combine functions of fold plus SUB (a - b)."""
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    integral = a.dtype.kind in "iu"
    if op in (0, 1, 7) and integral:  # ADD MUL SUB wrap
        fn = {0: np.add, 1: np.multiply, 7: np.subtract}[op]
        return _wrap_int(lambda u, v: fn(u, v), a, b)
    if op == 0:
        return a + b
    if op == 1:
        return a * b
    if op == 7:
        return a - b
    if op == 2:
        return np.fmin(a, b) if not integral else np.minimum(a, b)
    if op == 3:
        return np.fmax(a, b) if not integral else np.maximum(a, b)
    if op == 4:
        return a & b
    if op == 5:
        return a | b
    if op == 6:
        return a ^ b
    raise ValueError(f"combine {op}")


def generate(g, n, dtype, a, b=0):
    """generate (PTX/CodeGen/Generate.hs:37-62) with fill a or iota a + i*b."""
    dtype = np.dtype(dtype)
    if g == GEN_FILL:
        return np.full(n, a, dtype=dtype)
    if g == GEN_IOTA:
        if dtype.kind in "iu":
            u = np.dtype(f"u{dtype.itemsize}")
            au = np.asarray(a, dtype=dtype).astype(u)
            bu = np.asarray(b, dtype=dtype).astype(u)
            return (au + np.arange(n, dtype=np.uint64).astype(u) * bu).astype(u).view(dtype)
        t = dtype.type
        return t(a) + np.arange(n, dtype=np.int64).astype(dtype) * t(b)
    raise ValueError(f"generator {g}")


def transform(p, f, x, in_outer, in_inner, index=None, n=None, a=None, b=None):
    """transform / backpermute (PTX/CodeGen/Transform.hs:37-67): out[i] =
    f(x[p(i)]); a gather index outside x gives 0."""
    x = np.ascontiguousarray(x).reshape(-1)
    if p == IX_GATHER:
        idx = np.asarray(index, dtype=np.int64)
        ok = (idx >= 0) & (idx < x.size)
        src = np.zeros(idx.size, x.dtype)
        src[ok] = x[idx[ok]]
        y = map_unop(f, src, a=a, b=b)
        y[~ok] = 0
        return y
    if p == IX_REVERSE:
        return map_unop(f, x[::-1].copy(), a=a, b=b)
    if p == IX_TRANSPOSE:
        return map_unop(f, x.reshape(in_outer, in_inner).T.copy().reshape(-1), a=a, b=b)
    raise ValueError(f"index map {p}")


# ---- 3-D stencil (numpy restatement)
def _bound3(b, i, n):
    """The boundary map of LLVM/CodeGen/Stencil.hs:224-253 (clamp 0, mirror 1,
    wrap 2; constant 3 maps like clamp, the caller substitutes the value): one
    step of the map, then clamped into range (extents below the radius)."""
    i = np.asarray(i, dtype=np.int64)
    if b == 1:
        m = np.where(i < 0, -i, np.where(i >= n, 2 * n - 2 - i, i))
    elif b == 2:
        m = np.where(i < 0, i + n, np.where(i >= n, i - n, i))
    else:
        m = i
    return np.clip(m, 0, n - 1)


def _stencil3d(rz, ry, rx, weights, boundary, constant, vol, dtype):
    vol = np.asarray(vol)
    D, H, W = vol.shape
    zs, ys, xs = np.arange(-rz, D + rz), np.arange(-ry, H + ry), np.arange(-rx, W + rx)
    P = vol[np.ix_(_bound3(boundary, zs, D), _bound3(boundary, ys, H), _bound3(boundary, xs, W))].astype(dtype)
    if boundary == 3:
        out = ((zs < 0) | (zs >= D))[:, None, None] | ((ys < 0) | (ys >= H))[None, :, None] | \
              ((xs < 0) | (xs >= W))[None, None, :]
        P[out] = constant
    w = np.asarray(weights, dtype=dtype).reshape(2 * rz + 1, 2 * ry + 1, 2 * rx + 1)
    acc = None
    for dz in range(2 * rz + 1):
        for dy in range(2 * ry + 1):
            for dx in range(2 * rx + 1):
                t = w[dz, dy, dx] * P[dz:dz + D, dy:dy + H, dx:dx + W]
                acc = t if acc is None else acc + t
    return acc


def stencil3d(rz, ry, rx, weights, boundary, constant, vol):
    """stencil1 on a DIM3 array with the weighted-sum function, taps in tuple
    order, in the element type (integers wrap)."""
    vol = np.asarray(vol)
    return _stencil3d(rz, ry, rx, weights, boundary, constant, vol, vol.dtype)


def stencil3d_f64(rz, ry, rx, weights, boundary, constant, vol):
    """FP64 truth and the sum of |w * x| per output (for the rounding bound)."""
    t = _stencil3d(rz, ry, rx, weights, boundary, constant, vol, np.float64)
    ab = _stencil3d(rz, ry, rx, np.abs(np.asarray(weights, np.float64)), boundary,
                    abs(constant) if constant is not None else None, np.abs(np.asarray(vol, np.float64)), np.float64)
    return t, ab


# ---- FP64 truth for Float32 data
def fold_f64(x, seed=None):
    x = np.ascontiguousarray(x, dtype=np.float32)
    o, a = c_double(), c_double()
    keep, sp = _seed(np.float32, seed)
    lib().or_fold_f32_f64(_ptr(x), c_int64(x.size), sp, ctypes.byref(o), ctypes.byref(a))
    return o.value, a.value


def fold_seg_f64(x, offsets, seed=None):
    x = np.ascontiguousarray(x, dtype=np.float32)
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    nseg = off.size - 1
    out = np.empty(nseg, np.float64)
    ab = np.empty(nseg, np.float64)
    keep, sp = _seed(np.float32, seed)
    lib().or_fold_seg_f32_f64(_ptr(x), _ptr(off), c_int64(nseg), sp, _ptr(out), _ptr(ab))
    return out, ab


def scan_f64(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(x.size, np.float64)
    ab = np.empty(x.size, np.float64)
    lib().or_scan_f32_f64(_ptr(x), c_int64(x.size), _ptr(out), _ptr(ab))
    return out, ab


def stencil2d_f64(ry, rx, weights, boundary, constant, img):
    img = np.ascontiguousarray(img, dtype=np.float32)
    h, w = img.shape
    wt = np.ascontiguousarray(np.asarray(weights, dtype=np.float32).reshape(-1))
    out = np.empty((h, w), np.float64)
    ab = np.empty((h, w), np.float64)
    keep, cp = _seed(np.float32, constant if constant is not None else 0)
    lib().or_stencil2d_f32_f64(c_int(ry), c_int(rx), _ptr(wt), c_int(boundary), cp, _ptr(img), c_int64(h),
                               c_int64(w), _ptr(out), _ptr(ab))
    return out, ab


def _crop2(a, b):
    h, w = min(a.shape[0], b.shape[0]), min(a.shape[1], b.shape[1])
    return a[:h, :w], b[:h, :w]


def stencil2_2d(ry, rx, w1, b1, c1, img1, w2, b2, c2, img2):
    """stencil2 on DIM2 arrays with the weighted-sum function (stencil2Op,
    PTX/Execute.hs:669-684; NAT/CodeGen/Stencil.hs:75-107): each input's
    stencil over its own shape through its own boundary, both cut to the
    intersection of the shapes (PTX/Execute.hs:682) and added in the element
    type (integers wrap; floating sums round differently from a single
    accumulator: use stencil2_2d_f64 for them)."""
    s1, s2 = _crop2(stencil2d(ry, rx, w1, b1, c1, img1), stencil2d(ry, rx, w2, b2, c2, img2))
    return s1 + s2


def stencil2_2d_f64(ry, rx, w1, b1, c1, img1, w2, b2, c2, img2):
    """FP64 truth of stencil2_2d for Float32 inputs and the sum of |w * x| per
    output (for the rounding bound)."""
    t1, a1 = stencil2d_f64(ry, rx, w1, b1, c1, img1)
    t2, a2 = stencil2d_f64(ry, rx, w2, b2, c2, img2)
    t1, t2 = _crop2(t1, t2)
    a1, a2 = _crop2(a1, a2)
    return t1 + t2, a1 + a2


def divide_work1(splits, minsize, lo, hi):
    cap = 4 * max(splits, 1) + 16
    r = np.empty(2 * cap, np.int64)
    n = lib().or_divide_work1(c_int64(splits), c_int64(minsize), c_int64(lo), c_int64(hi), _ptr(r), c_int64(cap))
    return [(int(r[2 * i]), int(r[2 * i + 1])) for i in range(min(n, cap))]


# ---- native-schedule parallel versions (CPU baseline)
def native_fold_add_i32(threads, x, seed=0):
    x = np.ascontiguousarray(x, dtype=np.int32)
    return int(lib().or_native_fold_add_i32(c_int(threads), _ptr(x), c_int64(x.size), c_int32(seed)))


def native_scanl1_add_i32(threads, x, out=None):
    x = np.ascontiguousarray(x, dtype=np.int32)
    if out is None:
        out = np.empty_like(x)
    lib().or_native_scanl1_add_i32(c_int(threads), _ptr(x), c_int64(x.size), _ptr(out))
    return out


def native_fold1_seg_f32(threads, op, x, offsets, out=None):
    x = np.ascontiguousarray(x, dtype=np.float32)
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    if out is None:
        out = np.empty(off.size - 1, np.float32)
    lib().or_native_fold1_seg_f32(c_int(threads), c_int(op), _ptr(x), _ptr(off), c_int64(off.size - 1), _ptr(out))
    return out


def native_histogram_i32(threads, idx, m, bins=None):
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    if bins is None:
        bins = np.zeros(m, np.int32)
    lib().or_native_histogram_i32(c_int(threads), _ptr(idx), c_int64(idx.size), _ptr(bins), c_int64(m))
    return bins


def native_stencil2d_f32(threads, ry, rx, weights, boundary, img, out=None):
    img = np.ascontiguousarray(img, dtype=np.float32)
    h, w = img.shape
    wt = np.ascontiguousarray(np.asarray(weights, dtype=np.float32).reshape(-1))
    if out is None:
        out = np.empty_like(img)
    lib().or_native_stencil2d_f32(c_int(threads), c_int(ry), c_int(rx), _ptr(wt), c_int(boundary), _ptr(img),
                                  c_int64(h), c_int64(w), _ptr(out))
    return out


# ---- seeded generators
def gen_uniform_i32(seed, lo, hi, n):
    out = np.empty(n, np.int32)
    lib().or_gen_uniform_i32(c_uint64(seed), c_int32(lo), c_int32(hi), _ptr(out), c_int64(n))
    return out


def gen_uniform_f32(seed, lo, hi, n):
    out = np.empty(n, np.float32)
    lib().or_gen_uniform_f32(c_uint64(seed), c_float(lo), c_float(hi), _ptr(out), c_int64(n))
    return out


def gen_normal_f32(seed, n):
    out = np.empty(n, np.float32)
    lib().or_gen_normal_f32(c_uint64(seed), _ptr(out), c_int64(n))
    return out


def gen_zipf_i32(seed, s, nbins, n):
    out = np.empty(n, np.int32)
    lib().or_gen_zipf_i32(c_uint64(seed), c_double(s), c_int32(nbins), _ptr(out), c_int64(n))
    return out


def gen_segment_lengths(seed, nseg, maxlen, total):
    if nseg > total:
        raise ValueError("more segments than elements")
    out = np.empty(nseg, np.int64)
    lib().or_gen_segment_lengths(c_uint64(seed), c_int64(nseg), c_int64(maxlen), c_int64(total), _ptr(out))
    return out
