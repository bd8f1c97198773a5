"""GPU parity of the producers (map, zipWith, generate, transform and
backpermute) against the oracle's restatements, through the C ABI.

Every result is compared bit for bit (floating results by their bit
patterns): the element functions are exact or round once per step on both
sides.  NaNs appear only where a function passes them through unchanged
(id, negate, abs, min/max against a number), so NaN payloads never differ.
"""
import numpy as np
import pytest

import acc_hip as A
import oracle_py as O

pytestmark = pytest.mark.gpu

ADD, MUL, MIN, MAX, AND, OR, XOR, SUB = range(8)
TYPES = [np.int32, np.int64, np.uint32, np.uint64, np.float32, np.float64]
SIZES = [1, 5, 1027, (1 << 20) + 7]
AFF = {"i": (3, -7), "u": (3, 7), "f": (1.5, -0.25)}


def dev(stream, a):
    return A.DeviceArray.from_numpy(stream, np.ascontiguousarray(a))


def vals(dt, n, seed, specials=False):
    rng = np.random.default_rng(seed)
    dt = np.dtype(dt)
    if dt.kind == "f":
        x = (rng.standard_normal(n) * 100).astype(dt)
        if specials and n >= 8:
            x[:6] = [0.0, -0.0, np.inf, -np.inf, np.nan, -1.5]
        return x
    info = np.iinfo(dt)
    x = rng.integers(info.min, info.max, size=n, endpoint=True, dtype=dt)
    if n >= 4:
        x[:2] = [info.min, info.max]
    return x


def same(got, want):
    got, want = np.asarray(got), np.asarray(want)
    assert got.dtype == want.dtype and got.shape == want.shape
    if got.dtype.kind == "f":
        u = np.dtype(f"u{got.dtype.itemsize}")
        bad = np.flatnonzero(got.view(u) != want.view(u))
    else:
        bad = np.flatnonzero(got != want)
    assert bad.size == 0, f"{bad.size} differ, first at {bad[0]}: {got[bad[0]]!r} vs {want[bad[0]]!r}"


# ---------------------------------------------------------------------- map
@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("f", [A.UNOP_ID, A.UNOP_NEG, A.UNOP_ABS, A.UNOP_AFFINE])
@pytest.mark.parametrize("dt", TYPES)
def test_map(stream, dt, f, n):
    x = vals(dt, n, n + f, specials=f != A.UNOP_AFFINE)
    a, b = AFF[np.dtype(dt).kind] if f == A.UNOP_AFFINE else (None, None)
    o = A.DeviceArray(n, dt)
    A.map_(stream, f, dev(stream, x), n, o, a, b)
    same(o.to_numpy(stream), O.map_unop(f, x, a=a, b=b))


def _cast_source(src, dst, n):
    rng = np.random.default_rng(7)
    src, dst = np.dtype(src), np.dtype(dst)
    if src.kind == "f":
        # in range of every target; non-negative for unsigned targets
        lo = 0.0 if dst.kind == "u" else -1e6
        return rng.uniform(lo, 1e6, n).astype(src)
    return vals(src, n, 3)


@pytest.mark.parametrize("dst", TYPES)
@pytest.mark.parametrize("src", TYPES)
def test_map_conversion(stream, src, dst):
    n = 4099
    x = _cast_source(src, dst, n)
    o = A.DeviceArray(n, dst)
    A.map_(stream, A.UNOP_CAST, dev(stream, x), n, o)
    same(o.to_numpy(stream), O.map_unop(O.UNOP_CAST, x, dst))


def test_map_in_place_and_views(stream):
    n = 100003
    x = vals(np.int32, n + 3, 11)
    d = dev(stream, x)
    A.map_(stream, A.UNOP_NEG, d, n + 3, d)  # in == out
    same(d.to_numpy(stream), O.map_unop(O.UNOP_NEG, x))
    # unaligned input and output views (element-wise kernel)
    o = A.DeviceArray(n + 3, np.int32)
    A.map_(stream, A.UNOP_AFFINE, d.view(1, n), n, o.view(2, n), 5, 1)
    same(o.to_numpy(stream)[2:2 + n], O.map_unop(O.UNOP_AFFINE, O.map_unop(O.UNOP_NEG, x)[1:1 + n], a=5, b=1))
    # a conversion between element sizes on unaligned views
    f = A.DeviceArray(n + 1, np.float64)
    A.map_(stream, A.UNOP_CAST, d.view(1, n), n, f.view(1, n))
    same(f.to_numpy(stream)[1:], O.map_unop(O.UNOP_NEG, x)[1:1 + n].astype(np.float64))


# ------------------------------------------------------------------ zipWith
@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("op", [ADD, MUL, MIN, MAX, SUB, AND, OR, XOR])
@pytest.mark.parametrize("dt", TYPES)
def test_zipwith(stream, dt, op, n):
    if op in (AND, OR, XOR) and np.dtype(dt).kind == "f":
        pytest.skip("bitwise combines are for integral types")
    a = vals(dt, n, n, specials=op in (MIN, MAX))
    b = vals(dt, n, n + 1)
    o = A.DeviceArray(n, dt)
    A.zipwith(stream, op, dev(stream, a), dev(stream, b), n, o)
    same(o.to_numpy(stream), O.zipwith(op, a, b))


def test_zipwith_unaligned(stream):
    n = 5001
    a, b = vals(np.float32, n + 1, 1), vals(np.float32, n + 1, 2)
    da, db = dev(stream, a), dev(stream, b)
    o = A.DeviceArray(n, np.float32)
    A.zipwith(stream, SUB, da.view(1, n), db, n, o)
    same(o.to_numpy(stream), O.zipwith(SUB, a[1:], b[:n]))


# ----------------------------------------------------------------- generate
@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("dt", TYPES)
def test_generate(stream, dt, n):
    kind = np.dtype(dt).kind
    a, b = {"i": (-5, 3), "u": (5, 3), "f": (0.5, 0.125)}[kind]
    o = A.DeviceArray(n + 1, dt)
    for g in (A.GEN_FILL, A.GEN_IOTA):
        A.generate(stream, g, n, o, a, b)
        same(o.to_numpy(stream)[:n], O.generate(g, n, dt, a, b))
    A.generate(stream, A.GEN_IOTA, n, o.view(1, n), a, b)  # unaligned
    same(o.to_numpy(stream)[1:], O.generate(A.GEN_IOTA, n, dt, a, b))


def test_generate_float_index_rounding(stream):
    # indices past 2^24 round when converted to Float32
    n = (1 << 24) + 1000
    o = A.DeviceArray(n, np.float32)
    A.generate(stream, A.GEN_IOTA, n, o, 0.25, 3.0)
    same(o.to_numpy(stream), O.generate(O.GEN_IOTA, n, np.float32, 0.25, 3.0))


# --------------------------------------------------------- transform / backpermute
@pytest.mark.parametrize("f", [A.UNOP_ID, A.UNOP_NEG, A.UNOP_AFFINE])
@pytest.mark.parametrize("itype", [np.int32, np.int64])
@pytest.mark.parametrize("dt", TYPES)
def test_gather(stream, dt, itype, f):
    n_in, n = 70001, 100003
    x = vals(dt, n_in, 5)
    idx = np.random.default_rng(6).integers(-3, n_in + 3, n).astype(itype)  # a few out of range
    a, b = AFF[np.dtype(dt).kind] if f == A.UNOP_AFFINE else (None, None)
    o = A.DeviceArray(n, dt)
    A.transform(stream, A.IX_GATHER, f, dev(stream, x), 1, n_in, dev(stream, idx), n, o, a, b)
    same(o.to_numpy(stream), O.transform(O.IX_GATHER, f, x, 1, n_in, idx, n, a=a, b=b))
    # unaligned index view and output view (element-wise kernel)
    di = dev(stream, idx)
    o2 = A.DeviceArray(n, dt)
    A.transform(stream, A.IX_GATHER, f, dev(stream, x), 1, n_in, di.view(1, n - 1), n - 1, o2.view(1, n - 1), a, b)
    same(o2.to_numpy(stream)[1:], O.transform(O.IX_GATHER, f, x, 1, n_in, idx[1:], n - 1, a=a, b=b))


def test_gather_empty_source(stream):
    idx = np.arange(10, dtype=np.int32)
    o = dev(stream, np.full(10, 9, np.int64))
    A.transform(stream, A.IX_GATHER, A.UNOP_ID, None, 0, 0, dev(stream, idx), 10, o)
    same(o.to_numpy(stream), np.zeros(10, np.int64))


@pytest.mark.parametrize("n", [1, 4, 5, 1027, 4096, 1 << 20, (1 << 20) + 7])
@pytest.mark.parametrize("dt", TYPES)
def test_reverse(stream, dt, n):
    # n a multiple of the 16-byte chunk with aligned arrays: whole-vector reads
    x = vals(dt, n, n)
    o = A.DeviceArray(n, dt)
    A.transform(stream, A.IX_REVERSE, A.UNOP_ABS, dev(stream, x), 1, n, None, n, o)
    same(o.to_numpy(stream), O.transform(O.IX_REVERSE, O.UNOP_ABS, x, 1, n))
    if n >= 4:
        # the input one element in (not 16-byte aligned): element reads
        d = dev(stream, vals(dt, n + 1, n + 1))
        xs = d.to_numpy(stream)[1:]
        A.transform(stream, A.IX_REVERSE, A.UNOP_ID, d.view(1, n), 1, n, None, n, o)
        same(o.to_numpy(stream), O.transform(O.IX_REVERSE, O.UNOP_ID, xs, 1, n))


@pytest.mark.parametrize("shape", [(1, 1), (1, 100), (100, 1), (63, 65), (64, 64), (129, 257), (1000, 999),
                                   (128, 200), (100, 68), (192, 64), (256, 1024), (512, 256)])
@pytest.mark.parametrize("dt", [np.int32, np.float64, np.uint64, np.float32])
def test_transpose(stream, dt, shape):
    r, c = shape
    x = vals(dt, r * c, r + c)
    o = A.DeviceArray(r * c, dt)
    A.transform(stream, A.IX_TRANSPOSE, A.UNOP_ID, dev(stream, x), r, c, None, r * c, o)
    same(o.to_numpy(stream), O.transform(O.IX_TRANSPOSE, O.UNOP_ID, x, r, c))


def test_transpose_views_and_identity(stream):
    # a view one element in (rows no longer 16-byte aligned: element path)
    r, c = 128, 256
    x = vals(np.int32, r * c + 1, 3)
    d = dev(stream, x)
    o = A.DeviceArray(r * c, np.int32)
    A.transform(stream, A.IX_TRANSPOSE, A.UNOP_NEG, d.view(1, r * c), r, c, None, r * c, o)
    same(o.to_numpy(stream), O.transform(O.IX_TRANSPOSE, O.UNOP_NEG, x[1:], r, c))
    # transposing twice is the identity (vector path, 4096 x 2048 Float64)
    r, c = 4096, 2048
    y = vals(np.float64, r * c, 4)
    dy = dev(stream, y)
    t = A.DeviceArray(r * c, np.float64)
    A.transform(stream, A.IX_TRANSPOSE, A.UNOP_ID, dy, r, c, None, r * c, t)
    back = A.DeviceArray(r * c, np.float64)
    A.transform(stream, A.IX_TRANSPOSE, A.UNOP_ID, t, c, r, None, r * c, back)
    same(back.to_numpy(stream), y)
    same(t.to_numpy(stream), y.reshape(r, c).T.reshape(-1))


@pytest.mark.parametrize("esize", [1, 2, 4, 8])
def test_fill_async(stream, esize):
    # element heads up to a 16-byte boundary, 16-byte stores, element tails:
    # exactly [off, off + count) changes
    dt = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[esize]
    v = np.array([0xA5C3_5A3C_9616_6996 & ((1 << (8 * esize)) - 1)], dtype=dt)
    n = 5000
    for off, cnt in ((0, 0), (0, 1), (1, 3), (3, 17), (0, 4096), (5, 4000), (7, 4993), (2, 1)):
        d = dev(stream, np.zeros(n, dt))
        A.check(A.lib().acc_hip_fill_async(stream.handle, A.c_void_p(d.ptr + off * esize),
                                           v.ctypes.data_as(A.c_void_p), esize, cnt))
        want = np.zeros(n, dt)
        want[off:off + cnt] = v[0]
        same(d.to_numpy(stream), want)


def test_producer_errors(stream):
    d = A.DeviceArray(16, np.int32)
    f = A.DeviceArray(16, np.float32)
    with pytest.raises(A.AccError):
        A.map_(stream, A.UNOP_NEG, d, 16, f)  # only the conversion changes the type
    with pytest.raises(A.AccError):
        A.transform(stream, A.IX_REVERSE, A.UNOP_ID, d, 1, 16, None, 15, d)  # count must match
    with pytest.raises(A.AccError):
        A.zipwith(stream, AND, f, f, 16, f)  # bitwise combine on a floating type


def test_producers_full_size_properties(stream):
    # 2^28 Int32 through generate -> map -> reverse -> reverse, checked by
    # size-independent properties: the iota sum's closed form (mod 2^32) and
    # the double reverse's identity
    n = 1 << 28
    a, b = 11, 3
    x = A.DeviceArray(n, np.int32)
    y = A.DeviceArray(n, np.int32)
    A.generate(stream, A.GEN_IOTA, n, x, a, b)
    s = A.DeviceArray(1, np.int32)
    A.fold(stream, ADD, None, x, 1, n, s)
    want = (n * a + b * n * (n - 1) // 2) % 2**32
    assert int(s.to_numpy(stream)[0]) % 2**32 == want
    A.transform(stream, A.IX_REVERSE, A.UNOP_ID, x, 1, n, None, n, y)
    A.transform(stream, A.IX_REVERSE, A.UNOP_AFFINE, y, 1, n, None, n, x, 2, -a)  # 2x - a, reversed back
    ends = x.to_numpy(stream)
    last = (a + 2 * b * (n - 1) + 2**31) % 2**32 - 2**31  # Int32 wraps
    assert ends[0] == a and ends[1] == a + 2 * b and int(ends[-1]) == last
