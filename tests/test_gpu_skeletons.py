"""GPU parity of every skeleton entry point against the CPU oracle.

All calls go through the C ABI (ctypes -> libaccelerate_hip.so -> gfx950
kernels).  Integer results must be bit-exact.  Floating (+) results must be
within the normwise bound |y - s| <= 1e-5 * log2(n) * sum|x| of the FP64
truth (n = number of terms, at least 2); floating min/max are exact.
"""
import math

import numpy as np
import pytest

import acc_hip as A
import oracle_py as O

pytestmark = pytest.mark.gpu

ADD, MUL, MIN, MAX, AND, OR, XOR, SUB, EXCHANGE = range(9)
L2R, R2L = 0, 1
CLAMP, MIRROR, WRAP, CONSTANT, FUNCTION = range(5)
FP_REL = 1e-5


def dev(stream, a):
    return A.DeviceArray.from_numpy(stream, np.ascontiguousarray(a))


def rand(dt, n, seed):
    rng = np.random.default_rng(seed)
    if np.dtype(dt).kind == "f":
        return rng.standard_normal(n).astype(dt)
    info = np.iinfo(dt)
    lo, hi = max(info.min, -(1 << 40)), min(info.max, 1 << 40)
    return rng.integers(lo, hi, size=n, endpoint=True).astype(dt)


def fp_bound(nterms, abs_sum):
    return FP_REL * max(1.0, math.log2(max(nterms, 2))) * abs_sum


# -------------------------------------------------------------------- fold
FOLD_SIZES = [1, 2, 3, 5, 63, 64, 65, 1000, 4095, 4096, 4097, 65537, (1 << 20) + 3]


@pytest.mark.parametrize("n", FOLD_SIZES)
@pytest.mark.parametrize("op", [ADD, MUL, MIN, MAX, AND, OR, XOR])
def test_fold_int32(stream, n, op):
    x = rand(np.int32, n, n + op)
    d = dev(stream, x)
    o = A.DeviceArray(1, np.int32)
    for seed in (None, np.int32(17)):
        A.fold(stream, op, seed, d, 1, n, o)
        assert o.to_numpy(stream)[0] == O.fold(op, seed, x)[0], (n, op, seed)


@pytest.mark.parametrize("dt", [np.int64, np.uint32, np.uint64])
@pytest.mark.parametrize("op", [ADD, MIN, MAX, XOR])
@pytest.mark.parametrize("n", [7, 4097, 300001])
def test_fold_other_ints(stream, dt, op, n):
    x = rand(dt, n, 5)
    o = A.DeviceArray(1, dt)
    A.fold(stream, op, None, dev(stream, x), 1, n, o)
    assert o.to_numpy(stream)[0] == O.fold(op, None, x)[0]


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("n", [1, 33, 4096, 1000003])
def test_fold_float(stream, dt, n):
    x = rand(dt, n, 11)
    d = dev(stream, x)
    o = A.DeviceArray(1, dt)
    A.fold(stream, ADD, dt(0.5), d, 1, n, o)
    truth, ab = O.fold_f64(x.astype(np.float32) if dt == np.float32 else x, seed=None)
    got = float(o.to_numpy(stream)[0])
    assert abs(got - (0.5 + truth)) <= fp_bound(n + 1, ab + 0.5) + 1e-12
    for op in (MIN, MAX):
        A.fold(stream, op, None, d, 1, n, o)
        assert o.to_numpy(stream)[0] == O.fold(op, None, x)[0]


def test_fold_float_nan_propagation(stream):
    x = np.array([np.nan] * 5 + [1.0, 2.0] + [np.nan] * 100, np.float32)
    o = A.DeviceArray(1, np.float32)
    A.fold(stream, MAX, None, dev(stream, x), 1, x.size, o)
    assert o.to_numpy(stream)[0] == 2.0
    y = np.full(9, np.nan, np.float32)
    A.fold(stream, MAX, None, dev(stream, y), 1, y.size, o)
    assert np.isnan(o.to_numpy(stream)[0])


@pytest.mark.parametrize("offset", [1, 2, 3])
def test_fold_misaligned_view(stream, offset):
    x = rand(np.int32, 100000, 3)
    d = dev(stream, x)
    v = d.view(offset, x.size - offset - 5)
    o = A.DeviceArray(1, np.int32)
    A.fold(stream, ADD, None, v, 1, v.count, o)
    assert o.to_numpy(stream)[0] == O.fold(ADD, None, x[offset:offset + v.count])[0]


@pytest.mark.parametrize("shape", [(3, 1), (1000, 7), (64, 4096), (5, 100003), (20000, 3), (2, 1)])
@pytest.mark.parametrize("dt", [np.int32, np.int64])
def test_fold_rows(stream, shape, dt):
    outer, inner = shape
    x = rand(dt, outer * inner, outer)
    o = A.DeviceArray(outer, dt)
    A.fold(stream, ADD, dt(3), dev(stream, x), outer, inner, o)
    assert np.array_equal(o.to_numpy(stream), O.fold(ADD, dt(3), x, outer=outer))
    A.fold(stream, MAX, None, dev(stream, x), outer, inner, o)
    assert np.array_equal(o.to_numpy(stream), O.fold(MAX, None, x, outer=outer))


def test_fold_empty(stream):
    o = A.DeviceArray(4, np.int32)
    e = A.DeviceArray(0, np.int32)
    A.fold(stream, ADD, np.int32(9), e, 4, 0, o)
    assert o.to_numpy(stream).tolist() == [9, 9, 9, 9]
    with pytest.raises(A.EmptyArray):
        A.fold(stream, ADD, None, e, 4, 0, o)


def test_fold_known_answer_runq(stream):
    # sum (map (+1) [1..10000]) == 50015000 (accelerate-llvm-*/test/nofib/.../RunQ.hs:26-31)
    x = np.arange(2, 10002, dtype=np.int64)
    o = A.DeviceArray(1, np.int64)
    A.fold(stream, ADD, np.int64(0), dev(stream, x), 1, x.size, o)
    assert o.to_numpy(stream)[0] == 50015000


# -------------------------------------------------------------------- scan
# around the 16384-element Int32 tile (and the 8192-element 8-byte tile)
SCAN_SIZES = [1, 2, 7, 255, 4095, 4096, 4097, 12345, 16383, 16384, 16385, 32768 + 3, 65536 + 1, (1 << 20) + 17]


def gpu_scan(stream, op, direction, seed, x, outer=1):
    inner = x.size // outer
    on = inner + 1 if seed is not None else inner
    o = A.DeviceArray(outer * on, x.dtype)
    A.scan(stream, op, direction, seed, dev(stream, x), outer, inner, o)
    return o.to_numpy(stream)


def gpu_scan_prime(stream, op, direction, seed, x, outer=1):
    inner = x.size // outer
    o = A.DeviceArray(outer * inner, x.dtype)
    s = A.DeviceArray(outer, x.dtype)
    A.scan_prime(stream, op, direction, seed, dev(stream, x), outer, inner, o, s)
    return o.to_numpy(stream), s.to_numpy(stream)


@pytest.mark.parametrize("n", SCAN_SIZES)
@pytest.mark.parametrize("direction", [L2R, R2L])
def test_scan_int32_all_modes(stream, n, direction):
    x = rand(np.int32, n, n)
    for op in (ADD, MAX):
        assert np.array_equal(gpu_scan(stream, op, direction, None, x), O.scan(op, direction, None, x)), (op, "incl")
        z = np.int32(-12345)
        assert np.array_equal(gpu_scan(stream, op, direction, z, x), O.scan(op, direction, z, x)), (op, "excl")
        e, s = gpu_scan_prime(stream, op, direction, z, x)
        we, ws = O.scan_prime(op, direction, z, x)
        assert np.array_equal(e, we) and np.array_equal(s, ws), (op, "scan'")


@pytest.mark.parametrize("dt", [np.int64, np.uint32, np.uint64, np.float64])
@pytest.mark.parametrize("n", [5, 2049, 8191, 8192, 8193, 16385, 100001])
@pytest.mark.parametrize("direction", [L2R, R2L])
def test_scan_other_types(stream, dt, n, direction):
    x = rand(dt, n, 7)
    if np.dtype(dt).kind == "f":
        x = np.round(x * 64).astype(dt)  # small integers: exact in FP64 sums
    for op in (ADD, MIN):
        assert np.array_equal(gpu_scan(stream, op, direction, None, x), O.scan(op, direction, None, x))
        e, s = gpu_scan_prime(stream, op, direction, dt(1), x)
        we, ws = O.scan_prime(op, direction, dt(1), x)
        assert np.array_equal(e, we) and np.array_equal(s, ws)


@pytest.mark.parametrize("n", [100, 4096 * 3 + 5, 1 << 20])
def test_scan_float32_bound(stream, n):
    x = rand(np.float32, n, 21)
    y = gpu_scan(stream, ADD, L2R, None, x).astype(np.float64)
    truth, ab = O.scan_f64(x)
    k = np.arange(1, n + 1)
    bound = FP_REL * np.maximum(1.0, np.log2(np.maximum(k, 2))) * ab
    err = np.abs(y - truth)
    assert np.all(err <= bound + 1e-12), float((err / np.maximum(bound, 1e-30)).max())
    # max is exact
    assert np.array_equal(gpu_scan(stream, MAX, R2L, None, x), O.scan(MAX, R2L, None, x))


@pytest.mark.parametrize("offset", [1, 3])
def test_scan_misaligned(stream, offset):
    x = rand(np.int32, 50000, 4)
    d = dev(stream, x)
    v = d.view(offset, 40000)
    o = A.DeviceArray(40001, np.int32)
    A.scan(stream, ADD, L2R, np.int32(5), v, 1, 40000, o)
    assert np.array_equal(o.to_numpy(stream), O.scan(ADD, L2R, np.int32(5), x[offset:offset + 40000]))


@pytest.mark.parametrize("shape", [(3, 5000), (7, 1), (4, 4096)])
def test_scan_rows(stream, shape):
    outer, inner = shape
    x = rand(np.int32, outer * inner, 9)
    for direction in (L2R, R2L):
        assert np.array_equal(gpu_scan(stream, ADD, direction, None, x, outer), O.scan(ADD, direction, None, x, outer))
        assert np.array_equal(gpu_scan(stream, ADD, direction, np.int32(2), x, outer),
                              O.scan(ADD, direction, np.int32(2), x, outer))
        e, s = gpu_scan_prime(stream, ADD, direction, np.int32(2), x, outer)
        we, ws = O.scan_prime(ADD, direction, np.int32(2), x, outer)
        assert np.array_equal(e, we) and np.array_equal(s, ws)


ROW_SHAPES = [(1000, 3), (4096, 17), (300, 1), (9, 1000), (3, 100003), (2, 1 << 20), (70000, 2),
              (2048, 4), (999, 5), (777, 63), (513, 64), (300, 65), (33, 8191)]


@pytest.mark.parametrize("shape", ROW_SHAPES)
@pytest.mark.parametrize("direction", [L2R, R2L])
def test_scan_rows_all_modes(stream, shape, direction):
    # rank-2 scans along the innermost dimension (one fused launch below 2^20
    # columns, one scan per row above)
    outer, inner = shape
    x = rand(np.int32, outer * inner, outer + inner)
    for op in (ADD, MAX):
        assert np.array_equal(gpu_scan(stream, op, direction, None, x, outer),
                              O.scan(op, direction, None, x, outer)), (op, "incl")
        z = np.int32(-77)
        assert np.array_equal(gpu_scan(stream, op, direction, z, x, outer),
                              O.scan(op, direction, z, x, outer)), (op, "excl")
        e, sm = gpu_scan_prime(stream, op, direction, z, x, outer)
        we, ws = O.scan_prime(op, direction, z, x, outer)
        assert np.array_equal(e, we) and np.array_equal(sm, ws), (op, "scan'")


@pytest.mark.parametrize("dt", [np.int64, np.uint32, np.float64])
@pytest.mark.parametrize("shape", [(513, 31), (5, 40001)])
def test_scan_rows_other_types(stream, dt, shape):
    outer, inner = shape
    x = rand(dt, outer * inner, 11)
    for direction in (L2R, R2L):
        assert np.array_equal(gpu_scan(stream, MAX, direction, None, x, outer),
                              O.scan(MAX, direction, None, x, outer))
        if dt != np.float64:
            z = dt(3)
            assert np.array_equal(gpu_scan(stream, ADD, direction, z, x, outer),
                                  O.scan(ADD, direction, z, x, outer))


def test_scan_rows_float32_bound(stream):
    outer, inner = 2000, 777
    x = O.gen_normal_f32(5, outer * inner)
    got = gpu_scan(stream, ADD, L2R, None, x, outer).reshape(outer, inner).astype(np.float64)
    xr = x.reshape(outer, inner).astype(np.float64)
    truth = np.cumsum(xr, axis=1)
    ab = np.cumsum(np.abs(xr), axis=1)
    terms = np.arange(1, inner + 1)  # column c sums c + 1 terms
    bound = FP_REL * np.maximum(1.0, np.log2(np.maximum(terms, 2)))[None, :] * ab
    assert np.all(np.abs(got - truth) <= bound + 1e-30)


def test_scan_reference_examples(stream):
    x = np.arange(10, dtype=np.int64)
    assert gpu_scan(stream, ADD, L2R, np.int64(10), x).tolist() == [10, 10, 11, 13, 16, 20, 25, 31, 38, 46, 55]
    e, s = gpu_scan_prime(stream, ADD, R2L, np.int64(10), x)
    assert e.tolist() == [55, 54, 52, 49, 45, 40, 34, 27, 19, 10] and s.tolist() == [55]


def test_scan_empty(stream):
    e = A.DeviceArray(0, np.int32)
    o = A.DeviceArray(3, np.int32)
    A.scan(stream, ADD, L2R, np.int32(4), e, 3, 0, o)
    assert o.to_numpy(stream).tolist() == [4, 4, 4]
    s = A.DeviceArray(2, np.int32)
    A.scan_prime(stream, ADD, R2L, np.int32(6), e, 2, 0, A.DeviceArray(0, np.int32), s)
    assert s.to_numpy(stream).tolist() == [6, 6]


def test_scan_repeated_calls_reuse_workspace(stream):
    # the tile-status words are reused across launches via the epoch tag
    x = rand(np.int32, 300000, 99)
    want = O.scan(ADD, L2R, None, x)
    d = dev(stream, x)
    o = A.DeviceArray(x.size, np.int32)
    for _ in range(20):
        A.scan(stream, ADD, L2R, None, d, 1, x.size, o)
    assert np.array_equal(o.to_numpy(stream), want)


# ------------------------------------------------------------------ foldSeg
def make_offsets(nseg, maxlen, total, seed, itype=np.int64, empties=False):
    lens = O.gen_segment_lengths(seed, nseg, maxlen, total)
    if empties:
        lens[::7] = 0
    off = np.concatenate([[0], np.cumsum(lens)]).astype(itype)
    return off


@pytest.mark.parametrize("itype", [np.int32, np.int64])
@pytest.mark.parametrize("cfg", [(10, 100), (1000, 16), (3000, 4096), (200, 40000), (50, 1)])
def test_fold_seg_int(stream, itype, cfg):
    nseg, maxlen = cfg
    off = make_offsets(nseg, maxlen, nseg * max(1, maxlen // 3), 5, itype)
    n = int(off[-1])
    x = rand(np.int32, n, 6)
    d = dev(stream, x)
    do = dev(stream, off)
    o = A.DeviceArray(nseg, np.int32)
    for op in (ADD, MAX):
        A.fold_seg(stream, op, None, d, 1, n, do, nseg, o)
        assert np.array_equal(o.to_numpy(stream), O.fold_seg(op, None, x, off.astype(np.int64))), op
    A.fold_seg(stream, ADD, np.int32(100), d, 1, n, do, nseg, o)
    assert np.array_equal(o.to_numpy(stream), O.fold_seg(ADD, np.int32(100), x, off.astype(np.int64)))


def test_fold_seg_empty_segments(stream):
    off = make_offsets(500, 300, 50000, 8, empties=True)
    n = int(off[-1])
    x = rand(np.int64, n, 1)
    o = A.DeviceArray(500, np.int64)
    A.fold_seg(stream, ADD, np.int64(-7), dev(stream, x), 1, n, dev(stream, off), 500, o)
    assert np.array_equal(o.to_numpy(stream), O.fold_seg(ADD, np.int64(-7), x, off))


def test_fold_seg_one_huge_segment(stream):
    n = (1 << 22) + 5
    x = rand(np.int32, n, 2)
    off = np.array([0, n], np.int64)
    o = A.DeviceArray(1, np.int32)
    A.fold_seg(stream, ADD, None, dev(stream, x), 1, n, dev(stream, off), 1, o)
    assert o.to_numpy(stream)[0] == O.fold(ADD, None, x)[0]


def test_fold_seg_float32_bound_and_max(stream):
    off = make_offsets(1 << 12, 16384, 1 << 22, 42)
    n = int(off[-1])
    x = O.gen_normal_f32(42, n)
    d, do = dev(stream, x), dev(stream, off)
    o = A.DeviceArray(off.size - 1, np.float32)
    A.fold_seg(stream, ADD, None, d, 1, n, do, off.size - 1, o)
    truth, ab = O.fold_seg_f64(x, off)
    lens = np.diff(off)
    bound = FP_REL * np.maximum(1.0, np.log2(np.maximum(lens, 2))) * ab
    err = np.abs(o.to_numpy(stream).astype(np.float64) - truth)
    assert np.all(err <= bound), float((err / bound).max())
    A.fold_seg(stream, MAX, None, d, 1, n, do, off.size - 1, o)
    assert np.array_equal(o.to_numpy(stream), O.fold_seg(MAX, None, x, off))


def test_fold_seg_rows(stream):
    x = rand(np.int32, 4 * 1024, 3)
    off = np.array([0, 10, 10, 500, 1024], np.int64)
    o = A.DeviceArray(16, np.int32)
    A.fold_seg(stream, ADD, np.int32(0), dev(stream, x), 4, 1024, dev(stream, off), 4, o)
    assert np.array_equal(o.to_numpy(stream), O.fold_seg(ADD, np.int32(0), x, off, outer=4))


@pytest.mark.parametrize("offset", [1, 2, 3])
@pytest.mark.parametrize("dt", [np.int32, np.float32, np.int64])
def test_fold_seg_misaligned_view(stream, offset, dt):
    # the input starts off a 16-byte boundary: the tile grid is aligned below it
    off = make_offsets(300, 2000, 200000, 11)
    n = int(off[-1])
    x = rand(dt, n + offset, 12)
    d = dev(stream, x)
    o = A.DeviceArray(300, dt)
    A.fold_seg(stream, MAX, None, d.view(offset, n), 1, n, dev(stream, off), 300, o)
    assert np.array_equal(o.to_numpy(stream), O.fold_seg(MAX, None, x[offset:], off))
    if dt != np.float32:
        A.fold_seg(stream, ADD, dt(3), d.view(offset, n), 1, n, dev(stream, off), 300, o)
        assert np.array_equal(o.to_numpy(stream), O.fold_seg(ADD, dt(3), x[offset:], off))


def test_fold_seg_rows_unaligned(stream):
    # rows of 1001 Int32: every row after the first starts off a 16-byte boundary
    outer, inner = 5, 1001
    x = rand(np.int32, outer * inner, 13)
    off = np.array([0, 1, 1, 400, 999, 1001], np.int64)
    o = A.DeviceArray(outer * 5, np.int32)
    A.fold_seg(stream, ADD, np.int32(-1), dev(stream, x), outer, inner, dev(stream, off), 5, o)
    assert np.array_equal(o.to_numpy(stream), O.fold_seg(ADD, np.int32(-1), x, off, outer=outer))


@pytest.mark.parametrize("empties", [False, True])
@pytest.mark.parametrize("maxlen", [3, 40, 300, 70000])
@pytest.mark.parametrize("itype", [np.int32, np.int64])
def test_fold_seg_starts_per_tile(stream, maxlen, itype, empties):
    # ~8000, ~400, ~50 and < 1 segment starts per 8192-element Int64 tile: the
    # staged-from-the-offsets, handed-over and carry-across-tiles paths; with
    # distinct starts an end's segment is a count of heads, with empty
    # segments (repeated starts) a search over the starts in LDS or, past 512
    # starts in a tile, in the offsets
    total = 3 * (1 << 20) + 17
    nseg = max(2, total // max(1, maxlen // 2))
    off = make_offsets(nseg, maxlen, total, 21, itype, empties=empties)
    n = int(off[-1])
    x = rand(np.int64, n, 22)
    d, do = dev(stream, x), dev(stream, off)
    for op, seed in ((ADD, None), (MIN, np.int64(5)), (XOR, None)):
        # without a seed (fold1Seg) empty segments keep what the output held
        o = dev(stream, np.zeros(nseg, np.int64))
        A.fold_seg(stream, op, seed, d, 1, n, do, nseg, o)
        assert np.array_equal(o.to_numpy(stream), O.fold_seg(op, seed, x, off.astype(np.int64))), (op, maxlen)


@pytest.mark.parametrize("dt,n", [(np.int32, 1 << 26), (np.int64, (1 << 25) + 4099)])
def test_fold_seg_fixup_sizes(stream, dt, n):
    # 8192 Int32 tiles and 8194 Int64 tiles: fix-up workgroups of 256 tiles,
    # the last one full and one with two tiles
    lens = O.gen_segment_lengths(41, n // 1500, 3000, n)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    nseg = off.size - 1
    x = rand(dt, n, 42)
    d, do = dev(stream, x), dev(stream, off)
    o = A.DeviceArray(nseg, dt)
    for op, seed in ((ADD, None), (MAX, None), (ADD, dt(-3))):
        A.fold_seg(stream, op, seed, d, 1, n, do, nseg, o)
        assert np.array_equal(o.to_numpy(stream), O.fold_seg(op, seed, x, off)), (op, seed)


def test_fold_seg_float64(stream):
    off = make_offsets(2000, 5000, 3000000, 31)
    n = int(off[-1])
    x = rand(np.float64, n, 32)
    d, do = dev(stream, x), dev(stream, off)
    o = A.DeviceArray(2000, np.float64)
    A.fold_seg(stream, MAX, None, d, 1, n, do, 2000, o)
    assert np.array_equal(o.to_numpy(stream), O.fold_seg(MAX, None, x, off))
    A.fold_seg(stream, ADD, None, d, 1, n, do, 2000, o)
    truth = np.array([math.fsum(x[off[i]:off[i + 1]]) for i in range(2000)])
    ab = np.add.reduceat(np.abs(x), off[:-1])
    err = np.abs(o.to_numpy(stream) - truth)
    assert np.all(err <= 1e-12 * np.maximum(1.0, np.log2(np.maximum(np.diff(off), 2))) * ab + 1e-300)


@pytest.mark.parametrize("offset", [1, 3])
def test_fold_rows_misaligned_view(stream, offset):
    outer, inner = 7, 10003
    x = rand(np.int32, outer * inner + offset, 33)
    d = dev(stream, x)
    o = A.DeviceArray(outer, np.int32)
    A.fold(stream, ADD, np.int32(2), d.view(offset, outer * inner), outer, inner, o)
    want = O.fold(ADD, np.int32(2), x[offset:], outer=outer)
    assert np.array_equal(o.to_numpy(stream), want)


# ------------------------------------------------------------------ permute
def test_histogram_zipf(stream):
    n = 1 << 22
    idx = O.gen_zipf_i32(42, 1.1, 4096, n)
    bins = dev(stream, np.zeros(4096, np.int32))
    A.permute(stream, ADD, bins, bins, 4096, None, 1, dev(stream, idx), n)
    got = bins.to_numpy(stream)
    assert np.array_equal(got, np.bincount(idx, minlength=4096))
    assert got.sum() == n


@pytest.mark.parametrize("dt", [np.int32, np.int64, np.uint32, np.uint64])
@pytest.mark.parametrize("m", [1, 97, 4096, 100000])
def test_permute_int_ops(stream, dt, m):
    n = 200003
    rng = np.random.default_rng(m)
    idx = rng.integers(-3, m + 3, size=n).astype(np.int64)
    src = rand(dt, n, 4)
    defaults = rand(dt, m, 5)
    di, ds, dd = dev(stream, idx), dev(stream, src), dev(stream, defaults)
    for op in (ADD, SUB, MIN, MAX, AND, OR, XOR):
        out = A.DeviceArray(m, dt)
        A.permute(stream, op, dd, out, m, ds, None, di, n)  # clone
        assert np.array_equal(out.to_numpy(stream), O.permute(op, defaults, src, None, idx)), op
    assert np.array_equal(dd.to_numpy(stream), defaults)  # defaults untouched by a clone


def test_permute_in_place_int32_index(stream):
    m, n = 1000, 50000
    rng = np.random.default_rng(3)
    idx = rng.integers(0, m, size=n).astype(np.int32)
    src = rng.integers(-9, 9, size=n).astype(np.int32)
    d = dev(stream, np.zeros(m, np.int32))
    A.permute(stream, ADD, d, d, m, dev(stream, src), None, dev(stream, idx), n)
    assert np.array_equal(d.to_numpy(stream), O.permute(ADD, np.zeros(m, np.int32), src, None, idx))


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_permute_float(stream, dt):
    m, n = 4096, 300000
    rng = np.random.default_rng(8)
    idx = rng.integers(0, m, size=n).astype(np.int32)
    src = rng.integers(-64, 64, size=n).astype(dt) / 8  # exact in any order
    d0 = np.zeros(m, dt)
    for op in (ADD, MIN, MAX):
        out = A.DeviceArray(m, dt)
        A.permute(stream, op, dev(stream, d0), out, m, dev(stream, src), None, dev(stream, idx), n)
        assert np.array_equal(out.to_numpy(stream), O.permute(op, d0, src, None, idx)), op


def test_permute_exchange(stream):
    m = 10000
    perm = np.random.default_rng(1).permutation(m).astype(np.int64)
    src = np.arange(m, dtype=np.int32)
    d = np.full(m, -1, np.int32)
    out = A.DeviceArray(m, np.int32)
    A.permute(stream, EXCHANGE, dev(stream, d), out, m, dev(stream, src), None, dev(stream, perm), m)
    assert np.array_equal(out.to_numpy(stream), O.permute(EXCHANGE, d, src, None, perm))


# ------------------------------------------------------------------ stencil
W5 = np.outer([1, 4, 6, 4, 1], [1, 4, 6, 4, 1]).astype(np.float64) / 256
W3 = np.outer([1, 2, 1], [1, 2, 1]).astype(np.float64) / 16


@pytest.mark.parametrize("shape", [(1, 1), (3, 5), (64, 64), (100, 77), (257, 300), (1024, 1024)])
@pytest.mark.parametrize("case", [(2, 2, CLAMP), (1, 1, MIRROR), (1, 1, WRAP), (2, 2, CONSTANT), (1, 3, CLAMP),
                                  (4, 2, MIRROR)])
def test_stencil_float32(stream, shape, case):
    ry, rx, b = case
    h, w = shape
    if b == MIRROR and (h <= ry or w <= rx):
        pytest.skip("mirror needs extent > radius (the reference maps once)")
    img = O.gen_normal_f32(h * w + ry, h * w).reshape(h, w)
    wt = np.random.default_rng(ry * 10 + rx).standard_normal((2 * ry + 1, 2 * rx + 1)).astype(np.float32)
    out = A.DeviceArray(h * w, np.float32)
    A.stencil2d(stream, ry, rx, wt, b, 0.25, dev(stream, img), h, w, out)
    got = out.to_numpy(stream).reshape(h, w).astype(np.float64)
    truth, ab = O.stencil2d_f64(ry, rx, wt, b, 0.25, img)
    taps = (2 * ry + 1) * (2 * rx + 1)
    bound = taps * 2.0 ** -23 * ab + 1e-30
    assert np.all(np.abs(got - truth) <= bound)


@pytest.mark.parametrize("b", [CLAMP, MIRROR, WRAP, CONSTANT])
@pytest.mark.parametrize("dt", [np.int32, np.int64])
def test_stencil_int_exact(stream, b, dt):
    h, w = 130, 211
    img = np.random.default_rng(2).integers(-1000, 1000, size=(h, w)).astype(dt)
    wt = np.outer([1, 2, 1], [1, 2, 1]).astype(dt)
    out = A.DeviceArray(h * w, dt)
    A.stencil2d(stream, 1, 1, wt, b, dt(5), dev(stream, img), h, w, out)
    assert np.array_equal(out.to_numpy(stream).reshape(h, w), O.stencil2d(1, 1, wt, b, dt(5), img))


@pytest.mark.parametrize("dt", [np.int32, np.int64])
@pytest.mark.parametrize("r", [(1, 1), (2, 3), (4, 4), (3, 1)])
@pytest.mark.parametrize("shape", [(300, 512), (200, 516), (131, 515)])
def test_stencil_int_interior_tiles(stream, dt, r, shape):
    # widths with 16-byte rows take the vector-staged interior tiles (the
    # last shape's rows are not 16-byte aligned: element staging everywhere)
    ry, rx = r
    h, w = shape
    img = np.random.default_rng(h + w).integers(-1000, 1000, size=(h, w)).astype(dt)
    wt = np.random.default_rng(7).integers(-3, 4, size=(2 * ry + 1, 2 * rx + 1)).astype(dt)
    out = A.DeviceArray(h * w, dt)
    for b in (CLAMP, WRAP):
        A.stencil2d(stream, ry, rx, wt, b, None, dev(stream, img), h, w, out)
        assert np.array_equal(out.to_numpy(stream).reshape(h, w), O.stencil2d(ry, rx, wt, b, None, img)), b


@pytest.mark.parametrize("r", [(1, 1), (2, 2)])
def test_stencil_int_large_image(stream, r):
    # 2100 x 4096 Int32: 1922 interior tiles (XCD bands of 240-241 tiles) and a
    # frame of 190 border tiles; every boundary, calls back to back on one stream
    ry, rx = r
    h, w = 2100, 4096
    img = np.random.default_rng(h + ry).integers(-1000, 1000, size=(h, w)).astype(np.int32)
    wt = np.random.default_rng(9).integers(-3, 4, size=(2 * ry + 1, 2 * rx + 1)).astype(np.int32)
    d = dev(stream, img)
    out = A.DeviceArray(h * w, np.int32)
    for b in (CLAMP, MIRROR, WRAP, CONSTANT):
        A.stencil2d(stream, ry, rx, wt, b, np.int32(7), d, h, w, out)
        assert np.array_equal(out.to_numpy(stream).reshape(h, w), O.stencil2d(ry, rx, wt, b, np.int32(7), img)), b


# ---------------------------------------------------- function boundary
@pytest.mark.parametrize("dt", [np.int32, np.int64])
@pytest.mark.parametrize("r", [(1, 1), (2, 3), (4, 4)])
@pytest.mark.parametrize("shape", [(130, 211), (300, 512), (5, 7)])
def test_stencil_function_boundary_int(stream, shape, r, dt):
    # an index outside reads (c0 + cy*y) + cx*x; 512-wide rows take the
    # vector-staged interior tiles, the frame tiles evaluate f
    ry, rx = r
    h, w = shape
    rng = np.random.default_rng(h * 3 + ry)
    img = rng.integers(-1000, 1000, size=(h, w)).astype(dt)
    wt = rng.integers(-3, 4, size=(2 * ry + 1, 2 * rx + 1)).astype(dt)
    out = A.DeviceArray(h * w, dt)
    c = (5, -3, 7)
    A.stencil2d(stream, ry, rx, wt, FUNCTION, c, dev(stream, img), h, w, out)
    assert np.array_equal(out.to_numpy(stream).reshape(h, w), O.stencil2d(ry, rx, wt, FUNCTION, c, img))


@pytest.mark.parametrize("r", [(1, 1), (2, 2)])
@pytest.mark.parametrize("shape", [(131, 515), (256, 1024)])
def test_stencil_function_boundary_float32(stream, shape, r):
    ry, rx = r
    h, w = shape
    img = O.gen_normal_f32(h + w, h * w).reshape(h, w)
    wt = np.random.default_rng(ry).standard_normal((2 * ry + 1, 2 * rx + 1)).astype(np.float32)
    c = (0.5, 0.25, -0.125)
    out = A.DeviceArray(h * w, np.float32)
    A.stencil2d(stream, ry, rx, wt, FUNCTION, c, dev(stream, img), h, w, out)
    got = out.to_numpy(stream).reshape(h, w).astype(np.float64)
    truth, ab = O.stencil2d_f64(ry, rx, wt, FUNCTION, c, img)
    taps = (2 * ry + 1) * (2 * rx + 1)
    assert np.all(np.abs(got - truth) <= taps * 2.0 ** -23 * ab + 1e-30)


def test_stencil2_function_boundary_and_3d_refusal(stream):
    rng = np.random.default_rng(31)
    a = rng.integers(-1000, 1000, size=(70, 130)).astype(np.int32)
    b = rng.integers(-1000, 1000, size=(66, 140)).astype(np.int32)
    wa = rng.integers(-3, 4, size=(5, 3)).astype(np.int32)
    wb = rng.integers(-3, 4, size=(5, 3)).astype(np.int32)
    out = A.DeviceArray(66 * 130, np.int32)
    A.stencil2_2d(stream, 2, 1, wa, FUNCTION, (1, 2, 3), dev(stream, a), 70, 130, wb, FUNCTION, (-4, 5, -6),
                  dev(stream, b), 66, 140, out)
    want = O.stencil2_2d(2, 1, wa, FUNCTION, (1, 2, 3), a, wb, FUNCTION, (-4, 5, -6), b)
    assert np.array_equal(out.to_numpy(stream).reshape(66, 130), want)
    vol = A.DeviceArray(27, np.float32)
    with pytest.raises(A.AccError):
        A.stencil3d(stream, 1, 1, 1, np.ones(27, np.float32), FUNCTION, (1.0, 0.0, 0.0), vol, 3, 3, 3,
                    A.DeviceArray(27, np.float32))


# ------------------------------------------------------------- stencil2
S2_SHAPES = [((130, 211), (130, 211)), ((300, 512), (300, 512)), ((200, 516), (190, 640)),
             ((37, 70), (41, 66)), ((5, 3), (3, 7)), ((1, 1), (2, 2))]


@pytest.mark.parametrize("r", [(1, 1), (1, 2), (2, 1), (2, 2)])
@pytest.mark.parametrize("shapes", S2_SHAPES)
@pytest.mark.parametrize("dt", [np.int32, np.int64])
def test_stencil2_int_exact(stream, dt, shapes, r):
    # integer stencil2 is exact (wrapping): equal and unequal shapes, 16-byte
    # rows (vector-staged interior tiles) and others, every boundary pair
    ry, rx = r
    (h1, w1), (h2, w2) = shapes
    rng = np.random.default_rng(h1 * 7 + w2)
    a = rng.integers(-1000, 1000, size=(h1, w1)).astype(dt)
    b = rng.integers(-1000, 1000, size=(h2, w2)).astype(dt)
    wa = rng.integers(-3, 4, size=(2 * ry + 1, 2 * rx + 1)).astype(dt)
    wb = rng.integers(-3, 4, size=(2 * ry + 1, 2 * rx + 1)).astype(dt)
    h, w = min(h1, h2), min(w1, w2)
    da, db = dev(stream, a), dev(stream, b)
    out = A.DeviceArray(h * w, dt)
    for b1, b2 in ((CLAMP, WRAP), (MIRROR, CONSTANT), (CONSTANT, CLAMP), (WRAP, MIRROR)):
        if (b1 == MIRROR and (h1 <= ry or w1 <= rx)) or (b2 == MIRROR and (h2 <= ry or w2 <= rx)):
            continue  # mirror needs extent > radius (the reference maps once)
        A.stencil2_2d(stream, ry, rx, wa, b1, dt(5), da, h1, w1, wb, b2, dt(-9), db, h2, w2, out)
        want = O.stencil2_2d(ry, rx, wa, b1, dt(5), a, wb, b2, dt(-9), b)
        assert np.array_equal(out.to_numpy(stream).reshape(h, w), want), (b1, b2)


@pytest.mark.parametrize("r", [(1, 1), (2, 2)])
@pytest.mark.parametrize("shapes", [((256, 1024), (256, 1024)), ((131, 515), (140, 500))])
def test_stencil2_float32(stream, shapes, r):
    ry, rx = r
    (h1, w1), (h2, w2) = shapes
    a = O.gen_normal_f32(h1 + w1, h1 * w1).reshape(h1, w1)
    b = O.gen_normal_f32(h2 + w2 + 1, h2 * w2).reshape(h2, w2)
    rng = np.random.default_rng(ry * 10 + rx)
    wa = rng.standard_normal((2 * ry + 1, 2 * rx + 1)).astype(np.float32)
    wb = rng.standard_normal((2 * ry + 1, 2 * rx + 1)).astype(np.float32)
    h, w = min(h1, h2), min(w1, w2)
    out = A.DeviceArray(h * w, np.float32)
    A.stencil2_2d(stream, ry, rx, wa, CLAMP, None, dev(stream, a), h1, w1, wb, CONSTANT, 0.5, dev(stream, b), h2, w2,
                  out)
    got = out.to_numpy(stream).reshape(h, w).astype(np.float64)
    truth, ab = O.stencil2_2d_f64(ry, rx, wa, CLAMP, None, a, wb, CONSTANT, 0.5, b)
    taps = 2 * (2 * ry + 1) * (2 * rx + 1)
    assert np.all(np.abs(got - truth) <= taps * 2.0 ** -23 * ab + 1e-30)


def test_stencil2_large_and_errors(stream):
    h, w = 2048, 4096
    rng = np.random.default_rng(5)
    a = rng.integers(-1000, 1000, size=(h, w)).astype(np.int32)
    b = rng.integers(-1000, 1000, size=(h, w)).astype(np.int32)
    wa = np.outer([1, 2, 1], [1, 2, 1]).astype(np.int32)
    wb = np.outer([1, 4, 6, 4, 1], [1, 0, -1]).astype(np.int32).T.copy()  # 3 rows x 5 columns
    wb = wb.reshape(3, 5)
    wa5 = np.zeros((3, 5), np.int32)
    wa5[:, 1:4] = wa
    da, db = dev(stream, a), dev(stream, b)
    out = A.DeviceArray(h * w, np.int32)
    A.stencil2_2d(stream, 1, 2, wa5, MIRROR, None, da, h, w, wb, WRAP, None, db, h, w, out)
    assert np.array_equal(out.to_numpy(stream).reshape(h, w),
                          O.stencil2_2d(1, 2, wa5, MIRROR, None, a, wb, WRAP, None, b))
    with pytest.raises(A.AccError):  # radius 3
        A.stencil2_2d(stream, 3, 1, np.zeros(21, np.int32), CLAMP, None, da, h, w, np.zeros(21, np.int32), CLAMP,
                      None, db, h, w, out)
    with pytest.raises(A.AccError):  # the output aliases an input
        A.stencil2_2d(stream, 1, 1, wa, CLAMP, None, da, h, w, wa, CLAMP, None, db, h, w, da)
    with pytest.raises(A.AccError):  # constant boundary without a value
        A.stencil2_2d(stream, 1, 1, wa, CONSTANT, None, da, h, w, wa, CLAMP, None, db, h, w, out)


def test_stencil_float64_special(stream):
    h, w = 70, 90
    img = np.full((h, w), 3.0)
    out = A.DeviceArray(h * w, np.float64)
    A.stencil2d(stream, 2, 2, W5, CLAMP, None, dev(stream, img), h, w, out)
    assert np.allclose(out.to_numpy(stream), 3.0)
    imp = np.zeros((9, 9))
    imp[4, 4] = 1.0
    wt = np.arange(25, dtype=np.float64).reshape(5, 5)
    out = A.DeviceArray(81, np.float64)
    A.stencil2d(stream, 2, 2, wt, CONSTANT, 0.0, dev(stream, imp), 9, 9, out)
    assert np.array_equal(out.to_numpy(stream).reshape(9, 9)[2:7, 2:7], wt[::-1, ::-1])


# ------------------------------------------------------------------ 3-D stencil
# (6, 40, 300) and (5, 37, 520): rows of several blocks, 16-byte aligned
S3_SHAPES = [(1, 1, 1), (2, 3, 4), (5, 17, 70), (33, 40, 130), (70, 20, 64), (6, 40, 300), (5, 37, 520)]


@pytest.mark.parametrize("shape", S3_SHAPES)
@pytest.mark.parametrize("r", [(1, 1, 1), (2, 2, 2), (1, 2, 1), (2, 1, 2)])
@pytest.mark.parametrize("dt", [np.int32, np.int64])
def test_stencil3d_int_exact(stream, dt, r, shape):
    rz, ry, rx = r
    d, h, w = shape
    vol = np.random.default_rng(d * h + w).integers(-1000, 1000, size=shape).astype(dt)
    wt = np.random.default_rng(rz * 9 + ry * 3 + rx).integers(-3, 4, size=(2 * rz + 1, 2 * ry + 1, 2 * rx + 1))
    wt = wt.astype(dt)
    src = dev(stream, vol)
    out = A.DeviceArray(vol.size, dt)
    for b in (CLAMP, MIRROR, WRAP, CONSTANT):
        A.stencil3d(stream, rz, ry, rx, wt, b, dt(5), src, d, h, w, out)
        got = out.to_numpy(stream).reshape(shape)
        assert np.array_equal(got, O.stencil3d(rz, ry, rx, wt, b, dt(5), vol)), (b, shape)


@pytest.mark.parametrize("shape", [(9, 33, 65), (64, 64, 64), (9, 50, 520)])
@pytest.mark.parametrize("r", [(1, 1, 1), (2, 2, 2)])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_stencil3d_float_bound(stream, dt, r, shape):
    rz, ry, rx = r
    d, h, w = shape
    vol = O.gen_normal_f32(d + h + w, d * h * w).reshape(shape).astype(dt)
    wt = np.random.default_rng(3).standard_normal((2 * rz + 1, 2 * ry + 1, 2 * rx + 1)).astype(dt)
    out = A.DeviceArray(vol.size, dt)
    eps = 2.0 ** -23 if dt == np.float32 else 2.0 ** -52
    taps = wt.size
    for b in (CLAMP, MIRROR, WRAP, CONSTANT):
        A.stencil3d(stream, rz, ry, rx, wt, b, 0.25, dev(stream, vol), d, h, w, out)
        got = out.to_numpy(stream).reshape(shape).astype(np.float64)
        truth, ab = O.stencil3d_f64(rz, ry, rx, wt, b, 0.25, vol)
        assert np.all(np.abs(got - truth) <= taps * eps * ab + 1e-30), b


def test_stencil3d_large_volume(stream):
    # 256 x 256 x 384 Float32, 3x3x3 clamp: several z runs per column of
    # blocks, several blocks per row
    d, h, w = 256, 256, 384
    vol = O.gen_normal_f32(9, d * h * w).reshape(d, h, w)
    wt = np.full((3, 3, 3), 1.0 / 27.0, np.float32)
    out = A.DeviceArray(vol.size, np.float32)
    A.stencil3d(stream, 1, 1, 1, wt, CLAMP, None, dev(stream, vol), d, h, w, out)
    got = out.to_numpy(stream).reshape(d, h, w).astype(np.float64)
    truth, ab = O.stencil3d_f64(1, 1, 1, wt, CLAMP, None, vol)
    assert np.all(np.abs(got - truth) <= 27 * 2.0 ** -23 * ab + 1e-30)


# ------------------------------------------------------------------- runtime
@pytest.mark.parametrize("nbytes", [16, 4096, (1 << 20) + 16, (1 << 24) + 48, 17, 1000003])
@pytest.mark.parametrize("offset", [0, 4])
def test_device_copy(stream, nbytes, offset):
    # 16-byte aligned whole vectors take the streaming copy kernel; the rest
    # the runtime's copy
    src = np.random.default_rng(nbytes).integers(0, 256, nbytes + 16, dtype=np.uint8)
    d_src = dev(stream, src)
    d_dst = A.DeviceArray(nbytes + 16, np.uint8)
    A.check(A.lib().acc_hip_memcpy_d2d_async(stream.handle, A.c_void_p(d_dst.ptr + offset),
                                             A.c_void_p(d_src.ptr + offset), nbytes))
    got = d_dst.to_numpy(stream)
    assert np.array_equal(got[offset:offset + nbytes], src[offset:offset + nbytes])
