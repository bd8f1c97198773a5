"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each check comes from outside the oracle's own formulas: the known-answer
values written in the reference sources, numpy / scipy library routines,
closed forms and brute force.  A plausible slip in the restatement (wrong
direction, seed on the wrong side, off-by-one offsets, mirrored boundary)
fails at least one of them.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle_py as O

ADD, MUL, MIN, MAX, AND, OR, XOR, SUB, EXCHANGE = range(9)
L2R, R2L = 0, 1
CLAMP, MIRROR, WRAP, CONSTANT, FUNCTION = range(5)

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_examples.json")))


# ------------------------------------------------ reference known answers
def test_runq_known_answer():
    g = GOLDEN["runq_sum_map_plus1"]
    x = np.arange(1, 10001, dtype=np.int64) + 1  # map (+1) [1..10000]
    assert O.fold(ADD, 0, x)[0] == g["expect"]


def test_scanl_doc_example():
    g = GOLDEN["scanl_plus_10"]
    out = O.scan(ADD, L2R, g["seed"], np.array(g["input"], np.int64))
    assert out.tolist() == g["expect"]


def test_scanr_prime_code_example():
    g = GOLDEN["scanr_prime_plus_10"]
    out, s = O.scan_prime(ADD, R2L, g["seed"], np.array(g["input"], np.int64))
    assert out.tolist() == g["expect"] and s[0] == g["expect_sum"]


def test_scanl_prime_doc_example():
    g = GOLDEN["scanl_prime_plus_10"]
    out, s = O.scan_prime(ADD, L2R, g["seed"], np.array(g["input"], np.int64))
    assert out.tolist() == g["expect"] and s[0] == g["expect_sum"]


# ------------------------------------------------------- fold / scan vs numpy
@pytest.mark.parametrize("dt", [np.int32, np.int64, np.uint32, np.uint64])
def test_fold_add_wraps_like_numpy(dt):
    rng = np.random.default_rng(1)
    info = np.iinfo(dt)
    x = rng.integers(info.min // 2, info.max // 2, size=5000, dtype=dt, endpoint=True)
    assert O.fold(ADD, None, x)[0] == np.sum(x, dtype=dt)
    assert O.fold(ADD, dt(7), x)[0] == dt(np.sum(x, dtype=dt) + dt(7))
    assert O.fold(MAX, None, x)[0] == x.max()
    assert O.fold(MIN, None, x)[0] == x.min()
    assert O.fold(XOR, None, x)[0] == np.bitwise_xor.reduce(x)
    assert O.fold(AND, None, x)[0] == np.bitwise_and.reduce(x)
    assert O.fold(OR, None, x)[0] == np.bitwise_or.reduce(x)


@pytest.mark.parametrize("dt", [np.int32, np.int64, np.uint32, np.uint64])
def test_fold_mul_wraps_like_numpy(dt):
    # odd factors: their wrapped product is odd, so it cannot collapse to 0
    # (random factors make the product 0 mod 2^k after a few dozen terms)
    rng = np.random.default_rng(2)
    info = np.iinfo(dt)
    x = (rng.integers(info.min // 2, info.max // 2, size=4099, dtype=dt, endpoint=True) | dt(1)).astype(dt)
    want = np.multiply.reduce(x, dtype=dt)
    assert want % 2 == 1
    assert O.fold(MUL, None, x)[0] == want
    assert O.fold(MUL, dt(3), x)[0] == np.multiply.reduce(np.concatenate([np.array([3], dt), x]), dtype=dt)


def test_fold_rows_and_empty():
    x = np.arange(12, dtype=np.int32).reshape(3, 4)
    assert O.fold(ADD, 100, x, outer=3).tolist() == [106, 122, 138]
    assert O.fold(ADD, 5, np.zeros(0, np.int32), outer=1).tolist() == [5]
    with pytest.raises(ValueError):
        O.fold(ADD, None, np.zeros(0, np.int32), outer=1)


def test_fold1_closed_forms():
    n = 1 << 16
    assert O.fold(ADD, None, np.arange(1, n + 1, dtype=np.int64))[0] == n * (n + 1) // 2
    alt = np.where(np.arange(n) % 2 == 0, 1, -1).astype(np.int32)
    assert O.fold(ADD, None, alt)[0] == 0
    assert O.fold(ADD, None, np.full(n, 3, np.int32))[0] == 3 * n
    # wrap-around: 2^31 ones in int32 arithmetic would overflow; use (2^31-1) + 1
    assert O.fold(ADD, None, np.array([2**31 - 1, 1], np.int32))[0] == -(2**31)


def test_fold_float_minmax_nan_semantics():
    # fmaxf/fminf drop a NaN operand (C99 7.12.12)
    x = np.array([np.nan, 1.0, np.nan, -3.0], np.float32)
    assert O.fold(MAX, None, x)[0] == 1.0
    assert O.fold(MIN, None, x)[0] == -3.0
    assert np.isnan(O.fold(MAX, None, np.array([np.nan], np.float32))[0])


@pytest.mark.parametrize("dt", [np.int32, np.int64, np.uint64])
def test_scan_matches_cumsum(dt):
    rng = np.random.default_rng(2)
    info = np.iinfo(dt)
    x = rng.integers(info.min // 4, info.max // 4, size=3001, dtype=dt, endpoint=True)
    c = np.cumsum(x, dtype=dt)
    assert np.array_equal(O.scan(ADD, L2R, None, x), c)                       # scanl1
    assert np.array_equal(O.scan(ADD, L2R, dt(0), x), np.concatenate([np.zeros(1, dt), c]))  # scanl
    rc = np.cumsum(x[::-1], dtype=dt)[::-1]
    assert np.array_equal(O.scan(ADD, R2L, None, x), rc)                      # scanr1
    assert np.array_equal(O.scan(ADD, R2L, dt(0), x), np.concatenate([rc, np.zeros(1, dt)]))  # scanr
    e, s = O.scan_prime(ADD, L2R, dt(0), x)
    assert np.array_equal(e, np.concatenate([np.zeros(1, dt), c[:-1]])) and s[0] == c[-1]
    e, s = O.scan_prime(ADD, R2L, dt(0), x)
    assert np.array_equal(e, np.concatenate([rc[1:], np.zeros(1, dt)])) and s[0] == rc[0]


def test_scan_invariants_max():
    rng = np.random.default_rng(3)
    x = rng.integers(-1000, 1000, size=777, dtype=np.int32)
    s1 = O.scan(MAX, L2R, None, x)
    assert np.array_equal(s1, np.maximum.accumulate(x))
    assert s1[-1] == O.fold(MAX, None, x)[0]  # last (scanl1 f xs) == fold1 f xs
    # scanl z xs == scanl1 on (z : xs)
    z = np.int32(-5000)
    assert np.array_equal(O.scan(MAX, L2R, z, x), O.scan(MAX, L2R, None, np.concatenate([[z], x]).astype(np.int32)))


def test_scan_rows():
    x = np.arange(12, dtype=np.int32).reshape(3, 4)
    out = O.scan(ADD, L2R, None, x, outer=3).reshape(3, 4)
    assert np.array_equal(out, np.cumsum(x, axis=1, dtype=np.int32))
    out = O.scan(ADD, L2R, np.int32(1), x, outer=3).reshape(3, 5)
    assert out[:, 0].tolist() == [1, 1, 1] and np.array_equal(out[:, 1:], 1 + np.cumsum(x, axis=1))


def test_scan_empty():
    assert O.scan(ADD, L2R, np.int32(4), np.zeros(0, np.int32)).tolist() == [4]
    assert O.scan(ADD, L2R, None, np.zeros(0, np.int32)).tolist() == []
    e, s = O.scan_prime(ADD, R2L, np.int32(9), np.zeros(0, np.int32))
    assert e.tolist() == [] and s.tolist() == [9]


# ------------------------------------------------------------------ foldSeg
def test_fold_seg_matches_reduceat():
    rng = np.random.default_rng(4)
    lens = rng.integers(1, 50, size=300)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    x = rng.integers(-100, 100, size=int(off[-1]), dtype=np.int32)
    assert np.array_equal(O.fold_seg(ADD, None, x, off), np.add.reduceat(x, off[:-1]).astype(np.int32))
    assert np.array_equal(O.fold_seg(MAX, None, x, off), np.maximum.reduceat(x, off[:-1]))
    assert np.array_equal(O.fold_seg(ADD, np.int32(10), x, off), 10 + np.add.reduceat(x, off[:-1]))


def test_fold_seg_empty_segments_give_seed():
    x = np.arange(10, dtype=np.int64)
    off = np.array([0, 3, 3, 7, 10, 10], np.int64)
    assert O.fold_seg(ADD, np.int64(-1), x, off).tolist() == [2, -1, 17, 23, -1]


def test_fold_seg_rows():
    x = np.arange(20, dtype=np.int32).reshape(2, 10)
    off = np.array([0, 4, 10], np.int64)
    assert O.fold_seg(ADD, None, x, off, outer=2).tolist() == [6, 39, 46, 99]


def test_fp64_truth():
    rng = np.random.default_rng(5)
    x = rng.standard_normal(20000).astype(np.float32)
    v, a = O.fold_f64(x)
    assert v == pytest.approx(math.fsum(x.astype(np.float64)), abs=1e-12)
    assert a == pytest.approx(math.fsum(np.abs(x.astype(np.float64))), rel=1e-15)
    off = np.array([0, 5, 5, 19999, 20000], np.int64)
    sv, sa = O.fold_seg_f64(x, off, seed=np.float32(2.0))
    for s in range(4):
        assert sv[s] == pytest.approx(2.0 + math.fsum(x[off[s]:off[s + 1]].astype(np.float64)), abs=1e-12)
    c, _ = O.scan_f64(x[:1000])
    assert c[999] == pytest.approx(math.fsum(x[:1000].astype(np.float64)), abs=1e-12)


# ------------------------------------------------------------------ permute
def test_histogram_is_bincount():
    idx = O.gen_zipf_i32(42, 1.1, 4096, 200000)
    bins = O.permute(ADD, np.zeros(4096, np.int32), None, 1, idx)
    assert np.array_equal(bins, np.bincount(idx, minlength=4096).astype(np.int32))
    assert bins.sum() == idx.size


@pytest.mark.parametrize("dt", [np.int32, np.int64, np.float64])
def test_permute_ops_brute_force(dt):
    rng = np.random.default_rng(6)
    m, n = 97, 5000
    idx = rng.integers(-5, m + 5, size=n)  # out-of-range indices are Nothing
    src = rng.integers(-50, 50, size=n).astype(dt)
    d = rng.integers(-50, 50, size=m).astype(dt)
    ok = (idx >= 0) & (idx < m)
    for op, ufunc in ((ADD, np.add), (MIN, np.minimum), (MAX, np.maximum)):
        want = d.copy()
        ufunc.at(want, idx[ok], src[ok])
        assert np.array_equal(O.permute(op, d, src, None, idx), want)
    want = d.copy()
    np.subtract.at(want, idx[ok], src[ok])
    assert np.array_equal(O.permute(SUB, d, src, None, idx), want)


def test_permute_exchange_injective():
    idx = np.array([3, 0, 2, -1], np.int64)
    out = O.permute(EXCHANGE, np.full(4, 9, np.int32), np.array([10, 11, 12, 13], np.int32), None, idx)
    assert out.tolist() == [11, 9, 12, 10]


# ------------------------------------------------------------------ stencil
def _scipy_correlate(img, w, mode, cval=0.0):
    ndimage = pytest.importorskip("scipy.ndimage")
    # Accelerate boundaries as scipy modes: clamp = nearest (aaa|abc|ccc),
    # mirror = mirror (cb|abc|ba, edge not repeated), wrap = wrap
    m = {CLAMP: "nearest", MIRROR: "mirror", WRAP: "wrap", CONSTANT: "constant"}[mode]
    return ndimage.correlate(img.astype(np.float64), w.astype(np.float64), mode=m, cval=cval)


@pytest.mark.parametrize("mode", [CLAMP, MIRROR, WRAP, CONSTANT])
@pytest.mark.parametrize("r", [(1, 1), (2, 2), (1, 3)])
def test_stencil_matches_scipy(mode, r):
    ry, rx = r
    rng = np.random.default_rng(7)
    img = rng.standard_normal((37, 53))
    w = rng.standard_normal((2 * ry + 1, 2 * rx + 1))
    out = O.stencil2d(ry, rx, w, mode, 0.5, img)
    np.testing.assert_allclose(out, _scipy_correlate(img, w, mode, 0.5), rtol=1e-12, atol=1e-12)
    v, _ = O.stencil2d_f64(ry, rx, w.astype(np.float32), mode, 0.5, img.astype(np.float32))
    np.testing.assert_allclose(v, _scipy_correlate(img.astype(np.float32), w.astype(np.float32), mode, 0.5),
                               rtol=1e-9, atol=1e-9)


def test_stencil_special_cases():
    w5 = np.outer([1, 4, 6, 4, 1], [1, 4, 6, 4, 1]).astype(np.float64) / 256
    c = O.stencil2d(2, 2, w5, CLAMP, None, np.full((16, 16), 3.0))
    assert np.allclose(c, 3.0 * w5.sum())
    imp = np.zeros((9, 9))
    imp[4, 4] = 1.0
    w = np.arange(25, dtype=np.float64).reshape(5, 5)
    out = O.stencil2d(2, 2, w, CONSTANT, 0.0, imp)
    # correlation of an impulse gives the weights mirrored about the centre
    assert np.array_equal(out[2:7, 2:7], w[::-1, ::-1])


def test_boundary_maps_by_hand():
    # 1-D image [0,1,2,3,4] with a 3-point stencil picking only the left tap
    img = np.arange(5, dtype=np.float64).reshape(1, 5)
    wl = np.zeros((3, 3))
    wl[1, 0] = 1.0  # x - 1
    wr = np.zeros((3, 3))
    wr[1, 2] = 1.0  # x + 1
    assert O.stencil2d(1, 1, wl, CLAMP, None, img)[0, 0] == 0.0
    assert O.stencil2d(1, 1, wl, MIRROR, None, img)[0, 0] == 1.0   # -1 -> 1
    assert O.stencil2d(1, 1, wl, WRAP, None, img)[0, 0] == 4.0     # -1 -> n-1
    assert O.stencil2d(1, 1, wl, CONSTANT, 7.0, img)[0, 0] == 7.0
    assert O.stencil2d(1, 1, wr, CLAMP, None, img)[0, 4] == 4.0
    assert O.stencil2d(1, 1, wr, MIRROR, None, img)[0, 4] == 3.0   # n -> n-2
    assert O.stencil2d(1, 1, wr, WRAP, None, img)[0, 4] == 0.0


def test_integer_stencil_exact():
    rng = np.random.default_rng(8)
    img = rng.integers(-1000, 1000, size=(20, 30)).astype(np.int32)
    w = np.outer([1, 2, 1], [1, 2, 1]).astype(np.int32)
    out = O.stencil2d(1, 1, w, MIRROR, None, img)
    want = _scipy_correlate(img, w, MIRROR).astype(np.int64)
    assert np.array_equal(out.astype(np.int64), want)


def test_function_boundary_by_hand_and_reductions():
    # an index outside reads (c0 + cy*y) + cx*x of the index itself
    img = np.arange(5, dtype=np.float64).reshape(1, 5)
    c = (10.0, 2.0, 3.0)
    left = np.zeros((3, 3))
    left[1, 0] = 1.0   # (y, x - 1)
    up = np.zeros((3, 3))
    up[0, 1] = 1.0     # (y - 1, x)
    corner = np.zeros((3, 3))
    corner[0, 0] = 1.0  # (y - 1, x - 1)
    assert O.stencil2d(1, 1, left, FUNCTION, c, img)[0, 0] == 10.0 + 3.0 * -1   # (0, -1)
    assert O.stencil2d(1, 1, left, FUNCTION, c, img)[0, 3] == 2.0               # inside: x[0, 2]
    assert O.stencil2d(1, 1, up, FUNCTION, c, img)[0, 2] == 10.0 + 2.0 * -1 + 3.0 * 2
    assert O.stencil2d(1, 1, corner, FUNCTION, c, img)[0, 0] == 10.0 - 2.0 - 3.0
    # cy = cx = 0 is the constant boundary
    rng = np.random.default_rng(4)
    im = rng.integers(-1000, 1000, size=(19, 23)).astype(np.int32)
    w = rng.integers(-3, 4, size=(5, 3)).astype(np.int32)
    assert np.array_equal(O.stencil2d(2, 1, w, FUNCTION, (7, 0, 0), im), O.stencil2d(2, 1, w, CONSTANT, 7, im))


def test_function_boundary_against_padded_correlate():
    # the image padded with f at the outside indices, correlated by scipy
    ndimage = pytest.importorskip("scipy.ndimage")
    rng = np.random.default_rng(12)
    ry, rx, h, w = 2, 3, 17, 29
    im = rng.integers(-1000, 1000, size=(h, w)).astype(np.int64)
    wt = rng.integers(-3, 4, size=(2 * ry + 1, 2 * rx + 1)).astype(np.int64)
    c0, cy, cx = 5, -3, 7
    ys = np.arange(-ry, h + ry)[:, None]
    xs = np.arange(-rx, w + rx)[None, :]
    pad = (c0 + cy * ys + cx * xs).astype(np.float64)
    pad[ry:ry + h, rx:rx + w] = im
    want = ndimage.correlate(pad, wt.astype(np.float64), mode="constant")[ry:ry + h, rx:rx + w]
    got = O.stencil2d(ry, rx, wt, FUNCTION, (c0, cy, cx), im)
    assert np.array_equal(got.astype(np.float64), want)


def test_function_boundary_integer_wraps():
    # Int32: (c0 + cy*y) + cx*x wraps modulo 2^32
    img = np.zeros((2, 2), np.int32)
    up_left = np.zeros((3, 3), np.int32)
    up_left[0, 0] = 1  # (y - 1, x - 1) = (-1, -1) at the origin
    c = (2**31 - 5, 2**31 - 1, -(2**31))
    got = int(O.stencil2d(1, 1, up_left, FUNCTION, c, img)[0, 0])
    want = (c[0] + c[1] * -1 + c[2] * -1) % 2**32
    assert got % 2**32 == want


def test_stencil2_against_scipy_and_linearity():
    # stencil2 (two inputs, own boundaries, intersection of the shapes) against
    # scipy's correlate of each input, and its reductions: zero weights on one
    # side give the other side's stencil1; one array twice (same boundary)
    # gives stencil1 with the summed weights
    rng = np.random.default_rng(21)
    a = rng.integers(-1000, 1000, size=(23, 31)).astype(np.int32)
    b = rng.integers(-1000, 1000, size=(27, 29)).astype(np.int32)
    w1 = rng.integers(-3, 4, size=(3, 5)).astype(np.int32)
    w2 = rng.integers(-3, 4, size=(3, 5)).astype(np.int32)
    got = O.stencil2_2d(1, 2, w1, CLAMP, None, a, w2, WRAP, None, b)
    assert got.shape == (23, 29)
    want = (_scipy_correlate(a, w1, CLAMP)[:23, :29].astype(np.int64) +
            _scipy_correlate(b, w2, WRAP)[:23, :29].astype(np.int64))
    assert np.array_equal(got.astype(np.int64), want)
    z = np.zeros_like(w2)
    assert np.array_equal(O.stencil2_2d(1, 2, w1, MIRROR, None, a, z, CLAMP, None, b),
                          O.stencil2d(1, 2, w1, MIRROR, None, a)[:23, :29])
    assert np.array_equal(O.stencil2_2d(1, 2, w1, MIRROR, None, a, w2, MIRROR, None, a),
                          O.stencil2d(1, 2, w1 + w2, MIRROR, None, a))


def test_stencil2_each_input_its_own_extents():
    # a is 2 x 3, b is 3 x 2: the output is 2 x 2.  Below the output's last
    # row, b still has a row of its own (no boundary applies there); a is
    # clamped at its own last row
    a = np.arange(6, dtype=np.float64).reshape(2, 3)       # [[0,1,2],[3,4,5]]
    b = 10 * np.arange(6, dtype=np.float64).reshape(3, 2)  # [[0,10],[20,30],[40,50]]
    down = np.zeros((3, 3))
    down[2, 1] = 1.0  # the tap at y + 1
    z = np.zeros((3, 3))
    assert np.array_equal(O.stencil2_2d(1, 1, z, CLAMP, None, a, down, CLAMP, None, b),
                          np.array([[20.0, 30.0], [40.0, 50.0]]))
    assert np.array_equal(O.stencil2_2d(1, 1, down, CLAMP, None, a, z, CLAMP, None, b),
                          np.array([[3.0, 4.0], [3.0, 4.0]]))
    assert np.array_equal(O.stencil2_2d(1, 1, down, CONSTANT, -1.0, a, down, CONSTANT, -2.0, b),
                          np.array([[23.0, 34.0], [39.0, 49.0]]))


# --------------------------------------------------------------- schedule
def test_divide_work1_pieces():
    for splits, minsize, n in ((7, 4096, 1 << 20), (3, 8192, 100000), (15, 4096, 5000), (0, 4096, 123)):
        p = O.divide_work1(splits, minsize, 0, n)
        assert p[0][0] == 0 and p[-1][1] == n
        assert all(a[1] == b[0] for a, b in zip(p, p[1:]))
        assert len(p) <= splits + 1
        assert all(u % minsize == 0 for u, _ in p)
    assert O.divide_work1(1, 4096, 0, 10000) == [(0, 8192), (8192, 10000)]


def test_native_schedule_matches_sequential():
    x = O.gen_uniform_i32(42, -1000, 1000, 300000)
    assert O.native_fold_add_i32(4, x) == O.fold(ADD, 0, x)[0]
    assert np.array_equal(O.native_scanl1_add_i32(4, x), O.scan(ADD, L2R, None, x))
    idx = O.gen_zipf_i32(1, 1.1, 4096, 100000)
    assert np.array_equal(O.native_histogram_i32(4, idx, 4096), np.bincount(idx, minlength=4096))
    lens = O.gen_segment_lengths(3, 500, 1000, 100000)
    off = np.concatenate([[0], np.cumsum(lens)])
    v = O.gen_normal_f32(9, 100000)
    assert np.array_equal(O.native_fold1_seg_f32(4, MAX, v, off), O.fold_seg(MAX, None, v, off))
    img = O.gen_normal_f32(5, 64 * 80).reshape(64, 80)
    w = np.outer([1, 2, 1], [1, 2, 1]).astype(np.float32) / 16
    np.testing.assert_allclose(O.native_stencil2d_f32(4, 1, 1, w, MIRROR, img), O.stencil2d(1, 1, w, MIRROR, None, img),
                               rtol=1e-5, atol=1e-6)


def test_generators():
    lens = O.gen_segment_lengths(42, 1 << 12, 16384, 1 << 22)
    assert lens.sum() == 1 << 22 and lens.min() >= 1
    z = O.gen_zipf_i32(42, 1.1, 4096, 1 << 20)
    p1 = 1.0 / sum(k ** -1.1 for k in range(1, 4097))
    assert abs((z == 0).mean() - p1) < 0.01
    u = O.gen_uniform_i32(42, -1000, 1000, 1 << 16)
    assert u.min() == -1000 and u.max() == 1000
    assert np.array_equal(u, O.gen_uniform_i32(42, -1000, 1000, 1 << 16))


def test_gen_segment_lengths_dense_stay_positive():
    for nseg, maxlen, total in ((3145745, 3, 3145745), (1000, 5, 1200), (7, 1000, 7)):
        lens = O.gen_segment_lengths(21, nseg, maxlen, total)
        assert lens.min() >= 1 and int(lens.sum()) == total


# ------------------------------------------------------------------ producers
def test_iota_closed_form_and_rounding():
    # Σ (a + i*b) for i < n = n*a + b*n(n-1)/2, taken mod 2^32 for Int32
    n, a, b = 100003, -7, 3
    y = O.generate(O.GEN_IOTA, n, np.int32, a, b)
    want = (n * a + b * n * (n - 1) // 2) % 2**32
    assert int(y.astype(np.int64).sum()) % 2**32 == want
    assert y[0] == a and y[-1] == a + (n - 1) * b
    # wrapping: 2^31 - 1 + 1 is minBound
    assert O.generate(O.GEN_IOTA, 2, np.int32, 2**31 - 1, 1)[1] == -2**31
    # Float32: i is converted (rounded) before the product: (float)(2^24 + 1) == 2^24
    f = O.generate(O.GEN_IOTA, 2**24 + 2, np.float32, 0.0, 1.0)
    assert f[2**24 + 1] == np.float32(2**24) and f[12345] == np.float32(12345)
    assert np.array_equal(O.generate(O.GEN_IOTA, 4, np.float32, 0.5, 0.25), np.float32([0.5, 0.75, 1.0, 1.25]))
    assert np.array_equal(O.generate(O.GEN_FILL, 3, np.float64, 2.5), [2.5, 2.5, 2.5])


def test_negate_abs_known_answers():
    # LLVM/CodeGen/Arithmetic.hs:69-86: wrapping negate; abs minBound == minBound
    i32 = np.array([-2**31, -5, 0, 7, 2**31 - 1], np.int32)
    assert np.array_equal(O.map_unop(O.UNOP_NEG, i32), np.array([-2**31, 5, 0, -7, -2**31 + 1], np.int32))
    assert np.array_equal(O.map_unop(O.UNOP_ABS, i32), np.array([-2**31, 5, 0, 7, 2**31 - 1], np.int32))
    u = np.array([0, 1, 2**64 - 1], np.uint64)
    assert np.array_equal(O.map_unop(O.UNOP_ABS, u), u)
    assert np.array_equal(O.map_unop(O.UNOP_NEG, u), np.array([0, 2**64 - 1, 1], np.uint64))
    f = O.map_unop(O.UNOP_ABS, np.array([-0.0, -1.5, np.inf, -np.inf], np.float32))
    assert np.array_equal(f, [0.0, 1.5, np.inf, np.inf]) and not np.signbit(f[0])
    assert np.signbit(O.map_unop(O.UNOP_NEG, np.array([0.0], np.float64))[0])


def test_affine_rounds_each_step():
    # 3 * float32(1/3) rounds to 1.0, so (3x) - 1 is 0; a fused multiply-add
    # would give 3 * 0.333333343... - 1 = 2.98e-8
    x = np.array([np.float32(1.0) / np.float32(3.0)], np.float32)
    assert O.map_unop(O.UNOP_AFFINE, x, a=3.0, b=-1.0)[0] == 0.0
    # integers wrap: 2 * 2^30 is minBound
    assert O.map_unop(O.UNOP_AFFINE, np.array([2**30], np.int32), a=2, b=0)[0] == -2**31
    assert O.map_unop(O.UNOP_AFFINE, np.array([5], np.int64), a=-3, b=4)[0] == -11


def test_conversion_known_answers():
    # C conversions: truncation toward zero, wrapping narrowing, nearest-even rounding
    assert np.array_equal(O.map_unop(O.UNOP_CAST, np.float32([2.7, -2.7, 0.5]), np.int32), [2, -2, 0])
    assert O.map_unop(O.UNOP_CAST, np.array([2**32 + 5], np.int64), np.int32)[0] == 5
    assert O.map_unop(O.UNOP_CAST, np.array([-1], np.int32), np.uint32)[0] == 2**32 - 1
    assert O.map_unop(O.UNOP_CAST, np.array([2**53 + 1], np.int64), np.float64)[0] == 2.0**53
    assert O.map_unop(O.UNOP_CAST, np.array([16777217], np.int32), np.float32)[0] == np.float32(16777216)
    assert O.map_unop(O.UNOP_CAST, np.array([-1], np.int32), np.int64)[0] == -1


def test_zipwith_known_answers():
    a = np.array([-2**31, 5, 7], np.int32)
    b = np.array([1, -3, 7], np.int32)
    assert np.array_equal(O.zipwith(7, a, b), [2**31 - 1, 8, 0])  # SUB wraps
    assert np.array_equal(O.zipwith(0, a, b), [-2**31 + 1, 2, 14])
    assert np.array_equal(O.zipwith(1, np.array([2**16], np.int32), np.array([2**16], np.int32)), [0])
    f = O.zipwith(3, np.float32([np.nan, 1.0]), np.float32([2.0, np.nan]))  # fmaxf: NaN loses
    assert np.array_equal(f, [2.0, 1.0])
    assert np.array_equal(O.zipwith(6, np.array([12], np.uint32), np.array([10], np.uint32)), [6])


def test_transform_involutions_and_gather():
    x = np.arange(60, dtype=np.int64) * 3 - 17
    # reverse and transpose are involutions; transpose of a 1-row array is itself
    r = O.transform(O.IX_REVERSE, O.UNOP_ID, x, 1, 60)
    assert np.array_equal(O.transform(O.IX_REVERSE, O.UNOP_ID, r, 1, 60), x) and r[0] == x[-1]
    t = O.transform(O.IX_TRANSPOSE, O.UNOP_ID, x, 6, 10)
    assert np.array_equal(O.transform(O.IX_TRANSPOSE, O.UNOP_ID, t, 10, 6), x)
    assert t[1] == x[10] and t[6] == x[1]
    assert np.array_equal(O.transform(O.IX_TRANSPOSE, O.UNOP_ID, x, 1, 60), x)
    # gather: np.take for in-range indices, 0 for the others; f applies after the gather
    idx = np.array([5, 0, 59, -1, 60, 7], np.int32)
    g = O.transform(O.IX_GATHER, O.UNOP_NEG, x, 1, 60, idx, 6)
    assert np.array_equal(g[[0, 1, 2, 5]], -np.take(x, [5, 0, 59, 7])) and g[3] == 0 and g[4] == 0
    perm = np.random.default_rng(1).permutation(60)
    assert np.array_equal(O.transform(O.IX_GATHER, O.UNOP_ID, x, 1, 60, perm, 60), np.take(x, perm))


# ----------------------------------------------------------------- 3-D stencil
@pytest.mark.parametrize("b,mode", [(0, "nearest"), (1, "mirror"), (2, "wrap"), (3, "constant")])
def test_stencil3d_matches_scipy(b, mode):
    from scipy import ndimage
    rng = np.random.default_rng(5)
    vol = rng.standard_normal((7, 9, 11))
    w = rng.standard_normal((3, 5, 3))
    got = O.stencil3d(1, 2, 1, w, b, 0.5, vol)
    want = ndimage.correlate(vol, w, mode=mode, cval=0.5)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)


def test_stencil3d_central_plane_is_the_2d_stencil():
    # weights only in the central plane: each output plane is the (already
    # pinned) 2-D stencil of its input plane; integers exact
    rng = np.random.default_rng(6)
    vol = rng.integers(-1000, 1000, size=(5, 13, 17)).astype(np.int32)
    w2 = rng.integers(-3, 4, size=(3, 5)).astype(np.int32)
    w3 = np.zeros((5, 3, 5), np.int32)
    w3[2] = w2
    for b in (0, 1, 2, 3):
        got = O.stencil3d(2, 1, 2, w3, b, 7, vol)
        want = np.stack([O.stencil2d(1, 2, w2, b, np.int32(7), vol[z]) for z in range(5)])
        assert np.array_equal(got, want), b


def test_stencil3d_impulse_gives_mirrored_weights():
    vol = np.zeros((9, 9, 9))
    vol[4, 4, 4] = 1.0
    w = np.arange(27, dtype=np.float64).reshape(3, 3, 3)
    got = O.stencil3d(1, 1, 1, w, 3, 0.0, vol)
    assert np.array_equal(got[3:6, 3:6, 3:6], w[::-1, ::-1, ::-1])
