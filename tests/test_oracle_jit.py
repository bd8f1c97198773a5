"""Pins the oracle of the run-time compiled skeletons (oracle/oracle_jit.py)
before it is trusted, and checks that every skeleton node translates and
compiles for gfx950 (hiprtc needs no GPU).  Runs on the CPU.

The pins come from outside the interpreter's own formulas: Python's
arbitrary-precision integers (wrapping taken modulo 2^k by hand), the
textbook definitions of truncating and flooring division, known answers of
the reference's semantics (abs minBound, ordered NaN comparisons, round half
away from zero, fmin/fmax with NaN), scipy's maximum filter for stencils,
and sequential Python loops that restate each skeleton's definition element
by element on small inputs.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle_jit as J

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_examples.json")))

# the 2x2 integer matrix product: associative, not commutative
MATMUL = ("(fun (a b) (tuple (add (mul (prj 0 a) (prj 0 b)) (mul (prj 1 a) (prj 2 b)))"
          " (add (mul (prj 0 a) (prj 1 b)) (mul (prj 1 a) (prj 3 b)))"
          " (add (mul (prj 2 a) (prj 0 b)) (mul (prj 3 a) (prj 2 b)))"
          " (add (mul (prj 2 a) (prj 1 b)) (mul (prj 3 a) (prj 3 b)))))")


def wrap(v, bits, signed):
    v %= 1 << bits
    return v - (1 << bits) if signed and v >= 1 << (bits - 1) else v


def ints(dt, n, seed, lo=None, hi=None):
    info = np.iinfo(dt)
    rng = np.random.default_rng(seed)
    lo = info.min if lo is None else lo
    hi = info.max if hi is None else hi
    return rng.integers(lo, hi, size=n, endpoint=True, dtype=dt)


# ------------------------------------------------------------ primitives
@pytest.mark.parametrize("dt", [np.int8, np.int16, np.int32, np.int64, np.uint8, np.uint32, np.uint64])
def test_wrapping_arithmetic_against_python_integers(dt):
    t = np.dtype(dt)
    name = ("i" if t.kind == "i" else "u") + str(t.itemsize * 8)
    x, y = ints(dt, 300, 1), ints(dt, 300, 2)
    bits, signed = t.itemsize * 8, t.kind == "i"
    for op, f in (("add", lambda a, b: a + b), ("sub", lambda a, b: a - b), ("mul", lambda a, b: a * b)):
        got = J.evaluate(f"({op} x y)", {"x": x, "y": y})[0]
        want = [wrap(f(int(a), int(b)), bits, signed) for a, b in zip(x, y)]
        assert [int(v) for v in got] == want, (name, op)
    got = J.evaluate("(neg x)", {"x": x})[0]
    assert [int(v) for v in got] == [wrap(-int(a), bits, signed) for a in x]


@pytest.mark.parametrize("dt", [np.int8, np.int32, np.int64, np.uint16, np.uint64])
def test_division_against_the_textbook_definitions(dt):
    x, y = ints(dt, 500, 3), ints(dt, 500, 4)
    y = np.where(y == 0, 1, y).astype(dt)
    if np.dtype(dt).kind == "i":  # minBound `quot` (-1) is undefined
        bad = (x == np.iinfo(dt).min) & (y == -1)
        y = np.where(bad, 3, y).astype(dt)
    xs, ys = [int(v) for v in x], [int(v) for v in y]
    tq = [abs(a) // abs(b) * (1 if (a >= 0) == (b >= 0) else -1) for a, b in zip(xs, ys)]
    assert [int(v) for v in J.evaluate("(quot x y)", {"x": x, "y": y})[0]] == tq
    assert [int(v) for v in J.evaluate("(rem x y)", {"x": x, "y": y})[0]] == [a - q * b for a, b, q in zip(xs, ys, tq)]
    assert [int(v) for v in J.evaluate("(div x y)", {"x": x, "y": y})[0]] == [a // b for a, b in zip(xs, ys)]
    assert [int(v) for v in J.evaluate("(mod x y)", {"x": x, "y": y})[0]] == [a % b for a, b in zip(xs, ys)]
    q, r = J.evaluate("(quotrem x y)", {"x": x, "y": y})
    assert [int(v) for v in q] == tq
    d, m = J.evaluate("(divmod x y)", {"x": x, "y": y})
    assert [int(v) for v in m] == [a % b for a, b in zip(xs, ys)]


@pytest.mark.parametrize("dt", [np.int8, np.int32, np.uint32, np.int64, np.uint64])
def test_bit_operations_against_python_integers(dt):
    t = np.dtype(dt)
    bits, signed = t.itemsize * 8, t.kind == "i"
    x = ints(dt, 200, 5)
    s = np.random.default_rng(6).integers(0, bits, size=200).astype(np.int64)
    ux = [int(v) % (1 << bits) for v in x]
    sl = J.evaluate("(shiftl x s)", {"x": x, "s": s})[0]
    assert [int(v) for v in sl] == [wrap(u << int(k), bits, signed) for u, k in zip(ux, s)]
    sr = J.evaluate("(shiftr x s)", {"x": x, "s": s})[0]
    assert [int(v) for v in sr] == [int(a) >> int(k) for a, k in zip(x, s)]  # arithmetic for signed
    rl = J.evaluate("(rotatel x s)", {"x": x, "s": s})[0]
    m = (1 << bits) - 1
    assert [int(v) for v in rl] == [wrap(((u << int(k)) | (u >> (bits - int(k)))) & m, bits, signed)
                                    for u, k in zip(ux, s)]
    rr = J.evaluate("(rotater x s)", {"x": x, "s": s})[0]
    assert [int(v) for v in rr] == [wrap(((u >> int(k)) | (u << (bits - int(k)))) & m, bits, signed)
                                    for u, k in zip(ux, s)]
    assert [int(v) for v in J.evaluate("(popcount x)", {"x": x})[0]] == [bin(u).count("1") for u in ux]
    assert [int(v) for v in J.evaluate("(clz x)", {"x": x})[0]] == [bits - u.bit_length() for u in ux]
    ctz = [(u & -u).bit_length() - 1 if u else bits for u in ux]
    assert [int(v) for v in J.evaluate("(ctz x)", {"x": x})[0]] == ctz
    z = np.zeros(3, dt)
    assert list(J.evaluate("(clz x)", {"x": z})[0]) == [bits] * 3
    assert list(J.evaluate("(ctz x)", {"x": z})[0]) == [bits] * 3


def test_known_answers_of_the_reference_semantics():
    i32 = np.array([-2 ** 31, -5, 0, 7], np.int32)
    assert list(J.evaluate("(abs x)", {"x": i32})[0]) == [-2 ** 31, 5, 0, 7]  # abs minBound = minBound
    assert list(J.evaluate("(signum x)", {"x": i32})[0]) == [-1, -1, 0, 1]
    f = np.array([np.nan, -0.0, 2.5, -3.0], np.float32)
    assert list(J.evaluate("(signum x)", {"x": f})[0]) == [0.0, 0.0, 1.0, -1.0]
    nan = np.float32(np.nan)
    one = np.float32(1)
    for op, want in (("eq", False), ("neq", False), ("lt", False), ("gte", False)):  # ordered: false on NaN
        assert bool(J.evaluate(f"({op} x y)", {"x": nan, "y": one})[0]) is want, op
    assert bool(J.evaluate("(neq x y)", {"x": np.float32(2), "y": one})[0])
    assert float(J.evaluate("(max x y)", {"x": nan, "y": one})[0]) == 1.0  # fmaxf
    assert float(J.evaluate("(min x y)", {"x": one, "y": nan})[0]) == 1.0
    r = J.evaluate("(round i32 x)", {"x": np.array([2.5, -2.5, 0.49999997, 1.5, -0.5, 3.4999998], np.float32)})[0]
    assert list(r) == [3, -3, 0, 2, -1, 3]  # C round: half away from zero
    assert list(J.evaluate("(floor i32 x)", {"x": np.array([-1.5, 1.5], np.float32)})[0]) == [-2, 1]
    assert list(J.evaluate("(ceiling i32 x)", {"x": np.array([-1.5, 1.5], np.float32)})[0]) == [-1, 2]
    assert list(J.evaluate("(truncate i32 x)", {"x": np.array([-1.5, 1.5], np.float32)})[0]) == [-1, 1]
    neg0 = J.evaluate("(neg x)", {"x": np.float32(0.0)})[0]
    assert math.copysign(1.0, float(neg0)) == -1.0  # 0 * (-1) = -0
    big = J.evaluate("(tofloating f32 x)", {"x": np.int64(2 ** 24 + 1)})[0]
    assert float(big) == 2 ** 24  # nearest-even
    assert int(J.evaluate("(fromintegral i8 x)", {"x": np.int32(300)})[0]) == 44  # truncation wraps
    assert int(J.evaluate("(fromintegral i64 x)", {"x": np.uint32(2 ** 32 - 1)})[0]) == 2 ** 32 - 1  # zero extension
    assert int(J.evaluate("(fromintegral i64 x)", {"x": np.int32(-1)})[0]) == -1  # sign extension
    assert bool(J.evaluate("(isinfinite x)", {"x": np.float32(-np.inf)})[0])
    lb = J.evaluate("(logbase (const f64 2) x)", {"x": np.float64(1024.0)})[0]
    assert abs(float(lb) - 10.0) < 1e-12
    assert int(J.evaluate("(const i8 200)", {})[0]) == -56  # a literal wraps into its type
    assert float(J.evaluate("(const f32 0x1.8p1)", {})[0]) == 3.0


# ------------------------------------------------------------- skeletons
def seq_fold(f, xs, seed=None):
    acc = seed
    for x in xs:
        acc = x if acc is None else f(acc, x)
    return acc


def mat(e):
    return tuple(int(v) for v in e)


def matmul32(a, b):
    r = (a[0] * b[0] + a[1] * b[2], a[0] * b[1] + a[1] * b[3], a[2] * b[0] + a[3] * b[2], a[2] * b[1] + a[3] * b[3])
    return tuple(wrap(v, 32, True) for v in r)


def mats(n, seed):
    """n unimodular 2x2 matrices, [[1,a],[0,1]] or [[1,0],[b,1]]: their products have
    determinant 1, so a product never collapses to 0 mod 2^32, and they do not commute."""
    rng = np.random.default_rng(seed)
    up = rng.integers(0, 2, size=n).astype(bool)
    v = rng.integers(-1000, 1000, size=n, endpoint=True).astype(np.int32)
    one, zero = np.ones(n, np.int32), np.zeros(n, np.int32)
    return [one, np.where(up, v, zero).astype(np.int32), np.where(up, zero, v).astype(np.int32), one.copy()]


def test_fold_matrix_product_in_order_against_a_sequential_loop():
    for n in (1, 2, 3, 17, 1000):
        x = mats(n, n)
        seq = seq_fold(matmul32, [mat(v) for v in zip(*x)])
        got = J.fold("(fold (elt i32 i32 i32 i32) " + MATMUL + ")", x, 1, n)
        assert mat(v[0] for v in got) == seq, n
        # the reversed order gives another product: a fold out of order would be seen
        rev = seq_fold(matmul32, [mat(v) for v in zip(*x)][::-1])
        assert n < 17 or (rev != seq and seq != (0, 0, 0, 0))
    x = mats(5 * 7, 9)
    got = J.fold("(fold (elt i32 i32 i32 i32) " + MATMUL + " (seed (tuple (const i32 2) (const i32 0) "
                 "(const i32 0) (const i32 1))))", x, 5, 7)
    for r in range(5):
        row = [mat(v) for v in zip(*[f[r * 7:(r + 1) * 7] for f in x])]
        assert mat(v[r] for v in got) == seq_fold(matmul32, row, (2, 0, 0, 1))


def test_scan_variants_against_the_reference_examples():
    x = [np.arange(10, dtype=np.int64)]
    add = "(fun (a b) (add a b))"
    out, _ = J.scan("(scan (elt i64) " + add + " (seed (const i64 10)))", x, 1, 10)
    assert list(out[0]) == GOLDEN["scanl_plus_10"]["expect"]
    out, sums = J.scan("(scan (elt i64) " + add + " (seed (const i64 10)) (dir r2l) prime)", x, 1, 10)
    assert list(out[0]) == GOLDEN["scanr_prime_plus_10"]["expect"]
    assert int(sums[0][0]) == GOLDEN["scanr_prime_plus_10"]["expect_sum"]
    out, sums = J.scan("(scan (elt i64) " + add + " (seed (const i64 10)) prime)", x, 1, 10)
    assert list(out[0]) == GOLDEN["scanl_prime_plus_10"]["expect"]
    assert int(sums[0][0]) == GOLDEN["scanl_prime_plus_10"]["expect_sum"]
    got = J.fold("(fold (elt i64) " + add + " (seed (const i64 0)) (input (fun (i) (add i (const i64 2)))))", [], 1,
                 10000)
    assert int(got[0][0]) == GOLDEN["runq_sum_map_plus1"]["expect"]  # sum (map (+1) [1..10000]) as a fused producer


@pytest.mark.parametrize("shape", [(1, 1), (1, 37), (3, 20), (2, 0)])
def test_scan_matrix_product_every_variant_against_loops(shape):
    outer, inner = shape
    x = mats(outer * inner, 11)
    seed = (1, 1, 0, 1)
    sd = " (seed (tuple (const i32 1) (const i32 1) (const i32 0) (const i32 1)))"
    base = "(scan (elt i32 i32 i32 i32) " + MATMUL
    for variant in ("l1", "l", "r1", "r", "lp", "rp"):
        node = base + (sd if variant not in ("l1", "r1") else "") + (" (dir r2l)" if variant[0] == "r" else "") + \
            (" prime" if variant.endswith("p") else "") + ")"
        if inner == 0 and variant in ("l1", "r1"):
            continue
        out, sums = J.scan(node, x, outer, inner)
        for r in range(outer):
            row = [mat(v) for v in zip(*[f[r * inner:(r + 1) * inner] for f in x])]
            if variant == "l1":
                want = [seq_fold(matmul32, row[:k + 1]) for k in range(inner)]
            elif variant == "r1":
                want = [seq_fold(lambda a, b: matmul32(b, a), row[k:][::-1]) for k in range(inner)]
            elif variant == "l":
                want = [seq_fold(matmul32, row[:k], seed) for k in range(inner + 1)]
            elif variant == "r":
                want = [seq_fold(lambda a, b: matmul32(b, a), row[k:][::-1], seed) for k in range(inner)] + [seed]
            elif variant == "lp":
                want = [seq_fold(matmul32, row[:k], seed) for k in range(inner)]
                assert mat(s[r] for s in sums) == seq_fold(matmul32, row, seed)
            else:
                want = [seq_fold(lambda a, b: matmul32(b, a), row[k + 1:][::-1], seed) for k in range(inner)]
                assert mat(s[r] for s in sums) == seq_fold(lambda a, b: matmul32(b, a), row[::-1], seed)
            w = len(want)
            got = [mat(f[r * w + k] for f in out) for k in range(w)]
            assert got == want, (variant, shape)


def test_foldseg_against_a_loop_including_empty_segments():
    x = mats(60, 13)
    off = np.array([0, 0, 5, 6, 6, 30, 60], np.int64)
    node = "(foldseg (elt i32 i32 i32 i32) " + MATMUL + " (seed (tuple (const i32 1) (const i32 2) (const i32 0) (const i32 1))))"
    outs, have = J.foldseg(node, x, 1, 60, off)
    row = [mat(v) for v in zip(*x)]
    for s in range(6):
        want = seq_fold(matmul32, row[off[s]:off[s + 1]], (1, 2, 0, 1))
        assert mat(o[s] for o in outs) == want
    outs, have = J.foldseg("(foldseg (elt i32 i32 i32 i32) " + MATMUL + ")", x, 1, 60, off)
    assert list(have) == [False, True, True, False, True, True]


def test_permute_against_a_sequential_loop():
    rng = np.random.default_rng(3)
    n, m = 500, 17
    idx = rng.integers(-3, m + 3, size=n).astype(np.int32)
    v = rng.integers(-5, 6, size=n).astype(np.int32)
    node = "(permute (elt i32) (fun (new old) (mul new old)) (project (fun (i) (fromintegral i64 (read 0 i)))) (free i32))"
    d = np.full(m, 1, np.int32)
    got = J.permute(node, [d], [v], n, free=[idx])[0]
    want = [1] * m
    for i in range(n):
        if 0 <= idx[i] < m:
            want[idx[i]] = wrap(int(v[i]) * want[idx[i]], 32, True)
    assert list(got) == want
    # a two-field element (count, maximum): the lock path's kind of combine
    node2 = ("(permute (elt i32 f32) (fun (new old) (tuple (add (prj 0 new) (prj 0 old)) (max (prj 1 new) (prj 1 old))))"
             " (project (fun (i) (fromintegral i64 (read 0 i)))) (free i32))")
    w = rng.standard_normal(n).astype(np.float32)
    got = J.permute(node2, [np.zeros(m, np.int32), np.full(m, -np.inf, np.float32)], [np.ones(n, np.int32), w], n,
                    free=[idx])
    for j in range(m):
        sel = idx == j
        assert got[0][j] == sel.sum() and (got[1][j] == (w[sel].max() if sel.any() else -np.inf))


@pytest.mark.parametrize("b,mode", [("clamp", "nearest"), ("mirror", "mirror"), ("wrap", "wrap")])
def test_stencil_max_of_neighbours_against_scipy(b, mode):
    scipy_ndimage = pytest.importorskip("scipy.ndimage")
    img = np.random.default_rng(4).standard_normal((23, 31)).astype(np.float32)
    terms = " ".join(f"(nb {dy} {dx})" for dy in (-1, 0, 1) for dx in (-2, -1, 0, 1, 2))
    body = terms.split(" (nb")
    expr = "(nb -1 -2)"
    for dy in (-1, 0, 1):
        for dx in (-2, -1, 0, 1, 2):
            if (dy, dx) != (-1, -2):
                expr = f"(max {expr} (nb {dy} {dx}))"
    node = f"(stencil (in f32) (out f32) (radius 1 2) (boundary {b}) (fun () {expr}))"
    got = J.stencil(node, img)
    want = scipy_ndimage.maximum_filter(img, size=(3, 5), mode=mode)
    assert np.array_equal(got, want)


def test_stencil_constant_and_function_boundaries_by_hand():
    img = np.arange(12, dtype=np.int32).reshape(3, 4)
    node = "(stencil (in i32) (out i32) (radius 1 1) (boundary (constant (const i32 -7))) (fun () (nb -1 -1)))"
    got = J.stencil(node, img)
    assert got[0, 0] == -7 and got[1, 1] == img[0, 0] and got[2, 0] == -7
    node = ("(stencil (in i32) (out i32) (radius 1 1) (boundary (function (fun (y x) (fromintegral i32 "
            "(add (mul y (const i64 100)) x))))) (fun () (add (nb -1 0) (nb 0 1))))")
    got = J.stencil(node, img)
    assert got[0, 0] == -100 + 1  # above (-1, 0) -> f = -100; right neighbour img[0,1] = 1
    assert got[0, 3] == (-100 + 3) + 4  # above: f(-1, 3) = -97; right: f(0, 4) = 4


def test_generate_delayed_and_free_arrays():
    a = np.arange(10, dtype=np.float32)
    out = J.generate("(generate (elt f32 i64) (fun (i) (tuple (mul (read 0 i) (read 0 i)) (mul i i))) (free f32))",
                     1, 10, free=[a])
    assert list(out[0]) == list(a * a) and list(out[1]) == [k * k for k in range(10)]
    out = J.generate("(generate (elt i64) (fun (r c) (sub (mul r (const i64 10)) c)))", 3, 4)
    assert list(out[0]) == [r * 10 - c for r in range(3) for c in range(4)]
    # a delayed input fused into a fold: the dot product of two free arrays
    b = np.arange(10, dtype=np.float32) * 2
    got = J.fold("(fold (elt f32) (fun (x y) (add x y)) (input (fun (i) (mul (read 0 i) (read 1 i)))) (free f32 f32))",
                 [], 1, 10, free=[a, b])
    assert float(got[0][0]) == float((a * b).sum())


# ------------------------------------------- translation and compilation
NODES = [
    "(fold (elt i32 i32 i32 i32) " + MATMUL + ")",
    "(fold (elt f32) (fun (x y) (add x y)) (seed (const f32 0)) (input (fun (i) (mul (read 0 i) (read 1 i)))) (free f32 f32))",
    "(scan (elt i32 i32 i32 i32) " + MATMUL + " (dir r2l) (seed (tuple (const i32 1) (const i32 0) (const i32 0) (const i32 1))))",
    "(scan (elt f64) (fun (x y) (add x y)) prime (seed (const f64 0)))",
    "(foldseg (elt i32 i32 i32 i32) " + MATMUL + " (offsets i32))",
    "(permute (elt i32) (fun (n o) (mul n o)) (project (fun (i) (fromintegral i64 (read 0 i)))) (free i32))",
    "(permute (elt i32 f32) (fun (n o) (tuple (add (prj 0 n) (prj 0 o)) (max (prj 1 n) (prj 1 o)))) "
    "(project (fun (i) (read 0 i))) (free i64))",
    "(stencil (in f32) (out f32) (radius 1 1) (boundary mirror) (fun () (max (nb 0 0) (nb 1 1))))",
    "(stencil (in i32) (out bool) (radius 2 1) (boundary (function (fun (y x) (fromintegral i32 y)))) "
    "(fun () (gt (nb -2 1) (nb 2 -1))))",
    "(generate (elt u8 f64) (fun (r c) (tuple (fromintegral u8 r) (tofloating f64 c))))",
]


@pytest.fixture
def jit_cache(tmp_path, monkeypatch):
    monkeypatch.setenv("ACC_HIP_CACHE_DIR", str(tmp_path))
    return tmp_path


def test_every_node_kind_compiles_for_gfx950_and_caches(jit_cache):
    import acc_hip as A
    for node in NODES:
        src = A.node_source(node)
        assert "extern \"C\" __global__" in src
        assert A.precompile(node) == A.KERNEL_COMPILED, node
        assert A.precompile(node) == A.KERNEL_FROM_DISK_CACHE, node
    assert len(list(jit_cache.glob("*.gfx950.co"))) == len(NODES)


@pytest.mark.parametrize("bad,msg", [
    ("(fold (elt i32) (fun (a b) (add a c)))", "unbound variable c"),
    ("(fold (elt i32) (fun (a b) (add a (const f32 1))))", "different types"),
    ("(fold (elt i32) (fun (a) a))", "parameters"),
    ("(scan (elt i32) (fun (a b) (add a b)) prime)", "needs a seed"),
    ("(fold (elt i32 i32) (fun (a b) (prj 0 a)))", "result has type"),
    ("(generate (elt f32) (fun (i) (read 0 i)))", "no free array"),
    ("(fold (elt i32) (fun (a b) (frobnicate a b)))", "unknown operator"),
    ("(stencil (in f32) (out f32) (radius 1 1) (boundary clamp) (fun () (nb 2 0)))", "outside the stencil"),
    ("(fold (elt i32) (fun (a b) (add a b))", "unbalanced"),
])
def test_bad_nodes_are_rejected_with_a_message(bad, msg):
    import acc_hip as A
    with pytest.raises(A.AccError) as e:
        A.node_source(bad)
    assert msg in str(e.value) and e.value.status == A.INVALID_VALUE


def test_foreign_is_its_fallback_and_strings_parse():
    """(foreign "body" (fun (x ..) e) args ..) means the fallback applied to
    the arguments, in a closed environment; string literals keep escaped
    quotes and backslashes."""
    e = J.parse('(foreign "return x; // \\"q\\" \\\\n" (fun (x) x) i)')
    assert e[1] == 'return x; // "q" \\n'
    i = np.arange(-5, 6, dtype=np.int64)
    got = J.evaluate('(foreign "return x * x + 7;" (fun (x) (add (mul x x) (const i64 7))) i)', {"i": i})[0]
    assert np.array_equal(got, i * i + 7)
    # the fallback sees only its parameters: i inside it is the argument, not the outer i
    got = J.evaluate('(foreign "" (fun (i) (mul i (const i64 2))) (add i (const i64 1)))', {"i": i})[0]
    assert np.array_equal(got, (i + 1) * 2)
    # bit reversal by shifts and masks against Python
    x = np.array([1, 3, 0x80000000, 0x12345678, 0xffffffff], np.uint32)
    masks = ["0x55555555", "0x33333333", "0x0f0f0f0f", "0x00ff00ff"]
    body, v = "", "x"
    for k, m in enumerate(masks):
        s = 1 << k
        body += (f"(x{k + 1} (bor (band (shiftr {v} (const i64 {s})) (const u32 {m})) "
                 f"(shiftl (band {v} (const u32 {m})) (const i64 {s})))) ")
        v = f"x{k + 1}"
    rev = f"(let ({body}) (bor (shiftr x4 (const i64 16)) (shiftl x4 (const i64 16))))"
    got = J.evaluate(f'(foreign "" (fun (x) {rev}) y)', {"y": x})[0]
    want = np.array([int(f"{int(v):032b}"[::-1], 2) for v in x], np.uint32)
    assert np.array_equal(got, want)
