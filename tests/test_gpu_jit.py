"""GPU parity of the run-time compiled skeletons against their oracle
(oracle/oracle_jit.py), through the C ABI.

The combine of fold, scan and foldSeg is the 2x2 integer matrix product:
associative but not commutative, with no identity assumed, so any reduction
out of element order, any dropped or repeated element, and any use of a
padding value gives a different (bit-exact checked) result.  permute (*)
runs through compare-and-swap, a (count, maximum) pair through the per-slot
lock; stencils apply a max-of-neighbours function on every boundary.
Integer results are bit-exact; floating arithmetic that rounds correctly
(+ - * / sqrt, conversions, comparisons, min/max) is bit-exact too, device
math functions are within 4 ulp of the oracle's.
"""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

import acc_hip as A
import oracle_jit as J

pytestmark = pytest.mark.gpu

MATMUL = ("(fun (a b) (tuple (add (mul (prj 0 a) (prj 0 b)) (mul (prj 1 a) (prj 2 b)))"
          " (add (mul (prj 0 a) (prj 1 b)) (mul (prj 1 a) (prj 3 b)))"
          " (add (mul (prj 2 a) (prj 0 b)) (mul (prj 3 a) (prj 2 b)))"
          " (add (mul (prj 2 a) (prj 1 b)) (mul (prj 3 a) (prj 3 b)))))")
SEED4 = " (seed (tuple (const i32 1) (const i32 3) (const i32 0) (const i32 1)))"
M4 = "(elt i32 i32 i32 i32) "


def say(msg):
    print(f"[jit] {msg}", file=sys.stderr, flush=True)


def mats(n, seed):
    """Unimodular matrices [[1,a],[0,1]] / [[1,0],[b,1]]: products never collapse to 0 mod 2^32."""
    rng = np.random.default_rng(seed)
    up = rng.integers(0, 2, size=n).astype(bool)
    v = rng.integers(-1000, 1000, size=n, endpoint=True).astype(np.int32)
    one, zero = np.ones(n, np.int32), np.zeros(n, np.int32)
    return [one, np.where(up, v, zero).astype(np.int32), np.where(up, zero, v).astype(np.int32), one.copy()]


def dev(s, fields):
    return [A.DeviceArray.from_numpy(s, np.ascontiguousarray(f)) for f in fields]


def host(s, fields):
    return [f.to_numpy(s) for f in fields]


def same(got, want):
    return all(np.array_equal(g, np.asarray(w)) for g, w in zip(got, want))


@pytest.fixture(scope="module")
def kernels():
    cache = {}

    def get(node):
        if node not in cache:
            say(f"compile {node[:70]}")
            cache[node] = A.Kernel(node)
        return cache[node]
    return get


# ------------------------------------------------------------------ fold
@pytest.mark.parametrize("n", [1, 5, 1000, 1 << 16, (1 << 20) + 3])
@pytest.mark.parametrize("seeded", [False, True])
def test_fold_matrix_product(stream, kernels, n, seeded):
    node = "(fold " + M4 + MATMUL + (SEED4 if seeded else "") + ")"
    k = kernels(node)
    x = mats(n, n)
    d_in, d_out = dev(stream, x), [A.DeviceArray(1, np.int32) for _ in range(4)]
    k.run(stream, d_in, d_out, 1, n)
    assert same(host(stream, d_out), J.fold(node, x, 1, n)), (n, seeded)


@pytest.mark.parametrize("shape", [(7, 100003), (5000, 100), (3, 1), (4, 0)])
def test_fold_rows_matrix_product(stream, kernels, shape):
    outer, inner = shape
    node = "(fold " + M4 + MATMUL + SEED4 + ")"
    k = kernels(node)
    x = mats(outer * inner, 3)
    d_in = dev(stream, x) if inner else [A.DeviceArray(1, np.int32) for _ in range(4)]
    d_out = [A.DeviceArray(outer, np.int32) for _ in range(4)]
    k.run(stream, d_in, d_out, outer, inner)
    assert same(host(stream, d_out), J.fold(node, x, outer, inner)), shape


def test_fold1_of_empty_is_an_error(stream, kernels):
    k = kernels("(fold " + M4 + MATMUL + ")")
    d = [A.DeviceArray(4, np.int32) for _ in range(4)]
    with pytest.raises(A.EmptyArray):
        k.run(stream, d, d, 1, 0)


def test_fold_misaligned_views(stream, kernels):
    node = "(fold " + M4 + MATMUL + ")"
    k = kernels(node)
    n = 300001
    x = mats(n + 1, 8)
    d = dev(stream, x)
    views = [f.view(1, n) for f in d]  # 4-byte offset: element accesses, no vectors
    out = [A.DeviceArray(1, np.int32) for _ in range(4)]
    k.run(stream, views, out, 1, n)
    assert same(host(stream, out), J.fold(node, [f[1:] for f in x], 1, n))


def test_fold_delayed_producer_dot_product(stream, kernels):
    """A producer fused into the fold (LLVM/CodeGen/Sugar.hs:70-76): sum of a[i]*b[i]
    over Int64 free arrays, never materialised."""
    node = ("(fold (elt i64) (fun (x y) (add x y)) (seed (const i64 5)) "
            "(input (fun (i) (mul (read 0 i) (read 1 i)))) (free i64 i64))")
    k = kernels(node)
    n = (1 << 20) + 9
    a = np.random.default_rng(1).integers(-1 << 20, 1 << 20, size=n).astype(np.int64)
    b = np.random.default_rng(2).integers(-1 << 20, 1 << 20, size=n).astype(np.int64)
    da, db = A.DeviceArray.from_numpy(stream, a), A.DeviceArray.from_numpy(stream, b)
    out = A.DeviceArray(1, np.int64)
    k.run(stream, None, [out], 1, n, free=[da, db])
    want = J.fold(node, [], 1, n, free=[a, b])[0][0]
    assert out.to_numpy(stream)[0] == want == np.int64(5) + np.sum(a * b, dtype=np.int64)


# ------------------------------------------------------------------ scan
VARIANTS = {
    "scanl1": "", "scanl": SEED4, "scanr1": " (dir r2l)", "scanr": SEED4 + " (dir r2l)",
    "scanl'": SEED4 + " prime", "scanr'": SEED4 + " (dir r2l) prime",
}


def scan_outputs(variant, outer, inner):
    w = inner + 1 if variant in ("scanl", "scanr") else inner
    return [A.DeviceArray(max(1, outer * w), np.int32) for _ in range(4)], w


@pytest.mark.parametrize("variant", list(VARIANTS))
@pytest.mark.parametrize("n", [1, 4097, 100003, (1 << 21) + 5])
def test_scan_matrix_product_every_variant(stream, kernels, variant, n):
    node = "(scan " + M4 + MATMUL + VARIANTS[variant] + ")"
    k = kernels(node)
    x = mats(n, n + 1)
    d_in = dev(stream, x)
    out, w = scan_outputs(variant, 1, n)
    sums = [A.DeviceArray(1, np.int32) for _ in range(4)]
    k.run(stream, d_in, out, 1, n, sums=sums)
    want, wsums = J.scan(node, x, 1, n)
    got = [o.view(0, w).to_numpy(stream) for o in out]
    assert same(got, want), (variant, n)
    if wsums is not None:
        assert same(host(stream, sums), wsums), (variant, n)


@pytest.mark.parametrize("variant", list(VARIANTS))
@pytest.mark.parametrize("shape", [(3, 5000), (600, 17), (5, 0)])
def test_scan_rows_matrix_product(stream, kernels, variant, shape):
    outer, inner = shape
    if inner == 0 and variant in ("scanl1", "scanr1"):
        pytest.skip("no outputs")
    node = "(scan " + M4 + MATMUL + VARIANTS[variant] + ")"
    k = kernels(node)
    x = mats(outer * inner, 5)
    d_in = dev(stream, x) if inner else [A.DeviceArray(1, np.int32) for _ in range(4)]
    out, w = scan_outputs(variant, outer, inner)
    sums = [A.DeviceArray(outer, np.int32) for _ in range(4)]
    k.run(stream, d_in, out, outer, inner, sums=sums)
    want, wsums = J.scan(node, x, outer, inner)
    if outer * w:
        assert same([o.view(0, outer * w).to_numpy(stream) for o in out], want), (variant, shape)
    if wsums is not None:
        assert same(host(stream, sums), wsums), (variant, shape)


def test_scan_float_sum_within_bound(stream, kernels):
    node = "(scan (elt f32) (fun (a b) (add a b)))"
    k = kernels(node)
    n = (1 << 22) + 3
    x = np.random.default_rng(3).standard_normal(n).astype(np.float32)
    d, out = A.DeviceArray.from_numpy(stream, x), A.DeviceArray(n, np.float32)
    k.run(stream, [d], [out], 1, n)
    x64 = x.astype(np.float64)
    err = np.abs(out.to_numpy(stream).astype(np.float64) - np.cumsum(x64))
    bound = 1e-5 * np.maximum(1, np.log2(np.maximum(np.arange(1, n + 1), 2))) * np.cumsum(np.abs(x64))
    assert np.all(err <= bound)


# ------------------------------------------------------------- foldSeg
@pytest.mark.parametrize("itype", ["i32", "i64"])
@pytest.mark.parametrize("cfg", [(4096, 16384, 1 << 22), (20000, 64, 600000), (50, 1, 50)])
def test_fold_seg_matrix_product(stream, kernels, itype, cfg):
    m, maxlen, n = cfg
    rng = np.random.default_rng(m)
    lens = np.maximum(1, np.exp(rng.uniform(0, np.log(maxlen), size=m))).astype(np.int64)
    lens[rng.integers(0, m, size=max(1, m // 10))] = 0  # some empty segments
    lens = (lens * (n / max(1, lens.sum()))).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64 if itype == "i64" else np.int32)
    n = int(off[-1])
    x = mats(n, 7)
    for seeded in (True, False):
        node = "(foldseg " + M4 + MATMUL + (SEED4 if seeded else "") + f" (offsets {itype}))"
        k = kernels(node)
        d_in = dev(stream, x) if n else [A.DeviceArray(1, np.int32) for _ in range(4)]
        d_off = A.DeviceArray.from_numpy(stream, off)
        init = np.full(m, -123, np.int32)
        d_out = dev(stream, [init] * 4)
        k.run(stream, d_in, d_out, 1, n, nseg=m, offsets=d_off)
        want, have = J.foldseg(node, x, 1, n, off)
        got = host(stream, d_out)
        if seeded:
            assert same(got, want), (cfg, itype)
        else:  # fold1Seg leaves empty segments unwritten
            assert same([g[have] for g in got], [w[have] for w in want]) and all(np.all(g[~have] == -123) for g in got)


# -------------------------------------------------------------- permute
def test_permute_product_compare_and_swap(stream, kernels):
    """permute (*) of odd Int32 values into 1000 slots from a free index array,
    with indices outside the output (Nothing): compare-and-swap on the slot."""
    node = ("(permute (elt i32) (fun (new old) (mul new old)) "
            "(project (fun (i) (fromintegral i64 (read 0 i)))) (free i32))")
    k = kernels(node)
    assert k.info().lock_free_permute == 1
    n, m = 1 << 20, 1000
    rng = np.random.default_rng(9)
    idx = rng.integers(-50, m + 50, size=n).astype(np.int32)
    v = (rng.integers(-1 << 20, 1 << 20, size=n).astype(np.int32) | 1).astype(np.int32)
    d_idx, d_v = A.DeviceArray.from_numpy(stream, idx), A.DeviceArray.from_numpy(stream, v)
    defaults = np.full(m, 3, np.int32)
    d_def, d_out = A.DeviceArray.from_numpy(stream, defaults), A.DeviceArray(m, np.int32)
    k.run(stream, [d_v], [d_out], 1, n, nseg=m, defaults=[d_def], free=[d_idx])
    want = J.permute(node, [defaults], [v], n, free=[idx])[0]
    assert np.array_equal(d_out.to_numpy(stream), want)
    assert np.array_equal(d_def.to_numpy(stream), defaults)  # cloned, not updated
    # in place
    k.run(stream, [d_v], [d_def], 1, n, nseg=m, defaults=[d_def], free=[d_idx])
    assert np.array_equal(d_def.to_numpy(stream), want)


def test_permute_int64_compare_and_swap(stream, kernels):
    """The wrapping Int64 product of odd values (associative and commutative, as
    permute requires, PTX/CodeGen/Permute.hs:67-68) by 64-bit compare-and-swap."""
    node = ("(permute (elt i64) (fun (new old) (mul new old)) "
            "(project (fun (i) (rem i (const i64 777)))))")
    k = kernels(node)
    assert k.info().lock_free_permute == 1
    n, m = 500000, 777
    v = (np.random.default_rng(10).integers(-1 << 40, 1 << 40, size=n).astype(np.int64) | 1).astype(np.int64)
    d_v, d_out = A.DeviceArray.from_numpy(stream, v), A.DeviceArray.from_numpy(stream, np.ones(m, np.int64))
    k.run(stream, [d_v], [d_out], 1, n, nseg=m, defaults=[d_out])
    want = J.permute(node, [np.ones(m, np.int64)], [v], n)[0]
    assert np.array_equal(d_out.to_numpy(stream), want) and np.all(want % 2 == 1)


def test_permute_tuple_takes_the_lock_path(stream, kernels):
    """(count, maximum) per slot: a two-field element, updated under a per-slot
    spin lock (PTX/CodeGen/Permute.hs:219-258), Zipf-like contention on a few slots."""
    node = ("(permute (elt i32 f32) (fun (new old) (tuple (add (prj 0 new) (prj 0 old)) (max (prj 1 new) (prj 1 old)))) "
            "(project (fun (i) (read 0 i))) (input (fun (i) (tuple (const i32 1) (read 1 i)))) (free i64 f32))")
    k = kernels(node)
    assert k.info().lock_free_permute == 0
    n, m = 1 << 18, 64
    rng = np.random.default_rng(11)
    idx = np.minimum(rng.zipf(1.3, size=n) - 1, m + 5).astype(np.int64)  # hot slot 0, some Nothing
    w = rng.standard_normal(n).astype(np.float32)
    d_idx, d_w = A.DeviceArray.from_numpy(stream, idx), A.DeviceArray.from_numpy(stream, w)
    c0, m0 = np.zeros(m, np.int32), np.full(m, -np.inf, np.float32)
    d_c, d_m = A.DeviceArray.from_numpy(stream, c0), A.DeviceArray.from_numpy(stream, m0)
    k.run(stream, None, [d_c, d_m], 1, n, nseg=m, defaults=[d_c, d_m], free=[d_idx, d_w])
    want = J.permute(node, [c0, m0], [], n, free=[idx, w])
    assert np.array_equal(d_c.to_numpy(stream), want[0]) and np.array_equal(d_m.to_numpy(stream), want[1])
    assert int(want[0].sum()) == int((idx < m).sum())


# -------------------------------------------------------------- stencil
def max_of(ry, rx):
    e = f"(nb {-ry} {-rx})"
    for dy in range(-ry, ry + 1):
        for dx in range(-rx, rx + 1):
            if (dy, dx) != (-ry, -rx):
                e = f"(max {e} (nb {dy} {dx}))"
    return e


BOUNDS = ["clamp", "mirror", "wrap", "(constant (const f32 0.25))",
          "(function (fun (y x) (tofloating f32 (sub (mul y (const i64 3)) x))))"]


@pytest.mark.parametrize("b", BOUNDS)
@pytest.mark.parametrize("shape,r", [((300, 500), (1, 2)), ((2048, 2048), (2, 2)), ((5, 3), (4, 1))])
def test_stencil_max_of_neighbours(stream, kernels, b, shape, r):
    ry, rx = r
    if "mirror" in b and (shape[0] <= ry or shape[1] <= rx):
        pytest.skip("mirror of an extent not above the radius")
    node = f"(stencil (in f32) (out f32) (radius {ry} {rx}) (boundary {b}) (fun () {max_of(ry, rx)}))"
    k = kernels(node)
    img = np.random.default_rng(shape[0]).standard_normal(shape).astype(np.float32)
    d, out = A.DeviceArray.from_numpy(stream, img), A.DeviceArray(img.size, np.float32)
    k.run(stream, [d], [out], shape[0], shape[1])
    assert np.array_equal(out.to_numpy(stream).reshape(shape), J.stencil(node, img)), (b, shape)


def test_stencil_integer_to_bool_function(stream, kernels):
    node = ("(stencil (in i32) (out bool) (radius 1 1) (boundary wrap) "
            "(fun () (gt (add (nb -1 0) (nb 1 0)) (mul (const i32 2) (nb 0 0)))))")
    k = kernels(node)
    img = np.random.default_rng(4).integers(-100, 100, size=(257, 129)).astype(np.int32)
    d, out = A.DeviceArray.from_numpy(stream, img), A.DeviceArray(img.size, np.uint8)
    k.run(stream, [d], [out], 257, 129)
    assert np.array_equal(out.to_numpy(stream).reshape(257, 129), J.stencil(node, img))


# ------------------------------------------------------ generate / map
def test_generate_two_fields_from_the_index(stream, kernels):
    node = "(generate (elt i32 f64) (fun (r c) (tuple (fromintegral i32 (sub (mul r (const i64 1000)) c)) (tofloating f64 (mul r c)))))"
    k = kernels(node)
    outer, inner = 999, 1003
    out = [A.DeviceArray(outer * inner, np.int32), A.DeviceArray(outer * inner, np.float64)]
    k.run(stream, None, out, outer, inner)
    assert same(host(stream, out), J.generate(node, outer, inner))


INT_EXPRS = ["(add x y)", "(sub x y)", "(mul x y)", "(neg x)", "(abs x)", "(signum x)", "(quot x y)", "(rem x y)",
             "(div x y)", "(mod x y)", "(band x y)", "(bor x y)", "(bxor x y)", "(bnot x)",
             "(shiftl x (fromintegral i64 (band y (const T 7))))", "(shiftr x (fromintegral i64 (band y (const T 7))))",
             "(rotatel x (fromintegral i64 y))", "(rotater x (fromintegral i64 y))",
             "(fromintegral T (popcount x))", "(fromintegral T (clz x))", "(fromintegral T (ctz x))",
             "(min x y)", "(max x y)", "(cond (lt x y) (const T 1) (const T 0))",
             "(cond (land (gte x y) (neq x (const T 0))) x y)"]
FLT_EXACT = ["(add x y)", "(sub x y)", "(mul x y)", "(fdiv x y)", "(neg x)", "(abs x)", "(signum x)",
             "(min x y)", "(max x y)", "(recip x)", "(sqrt (abs x))",
             "(cond (neq x y) x y)", "(cond (lte x y) x y)", "(tofloating T (round i32 x))",
             "(tofloating T (floor i32 x))", "(tofloating T (ceiling i32 x))", "(tofloating T (truncate i32 x))",
             "(cond (isnan (fdiv x (const T 0))) (const T 7) x)", "(cond (isinfinite (fdiv x (const T 0))) (const T 7) x)"]
FLT_MATH = ["(sin x)", "(cos x)", "(tan x)", "(atan x)", "(exp (fdiv x (const T 4)))", "(log (abs x))",
            "(atan2 x y)", "(fpow (abs x) (fdiv y (const T 4)))", "(tanh x)", "(sinh (fdiv x (const T 4)))",
            "(asin (fdiv x (const T 5)))", "(logbase (const T 2) (abs x))", "(logbase (const T 3) (abs x))"]


def groups(exprs, k=4):
    return [exprs[g::k] for g in range(k)]


def xor_all(es, t):
    acc = es[0]
    for e in es[1:]:
        acc = f"(bxor {acc} {e})"
    return acc


@pytest.mark.parametrize("t", ["i8", "i32", "i64", "u16", "u32", "u64"])
def test_integer_primitives_on_the_device(stream, kernels, t):
    """Every integer primitive on device data, bit-exact against the oracle:
    four outputs, each the XOR of a quarter of the primitives' results."""
    dt = J.DT[t]
    n = 8192
    rng = np.random.default_rng(12)
    info = np.iinfo(dt)
    x = rng.integers(info.min, info.max, size=n, endpoint=True, dtype=dt)
    y = rng.integers(info.min, info.max, size=n, endpoint=True, dtype=dt)
    y = np.where(y == 0, 5, y).astype(dt)
    if info.min < 0:
        x = np.where((x == info.min) & (y == -1), 1, x).astype(dt)
    dx, dy = A.DeviceArray.from_numpy(stream, x), A.DeviceArray.from_numpy(stream, y)
    exprs = [e.replace("const T", f"const {t}").replace("fromintegral T", f"fromintegral {t}") for e in INT_EXPRS]
    outs = " ".join(xor_all(g, t) for g in groups(exprs))
    node = f"(generate (elt {t} {t} {t} {t}) (fun (i) (let ((x (read 0 i)) (y (read 1 i))) (tuple {outs}))) (free {t} {t}))"
    k = kernels(node)
    out = [A.DeviceArray(n, dt) for _ in range(4)]
    k.run(stream, None, out, 1, n, free=[dx, dy])
    want = J.generate(node, 1, n, free=[x, y])
    for g, (o, w) in enumerate(zip(host(stream, out), want)):
        assert np.array_equal(o, w), (t, groups(exprs)[g])


@pytest.mark.parametrize("t", ["f32", "f64"])
def test_floating_primitives_on_the_device(stream, kernels, t):
    """Correctly rounded operations bit-exact (NaN results compared as NaN);
    device math functions within 8 ulp of a float64 truth (f32) or of numpy (f64)."""
    dt = J.DT[t]
    it = "u32" if t == "f32" else "u64"
    n = 8192
    rng = np.random.default_rng(13)
    x = (rng.standard_normal(n) * 10).astype(dt)
    y = (rng.standard_normal(n) * 10).astype(dt)
    x[:8] = [0.0, -0.0, 2.5, -2.5, np.inf, -np.inf, np.nan, 1e-30]
    xf = np.where(np.isfinite(x), x, 0).astype(dt)  # float-to-int conversions of inf/NaN are undefined
    dx, dxf, dy = (A.DeviceArray.from_numpy(stream, v) for v in (x, xf, y))
    sub = lambda e: e.replace("const T", f"const {t}").replace("tofloating T", f"tofloating {t}")
    # exact: each result with NaN made 7, as bits, XOR-combined per output
    canon = lambda e: f"(bitcast {it} (let ((v {sub(e)})) (cond (isnan v) (const {t} 7) v)))"
    conv = [e for e in FLT_EXACT if "i32 x" in e]
    plain = [e for e in FLT_EXACT if "i32 x" not in e]
    outs = " ".join(xor_all([canon(e) for e in g], it) for g in groups(plain, 3))
    conv_out = xor_all([canon(e) for e in conv], it)
    node = (f"(generate (elt {it} {it} {it} {it}) (fun (i) (let ((x (read 0 i)) (y (read 1 i)) (xf (read 2 i))) "
            f"(tuple {outs} {conv_out.replace(' x)', ' xf)')}))) (free {t} {t} {t}))")
    k = kernels(node)
    out = [A.DeviceArray(n, J.DT[it]) for _ in range(4)]
    k.run(stream, None, out, 1, n, free=[dx, dy, dxf])
    with np.errstate(all="ignore"):
        want = J.generate(node, 1, n, free=[x, y, xf])
    for g, (o, w) in enumerate(zip(host(stream, out), want)):
        assert np.array_equal(o, w), (t, g)
    # device math functions, four per kernel
    eps = np.finfo(dt).eps
    for grp in groups([sub(e) for e in FLT_MATH]):
        body = " ".join(grp)
        node = f"(generate (elt {t} {t} {t} {t}) (fun (i) (let ((x (read 0 i)) (y (read 1 i))) (tuple {body}))) (free {t} {t}))"
        while len(grp) < 4:  # pad to four fields
            grp = grp + [grp[-1]]
            body = " ".join(grp)
            node = f"(generate (elt {t} {t} {t} {t}) (fun (i) (let ((x (read 0 i)) (y (read 1 i))) (tuple {body}))) (free {t} {t}))"
        k = kernels(node)
        out = [A.DeviceArray(n, dt) for _ in range(4)]
        k.run(stream, None, out, 1, n, free=[dx, dy])
        with np.errstate(all="ignore"):
            truth = J.generate(node.replace(t, "f64"), 1, n, free=[x.astype(np.float64), y.astype(np.float64)])
        for e, o, w in zip(grp, host(stream, out), truth):
            w = w.astype(dt) if t == "f32" else w
            fin = np.isfinite(w) & np.isfinite(o)
            err = np.abs(o[fin].astype(np.float64) - w[fin].astype(np.float64))
            tol = 8 * eps * np.maximum(np.abs(w[fin].astype(np.float64)), np.finfo(dt).tiny)
            assert np.all(err <= tol), (t, e, float((err / tol).max()))
            assert np.array_equal(np.isnan(o), np.isnan(w)), (t, e)
            assert np.array_equal(np.isinf(o), np.isinf(w)), (t, e)


# ---------------------------------------------- cache, images, embedding
NODE_C = "(fold (elt i64) (fun (a b) (add a b)) (seed (const i64 1)))"


def run_fold_in_child(tmp, cache_dir, image_path):
    """A fresh process that loads NODE_C (from the cache dir, or from an image
    only), runs it and prints the origin and result."""
    code = f"""
import sys, numpy as np
sys.path.insert(0, {os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'accelerate-llvm_amd')!r})
import acc_hip as A
s = A.Stream(0)
img = open({image_path!r}, 'rb').read() if {image_path!r} else None
k = A.Kernel({NODE_C!r}, image=img)
x = A.DeviceArray.from_numpy(s, np.arange(1, 101, dtype=np.int64))
o = A.DeviceArray(1, np.int64)
k.run(s, [x], [o], 1, 100)
print(k.info().origin, int(o.to_numpy(s)[0]))
"""
    env = dict(os.environ, ACC_HIP_CACHE_DIR=str(cache_dir))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=180)
    assert r.returncode == 0, r.stderr[-2000:]
    origin, value = r.stdout.split()
    return int(origin), int(value)


def test_code_object_cache_and_embedded_images(stream, tmp_path, monkeypatch):
    """The first compile stores the code object (LLVM/Compile/Cache.hs:57-91); a
    second handle in the process reuses the loaded module; a fresh process
    reads the cache; an image carried by a program loads with an empty cache
    and compiles nothing (PTX/Embed.hs:50-79)."""
    cache = tmp_path / "cache"
    monkeypatch.setenv("ACC_HIP_CACHE_DIR", str(cache))
    node = NODE_C + " "  # another text of the same node is the same code object
    k1 = A.Kernel(node)
    assert k1.info().origin in (A.KERNEL_COMPILED, A.KERNEL_LOADED)
    k2 = A.Kernel(NODE_C)
    assert k2.info().origin == A.KERNEL_LOADED
    want = 1 + 5050  # seed (+) sum [1..100]
    x = A.DeviceArray.from_numpy(stream, np.arange(1, 101, dtype=np.int64))
    o = A.DeviceArray(1, np.int64)
    k2.run(stream, [x], [o], 1, 100)
    assert int(o.to_numpy(stream)[0]) == want
    image = k2.image()
    img_path = tmp_path / "node.img"
    img_path.write_bytes(image)
    if k1.info().origin != A.KERNEL_COMPILED:  # loaded by an earlier test: store it now
        assert A.precompile(NODE_C) in (A.KERNEL_COMPILED, A.KERNEL_FROM_DISK_CACHE)
    assert run_fold_in_child(tmp_path, cache, "") == (A.KERNEL_FROM_DISK_CACHE, want)
    empty = tmp_path / "empty"
    assert run_fold_in_child(tmp_path, empty, str(img_path)) == (A.KERNEL_FROM_IMAGE, want)
    assert not empty.exists() or not list(empty.glob("*.co"))  # nothing compiled or stored
    with pytest.raises(A.AccError, match="another node"):
        A.Kernel("(fold (elt i64) (fun (a b) (add a b)))", image=image)


# ------------------------------------------------------------- foreign functions
# (foreign "body" (fun (x ..) fallback) args ..): ForeignExp (PTX/Foreign.hs:74-84).
# The device runs the body; the oracle evaluates the fallback; they must agree.
BITREV32 = ("(let ((x1 (bor (band (shiftr x (const i64 1)) (const u32 0x55555555)) "
            "(shiftl (band x (const u32 0x55555555)) (const i64 1)))) "
            "(x2 (bor (band (shiftr x1 (const i64 2)) (const u32 0x33333333)) "
            "(shiftl (band x1 (const u32 0x33333333)) (const i64 2)))) "
            "(x3 (bor (band (shiftr x2 (const i64 4)) (const u32 0x0f0f0f0f)) "
            "(shiftl (band x2 (const u32 0x0f0f0f0f)) (const i64 4)))) "
            "(x4 (bor (band (shiftr x3 (const i64 8)) (const u32 0x00ff00ff)) "
            "(shiftl (band x3 (const u32 0x00ff00ff)) (const i64 8))))) "
            "(bor (shiftr x4 (const i64 16)) (shiftl x4 (const i64 16))))")


def test_foreign_scalar_function_in_a_generator(stream, kernels):
    """A device intrinsic outside the primitive set (bit reversal) as a foreign
    function, with a fallback built from shifts and masks."""
    node = ("(generate (elt u32 i64) (fun (i) (tuple "
            "(foreign \"return __builtin_bitreverse32(x);\" (fun (x) " + BITREV32 + ") (fromintegral u32 i)) "
            "(foreign \"return x * x + 7;\" (fun (x) (add (mul x x) (const i64 7))) i))))")
    assert "__builtin_bitreverse32" in A.node_source(node)
    k = kernels(node)
    n = (1 << 20) + 77
    out = [A.DeviceArray(n, np.uint32), A.DeviceArray(n, np.int64)]
    k.run(stream, None, out, 1, n)
    got, want = host(stream, out), J.generate(node, 1, n)
    assert same(got, want)
    assert got[0][1] == 1 << 31 and got[0][3] == 3 << 30  # reversal of 1 and 3


def test_foreign_function_in_a_fused_fold_input(stream, kernels):
    """The same foreign function inside a delayed producer fused into a fold,
    with a second foreign function of two arguments as the combine's operand."""
    node = ("(fold (elt u32) (fun (a b) (add a b)) (seed (const u32 3)) (input (fun (i) "
            "(foreign \"return (x < y) ? y : x;\" (fun (x y) (max x y)) "
            "(foreign \"return __builtin_bitreverse32(x);\" (fun (x) " + BITREV32 + ") (fromintegral u32 i)) "
            "(read 0 i)))) (free u32))")
    k = kernels(node)
    n = (1 << 22) + 5
    f = np.random.default_rng(4).integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32)
    df = A.DeviceArray.from_numpy(stream, f)
    out = A.DeviceArray(1, np.uint32)
    k.run(stream, None, [out], 1, n, free=[df])
    assert out.to_numpy(stream)[0] == J.fold(node, [], 1, n, free=[f])[0][0]


def test_foreign_function_errors():
    with pytest.raises(A.AccError, match="foreign"):  # fallback arity
        A.node_source("(generate (elt i64) (fun (i) (foreign \"return x;\" (fun (x y) x) i)))")
    with pytest.raises(A.AccError, match="foreign"):  # not a string body
        A.node_source("(generate (elt i64) (fun (i) (foreign body (fun (x) x) i)))")
    with pytest.raises(A.AccError, match="foreign"):  # Bool
        A.node_source("(generate (elt bool) (fun (i) (foreign \"return x;\" (fun (x) x) (lt i (const i64 3)))))")
