"""CPU oracle for the run-time compiled skeletons (test infrastructure only).

An interpreter of the skeleton-node expression language over numpy arrays,
with the reference's scalar semantics restated independently of the GPU
code generator:

  * accelerate-llvm/src/Data/Array/Accelerate/LLVM/CodeGen/Arithmetic.hs
    (add/sub/mul wrap :60-67; negate = x*(-1) :69-73; abs :75-86; signum
    :88-115; quot/rem truncate :121-125; div/mod floor :145-191; shifts
    :239-258; rotates :261-279; popCount/clz/ctz :281-315; floating ops
    :318-392; isNaN/isInfinite :395-410; truncate/round/floor/ceiling
    :413-431; ordered comparisons, LLVM/AST/Type/Instruction.hs:518-524;
    min/max :456-472; fromIntegral/toFloating :501-535),

and the skeletons' element order (native backend, the oracle of the GPU
path):

  * fold: seed on the left, elements in order (NAT/CodeGen/Fold.hs:126-228)
  * scan: scanl1/scanl/scanr1/scanr/scan' element order and output sizes
    (NAT/CodeGen/Scan.hs:54-270)
  * foldSeg over offsets (NAT/CodeGen/FoldSeg.hs:103-141)
  * permute: out[p(i)] = f(src[i], out[p(i)]) (NAT/CodeGen/Permute.hs:85-284)
  * stencil with the boundary maps of LLVM/CodeGen/Stencil.hs:224-253

For an associative combine a pairwise tree (fold) or a Hillis-Steele prefix
(scan) re-associates but keeps the element order, so integer results equal
the sequential definition; floating results are compared against an FP64
truth with a bound by the tests.  Division by zero, minBound `quot` (-1) and
out-of-range float-to-int conversions are undefined in the reference and
are not evaluated by the tests.
"""
from __future__ import annotations

import math

import numpy as np

DT = {
    "i8": np.int8, "i16": np.int16, "i32": np.int32, "i64": np.int64,
    "u8": np.uint8, "u16": np.uint16, "u32": np.uint32, "u64": np.uint64,
    "f32": np.float32, "f64": np.float64, "bool": np.bool_,
    "int": np.int64, "word": np.uint64, "float": np.float32, "double": np.float64,
}
STORE = {np.dtype(np.bool_): np.uint8}


def storage(t):
    d = np.dtype(DT[t])
    return STORE.get(d, d)


# ------------------------------------------------------------------ parser
class Quoted(str):
    """A string literal of a node (the body of a foreign function)."""


def parse(text):
    toks, i = [], 0
    while i < len(text):
        c = text[i]
        if c.isspace():
            i += 1
        elif c == '"':
            j, lit = i + 1, []
            while j < len(text) and text[j] != '"':
                if text[j] == "\\" and j + 1 < len(text) and text[j + 1] in '"\\':
                    j += 1
                lit.append(text[j])
                j += 1
            assert j < len(text), "unterminated string literal"
            toks.append(Quoted("".join(lit)))
            i = j + 1
        elif c == ";":
            while i < len(text) and text[i] != "\n":
                i += 1
        elif c in "()":
            toks.append(c)
            i += 1
        else:
            j = i
            while j < len(text) and not text[j].isspace() and text[j] not in "();":
                j += 1
            toks.append(text[i:j])
            i = j
    pos = 0

    def read():
        nonlocal pos
        t = toks[pos]
        pos += 1
        if t == "(" and not isinstance(t, Quoted):
            out = []
            while toks[pos] != ")":
                out.append(read())
            pos += 1
            return out
        return t

    e = read()
    assert pos == len(toks), "trailing tokens"
    return e


def _is(e, head):
    return isinstance(e, list) and e and e[0] == head


# ---------------------------------------------------------- scalar helpers
def _bits(dt):
    return np.dtype(dt).itemsize * 8


def _unsigned(dt):
    return np.dtype("u%d" % np.dtype(dt).itemsize)


def _signed(dt):
    return np.dtype(dt).kind == "i"


def _isf(dt):
    return np.dtype(dt).kind == "f"


def _lit(t, s):
    dt = np.dtype(DT[t])
    if dt == np.bool_:
        return np.bool_(s in ("true", "1"))
    if dt.kind == "f":
        return dt.type(float.fromhex(s) if "0x" in s.lower() and "p" in s.lower() else float(s))
    v = int(s, 0)
    return np.array(v % (1 << _bits(dt)), dtype=_unsigned(dt)).view(dt)[()]


def _vec_quot(a, b):
    # truncation toward zero from floor division: adjust where signs differ and the division is inexact
    with np.errstate(all="ignore"):
        if not _signed(a.dtype):
            return (a // b).astype(a.dtype)
        q = a // b
        r = a - q * b
        fix = (r != 0) & ((a < 0) != (b < 0))
        return (q + fix.astype(a.dtype)).astype(a.dtype)


def _shamt(s, dt):
    return (np.asarray(s).astype(np.int64).astype(dt).astype(np.int64) & (_bits(dt) - 1)).astype(np.int64)


def _popcount(x):
    u = x.astype(_unsigned(x.dtype))
    b = np.unpackbits(u.view(np.uint8).reshape(u.shape + (u.dtype.itemsize,)), axis=-1)
    return b.sum(axis=-1).astype(np.int64)


def _clz(x):
    u = x.astype(_unsigned(x.dtype)).astype(np.uint64)
    nb = _bits(x.dtype)
    out = np.full(u.shape, nb, np.int64)
    for k in range(nb):  # highest set bit
        hit = ((u >> np.uint64(k)) & np.uint64(1)) != 0
        out = np.where(hit, nb - 1 - k, out)
    return out


def _ctz(x):
    u = x.astype(_unsigned(x.dtype)).astype(np.uint64)
    nb = _bits(x.dtype)
    out = np.full(u.shape, nb, np.int64)
    for k in range(nb - 1, -1, -1):  # lowest set bit
        hit = ((u >> np.uint64(k)) & np.uint64(1)) != 0
        out = np.where(hit, k, out)
    return out


def _round_half_away(x):
    t = np.trunc(x)
    frac = x - t  # exact
    return t + np.where(np.abs(frac) >= 0.5, np.sign(x), 0).astype(x.dtype)


class Interp:
    """Evaluates expressions over numpy arrays (one array per tuple field)."""

    def __init__(self, free=(), nb=None, free_types=()):
        self.free = [np.asarray(f) for f in free]
        self.free_bool = [t == "bool" for t in free_types]
        self.nb = nb  # stencil: dict (dy, dx) -> neighbour array

    def ev(self, e, env):
        if isinstance(e, str):
            if e not in env:
                raise KeyError("unbound " + e)
            return env[e]
        op = e[0]
        if op == "const":
            return [_lit(e[1], e[2])]
        if op == "let":
            env = dict(env)
            for name, ex in e[1]:
                env[name] = self.ev(ex, env)
            return self.ev(e[2], env)
        if op == "cond":
            c = self.ev(e[1], env)[0]
            a, b = self.ev(e[2], env), self.ev(e[3], env)
            return [np.where(c, x, y).astype(np.result_type(x, y)) for x, y in zip(a, b)]
        if op == "tuple":
            out = []
            for x in e[1:]:
                out += self.ev(x, env)
            return out
        if op == "prj":
            return [self.ev(e[2], env)[int(e[1])]]
        if op == "read":
            arr = self.free[int(e[1])]
            i = np.asarray(self.ev(e[2], env)[0]).astype(np.int64)
            v = arr[np.clip(i, 0, len(arr) - 1)] if len(arr) else np.zeros(np.shape(i), arr.dtype)
            k = int(e[1])
            return [v.astype(np.bool_) if k < len(self.free_bool) and self.free_bool[k] else v]
        if op == "foreign":
            # a foreign scalar function: its meaning is the fallback, applied
            # to the arguments in a closed environment (the body is device
            # code the oracle does not run)
            fun = e[2]
            args = [self.ev(x, env) for x in e[3:]]
            return self.ev(fun[2], dict(zip(fun[1], args)))
        if op == "nb":
            return [self.nb[(int(e[1]), int(e[2]))]]
        conv = op in ("fromintegral", "tofloating", "bitcast", "truncate", "round", "floor", "ceiling")
        if conv:
            tgt = np.dtype(DT[e[1]])
            x = np.asarray(self.ev(e[2], env)[0])
            with np.errstate(all="ignore"):
                if op == "bitcast":
                    return [np.ascontiguousarray(x).view(tgt)]
                if op in ("fromintegral", "tofloating"):
                    return [x.astype(tgt)]
                f = {"truncate": np.trunc, "round": _round_half_away, "floor": np.floor, "ceiling": np.ceil}[op]
                return [f(x).astype(tgt)]
        args = [np.asarray(self.ev(x, env)[0]) for x in e[1:]]
        with np.errstate(all="ignore"):
            return self.prim(op, args)

    def prim(self, op, a):
        x = a[0]
        dt = x.dtype
        y = a[1] if len(a) > 1 else None
        if op in ("add", "sub", "mul"):
            return [{"add": np.add, "sub": np.subtract, "mul": np.multiply}[op](x, y).astype(dt)]
        if op == "neg":  # x * (-1), the -1 wrapped into the type
            return [(x * np.array(-1, np.int64).astype(dt)[()]).astype(dt)]
        if op == "abs":
            if _isf(dt):
                return [np.abs(x)]
            if not _signed(dt):
                return [x]
            return [np.where(x < 0, (x * np.array(-1, np.int64).astype(dt)[()]).astype(dt), x).astype(dt)]
        if op == "signum":
            if _isf(dt):
                return [((x > 0).astype(dt) - (x < 0).astype(dt)).astype(dt)]
            return [((x > 0).astype(np.int64) - (x < 0).astype(np.int64)).astype(dt)] if _signed(dt) else \
                [(x != 0).astype(dt)]
        if op == "quot":
            return [_vec_quot(x, y)]
        if op == "rem":
            return [(x - (_vec_quot(x, y) * y)).astype(dt)]
        if op == "quotrem":
            q = _vec_quot(x, y)
            return [q, (x - q * y).astype(dt)]
        if op in ("div", "idiv"):
            return [(x // y).astype(dt)]  # floor division
        if op == "mod":
            return [(x % y).astype(dt)]  # sign of the divisor
        if op == "divmod":
            return [(x // y).astype(dt), (x % y).astype(dt)]
        if op in ("band", "bor", "bxor"):
            return [{"band": np.bitwise_and, "bor": np.bitwise_or, "bxor": np.bitwise_xor}[op](x, y)]
        if op == "bnot":
            return [np.invert(x)]
        if op in ("shiftl", "shiftr"):
            s = _shamt(y, dt)
            u = x.astype(_unsigned(dt))
            if op == "shiftl":
                return [np.left_shift(u, s.astype(u.dtype)).astype(_unsigned(dt)).view(dt)]
            return [np.right_shift(x, s.astype(dt) if _signed(dt) else s.astype(u.dtype))]
        if op in ("rotatel", "rotater"):
            nb = _bits(dt)
            s = (np.asarray(y).astype(np.int64).astype(dt).astype(_unsigned(dt)).astype(np.uint64) % np.uint64(nb))
            u = x.astype(_unsigned(dt)).astype(np.uint64)
            mask = np.uint64((1 << nb) - 1)
            if op == "rotater":
                s = (np.uint64(nb) - s) % np.uint64(nb)
            r = ((u << s) | (u >> ((np.uint64(nb) - s) % np.uint64(nb)))) & mask
            r = np.where(s == 0, u, r)
            return [r.astype(_unsigned(dt)).view(dt)]
        if op == "popcount":
            return [_popcount(x)]
        if op == "clz":
            return [_clz(x)]
        if op == "ctz":
            return [_ctz(x)]
        if op == "fdiv":
            return [(x / y).astype(dt)]
        if op == "recip":
            return [(dt.type(1) / x).astype(dt)]
        f1 = {"sin": np.sin, "cos": np.cos, "tan": np.tan, "asin": np.arcsin, "acos": np.arccos, "atan": np.arctan,
              "sinh": np.sinh, "cosh": np.cosh, "tanh": np.tanh, "asinh": np.arcsinh, "acosh": np.arccosh,
              "atanh": np.arctanh, "exp": np.exp, "sqrt": np.sqrt, "log": np.log}
        if op in f1:
            return [f1[op](x).astype(dt)]
        if op == "atan2":
            return [np.arctan2(x, y).astype(dt)]
        if op == "fpow":
            return [np.power(x, y).astype(dt)]
        if op == "logbase":
            return [(np.log(y) / np.log(x)).astype(dt)]
        if op == "isnan":
            return [np.isnan(x)]
        if op == "isinfinite":
            return [np.isinf(x)]
        if op in ("lt", "gt", "lte", "gte", "eq"):
            return [{"lt": np.less, "gt": np.greater, "lte": np.less_equal, "gte": np.greater_equal,
                     "eq": np.equal}[op](x, y)]
        if op == "neq":
            return [(x < y) | (x > y)] if _isf(dt) else [x != y]
        if op == "min":
            return [np.fmin(x, y)] if _isf(dt) else [np.where(x <= y, x, y)]
        if op == "max":
            return [np.fmax(x, y)] if _isf(dt) else [np.where(x >= y, x, y)]
        if op == "land":
            return [x & y]
        if op == "lor":
            return [x | y]
        if op == "lnot":
            return [~x]
        raise KeyError("unknown operator " + op)


# --------------------------------------------------------------- functions
def fun_params(fun):
    assert _is(fun, "fun")
    return fun[1], fun[2]


def node_items(node):
    kind = node[0]
    items = {}
    flags = set()
    fun = None
    for it in node[1:]:
        if isinstance(it, str):
            flags.add(it)
        elif it and it[0] == "fun":
            fun = it
        elif len(it) == 1:
            flags.add(it[0])
        else:
            items[it[0]] = it
    return kind, items, flags, fun


def _combiner(interp, fun, nf):
    ps, body = fun_params(fun)

    def f(a, b):
        env = {ps[0]: a, ps[1]: b}
        r = interp.ev(body, env)
        assert len(r) == nf
        return r
    return f


def _cast(fields, types, n=None):
    """Fields in their storage types; with n, constant (0-d) fields broadcast to n elements."""
    out = [np.asarray(v).astype(storage(t)) for v, t in zip(fields, types)]
    if n is not None:
        out = [np.broadcast_to(v, (n,)).copy() if v.ndim == 0 else v for v in out]
    return out


def _seed(interp, items, types):
    if "seed" not in items:
        return None
    return _cast(interp.ev(items["seed"][1], {}), types)


def _input(interp, items, inputs, n, types):
    """Element fields of the input: the manifest arrays, or the delayed producer at 0..n-1."""
    if "input" in items:
        ps, body = fun_params(items["input"][1])
        i = np.arange(n, dtype=np.int64)
        return _cast(interp.ev(body, {ps[0]: [i]}), types, n)
    return [np.asarray(x).astype(storage(t)) for x, t in zip(inputs, types)]


def _tree(f, xs):
    """In-order pairwise reduction of the rows' last axis: fields are (..., n) arrays."""
    while xs[0].shape[-1] > 1:
        n = xs[0].shape[-1]
        h = n // 2
        a = [x[..., 0:2 * h:2] for x in xs]
        b = [x[..., 1:2 * h:2] for x in xs]
        c = f(a, b)
        if n % 2:
            c = [np.concatenate([cc, x[..., -1:]], axis=-1) for cc, x in zip(c, xs)]
        xs = c
    return [x[..., 0] for x in xs]


def fold(node_text, inputs, outer, inner, free=()):
    node = parse(node_text)
    kind, items, flags, fun = node_items(node)
    types = items["elt"][1:]
    it = Interp(free)
    f = _combiner(it, fun, len(types))
    fs = lambda a, b: _cast(f(a, b), types)
    x = _input(it, items, inputs, outer * inner, types)
    seed = _seed(it, items, types)
    if inner == 0:
        assert seed is not None, "fold1 of an empty array"
        return [np.repeat(s, outer) for s in seed]
    x = [v.reshape(outer, inner) for v in x]
    r = _tree(fs, x)
    if seed is not None:
        r = fs([np.broadcast_to(s, r[0].shape) for s in seed], r)
    return r


def _hillis_steele(fs, xs, seg=None):
    """In-order inclusive prefix of an associative function along the last axis
    (a segmented one when seg gives each element's segment id)."""
    xs = [x.copy() for x in xs]
    n = xs[0].shape[-1]
    d = 1
    while d < n:
        a = [x[..., :-d] for x in xs]
        b = [x[..., d:] for x in xs]
        c = fs(a, b)
        if seg is not None:
            same = seg[..., d:] == seg[..., :-d]
            c = [np.where(same, cc, bb) for cc, bb in zip(c, b)]
        xs = [np.concatenate([x[..., :d], cc], axis=-1) for x, cc in zip(xs, c)]
        d *= 2
    return xs


def scan(node_text, inputs, outer, inner, free=()):
    """Returns (out fields, sums fields or None) with the reference's output sizes."""
    node = parse(node_text)
    kind, items, flags, fun = node_items(node)
    types = items["elt"][1:]
    it = Interp(free)
    f = _combiner(it, fun, len(types))
    rev = "dir" in items and items["dir"][1] == "r2l"
    prime = "prime" in flags
    fs = (lambda a, b: _cast(f(b, a), types)) if rev else (lambda a, b: _cast(f(a, b), types))
    seed = _seed(it, items, types)
    x = [v.reshape(outer, inner) for v in _input(it, items, inputs, outer * inner, types)]
    if rev:
        x = [v[:, ::-1] for v in x]
    if seed is not None:
        x = [np.concatenate([np.broadcast_to(s, (outer, 1)).astype(v.dtype), v], axis=1) for s, v in zip(seed, x)]
    inc = _hillis_steele(fs, x) if x[0].shape[-1] else x
    if seed is None:  # scanl1 / scanr1
        out = inc
        sums = None
    elif prime:  # scan': exclusive values, the total apart
        out = [v[:, :-1] for v in inc]
        sums = [v[:, -1] for v in inc]
    elif not rev:  # scanl z: n + 1 outputs, the total last
        out, sums = inc, None
    else:  # scanr z: traversal-inclusive values, then z
        out = [np.concatenate([v[:, 1:], v[:, :1]], axis=1) for v in inc]
        sums = None
    if rev:
        if seed is None or prime:
            out = [v[:, ::-1] for v in out]
        else:
            out = [np.concatenate([v[:, :-1][:, ::-1], v[:, -1:]], axis=1) for v in out]
    return [v.reshape(-1) for v in out], sums


def foldseg(node_text, inputs, outer, inner, offsets, free=()):
    node = parse(node_text)
    kind, items, flags, fun = node_items(node)
    types = items["elt"][1:]
    it = Interp(free)
    f = _combiner(it, fun, len(types))
    fs = lambda a, b: _cast(f(a, b), types)
    seed = _seed(it, items, types)
    offsets = np.asarray(offsets, np.int64)
    m = len(offsets) - 1
    x = _input(it, items, inputs, outer * inner, types)
    lens = np.diff(offsets)
    seg = np.repeat(np.arange(m), lens)
    outs = [np.zeros(outer * m, storage(t)) for t in types]
    have = np.zeros(outer * m, bool)
    for r in range(outer):
        xr = [v[r * inner + offsets[0]: r * inner + offsets[-1]] for v in x]
        if len(seg):
            inc = _hillis_steele(fs, xr, seg)
            ends = offsets[1:] - offsets[0] - 1
            ne = lens > 0
            for o, v in zip(outs, inc):
                o[r * m + np.nonzero(ne)[0]] = v[ends[ne]]
            have[r * m + np.nonzero(ne)[0]] = True
    if seed is not None:
        base = [np.broadcast_to(s, (outer * m,)) for s in seed]
        comb = fs(base, outs)
        outs = [np.where(have, c, b).astype(c.dtype) for c, b in zip(comb, base)]
    return outs, have


def permute(node_text, defaults, inputs, nsrc, free=()):
    """out[j] = f(src[i], out[j]) for every i whose projection j is in range:
    the sources of one slot combined in order, then with the default (equal to
    any order for an associative and commutative f)."""
    node = parse(node_text)
    kind, items, flags, fun = node_items(node)
    types = items["elt"][1:]
    it = Interp(free)
    f = _combiner(it, fun, len(types))
    fs = lambda a, b: _cast(f(a, b), types)
    m = len(defaults[0])
    src = _input(it, items, inputs, nsrc, types)
    ps, body = fun_params(items["project"][1])
    j = np.asarray(it.ev(body, {ps[0]: [np.arange(nsrc, dtype=np.int64)]})[0]).astype(np.int64)
    keep = (j >= 0) & (j < m)
    j = j[keep]
    src = [v[keep] for v in src]
    order = np.argsort(j, kind="stable")
    j = j[order]
    src = [v[order] for v in src]
    out = [np.array(d, copy=True).astype(storage(t)) for d, t in zip(defaults, types)]
    if len(j):
        inc = _hillis_steele(fs, src, j)
        last = np.nonzero(np.append(j[1:] != j[:-1], True))[0]
        slots = j[last]
        agg = [v[last] for v in inc]
        cur = [o[slots] for o in out]
        new = fs(agg, cur)
        for o, v in zip(out, new):
            o[slots] = v
    return out


def _bmap(b, i, n):
    if b == "clamp":
        i = np.where(i < 0, 0, np.where(i >= n, n - 1, i))
    elif b == "mirror":
        i = np.where(i < 0, -i, np.where(i >= n, n - (i - n + 2), i))
    elif b == "wrap":
        i = np.where(i < 0, n + i, np.where(i >= n, i - n, i))
    return np.clip(i, 0, n - 1)


def stencil(node_text, img, free=()):
    node = parse(node_text)
    kind, items, flags, fun = node_items(node)
    ti, to = items["in"][1], items["out"][1]
    ry, rx = int(items["radius"][1]), int(items["radius"][2])
    b = items["boundary"][1]
    img = np.asarray(img).astype(storage(ti))
    H, W = img.shape
    yy, xx = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    nb = {}
    it0 = Interp(free)
    for dy in range(-ry, ry + 1):
        for dx in range(-rx, rx + 1):
            y, x = yy + dy, xx + dx
            inside = (y >= 0) & (y < H) & (x >= 0) & (x < W)
            if isinstance(b, str):
                v = img[_bmap(b, y, H), _bmap(b, x, W)]
            elif b[0] == "constant":
                c = np.asarray(it0.ev(b[1], {})[0]).astype(img.dtype)
                v = np.where(inside, img[np.clip(y, 0, H - 1), np.clip(x, 0, W - 1)], c)
            else:
                ps, body = fun_params(b[1])
                fv = np.asarray(it0.ev(body, {ps[0]: [y.astype(np.int64)], ps[1]: [x.astype(np.int64)]})[0])
                v = np.where(inside, img[np.clip(y, 0, H - 1), np.clip(x, 0, W - 1)], fv.astype(img.dtype))
            nb[(dy, dx)] = v
    it = Interp(free, nb)
    ps, body = fun_params(fun)
    return np.asarray(it.ev(body, {})[0]).astype(storage(to))


def generate(node_text, outer, inner, free=()):
    node = parse(node_text)
    kind, items, flags, fun = node_items(node)
    types = items["elt"][1:]
    it = Interp(free)
    ps, body = fun_params(fun)
    i = np.arange(outer * inner, dtype=np.int64)
    env = {ps[0]: [i]} if len(ps) == 1 else {ps[0]: [i // max(inner, 1)], ps[1]: [i % max(inner, 1)]}
    return _cast(it.ev(body, env), types, outer * inner)


def evaluate(expr_text, env_arrays, free=()):
    """One expression over named arrays (for the primitive tests)."""
    return Interp(free).ev(parse(expr_text), {k: [np.asarray(v)] for k, v in env_arrays.items()})
