"""Batched execution: many uneven scans, folds, segmented folds, histograms
and stencils in one persistent launch.

Each problem's result must equal its standalone result (integers bit-exact,
floating sums within the FP32 bound of the FP64 truth), for every shape
(workgroup size x vectors per thread x tiles per ticket), for the autotuned
batch and for replays of a captured graph.
"""
import math

import numpy as np
import pytest

import acc_hip as A
import oracle_py as O

pytestmark = pytest.mark.gpu

ADD, MAX = A.ADD, A.MAX
TYPES = [np.int32, np.int64, np.float32, np.float64]


def make_problems(stream, count, seed, max_log2=18, kinds=(A.BATCH_SCAN, A.BATCH_FOLD)):
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(0, math.log(2 ** max_log2), size=count)).astype(np.int64) + 1
    probs = []
    for i, n in enumerate(sizes):
        dt = TYPES[rng.integers(0, 4)]
        kind = kinds[rng.integers(0, len(kinds))]
        op = ADD if (rng.integers(0, 3) or kind == A.BATCH_HISTOGRAM) else MAX
        seed_v = None if rng.integers(0, 2) == 0 else dt(rng.integers(-5, 5))
        if np.dtype(dt).kind == "f":
            x = (rng.integers(-64, 64, size=n) / 16).astype(dt) if rng.integers(0, 2) else rng.standard_normal(n).astype(dt)
        else:
            x = rng.integers(-1000, 1000, size=n).astype(dt)
        p = dict(kind=kind, op=op, seed=seed_v, x=x, n=int(n), dt=dt, aux=None, m=0, fill=0)
        if kind == A.BATCH_FOLD_SEG:
            m = int(rng.integers(1, max(2, n // 8)))
            lens = rng.multinomial(int(n), np.ones(m) / m)  # some segments empty
            off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64 if rng.integers(0, 2) else np.int32)
            p.update(aux=A.DeviceArray.from_numpy(stream, off), off=off, m=m)
            on = m
        elif kind == A.BATCH_HISTOGRAM:
            m = int(rng.integers(1, 4097))
            idx = rng.integers(-10, m + 10, size=n).astype(np.int32 if rng.integers(0, 2) else np.int64)
            p.update(x=idx, m=m, fill=dt(rng.integers(1, 4)), seed=dt(rng.integers(-5, 5)))
            on = m
        else:
            on = 1 if kind == A.BATCH_FOLD else (n + 1 if seed_v is not None else n)
        p["src"] = A.DeviceArray.from_numpy(stream, p["x"])
        p["out"] = A.DeviceArray(on, dt)
        if kind == A.BATCH_FOLD_SEG:  # fold1Seg leaves empty segments: start from a known value
            p["out"] = A.DeviceArray.from_numpy(stream, np.full(on, 7, dt))
            p["init"] = 7
        probs.append(p)
    return probs


def expected(p):
    x, op, seed = p["x"], p["op"], p["seed"]
    if p["kind"] == A.BATCH_FOLD:
        return O.fold(op, seed, x)
    if p["kind"] == A.BATCH_FOLD_SEG:
        want = O.fold_seg(op, seed, x, p["off"].astype(np.int64))
        if seed is None:
            want = np.where(np.diff(p["off"].astype(np.int64)) == 0, p["dt"](p["init"]), want)
        return want
    if p["kind"] == A.BATCH_HISTOGRAM:
        m = p["m"]
        keep = (x >= 0) & (x < m)
        return (p["seed"] + np.bincount(x[keep], minlength=m).astype(np.float64) * float(p["fill"])).astype(p["dt"])
    return O.scan(op, 0, seed, x)


def check(stream, p, got):
    want = expected(p)
    if np.dtype(p["dt"]).kind != "f" or p["op"] == MAX or p["kind"] == A.BATCH_HISTOGRAM:
        # histogram bins of small integer counts x fill are exact in floating point too
        assert np.array_equal(got, want), (p["kind"], p["dt"], p["op"], p["n"])
        return
    if p["kind"] == A.BATCH_FOLD_SEG:
        off = p["off"].astype(np.int64)
        truth, ab = O.fold_seg_f64(p["x"], off, None if p["seed"] is None else float(p["seed"]))
        lens = np.diff(off)
        bound = 1e-5 * np.maximum(1, np.log2(np.maximum(lens, 2))) * ab + 1e-9
        ok = np.abs(got.astype(np.float64) - truth) <= bound
        if p["seed"] is None:
            ok |= lens == 0
        assert np.all(ok), (p["kind"], p["dt"], p["n"])
        return
    # floating (+): normwise bound against the FP64 truth
    x = p["x"].astype(np.float64)
    s0 = 0.0 if p["seed"] is None else float(p["seed"])
    if p["kind"] == A.BATCH_FOLD:
        truth, ab = s0 + math.fsum(x), abs(s0) + math.fsum(np.abs(x))
        assert abs(float(got[0]) - truth) <= 1e-5 * max(1, math.log2(max(p["n"], 2))) * ab + 1e-9
    else:
        c = np.cumsum(x) + s0
        ab = np.cumsum(np.abs(x)) + abs(s0)
        if p["seed"] is not None:  # exclusive scanl: [seed, ...] shifted
            c = np.concatenate([[s0], c])
            ab = np.concatenate([[abs(s0)], ab])
        k = np.arange(1, c.size + 1)
        bound = 1e-5 * np.maximum(1.0, np.log2(np.maximum(k, 2))) * ab + 1e-9
        assert np.all(np.abs(got.astype(np.float64) - c) <= bound)


def batch_of(probs):
    return [(p["kind"], p["op"], p["seed"], p["src"], p["n"], p["out"], p["aux"], p["m"], p["fill"]) for p in probs]


ALL_KINDS = (A.BATCH_SCAN, A.BATCH_FOLD, A.BATCH_FOLD_SEG, A.BATCH_HISTOGRAM)


@pytest.mark.parametrize("tile", [2, 4, 8])
def test_batch_matches_oracle(stream, tile):
    probs = make_problems(stream, 150, seed=tile, kinds=ALL_KINDS)
    b = A.Batch(0, batch_of(probs), tile=tile)
    assert b.tile()[0] == tile
    b.launch(stream)
    stream.synchronize()
    for p in probs:
        check(stream, p, p["out"].to_numpy(stream))


def test_batch_launch_right_after_create(stream):
    """Creating a batch zeroes its workspace (tickets, status words); the first
    launch follows at once on a non-blocking stream and must see that zeroing
    complete.  A batch large enough that the zeroing is not instantaneous."""
    probs = make_problems(stream, 1200, seed=29, max_log2=18)
    for tile in (8, 4):
        b = A.Batch(0, batch_of(probs), tile=tile)
        b.launch(stream)
        stream.synchronize()
        for p in probs:
            check(stream, p, p["out"].to_numpy(stream))
        b.destroy()


def test_batch_same_alone_and_in_mix(stream):
    """Every problem's batched result against the same problem through the plain
    C calls (acc_hip_batch_launch_plain): bit-exact where the arithmetic is."""
    probs = make_problems(stream, 80, seed=11, kinds=ALL_KINDS)
    b = A.Batch(0, batch_of(probs), tile=4)
    b.launch(stream)
    stream.synchronize()
    mixed = [p["out"].to_numpy(stream) for p in probs]
    b.launch_plain(stream)
    stream.synchronize()
    for p, mx in zip(probs, mixed):
        alone = p["out"].to_numpy(stream)
        if np.dtype(p["dt"]).kind != "f" or p["op"] == MAX or p["kind"] == A.BATCH_HISTOGRAM:
            assert np.array_equal(mx, alone), (p["kind"], p["dt"], p["n"])
        check(stream, p, mx)


def test_batch_autotuned(stream, tmp_path, monkeypatch):
    monkeypatch.setenv("ACC_HIP_TUNE_CACHE", str(tmp_path / "tune.txt"))
    probs = make_problems(stream, 80, seed=5, kinds=ALL_KINDS)
    b = A.Batch(0, batch_of(probs), tile=0)
    block, k, run, tickets = b.shape()
    assert block in (256, 512) and k in (2, 4, 8) and run in (1, 4, 16) and tickets > 0
    assert (tmp_path / "tune.txt").read_text().strip().endswith(f"{block} {k} {run}")
    b.launch(stream)
    stream.synchronize()
    for p in probs:
        check(stream, p, p["out"].to_numpy(stream))
    # a second batch with the same signature reuses the stored choice
    b2 = A.Batch(0, batch_of(probs), tile=0)
    assert b2.shape() == (block, k, run, tickets)


def test_batch_graph_replays(stream):
    probs = make_problems(stream, 40, seed=21)
    b = A.Batch(0, batch_of(probs), tile=2)
    g = A.Graph.capture(stream, lambda: b.launch(stream))
    other = A.DeviceArray.from_numpy(stream, np.arange(100000, dtype=np.int32))
    other_out = A.DeviceArray(100000, np.int32)
    for rep in range(4):
        for p in probs:  # clear outputs eagerly between replays
            A.check(A.lib().acc_hip_fill_async(stream.handle, A.c_void_p(p["out"].ptr),
                                               np.zeros(1, np.uint64).ctypes.data_as(A.c_void_p),
                                               p["out"].dtype.itemsize, p["out"].count))
        A.scan(stream, ADD, A.LEFT_TO_RIGHT, None, other, 1, 100000, other_out)  # eager work in between
        g.launch(stream)
        stream.synchronize()
        for p in probs:
            check(stream, p, p["out"].to_numpy(stream))
    assert np.array_equal(other_out.to_numpy(stream), np.cumsum(np.arange(100000, dtype=np.int32), dtype=np.int32))


def test_batch_rejects_bad_problems(stream):
    x = A.DeviceArray.from_numpy(stream, np.arange(10, dtype=np.int32))
    o = A.DeviceArray(10, np.int32)
    with pytest.raises(A.AccError):
        A.Batch(0, [(A.BATCH_SCAN, A.MUL, None, x, 10, o)], tile=2)
    with pytest.raises(A.AccError):
        A.Batch(0, [(A.BATCH_SCAN, ADD, None, x.view(1, 9), 9, o)], tile=2)  # misaligned input
    with pytest.raises(A.AccError):
        A.Batch(0, [(A.BATCH_HISTOGRAM, ADD, 0, x, 10, o, None, 5000, 1)], tile=2)  # too many bins


def make_stencils(stream, count, seed):
    """3 x 3 stencils of every boundary and element type, on images from
    1 x 1 to a few hundred rows and columns (wider than one 128-column strip,
    taller than one run, ragged at both edges)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        dt = TYPES[i % 4]
        r = 1
        bnd = [A.CLAMP, A.MIRROR, A.WRAP, A.CONSTANT][(i // 4) % 4]
        h = int(rng.integers(1, 300)) if i % 5 else int(rng.integers(1, 4))
        w = int(rng.integers(1, 700)) if i % 7 else int(rng.integers(1, 4))
        k = 2 * r + 1
        if np.dtype(dt).kind == "f":
            img = rng.standard_normal((h, w)).astype(dt)
            wt = rng.standard_normal((k, k)).astype(dt)
            cval = dt(rng.standard_normal())
        else:
            img = rng.integers(-1000, 1000, size=(h, w)).astype(dt)
            wt = rng.integers(-9, 10, size=(k, k)).astype(dt)
            cval = dt(rng.integers(-50, 50))
        out.append(dict(kind=A.BATCH_STENCIL, op=ADD, seed=None, x=img, n=h, m=w, dt=dt, r=r, bnd=bnd, wt=wt,
                        cval=cval, src=A.DeviceArray.from_numpy(stream, img),
                        aux=A.DeviceArray.from_numpy(stream, wt.reshape(-1)), fill=cval,
                        params=r | (r << 8) | (bnd << 16), out=A.DeviceArray(h * w, dt)))
    return out


def stencil_ok(p, got):
    """Integers wrap exactly; floating results within taps * eps * sum |w x|
    of the float64 oracle (the oracle sums in tuple order; a fused
    multiply-add may round differently)."""
    r, bnd, wt, img = p["r"], p["bnd"], p["wt"], p["x"]
    want = O.stencil2d(r, r, wt, bnd, p["cval"], img).reshape(-1)
    if np.dtype(p["dt"]).kind != "f":
        return np.array_equal(got, want)
    w64, x64 = wt.astype(np.float64), img.astype(np.float64)
    truth = O.stencil2d(r, r, w64, bnd, float(p["cval"]), x64).reshape(-1)
    ab = O.stencil2d(r, r, np.abs(w64), bnd, abs(float(p["cval"])), np.abs(x64)).reshape(-1)
    eps = 2.0 ** -23 if p["dt"] == np.float32 else 2.0 ** -52
    taps = (2 * r + 1) ** 2
    return np.all(np.abs(got.astype(np.float64) - truth) <= 2 * taps * eps * ab + 1e-300)


def stencil_tuple(p):
    return (p["kind"], p["op"], None, p["src"], p["n"], p["out"], p["aux"], p["m"], p["fill"], p["params"])


@pytest.mark.parametrize("tile", [2, 8])
def test_batch_stencils_match_oracle_and_plain(stream, tile):
    """Stencils alone and mixed with the other kinds in one batch; each result
    against the oracle, integers also bit-exact against the plain
    acc_hip_stencil2d call on the same problem."""
    st = make_stencils(stream, 64, seed=100 + tile)
    others = make_problems(stream, 40, seed=200 + tile, kinds=ALL_KINDS)
    b = A.Batch(0, batch_of(others) + [stencil_tuple(p) for p in st], tile=tile)
    b.launch(stream)
    stream.synchronize()
    mixed = [p["out"].to_numpy(stream) for p in st]
    for p, got in zip(st, mixed):
        assert stencil_ok(p, got), (np.dtype(p["dt"]).name, p["r"], p["bnd"], p["n"], p["m"])
    for p in others:
        check(stream, p, p["out"].to_numpy(stream))
    b.launch_plain(stream)
    stream.synchronize()
    for p, got in zip(st, mixed):
        if np.dtype(p["dt"]).kind != "f":
            assert np.array_equal(got, p["out"].to_numpy(stream)), (np.dtype(p["dt"]).name, p["r"], p["bnd"])


def test_batch_stencil_rejects_unsupported(stream):
    x = A.DeviceArray.from_numpy(stream, np.zeros(64, np.float32))
    o = A.DeviceArray(64, np.float32)
    wt = A.DeviceArray.from_numpy(stream, np.ones(15, np.float32))
    with pytest.raises(A.AccError):  # 3 x 5
        A.Batch(0, [(A.BATCH_STENCIL, ADD, None, x, 8, o, wt, 8, 0.0, 1 | (2 << 8) | (A.CLAMP << 16))], tile=2)
    wt5 = A.DeviceArray.from_numpy(stream, np.ones(25, np.float32))
    with pytest.raises(A.AccError):  # 5 x 5: the plain stencil's
        A.Batch(0, [(A.BATCH_STENCIL, ADD, None, x, 8, o, wt5, 8, 0.0, 2 | (2 << 8) | (A.CLAMP << 16))], tile=2)
    with pytest.raises(A.AccError):  # the Function boundary
        A.Batch(0, [(A.BATCH_STENCIL, ADD, None, x, 8, o, wt, 8, 0.0, 1 | (1 << 8) | (A.FUNCTION << 16))], tile=2)
