"""Batched execution: many uneven scans/folds in one persistent launch.

Each problem's result must equal its standalone result (integers bit-exact,
floating sums within the FP32 bound of the FP64 truth), for every tile size,
for the autotuned batch and for replays of a captured graph.
"""
import math

import numpy as np
import pytest

import acc_hip as A
import oracle_py as O

pytestmark = pytest.mark.gpu

ADD, MAX = A.ADD, A.MAX
TYPES = [np.int32, np.int64, np.float32, np.float64]


def make_problems(stream, count, seed, max_log2=18):
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(0, math.log(2 ** max_log2), size=count)).astype(np.int64) + 1
    probs = []
    for i, n in enumerate(sizes):
        dt = TYPES[rng.integers(0, 4)]
        kind = A.BATCH_SCAN if rng.integers(0, 2) == 0 else A.BATCH_FOLD
        op = ADD if rng.integers(0, 3) else MAX
        seed_v = None if rng.integers(0, 2) == 0 else dt(rng.integers(-5, 5))
        if np.dtype(dt).kind == "f":
            x = (rng.integers(-64, 64, size=n) / 16).astype(dt) if rng.integers(0, 2) else rng.standard_normal(n).astype(dt)
        else:
            x = rng.integers(-1000, 1000, size=n).astype(dt)
        src = A.DeviceArray.from_numpy(stream, x)
        on = 1 if kind == A.BATCH_FOLD else (n + 1 if seed_v is not None else n)
        out = A.DeviceArray(on, dt)
        probs.append(dict(kind=kind, op=op, seed=seed_v, x=x, src=src, out=out, n=int(n), dt=dt))
    return probs


def expected(p):
    x, op, seed = p["x"], p["op"], p["seed"]
    if p["kind"] == A.BATCH_FOLD:
        return O.fold(op, seed, x)
    return O.scan(op, 0, seed, x)


def check(stream, p, got):
    want = expected(p)
    if np.dtype(p["dt"]).kind != "f" or p["op"] == MAX:
        assert np.array_equal(got, want), (p["kind"], p["dt"], p["op"], p["n"])
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
    return [(p["kind"], p["op"], p["seed"], p["src"], p["n"], p["out"]) for p in probs]


@pytest.mark.parametrize("tile", [1, 2, 4])
def test_batch_matches_oracle(stream, tile):
    probs = make_problems(stream, 150, seed=tile)
    b = A.Batch(0, batch_of(probs), tile=tile)
    assert b.tile()[0] == tile
    b.launch(stream)
    stream.synchronize()
    for p in probs:
        check(stream, p, p["out"].to_numpy(stream))


def test_batch_same_alone_and_in_mix(stream):
    probs = make_problems(stream, 60, seed=11)
    b = A.Batch(0, batch_of(probs), tile=2)
    b.launch(stream)
    stream.synchronize()
    for p in probs:
        mixed = p["out"].to_numpy(stream)
        alone = A.DeviceArray(p["out"].count, p["dt"])
        if p["kind"] == A.BATCH_FOLD:
            A.fold(stream, p["op"], p["seed"], p["src"], 1, p["n"], alone)
        else:
            A.scan(stream, p["op"], A.LEFT_TO_RIGHT, p["seed"], p["src"], 1, p["n"], alone)
        alone = alone.to_numpy(stream)
        if np.dtype(p["dt"]).kind != "f" or p["op"] == MAX:
            assert np.array_equal(mixed, alone)
        check(stream, p, mixed)


def test_batch_autotuned(stream, tmp_path, monkeypatch):
    monkeypatch.setenv("ACC_HIP_TUNE_CACHE", str(tmp_path / "tune.txt"))
    probs = make_problems(stream, 80, seed=5)
    b = A.Batch(0, batch_of(probs), tile=0)
    k, tiles = b.tile()
    assert k in (1, 2, 4) and tiles > 0
    assert (tmp_path / "tune.txt").read_text().strip().endswith(str(k))
    b.launch(stream)
    stream.synchronize()
    for p in probs:
        check(stream, p, p["out"].to_numpy(stream))
    # a second batch with the same signature reuses the stored choice
    b2 = A.Batch(0, batch_of(probs), tile=0)
    assert b2.tile()[0] == k


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
        A.Batch(0, [(A.BATCH_SCAN, A.MUL, None, x, 10, o)], tile=1)
    with pytest.raises(A.AccError):
        A.Batch(0, [(A.BATCH_SCAN, ADD, None, x.view(1, 9), 9, o)], tile=1)  # misaligned input
