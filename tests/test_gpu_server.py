"""Serving loop: requests admitted as slots and memory free up, overlapping
on the device; each result equals the same request run alone."""
import time

import numpy as np
import pytest

import acc_hip as A
import oracle_py as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pool(stream):
    """Device-resident inputs shared by the requests."""
    n = 1 << 20
    p = {}
    p["f32"] = A.DeviceArray.from_numpy(stream, O.gen_normal_f32(1, n))
    p["i32"] = A.DeviceArray.from_numpy(stream, O.gen_uniform_i32(2, -1000, 1000, n))
    p["i64"] = A.DeviceArray.from_numpy(stream, O.gen_uniform_i32(3, -1000, 1000, n).astype(np.int64))
    lens = O.gen_segment_lengths(4, n // 1024, 4096, n)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    p["off"], p["off_host"] = A.DeviceArray.from_numpy(stream, off), off
    p["idx_host"] = O.gen_zipf_i32(5, 1.1, 4096, n)
    p["idx"] = A.DeviceArray.from_numpy(stream, p["idx_host"])
    p["host"] = {k: p[k].to_numpy(stream) for k in ("f32", "i32", "i64")}
    return p


def mix(pool, count, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        key = ["f32", "i32", "i64"][rng.integers(0, 3)]
        dt = {"f32": np.float32, "i32": np.int32, "i64": np.int64}[key]
        src = pool[key]
        kind = rng.choice(4, p=[0.35, 0.30, 0.20, 0.15])
        n = int(np.exp(rng.uniform(np.log(1 << 10), np.log(1 << 20))))
        if kind == 0:
            reqs.append(("scan", dt, key, n, A.Server.request(A.REQ_SCAN, dt, A.ADD, src=src, n=n)))
        elif kind == 1:
            off = pool["off_host"]
            nseg = max(1, int(np.searchsorted(off, n, side="right")) - 1)
            ne = int(off[nseg])
            reqs.append(("fold1seg", dt, key, (ne, nseg),
                         A.Server.request(A.REQ_FOLD_SEG, dt, A.ADD, src=src, n=ne, m=nseg, aux=pool["off"])))
        elif kind == 2:
            reqs.append(("hist", dt, key, n,
                         A.Server.request(A.REQ_PERMUTE, dt, A.ADD, seed=0, n=n, m=4096, aux=pool["idx"], fill=1)))
        else:
            side = max(3, int(np.sqrt(n)))
            w = (np.outer([1, 2, 1], [1, 2, 1]) / (16 if key == "f32" else 1)).astype(dt)
            reqs.append(("stencil", dt, key, side,
                         A.Server.request(A.REQ_STENCIL2D, dt, src=src, n=side, m=side, ry=1, rx=1,
                                          boundary=A.MIRROR, weights=w)))
    return reqs


def expected(pool, kind, dt, key, size):
    x = pool["host"][key]
    if kind == "scan":
        return O.scan(A.ADD, 0, None, x[:size])
    if kind == "scanr_max":
        return O.scan(A.MAX, 1, None, x[:size])
    if kind == "fold1seg_min":
        ne, nseg = size
        return O.fold_seg(A.MIN, None, x[:ne], pool["off_host"][:nseg + 1])
    if kind == "fold1seg":
        ne, nseg = size
        return O.fold_seg(A.ADD, None, x[:ne], pool["off_host"][:nseg + 1])
    if kind == "hist":
        return np.bincount(pool["idx_host"][:size], minlength=4096).astype(dt)
    side = size
    w = (np.outer([1, 2, 1], [1, 2, 1]) / (16 if key == "f32" else 1)).astype(dt)
    return O.stencil2d(1, 1, w, A.MIRROR, None, x[:side * side].reshape(side, side)).reshape(-1)


def agree(got, want, kind, dt, pool, key, size):
    if np.dtype(dt).kind != "f" or kind in ("hist", "scanr_max", "fold1seg_min"):
        return np.array_equal(got, want)
    x = pool["host"][key]
    g = got.astype(np.float64)
    if kind == "scan":
        truth, ab = O.scan_f64(x[:size])
        k = np.arange(1, size + 1)
        return np.all(np.abs(g - truth) <= 1e-5 * np.maximum(1, np.log2(np.maximum(k, 2))) * ab + 1e-12)
    if kind == "fold1seg":
        ne, nseg = size
        off = pool["off_host"][:nseg + 1]
        truth, ab = O.fold_seg_f64(x[:ne], off)
        lens = np.diff(off)
        return np.all(np.abs(g - truth) <= 1e-5 * np.maximum(1, np.log2(np.maximum(lens, 2))) * ab + 1e-12)
    side = size
    w = (np.outer([1, 2, 1], [1, 2, 1]) / 16).astype(np.float32)
    truth, ab = O.stencil2d_f64(1, 1, w, A.MIRROR, None, x[:side * side].reshape(side, side))
    return np.all(np.abs(g - truth.reshape(-1)) <= 9 * 2.0 ** -23 * ab.reshape(-1) + 1e-30)


def test_server_results_match_alone(stream, pool):
    srv = A.Server(0, slots=4)
    reqs = mix(pool, 60, seed=1)
    ids = [srv.submit(r, keep=True) for (_, _, _, _, r) in reqs]
    for rid, (kind, dt, key, size, r) in zip(ids, reqs):
        ptr, lat = srv.wait(rid)
        count = {"scan": size, "hist": 4096}.get(kind, None)
        if kind == "fold1seg":
            count = size[1]
        if kind == "stencil":
            count = size * size
        got = A.copy_to_host(stream, ptr, count, dt)
        want = expected(pool, kind, dt, key, size)
        assert agree(got, want, kind, dt, pool, key, size), (kind, np.dtype(dt).name, size)
        assert lat > 0
        srv.release(rid)
    st = srv.stats()
    assert st.completed == 60 and st.failed == 0 and st.p99_us >= st.p50_us > 0
    srv.destroy()


def test_server_admission_waits_for_memory(stream, pool):
    # an arena that holds only a few outputs at a time: requests must queue for memory
    srv = A.Server(0, slots=8, hbm_budget=3 << 20)
    n = 1 << 18  # 1 MiB of Int32 output per scan
    x = pool["i32"]
    ids = [srv.submit(A.Server.request(A.REQ_SCAN, np.int32, A.ADD, src=x, n=n), keep=True) for _ in range(12)]
    want = O.scan(A.ADD, 0, None, pool["host"]["i32"][:n])
    for rid in ids:
        ptr, _ = srv.wait(rid)
        assert np.array_equal(A.copy_to_host(stream, ptr, n, np.int32), want)
        srv.release(rid)
    assert srv.stats().completed == 12
    with pytest.raises(A.OutOfMemory):
        srv.submit(A.Server.request(A.REQ_SCAN, np.int32, A.ADD, src=x, n=1 << 20), keep=True)
    srv.destroy()


def test_server_future_arrivals_and_drain(stream, pool):
    srv = A.Server(0, slots=2)
    x = pool["f32"]
    for k in range(20):
        srv.submit(A.Server.request(A.REQ_FOLD, np.float32, A.MAX, src=x, n=1000 + k), arrival_offset_ns=k * 200_000)
    srv.drain()
    st = srv.stats()
    assert st.completed == 20 and st.span_s >= 0.0038
    srv.destroy()


def test_server_wait_on_a_request_that_is_not_kept(stream, pool):
    """A waiter on a keep=False request must not touch the record the worker
    retires on completion: it returns success and no output address."""
    srv = A.Server(0, slots=2)
    x = pool["i32"]
    ids = [srv.submit(A.Server.request(A.REQ_SCAN, np.int32, A.ADD, src=x, n=(1 << 20) - k)) for k in range(8)]
    for rid in ids:
        try:
            ptr, lat = srv.wait(rid)
        except A.AccError as e:
            # already completed and retired before the wait: the id is unknown
            assert e.status == A.INVALID_VALUE
            continue
        assert ptr is None and lat > 0
    srv.drain()
    assert srv.stats().completed == 8
    srv.destroy()


def out_count(kind, size):
    if kind in ("fold1seg", "fold1seg_min"):
        return size[1]
    if kind == "hist":
        return 4096
    if kind == "stencil":
        return size * size
    return size


def test_fused_rounds_same_alone_and_in_the_mix(stream, pool):
    """All requests arrive at once, so slots take them in fused rounds (one
    batched launch per round); requests without a batched form (a right-to-
    left scan, a min fold1Seg) run through the plain calls beside them.
    Every result agrees with the oracle, and each request run alone (a round
    of its own) gives the same result: bit-exact for integers, histograms and
    stencils, within the FP32 bound for floating (+) scans and folds (the
    look-back may associate the tile sums differently from run to run)."""
    reqs = mix(pool, 90, seed=2)
    off = pool["off_host"]
    for k, key in enumerate(("f32", "i32", "i64")):
        dt = {"f32": np.float32, "i32": np.int32, "i64": np.int64}[key]
        n = 5000 + 777 * k
        reqs.append(("scanr_max", dt, key, n,
                     A.Server.request(A.REQ_SCAN, dt, A.MAX, src=pool[key], n=n, direction=1)))
        nseg = 300 + k
        reqs.append(("fold1seg_min", dt, key, (int(off[nseg]), nseg),
                     A.Server.request(A.REQ_FOLD_SEG, dt, A.MIN, src=pool[key], n=int(off[nseg]), m=nseg,
                                      aux=pool["off"])))
    order = np.random.default_rng(3).permutation(len(reqs))
    reqs = [reqs[i] for i in order]
    srv = A.Server(0, slots=2)
    # one arrival instant for all (30 ms ahead), so the rounds are formed
    # from the whole set rather than from whatever the submit loop has reached
    t_arrive = time.perf_counter_ns() + 30_000_000
    ids = [srv.submit(r, keep=True, arrival_offset_ns=t_arrive - time.perf_counter_ns()) for (*_, r) in reqs]
    mixed = []
    for rid, (kind, dt, key, size, r) in zip(ids, reqs):
        ptr, _ = srv.wait(rid)
        got = A.copy_to_host(stream, ptr, out_count(kind, size), dt)
        assert agree(got, expected(pool, kind, dt, key, size), kind, dt, pool, key, size), (kind, dt, size)
        mixed.append(got)
        srv.release(rid)
    st = srv.stats()
    fusable = sum(k in ("scan", "fold1seg", "hist", "stencil") for (k, *_) in reqs)
    assert st.completed == len(reqs) and st.failed == 0
    assert st.fused == fusable and 1 <= st.rounds < fusable, (st.rounds, st.fused, fusable)
    srv.destroy()

    alone = A.Server(0, slots=1)
    for (kind, dt, key, size, r), m in zip(reqs, mixed):
        rid = alone.submit(r, keep=True)
        ptr, _ = alone.wait(rid)
        got = A.copy_to_host(stream, ptr, out_count(kind, size), dt)
        alone.release(rid)
        if np.dtype(dt).kind != "f" or kind not in ("scan", "fold1seg"):
            assert np.array_equal(got, m), (kind, np.dtype(dt).name, size)
        else:
            assert agree(got, expected(pool, kind, dt, key, size), kind, dt, pool, key, size)
    st = alone.stats()
    assert st.fused == fusable and st.rounds == fusable
    alone.destroy()


def test_unfused_server_is_the_plain_path(stream, pool):
    """fuse=False: every request through the plain entry points."""
    reqs = mix(pool, 30, seed=4)
    srv = A.Server(0, slots=2, fuse=False)
    ids = [srv.submit(r, keep=True) for (*_, r) in reqs]
    for rid, (kind, dt, key, size, r) in zip(ids, reqs):
        ptr, _ = srv.wait(rid)
        got = A.copy_to_host(stream, ptr, out_count(kind, size), dt)
        assert agree(got, expected(pool, kind, dt, key, size), kind, dt, pool, key, size)
        srv.release(rid)
    st = srv.stats()
    assert st.completed == 30 and st.rounds == 0 and st.fused == 0
    with pytest.raises(A.AccError):
        srv.configure(shape=(128, 8, 4))  # no such batched kernel
    srv.destroy()
