#!/usr/bin/env python3
"""Benchmark of the MI355X Accelerate backend's hot path.

One step = one pass of the segmented-reduction workload over resident data:
  Int64 scanl (+) 0 of 2^16 segment lengths (the offsets),
  Float32 fold1Seg (+) and fold1Seg max over 2^26 N(0,1) values in those
  2^16 segments (lengths log-uniform in [1, 16384], normalised to 2^26),
  both on one tile plan of the offsets (built with the offsets, one launch).
`value` is the step's algorithmic HBM bytes / step time, in GB/s.

Also reported: the dominant kernel's roofline (the fold1Seg (+) call on the
plan, timed with HIP events on its stream; the plan build and the unplanned
call beside it), each skeleton at its target size (GB/s and fraction of
the 8 TB/s HBM peak), a device-to-device copy ceiling, and the CPU baseline
(the oracle's restatement of the native backend's schedule on the host's
cores, on a bounded sample).

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "accelerate-llvm_amd"))

import acc_hip as A  # noqa: E402
from acc_hip import workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "achieved HBM GB/s (% of peak) per skeleton; served req/s + p99 latency"
SEED = 42
N_VALUES = 1 << 26
N_SEG = 1 << 16
MAX_LEN = 16384


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def timed(stream, fn, reps=20, warmup=2, stat="median", warm_ms=20.0):
    """Median (or mean) milliseconds of `reps` calls of fn() on `stream`, each
    bracketed by its own pair of HIP events on that stream (calls stay back to
    back).  Before them: `warmup` calls, one timed call, then as many more as
    make about `warm_ms` of device time (at most 500).  The clocks ramp up over
    several milliseconds of work after an idle pause, such as the host-side
    setup between workloads (8192^2 Float32 stencils: 92-106 us with two
    warm-up calls, still up to 103 us after 5 ms, 85-89 us after 18 ms), so
    the timed calls see the steady state."""
    for _ in range(warmup):
        fn()
    e0, e1 = A.Event(stream.device), A.Event(stream.device)
    e0.record(stream)
    fn()
    e1.record(stream)
    e1.synchronize()
    per_ms = max(e0.elapsed_ms(e1), 1e-3)
    for _ in range(min(500, int(warm_ms / per_ms))):
        fn()
    ev = [A.Event(stream.device) for _ in range(reps + 1)]
    stream.synchronize()
    ev[0].record(stream)
    for i in range(reps):
        fn()
        ev[i + 1].record(stream)
    ev[-1].synchronize()
    t = [ev[i].elapsed_ms(ev[i + 1]) for i in range(reps)]
    return float(np.mean(t) if stat == "mean" else np.median(t))


def timed_run(stream, fn, reps=50, warmup=3):
    """Mean milliseconds per call of `reps` back-to-back calls of fn() between
    one pair of HIP events on `stream` (no event between the calls)."""
    for _ in range(warmup):
        fn()
    e0, e1 = A.Event(stream.device), A.Event(stream.device)
    stream.synchronize()
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_ms(e1) / reps


def seg_bytes(n, m, b=4, ib=8):
    return n * b + (m + 1) * ib + m * b


def bench_batched(s, args):
    """10,000 problems: even -> Float32 scanl1 (+) on U(-1,1), odd -> Int32 fold (+) 0
    on U[-100,100]; sizes floor(exp(U(ln 2^8, ln 2^20))).  One launch per problem
    (the plain path) against one persistent batched launch replayed from a
    captured hipGraph."""
    import math
    rng = np.random.default_rng(SEED)
    P = args.batch_problems
    sizes = np.floor(np.exp(rng.uniform(math.log(2 ** 8), math.log(2 ** 20), size=P))).astype(np.int64)
    pad = lambda k: (k + 63) // 64 * 64  # 256-byte aligned sub-arrays
    scan_idx = np.arange(0, P, 2)
    fold_idx = np.arange(1, P, 2)
    s_off = np.concatenate([[0], np.cumsum([pad(int(sizes[i])) for i in scan_idx])])
    f_off = np.concatenate([[0], np.cumsum([pad(int(sizes[i])) for i in fold_idx])])
    log(f"batched: {P} problems, {int(sizes.sum())} elements; generating inputs")
    xs = W.uniform_f32(SEED, -1.0, 1.0, int(s_off[-1]))
    xf = W.uniform_i32(SEED + 1, -100, 100, int(f_off[-1]))
    d_xs, d_xf = A.DeviceArray.from_numpy(s, xs), A.DeviceArray.from_numpy(s, xf)
    d_ys = A.DeviceArray(int(s_off[-1]), np.float32)
    d_yf = A.DeviceArray(len(fold_idx) * 64, np.int32)
    probs = []
    for j, i in enumerate(scan_idx):
        n = int(sizes[i])
        probs.append((A.BATCH_SCAN, A.ADD, None, d_xs.view(int(s_off[j]), n), n, d_ys.view(int(s_off[j]), n)))
    for j, i in enumerate(fold_idx):
        n = int(sizes[i])
        probs.append((A.BATCH_FOLD, A.ADD, np.int32(0), d_xf.view(int(f_off[j]), n), n, d_yf.view(j * 64, 1)))

    def per_call_python():
        for (kind, op, seed, src, n, out) in probs:
            if kind == A.BATCH_SCAN:
                A.scan(s, op, A.LEFT_TO_RIGHT, None, src, 1, n, out)
            else:
                A.fold(s, op, seed, src, 1, n, out)

    log("batched: inputs on the device; per-call loop from Python")
    per_call_python()
    s.synchronize()
    e0, e1 = A.Event(0), A.Event(0)
    w0 = time.perf_counter()
    e0.record(s)
    per_call_python()
    e1.record(s)
    e1.synchronize()
    percall_py_wall = time.perf_counter() - w0
    percall_py_ms = e0.elapsed_ms(e1)
    ref_fold = d_yf.to_numpy(s)[::64].copy()
    log("batched: creating the batch (autotuned shape)")
    t0 = time.perf_counter()
    b = A.Batch(0, probs, tile=0)
    tune_s = time.perf_counter() - t0
    log(f"batched: shape {b.shape()} chosen in {tune_s:.2f} s; timing the per-call path from C")
    k, tiles = b.tile()
    shape = b.shape()
    # the per-call path from C: one plain entry-point call per problem, in
    # order on the stream (no Python between the calls)
    percall_ms = timed(s, lambda: b.launch_plain(s), reps=5, warmup=1)
    assert np.array_equal(d_yf.to_numpy(s)[::64], ref_fold), "plain calls from C differ from the Python loop"
    log("batched: capturing and checking the batched launch")
    g = A.Graph.capture(s, lambda: b.launch(s))
    g.launch(s)
    s.synchronize()
    # every problem equals its standalone result (same algorithm per element for
    # the integer folds; float scans within the FP32 bound of each other)
    got_fold = d_yf.to_numpy(s)[::64]
    assert np.array_equal(got_fold, ref_fold), "batched fold differs from the plain calls"
    got_scan = d_ys.to_numpy(s)
    sample = scan_idx[:: max(1, len(scan_idx) // 50)]
    for i in sample:
        j = int(np.searchsorted(scan_idx, i))
        n = int(sizes[i])
        x64 = xs[int(s_off[j]):int(s_off[j]) + n].astype(np.float64)
        a = got_scan[int(s_off[j]):int(s_off[j]) + n].astype(np.float64)
        kk = np.arange(1, n + 1)
        bound = 1e-5 * np.maximum(1, np.log2(np.maximum(kk, 2))) * np.cumsum(np.abs(x64)) + 1e-12
        assert np.all(np.abs(a - np.cumsum(x64)) <= bound), "batched scan outside the FP32 bound"
    want = np.sum(xf[int(f_off[0]):int(f_off[0]) + int(sizes[fold_idx[0]])], dtype=np.int32)
    assert got_fold[0] == want, "batched fold"
    batched_ms = timed(s, lambda: g.launch(s), reps=20)
    nbytes = sum(int(sizes[i]) * 8 for i in scan_idx) + sum(int(sizes[i]) * 4 + 4 for i in fold_idx)
    out = {
        "problems": P,
        "elements": int(sizes.sum()),
        "bytes": nbytes,
        "per_call_ms": round(percall_ms, 3),
        "per_call_python_ms": round(percall_py_ms, 3),
        "per_call_python_wall_ms": round(percall_py_wall * 1e3, 3),
        "batched_ms": round(batched_ms, 3),
        "speedup": round(percall_ms / batched_ms, 2),
        "batched_gbps": round(nbytes / (batched_ms * 1e-3) / 1e9, 1),
        "shape": {"block": shape[0], "vectors_per_thread": shape[1], "tiles_per_run": shape[2],
                  "tickets": shape[3]},
        "autotune_s": round(tune_s, 2),
    }
    log(f"batched: per-call from C {percall_ms:.2f} ms (Python loop {percall_py_ms:.2f} ms), batched "
        f"{batched_ms:.3f} ms -> {out['speedup']}x, {out['batched_gbps']} GB/s (shape {shape})")
    g.destroy()
    b.destroy()
    return out


def bench_ceilings(s):
    """The device's HBM ceilings: read-only, copy and write-only probe kernels
    over 2 GiB (1 GiB each way for the copy) for vectors in flight per lane x
    workgroups per CU x non-temporal hint; the best of each kind."""
    nbytes = 1 << 31
    src, dst = A.DeviceArray(nbytes, np.uint8), A.DeviceArray(nbytes, np.uint8)
    A.check(A.lib().acc_hip_fill_async(s.handle, A.c_void_p(src.ptr), np.ones(1, np.uint32).ctypes.data_as(A.c_void_p),
                                       4, nbytes // 4))
    table = []
    for kind, name in ((A.PROBE_READ, "read"), (A.PROBE_COPY, "copy"), (A.PROBE_WRITE, "write")):
        nb = nbytes // 2 if kind == A.PROBE_COPY else nbytes
        moved = 2 * nb if kind == A.PROBE_COPY else nb
        for vec in (4, 8, 16):
            for bpc in (1, 2, 4, 8):
                for nt in (False, True):
                    us = A.bandwidth_probe(s, kind, vec, bpc, nt, src if kind != A.PROBE_WRITE else None,
                                           dst if kind != A.PROBE_READ else None, nb, reps=10)
                    table.append({"kind": name, "vectors": vec, "blocks_per_cu": bpc, "nontemporal": nt,
                                  "us": round(us, 1), "gbps": round(moved / (us * 1e-6) / 1e9, 1)})
    best = {}
    for r in table:
        if r["kind"] not in best or r["gbps"] > best[r["kind"]]["gbps"]:
            best[r["kind"]] = r
    for k, r in best.items():
        log(f"ceiling {k}: {r['gbps']:.0f} GB/s ({100 * r['gbps'] / HBM_PEAK_GBS:.1f}% of peak) with "
            f"{r['vectors']} vectors/lane, {r['blocks_per_cu']} workgroups/CU, non-temporal {r['nontemporal']}")
    src.free()
    dst.free()
    return {"best": best, "table": table}


def bench_serving(s, args):
    """Poisson arrivals of mixed requests (35% scanl1 (+), 30% fold1Seg (+), 20%
    histogram permute into 4096 bins, 15% 3x3 mirror stencil), sizes
    log-uniform in [2^10, 2^24], element type uniform over Float32/Int32/Int64,
    inputs resident on the device, admission by free HBM."""
    import math
    R = args.serve_requests
    rng = np.random.default_rng(SEED + 7)
    N = 1 << 24
    pool = {
        "f32": A.DeviceArray.from_numpy(s, W.uniform_f32(SEED, -1.0, 1.0, N)),
        "i32": A.DeviceArray.from_numpy(s, W.uniform_i32(SEED, -100, 100, N)),
        "i64": A.DeviceArray.from_numpy(s, W.uniform_i32(SEED, -100, 100, N).astype(np.int64)),
    }
    lens = W.segment_lengths(SEED, N // 1024, 16384, N)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    d_off = A.DeviceArray.from_numpy(s, off)
    d_idx = A.DeviceArray.from_numpy(s, W.zipf_i32(SEED, 1.1, 4096, N))
    dts = {"f32": np.float32, "i32": np.int32, "i64": np.int64}
    reqs, nbytes = [], 0.0
    for i in range(R):
        key = ["f32", "i32", "i64"][rng.integers(0, 3)]
        dt, src = dts[key], pool[key]
        kind = rng.choice(4, p=[0.35, 0.30, 0.20, 0.15])
        n = int(math.exp(rng.uniform(math.log(2 ** 10), math.log(2 ** 24))))
        B = np.dtype(dt).itemsize
        if kind == 0:
            reqs.append(A.Server.request(A.REQ_SCAN, dt, A.ADD, src=src, n=n))
            nbytes += 2 * n * B
        elif kind == 1:
            nseg = max(1, int(np.searchsorted(off, n, side="right")) - 1)
            ne = int(off[nseg])
            reqs.append(A.Server.request(A.REQ_FOLD_SEG, dt, A.ADD, src=src, n=ne, m=nseg, aux=d_off))
            nbytes += ne * B + (nseg + 1) * 8 + nseg * B
        elif kind == 2:
            reqs.append(A.Server.request(A.REQ_PERMUTE, dt, A.ADD, seed=0, n=n, m=4096, aux=d_idx, fill=1))
            nbytes += n * 4 + 2 * 4096 * B
        else:
            side = max(3, int(math.sqrt(n)))
            w = (np.outer([1, 2, 1], [1, 2, 1]) / (16 if key == "f32" else 1)).astype(dt)
            reqs.append(A.Server.request(A.REQ_STENCIL2D, dt, src=src, n=side, m=side, ry=1, rx=1,
                                         boundary=A.MIRROR, weights=w))
            nbytes += 2 * side * side * B

    # the plain per-call path, from C: one slot, no fusion -- each request runs
    # alone through the plain entry points and completes before the next
    # starts (the client waits for each)
    def capacity_of(srv):
        # one arrival instant for all, far enough ahead that the submit loop
        # (a few microseconds per request from Python) has finished by then
        lead = max(50_000_000, 40_000 * len(reqs))
        t_arrive = time.perf_counter_ns() + lead
        for q in reqs:
            srv.submit(q, arrival_offset_ns=t_arrive - time.perf_counter_ns())
        assert time.perf_counter_ns() < t_arrive, "submit loop slower than the arrival lead time"
        srv.drain()
        st = srv.stats()
        srv.reset_stats()
        return st

    log(f"serving: {R} requests; plain per-call path (one slot, no fusion)")
    plain_srv = A.Server(0, slots=1, fuse=False)
    capacity_of(plain_srv)  # untimed pass: every kernel of the mix loaded and warm
    plain = capacity_of(plain_srv).req_per_s
    plain_srv.destroy()
    # and from Python (ctypes call + synchronize per request), for reference
    plain_t0 = time.perf_counter()
    outs = A.DeviceArray(N * 8, np.uint8)
    for (r, w) in reqs:
        run_plain(s, r, w, outs)
        s.synchronize()
    plain_py = R / (time.perf_counter() - plain_t0)

    log(f"serving: plain {plain:.0f} req/s; unfused and fused servers")
    unfused_srv = A.Server(0, slots=args.serve_slots, fuse=False)
    unfused = capacity_of(unfused_srv).req_per_s
    unfused_srv.destroy()

    srv = A.Server(0, slots=args.serve_slots, round_requests=args.serve_round_requests,
                   round_bytes=args.serve_round_mib << 20, shape=(256, args.serve_vectors, 4))
    # capacity: every request arrives at once
    cap = capacity_of(srv)
    capacity = cap.req_per_s
    # open-loop Poisson trace at 80% of capacity: arrival times are measured
    # from one origin (each submit gives its offset from the moment it is made)
    lam = 0.8 * capacity
    arrivals = np.cumsum(rng.exponential(1.0 / lam, size=len(reqs))) + 2e-3
    t0 = time.perf_counter()
    for q, at in zip(reqs, arrivals):
        srv.submit(q, arrival_offset_ns=max(0, int((at - (time.perf_counter() - t0)) * 1e9)))
    srv.drain()
    st = srv.stats()
    srv.destroy()
    out = {
        "requests": R,
        "slots": args.serve_slots,
        "capacity_req_per_s": round(capacity, 1),
        "arrival_rate_req_per_s": round(lam, 1),
        "req_per_s": round(st.req_per_s, 1),
        "p50_ms": round(st.p50_us / 1e3, 3),
        "p99_ms": round(st.p99_us / 1e3, 3),
        "mean_ms": round(st.mean_us / 1e3, 3),
        "gbps": round(st.bytes_per_s / 1e9, 1),
        "fused_rounds_at_capacity": int(cap.rounds),
        "fused_requests_at_capacity": int(cap.fused),
        "plain_req_per_s": round(plain, 1),
        "plain_python_req_per_s": round(plain_py, 1),
        "unfused_slots_req_per_s": round(unfused, 1),
        "capacity_vs_plain": round(capacity / plain, 2),
    }
    log(f"serving: capacity {capacity:.0f} req/s in {cap.rounds} fused rounds ({cap.fused} fused requests); "
        f"plain per-call from C {plain:.0f} req/s (Python {plain_py:.0f}), unfused {args.serve_slots} slots "
        f"{unfused:.0f} req/s; at {lam:.0f} req/s p50 {out['p50_ms']} ms p99 {out['p99_ms']} ms, "
        f"{out['req_per_s']} req/s")
    return out


def run_plain(s, r, w, out):
    """One request through the plain C-ABI call (what the server runs per request)."""
    L = A.lib()
    seed = A.ctypes.c_uint64(r.seed)
    fill = A.ctypes.c_uint64(r.fill)
    sp = A.ctypes.byref(seed) if r.has_seed else None
    if r.skeleton == A.REQ_SCAN:
        A.check(L.acc_hip_scan(s.handle, r.type, r.op, r.direction, sp, r.in_, 1, r.n, A.c_void_p(out.ptr)))
    elif r.skeleton == A.REQ_FOLD_SEG:
        A.check(L.acc_hip_fold_seg(s.handle, r.type, r.op, sp, r.in_, 1, r.n, r.aux_type, r.aux, r.m,
                                   A.c_void_p(out.ptr)))
    elif r.skeleton == A.REQ_PERMUTE:
        sz = 8 if r.type in (A.INT64, A.FLOAT64, A.WORD64) else 4
        A.check(L.acc_hip_fill_async(s.handle, A.c_void_p(out.ptr), A.ctypes.byref(seed), sz, r.m))
        A.check(L.acc_hip_permute(s.handle, r.type, r.op, A.c_void_p(out.ptr), A.c_void_p(out.ptr), r.m, r.in_,
                                  A.ctypes.byref(fill), r.aux_type, r.aux, r.n))
    else:
        A.check(L.acc_hip_stencil2d(s.handle, r.type, r.ry, r.rx, w.ctypes.data_as(A.c_void_p), r.boundary,
                                    A.ctypes.byref(fill), r.in_, r.n, r.m, A.c_void_p(out.ptr)))


def kernel_only():
    """The dominant kernel (fold1Seg (+) on the step's inputs) five times and a
    1 GiB device copy three times: the program the counter runs profile."""
    s = A.Stream(0)
    lens = W.segment_lengths(SEED, N_SEG, MAX_LEN, N_VALUES)
    off = W.offsets(lens)
    d_vals = A.DeviceArray.from_numpy(s, W.normal_f32(SEED, N_VALUES))
    d_off = A.DeviceArray.from_numpy(s, off)
    d_sum = A.DeviceArray(N_SEG, np.float32)
    plan = A.SegPlan(0)
    plan.build(s, np.float32, d_vals, N_VALUES, d_off, N_SEG)
    for _ in range(5):
        plan.fold_seg(s, A.ADD, None, d_vals, 1, d_sum)
    src, dst = A.DeviceArray(1 << 28, np.int32), A.DeviceArray(1 << 28, np.int32)
    for _ in range(3):
        A.check(A.lib().acc_hip_memcpy_d2d_async(s.handle, A.c_void_p(dst.ptr), A.c_void_p(src.ptr), 1 << 30))
    s.synchronize()
    return 0


# the kernels of one fold1Seg call on a built tile plan: the tiles, the
# carries between tiles
FOLD_SEG_KERNELS = ("k_fold_seg", "k_seg_fix")


def measure_traffic():
    """HBM bytes per fold1Seg (+) call (its two kernels) from two rocprofv3
    counter runs (FETCH_SIZE, WRITE_SIZE; KiB per dispatch), each scaled so
    that the 1 GiB device copy in the same run reads/writes exactly 1 GiB.
    None on failure."""
    import csv
    import shutil
    import signal
    import subprocess
    import tempfile
    vals = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="acc_pmc_")
        cmd = ["rocprofv3", "--pmc", counter, "-T", "-f", "csv", "-d", d, "-o", "k", "--",
               sys.executable, os.path.abspath(__file__), "--kernel-only"]
        log(f"counter run {counter}")
        try:
            p = subprocess.Popen(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, start_new_session=True)
        except OSError:
            return None
        try:
            p.wait(timeout=120)
        except subprocess.TimeoutExpired:
            os.killpg(p.pid, signal.SIGKILL)
            p.wait()
            log(f"counter run {counter} timed out")
            return None
        files = [os.path.join(r, f) for r, _, fs in os.walk(d) for f in fs if f.endswith("counter_collection.csv")]
        if not files:
            return None
        seg, cp = {k: [] for k in FOLD_SEG_KERNELS}, []
        for row in csv.DictReader(open(files[0])):
            if row["Counter_Name"] != counter:
                continue
            name = row["Kernel_Name"].split("<")[0].split("(")[0].split("::")[-1]
            if name in seg:
                seg[name].append(float(row["Counter_Value"]))
            elif ("k_copy16" in row["Kernel_Name"] or "copyBuffer" in row["Kernel_Name"]) and \
                    int(row["Grid_Size"]) >= 65536:
                cp.append(float(row["Counter_Value"]))
        shutil.rmtree(d, ignore_errors=True)
        if not all(seg.values()) or not cp:
            return None
        scale = (1 << 30) / (sum(cp) / len(cp) * 1024.0)
        kib = sum(sum(v) / len(v) for v in seg.values())  # one fold1Seg call: its kernels
        vals[counter] = kib * 1024.0 * scale
        log(f"{counter}: {kib:.0f} KiB per fold1Seg call, scale {scale:.3f}")
    return int(vals["FETCH_SIZE"] + vals["WRITE_SIZE"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=60, help="untimed steps right before the timed ones (the device settles over the first few ms of work)")
    ap.add_argument("--skeleton-reps", type=int, default=20, help="timed calls per skeleton (median reported)")
    ap.add_argument("--no-skeletons", action="store_true", help="skip the per-skeleton lines")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--batch-problems", type=int, default=10000)
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-slots", type=int, default=8, help="server execution slots (one worker thread each)")
    ap.add_argument("--serve-round-requests", type=int, default=256, help="requests per fused round at most")
    ap.add_argument("--serve-round-mib", type=int, default=1024, help="algorithmic MiB per fused round at most")
    ap.add_argument("--serve-vectors", type=int, default=8, help="vectors per thread of the fused launch (2, 4, 8)")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter runs")
    ap.add_argument("--no-ceilings", action="store_true", help="skip the HBM ceiling probes")
    ap.add_argument("--kernel-only", action="store_true",
                    help="(internal) run only the dominant kernel and a 1 GiB copy, for counter collection")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design note, not a measured path)", file=sys.stderr)
        return 2
    if args.kernel_only:
        return kernel_only()
    # the counter runs are child processes; they run before this process
    # touches the GPU
    traffic = None if args.no_traffic else measure_traffic()

    if A.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        return 1

    s = A.Stream(0)
    props = A.device_props(0)
    log(f"device {props.name.decode()} arch {props.arch.decode()} CUs {props.compute_units}")

    # ---------------------------------------------------------------- inputs
    t0 = time.time()
    lens = W.segment_lengths(SEED, N_SEG, MAX_LEN, N_VALUES)
    vals = W.normal_f32(SEED, N_VALUES)
    d_lens = A.DeviceArray.from_numpy(s, lens)
    d_vals = A.DeviceArray.from_numpy(s, vals)
    d_off = A.DeviceArray(N_SEG + 1, np.int64)
    d_sum = A.DeviceArray(N_SEG, np.float32)
    d_max = A.DeviceArray(N_SEG, np.float32)
    log(f"inputs ready ({time.time() - t0:.1f}s)")

    # The step: the offsets (`scanl (+) 0` of the lengths) together with their
    # tile plan (one launch; the plan is shared by the two folds over those
    # segments), then the two folds in one call (their fix-ups share a launch).
    plan = A.SegPlan(0)

    def step():
        plan.build_lengths(s, np.float32, d_vals, N_VALUES, d_lens, N_SEG, d_off)
        plan.fold_seg_multi(s, [(A.ADD, None, d_vals, d_sum), (A.MAX, None, d_vals, d_max)])

    step_bytes = (N_SEG * 8 + (N_SEG + 1) * 8) + 2 * seg_bytes(N_VALUES, N_SEG)

    step()  # one step to check the outputs of
    s.synchronize()

    # sanity of the step's outputs with numpy (the parity tests cover the rest)
    off = W.offsets(lens)
    assert np.array_equal(d_off.to_numpy(s), off), "offsets"
    assert np.array_equal(d_max.to_numpy(s), np.maximum.reduceat(vals, off[:-1])), "fold1Seg max"
    truth = np.add.reduceat(vals.astype(np.float64), off[:-1])
    ab = np.add.reduceat(np.abs(vals.astype(np.float64)), off[:-1])
    bound = 1e-5 * np.maximum(1.0, np.log2(np.maximum(np.diff(off), 2))) * ab
    err = np.abs(d_sum.to_numpy(s).astype(np.float64) - truth)
    assert np.all(err <= bound), "fold1Seg (+) outside the FP32 bound"
    rel = err / np.maximum(ab, 1e-300)
    fp_error = {
        "result": "fold1Seg (+) Float32, one value per segment, against the FP64 sum",
        "results": int(N_SEG),
        "max_rel": float(rel.max()),
        "p99_rel": float(np.quantile(rel, 0.99)),
        "median_rel": float(np.median(rel)),
        "rel": "|y - s| / sum|x| per segment",
        "bound": "1e-5 * max(1, log2(len)) * sum|x|",
        "max_err_over_bound": float((err / bound).max()),
    }
    log(f"step outputs checked (fold1Seg (+) per segment |y-s|/sum|x|: max {fp_error['max_rel']:.3e} "
        f"p99 {fp_error['p99_rel']:.3e}; max error/bound {fp_error['max_err_over_bound']:.3f})")

    # ------------------------------------------------------------ timed steps
    # the untimed warm-up steps right before the timed ones (the checks above
    # left the device idle, and its clocks settle over the first ms of work)
    for _ in range(args.warmup):
        step()
    e0, e1 = A.Event(0), A.Event(0)
    s.synchronize()
    w0 = time.perf_counter()
    e0.record(s)
    for _ in range(args.steps):
        step()
    e1.record(s)
    e1.synchronize()
    wall = time.perf_counter() - w0
    ms_step = e0.elapsed_ms(e1) / args.steps
    value = step_bytes / (ms_step * 1e-3) / 1e9
    log(f"step {ms_step * 1e3:.1f} us ({value:.0f} GB/s); wall {wall / args.steps * 1e3:.3f} ms/step")

    # ---------------------------------------------- dominant kernel roofline
    k_bytes = seg_bytes(N_VALUES, N_SEG)
    k_fn = lambda: plan.fold_seg(s, A.ADD, None, d_vals, 1, d_sum)
    u_fn = lambda: A.fold_seg(s, A.ADD, None, d_vals, 1, N_VALUES, d_off, N_SEG, d_sum)  # builds its own plan
    # back-to-back calls, the planned and the unplanned call in alternating
    # runs of 25 (the device's clocks drift over a run of tens of ms: the same
    # kernel slows by up to ~15%, so an order effect would bias either one)
    k_runs, u_runs = [], []
    for _ in range(3):
        k_runs.append(timed_run(s, k_fn, 25))
        u_runs.append(timed_run(s, u_fn, 25))
    k_ms, u_ms = float(np.mean(k_runs)), float(np.mean(u_runs))
    k_med = timed(s, k_fn, 20)     # median of 20 calls, each between its own two events
    k_gbs = k_bytes / (k_ms * 1e-3) / 1e9
    d_off2 = A.DeviceArray(N_SEG + 1, np.int64)
    p2 = A.SegPlan(0)
    b_ms = timed_run(s, lambda: p2.build(s, np.float32, d_vals, N_VALUES, d_off, N_SEG), 50)
    bl_ms = timed_run(s, lambda: p2.build_lengths(s, np.float32, d_vals, N_VALUES, d_lens, N_SEG, d_off2), 50)
    sc_ms = timed_run(s, lambda: A.scan(s, A.ADD, A.LEFT_TO_RIGHT, np.int64(0), d_lens, 1, N_SEG, d_off2), 50)
    p2.destroy()
    roofline = {
        "kernel": "fold1Seg (+) Float32 call on the offsets' tile plan: k_fold_seg + k_seg_fix",
        "bound": "hbm",
        "achieved": round(k_gbs, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(k_gbs / HBM_PEAK_GBS, 4),
        "traffic": None,
        "algorithmic_bytes": k_bytes,
        "avg_us": round(k_ms * 1e3, 2),
        "median_us": round(k_med * 1e3, 2),
        "unplanned_call_us": round(u_ms * 1e3, 2),
        "offsets_and_plan_us": round(bl_ms * 1e3, 2),
        "offsets_scan_us": round(sc_ms * 1e3, 2),
        "plan_from_offsets_us": round(b_ms * 1e3, 2),
    }
    log(f"fold1Seg (+) call on a plan {k_ms * 1e3:.1f} us -> {k_gbs:.0f} GB/s; unplanned call {u_ms * 1e3:.1f} us; "
        f"offsets + plan {bl_ms * 1e3:.1f} us (offsets scan alone {sc_ms * 1e3:.1f}, plan alone {b_ms * 1e3:.1f})")
    if traffic:
        roofline["traffic"] = traffic
        if roofline["traffic"]:
            log(f"fold1Seg (+) HBM traffic {roofline['traffic']} B per launch "
                f"({roofline['traffic'] / k_bytes:.3f}x algorithmic)")

    # --------------------------------------------------- per-skeleton lines
    skel = {}
    if not args.no_skeletons:
        R = args.skeleton_reps

        def record(name, nbytes, ms, **extra):
            gbs = nbytes / (ms * 1e-3) / 1e9
            skel[name] = {"gbps": round(gbs, 1), "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4),
                          "us": round(ms * 1e3, 2), "bytes": int(nbytes), **extra}
            log(f"{name}: {ms * 1e3:.1f} us, {gbs:.0f} GB/s ({100 * gbs / HBM_PEAK_GBS:.1f}% of peak)")

        n28 = 1 << 28
        x = W.uniform_i32(SEED, -1000, 1000, n28)
        dx = A.DeviceArray.from_numpy(s, x)
        del x
        dy = A.DeviceArray(n28, np.int32)
        d1 = A.DeviceArray(1, np.int32)
        record("copy_d2d_1GiB", 2 * n28 * 4,
               timed(s, lambda: A.check(A.lib().acc_hip_memcpy_d2d_async(s.handle, A.c_void_p(dy.ptr),
                                                                       A.c_void_p(dx.ptr), n28 * 4)), R))
        record("fold_add_i32_2^28", n28 * 4 + 4, timed(s, lambda: A.fold(s, A.ADD, np.int32(0), dx, 1, n28, d1), R))
        record("scanl1_add_i32_2^28", 2 * n28 * 4,
               timed(s, lambda: A.scan(s, A.ADD, A.LEFT_TO_RIGHT, None, dx, 1, n28, dy), R))
        record("scanl_add_i32_2^28", n28 * 4 + (n28 + 1) * 4,
               timed(s, lambda: A.scan(s, A.ADD, A.LEFT_TO_RIGHT, np.int32(0), dx, 1, n28 - 1, dy), R))
        record("scanl1_rows_add_i32_2^18x1024", 2 * n28 * 4,
               timed(s, lambda: A.scan(s, A.ADD, A.LEFT_TO_RIGHT, None, dx, 1 << 18, 1024, dy), R))
        # producers
        dz = A.DeviceArray(n28, np.int32)
        record("generate_iota_i32_2^28", n28 * 4, timed(s, lambda: A.generate(s, A.GEN_IOTA, n28, dz, 1, 3), R))
        record("map_affine_i32_2^28", 2 * n28 * 4, timed(s, lambda: A.map_(s, A.UNOP_AFFINE, dx, n28, dy, 3, -7), R))
        record("zipWith_add_i32_2^28", 3 * n28 * 4, timed(s, lambda: A.zipwith(s, A.ADD, dx, dz, n28, dy), R))
        record("reverse_i32_2^28", 2 * n28 * 4,
               timed(s, lambda: A.transform(s, A.IX_REVERSE, A.UNOP_ID, dx, 1, n28, None, n28, dy), R))
        record("transpose_i32_16384^2", 2 * n28 * 4,
               timed(s, lambda: A.transform(s, A.IX_TRANSPOSE, A.UNOP_ID, dx, 1 << 14, 1 << 14, None, n28, dy), R))
        k26 = 1 << 26
        d_perm = A.DeviceArray.from_numpy(s, np.random.default_rng(SEED).permutation(k26).astype(np.int32))
        record("backpermute_random_i32_2^26", 3 * k26 * 4,
               timed(s, lambda: A.transform(s, A.IX_GATHER, A.UNOP_ID, dx, 1, k26, d_perm, k26, dy), R))
        del dz, d_perm
        del dx, dy
        record("fold1Seg_max_f32_2^26", k_bytes,
               timed(s, lambda: A.fold_seg(s, A.MAX, None, d_vals, 1, N_VALUES, d_off, N_SEG, d_max), R))

        idx = W.zipf_i32(SEED, 1.1, 4096, 1 << 26)
        d_idx = A.DeviceArray.from_numpy(s, idx)
        del idx
        bins = A.DeviceArray(4096, np.int32)
        zero = A.DeviceArray.from_numpy(s, np.zeros(4096, np.int32))
        record("histogram_zipf_2^26_4096", (1 << 26) * 4 + 2 * 4096 * 4,
               timed(s, lambda: A.permute(s, A.ADD, zero, bins, 4096, None, 1, d_idx, 1 << 26), R))
        del d_idx

        H = WD = 8192
        d_img = d_vals  # 2^26 N(0,1) values viewed as an 8192 x 8192 image
        d_out = A.DeviceArray(H * WD, np.float32)
        w5 = (np.outer([1, 4, 6, 4, 1], [1, 4, 6, 4, 1]) / 256).astype(np.float32)
        w3 = (np.outer([1, 2, 1], [1, 2, 1]) / 16).astype(np.float32)
        record("stencil5x5_clamp_f32_8192^2", 2 * H * WD * 4,
               timed(s, lambda: A.stencil2d(s, 2, 2, w5, A.CLAMP, None, d_img, H, WD, d_out), R))
        record("stencil3x3_mirror_f32_8192^2", 2 * H * WD * 4,
               timed(s, lambda: A.stencil2d(s, 1, 1, w3, A.MIRROR, None, d_img, H, WD, d_out), R))
        # stencil2: the image (3x3, clamp) minus its blur (3x3, wrap)
        d_out2 = A.DeviceArray(H * WD, np.float32)
        record("stencil2_3x3_clamp+wrap_f32_8192^2", 3 * H * WD * 4,
               timed(s, lambda: A.stencil2_2d(s, 1, 1, w3, A.CLAMP, None, d_img, H, WD, -w3, A.WRAP, None, d_out,
                                              H, WD, d_out2), R))
        del d_out2
        # 3-D: 3x3x3 box, clamp, on 512^3 Float32 (2^27 elements)
        n3 = 512
        d_vol = A.DeviceArray(n3 ** 3, np.float32)
        d_vout = A.DeviceArray(n3 ** 3, np.float32)
        A.generate(s, A.GEN_IOTA, n3 ** 3, d_vol, 0.0, 1e-6)
        w27 = np.full((3, 3, 3), 1.0 / 27.0, np.float32)
        record("stencil3d_3x3x3_clamp_f32_512^3", 2 * n3 ** 3 * 4,
               timed(s, lambda: A.stencil3d(s, 1, 1, 1, w27, A.CLAMP, None, d_vol, n3, n3, n3, d_vout), R))
        del d_vol, d_vout
        del d_out

    # ------------------------------------------------------- HBM ceilings
    ceilings = None if args.no_ceilings else bench_ceilings(s)
    # ------------------------------------------------------- batched (C4)
    batched = None if args.no_batched else bench_batched(s, args)
    # ------------------------------------------------------- serving (C5)
    serving = None if args.no_serving else bench_serving(s, args)

    # ------------------------------------------------------------ CPU baseline
    cpu = None
    if not args.no_cpu_baseline:
        # the CPU baseline is the oracle's restatement of the native backend's
        # schedule, timed on this host's cores
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O
        cores = len(os.sched_getaffinity(0))
        out_s = np.empty(N_SEG, np.float32)
        out_m = np.empty(N_SEG, np.float32)
        reps, t_cpu = 0, 0.0
        while t_cpu < args.cpu_seconds:
            c0 = time.perf_counter()
            offc = np.empty(N_SEG + 1, np.int64)
            offc[0] = 0
            np.cumsum(lens, out=offc[1:])  # scanl (+) 0 of 2^16 lengths (sequential, n < 8192 per worker rule aside)
            O.native_fold1_seg_f32(cores, A.ADD, vals, offc, out_s)
            O.native_fold1_seg_f32(cores, A.MAX, vals, offc, out_m)
            t_cpu += time.perf_counter() - c0
            reps += 1
        cpu_gbs = step_bytes * reps / t_cpu / 1e9
        model = "unknown CPU"
        try:
            with open("/proc/cpuinfo") as f:
                model = next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
        except (OSError, StopIteration):
            pass
        cpu = {"value": round(cpu_gbs, 2), "unit": "GB/s", "cores": cores, "kind": "port",
               "sample": f"the full step ({reps} repetitions, {t_cpu:.1f} s) with the native backend's schedule "
                         f"(divideWork over segments, {cores} threads)",
               "cpu": model, "build": O.build_info()}
        log(f"CPU baseline: {cpu_gbs:.2f} GB/s on {cores} threads")

    result = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "Float32 fold1Seg (+) and fold1Seg max over 2^26 N(0,1) values in 2^16 segments "
                        "(lengths log-uniform in [1,16384], sum 2^26) + Int64 scanl (+) 0 of the lengths "
                        "(the offsets and their tile plan in one launch per step; the two folds in one call on that plan, their tiles in one launch)",
            "elements": N_VALUES,
            "segments": N_SEG,
            "seed": SEED,
            "step_bytes": step_bytes,
        },
        "frac_of_hbm_peak": round(value / HBM_PEAK_GBS, 4),
        "roofline": roofline,
        "fp_error": fp_error,
        "cpu_baseline": cpu,
        "skeletons": skel,
        "ceilings": ceilings,
        "batched": batched,
        "serving": serving,
        "device": {"name": props.name.decode(), "arch": props.arch.decode(), "cus": props.compute_units},
    }
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
