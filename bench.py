#!/usr/bin/env python3
"""Benchmark of the MI355X Accelerate backend's hot path.

One step = one pass of the segmented-reduction workload over resident data:
  Int64 scanl (+) 0 of 2^16 segment lengths (the offsets),
  Float32 fold1Seg (+) and fold1Seg max over 2^26 N(0,1) values in those
  2^16 segments (lengths log-uniform in [1, 16384], normalised to 2^26).
`value` is the step's algorithmic HBM bytes / step time, in GB/s.

Also reported: the dominant kernel's roofline (fold1Seg (+), timed with HIP
events on its stream), each skeleton at its target size (GB/s and fraction of
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

HBM_PEAK_GBS = 8000.0
METRIC = "achieved HBM GB/s (% of peak) per skeleton; served req/s + p99 latency"
SEED = 42
N_VALUES = 1 << 26
N_SEG = 1 << 16
MAX_LEN = 16384


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def timed(stream, fn, reps, warmup=2):
    """Average milliseconds per call of fn() on `stream`, by HIP events."""
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--skeleton-reps", type=int, default=10)
    ap.add_argument("--no-skeletons", action="store_true", help="skip the per-skeleton lines")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design note, not a measured path)", file=sys.stderr)
        return 2
    if A.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        return 1

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O  # input generators; the oracle is also the CPU baseline below

    s = A.Stream(0)
    props = A.device_props(0)
    log(f"device {props.name.decode()} arch {props.arch.decode()} CUs {props.compute_units}")

    # ---------------------------------------------------------------- inputs
    t0 = time.time()
    lens = O.gen_segment_lengths(SEED, N_SEG, MAX_LEN, N_VALUES)
    vals = O.gen_normal_f32(SEED, N_VALUES)
    d_lens = A.DeviceArray.from_numpy(s, lens)
    d_vals = A.DeviceArray.from_numpy(s, vals)
    d_off = A.DeviceArray(N_SEG + 1, np.int64)
    d_sum = A.DeviceArray(N_SEG, np.float32)
    d_max = A.DeviceArray(N_SEG, np.float32)
    log(f"inputs ready ({time.time() - t0:.1f}s)")

    def step():
        A.scan(s, A.ADD, A.LEFT_TO_RIGHT, np.int64(0), d_lens, 1, N_SEG, d_off)
        A.fold_seg(s, A.ADD, None, d_vals, 1, N_VALUES, d_off, N_SEG, d_sum)
        A.fold_seg(s, A.MAX, None, d_vals, 1, N_VALUES, d_off, N_SEG, d_max)

    step_bytes = (N_SEG * 8 + (N_SEG + 1) * 8) + 2 * seg_bytes(N_VALUES, N_SEG)

    for _ in range(args.warmup):
        step()
    s.synchronize()

    # parity of the step's outputs (cheap: the checker runs once)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    assert np.array_equal(d_off.to_numpy(s), off), "offsets"
    assert np.array_equal(d_max.to_numpy(s), O.fold_seg(A.MAX, None, vals, off)), "fold1Seg max"
    truth, ab = O.fold_seg_f64(vals, off)
    bound = 1e-5 * np.maximum(1.0, np.log2(np.maximum(np.diff(off), 2))) * ab
    err = np.abs(d_sum.to_numpy(s).astype(np.float64) - truth)
    assert np.all(err <= bound), "fold1Seg (+) outside the FP32 bound"
    log(f"step outputs match the oracle (max rel-to-bound {float((err / bound).max()):.3f})")

    # ------------------------------------------------------------ timed steps
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
    k_ms = timed(s, lambda: A.fold_seg(s, A.ADD, None, d_vals, 1, N_VALUES, d_off, N_SEG, d_sum), 50)
    k_gbs = k_bytes / (k_ms * 1e-3) / 1e9
    roofline = {
        "kernel": "k_fold_seg<float, OpAdd, OffsetStarts<long>>",
        "bound": "hbm",
        "achieved": round(k_gbs, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(k_gbs / HBM_PEAK_GBS, 4),
        "traffic": None,
        "algorithmic_bytes": k_bytes,
        "avg_us": round(k_ms * 1e3, 2),
    }
    log(f"fold1Seg (+) kernel {k_ms * 1e3:.1f} us -> {k_gbs:.0f} GB/s")

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
        x = O.gen_uniform_i32(SEED, -1000, 1000, n28)
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
        del dx, dy
        record("fold1Seg_max_f32_2^26", k_bytes,
               timed(s, lambda: A.fold_seg(s, A.MAX, None, d_vals, 1, N_VALUES, d_off, N_SEG, d_max), R))

        idx = O.gen_zipf_i32(SEED, 1.1, 4096, 1 << 26)
        d_idx = A.DeviceArray.from_numpy(s, idx)
        del idx
        bins = A.DeviceArray(4096, np.int32)
        zero = A.DeviceArray.from_numpy(s, np.zeros(4096, np.int32))
        record("histogram_zipf_2^26_4096", (1 << 26) * 4 + 2 * 4096 * 4,
               timed(s, lambda: A.permute(s, A.ADD, zero, bins, 4096, None, 1, d_idx, 1 << 26), R))
        del d_idx

        H = W = 8192
        d_img = d_vals  # 2^26 N(0,1) values viewed as an 8192 x 8192 image
        d_out = A.DeviceArray(H * W, np.float32)
        w5 = (np.outer([1, 4, 6, 4, 1], [1, 4, 6, 4, 1]) / 256).astype(np.float32)
        w3 = (np.outer([1, 2, 1], [1, 2, 1]) / 16).astype(np.float32)
        record("stencil5x5_clamp_f32_8192^2", 2 * H * W * 4,
               timed(s, lambda: A.stencil2d(s, 2, 2, w5, A.CLAMP, None, d_img, H, W, d_out), R))
        record("stencil3x3_mirror_f32_8192^2", 2 * H * W * 4,
               timed(s, lambda: A.stencil2d(s, 1, 1, w3, A.MIRROR, None, d_img, H, W, d_out), R))
        del d_out

    # ------------------------------------------------------------ CPU baseline
    cpu = None
    if not args.no_cpu_baseline:
        cores = len(os.sched_getaffinity(0))
        out_s = np.empty(N_SEG, np.float32)
        out_m = np.empty(N_SEG, np.float32)
        reps, t_cpu = 0, 0.0
        while t_cpu < args.cpu_seconds and reps < 100:
            c0 = time.perf_counter()
            offc = np.empty(N_SEG + 1, np.int64)
            offc[0] = 0
            np.cumsum(lens, out=offc[1:])  # scanl (+) 0 of 2^16 lengths (sequential, n < 8192 per worker rule aside)
            O.native_fold1_seg_f32(cores, A.ADD, vals, offc, out_s)
            O.native_fold1_seg_f32(cores, A.MAX, vals, offc, out_m)
            t_cpu += time.perf_counter() - c0
            reps += 1
        cpu_gbs = step_bytes * reps / t_cpu / 1e9
        cpu = {"value": round(cpu_gbs, 2), "unit": "GB/s", "cores": cores, "kind": "port",
               "sample": f"the full step ({reps} repetitions, {t_cpu:.1f} s) with the native backend's schedule "
                         f"(divideWork over segments, {cores} threads)"}
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
                        "(lengths log-uniform in [1,16384], sum 2^26) + Int64 scanl (+) 0 of the lengths",
            "elements": N_VALUES,
            "segments": N_SEG,
            "seed": SEED,
            "step_bytes": step_bytes,
        },
        "frac_of_hbm_peak": round(value / HBM_PEAK_GBS, 4),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "skeletons": skel,
        "device": {"name": props.name.decode(), "arch": props.arch.decode(), "cus": props.compute_units},
    }
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
