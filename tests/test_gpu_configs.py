"""Every benchmark configuration at its stated size, checked against the oracle.

C1  scanl1 / scanl / scanr1 (+) and fold (+) on Int32 U[-1000,1000] (seed 42)
    at 2^24 and 2^28, bit-exact; Int64, Float32 and Float64 scans above the
    scan kernel's persistent grid (several tiles per workgroup, look-backs
    past the first window); a rank-2 scan whose rows do not divide the tile.
C2  Float32 fold1Seg (+) / max over 2^26 N(0,1) values in 2^16 segments, with
    the Int64 offsets from `scanl (+) 0` on the GPU; the FP32 error of every
    segment against the FP64 oracle.
C3  the 2^26 Zipf(1.1) histogram into 4096 bins (bit-exact); the 5x5 clamp and
    3x3 mirror stencils on an 8192^2 N(0,1) image against the FP64 oracle.
C4  all 10,000 batched problems, each against the oracle.
C5  the served request mix with sizes up to 2^24, every result checked.

Integer results bit-exact; floating (+) within |y - s| <= 1e-5 log2(n) sum|x|
of the FP64 truth; stencils within taps * 2^-23 * sum|w_i x_i|; max exact.
"""
import math
import sys

import numpy as np
import pytest

import acc_hip as A
import oracle_py as O

pytestmark = pytest.mark.gpu

ADD, MUL, MIN, MAX = 0, 1, 2, 3
L2R, R2L = 0, 1
SEED = 42


def say(msg):
    print(f"[configs] {msg}", file=sys.stderr, flush=True)


def fp_bound(nterms, abs_sum):
    return 1e-5 * np.maximum(1.0, np.log2(np.maximum(nterms, 2))) * abs_sum


@pytest.fixture(scope="module")
def c1_input(stream):
    x = O.gen_uniform_i32(SEED, -1000, 1000, 1 << 28)
    return x, A.DeviceArray.from_numpy(stream, x)


# ---------------------------------------------------------------------- C1
@pytest.mark.parametrize("n", [1 << 24, 1 << 28])
def test_c1_int32_scans_and_fold(stream, c1_input, n):
    """PTX/Execute.hs:422-468 (scanAllOp), :267-318 (foldAllOp): every output
    of the 2^24 and 2^28 scans equals the oracle's, in both directions."""
    x_all, d_all = c1_input
    x = x_all[:n]
    out = A.DeviceArray(n + 1, np.int32)
    say(f"C1 n={n} scanl1")
    A.scan(stream, ADD, L2R, None, d_all, 1, n, out)
    want = np.cumsum(x, dtype=np.int32)
    got = out.view(0, n).to_numpy(stream)
    assert np.array_equal(got, want), "scanl1 (+)"
    assert np.array_equal(want, O.scan(ADD, L2R, None, x)), "oracle agrees with numpy"
    say(f"C1 n={n} scanl 0")
    A.scan(stream, ADD, L2R, np.int32(0), d_all, 1, n, out)
    got = out.to_numpy(stream)
    assert got[0] == 0 and np.array_equal(got[1:], want), "scanl (+) 0"
    del want
    say(f"C1 n={n} scanr1")
    A.scan(stream, ADD, R2L, None, d_all, 1, n, out)
    got = out.view(0, n).to_numpy(stream)
    assert np.array_equal(got, np.cumsum(x[::-1], dtype=np.int32)[::-1]), "scanr1 (+)"
    say(f"C1 n={n} fold")
    o = A.DeviceArray(1, np.int32)
    A.fold(stream, ADD, np.int32(0), d_all, 1, n, o)
    assert o.to_numpy(stream)[0] == np.sum(x, dtype=np.int32) == O.fold(ADD, np.int32(0), x)[0]


@pytest.mark.parametrize("dt", [np.int64, np.float32, np.float64])
@pytest.mark.parametrize("direction", [L2R, R2L])
def test_c1_scans_beyond_the_grid(stream, dt, direction):
    """2^24 + 17 elements: 1025+ tiles of the persistent scan (grid = CUs x
    occupancy), so workgroups take several tickets and look-backs pass the
    first 192-tile window; the ragged tile is last (L2R) or first (R2L)."""
    n = (1 << 24) + 17
    if np.dtype(dt).kind == "f":
        x = O.gen_normal_f32(7, n).astype(dt)
    else:
        x = O.gen_uniform_i32(7, -1 << 30, (1 << 30) - 1, n).astype(dt) * 3
    d = A.DeviceArray.from_numpy(stream, x)
    out = A.DeviceArray(n, dt)
    say(f"beyond-grid scan {np.dtype(dt).name} dir={direction}")
    A.scan(stream, ADD, direction, None, d, 1, n, out)
    got = out.to_numpy(stream)
    xs = x if direction == L2R else x[::-1]
    if np.dtype(dt).kind == "i":
        want = np.cumsum(xs, dtype=dt)
        assert np.array_equal(got if direction == L2R else got[::-1], want)
    else:
        x64 = xs.astype(np.float64)
        truth = np.cumsum(x64)
        ab = np.cumsum(np.abs(x64))
        g = (got if direction == L2R else got[::-1]).astype(np.float64)
        bound = fp_bound(np.arange(1, n + 1), ab)
        assert np.all(np.abs(g - truth) <= bound + 1e-300)


def test_c1_scan_prime_beyond_the_grid(stream):
    n = (1 << 24) + 5
    x = O.gen_uniform_i32(8, -1000, 1000, n)
    d = A.DeviceArray.from_numpy(stream, x)
    out, sums = A.DeviceArray(n, np.int32), A.DeviceArray(1, np.int32)
    for direction in (L2R, R2L):
        A.scan_prime(stream, ADD, direction, np.int32(10), d, 1, n, out, sums)
        want, wsum = O.scan_prime(ADD, direction, np.int32(10), x)
        assert np.array_equal(out.to_numpy(stream), want) and sums.to_numpy(stream)[0] == wsum[0]


@pytest.mark.parametrize("direction", [L2R, R2L])
def test_c1_rank2_rows_not_dividing_the_tile(stream, direction):
    """(40, 999999): rows that do not divide the tile, more tiles than the
    grid, so the rows' segmented scan takes tiles by ticket."""
    outer, inner = 40, 999999
    x = O.gen_uniform_i32(9, -1000, 1000, outer * inner)
    d = A.DeviceArray.from_numpy(stream, x)
    out = A.DeviceArray(outer * (inner + 1), np.int32)
    A.scan(stream, ADD, direction, None, d, outer, inner, out)
    got = out.view(0, outer * inner).to_numpy(stream)
    assert np.array_equal(got, O.scan(ADD, direction, None, x, outer))
    A.scan(stream, ADD, direction, np.int32(3), d, outer, inner, out)
    assert np.array_equal(out.to_numpy(stream), O.scan(ADD, direction, np.int32(3), x, outer))


def test_c1_fold_mul_wrapping_pinned(stream):
    """fold (*) over odd Int32 values: the wrapped product of odd numbers is
    odd, so it cannot collapse to 0; numpy's multiply.reduce wraps the same
    way (two's complement, no nsw)."""
    n = (1 << 20) + 3
    x = (O.gen_uniform_i32(11, -1 << 30, (1 << 30) - 1, n) | 1).astype(np.int32)
    d = A.DeviceArray.from_numpy(stream, x)
    o = A.DeviceArray(1, np.int32)
    A.fold(stream, MUL, np.int32(3), d, 1, n, o)
    want = np.int32(np.multiply.reduce(np.concatenate([[3], x]).astype(np.int32), dtype=np.int32))
    assert want % 2 != 0
    assert o.to_numpy(stream)[0] == want == O.fold(MUL, np.int32(3), x)[0]


# ---------------------------------------------------------------------- C2
def test_c2_fold1seg_full_size_with_per_segment_error(stream):
    n, m = 1 << 26, 1 << 16
    lens = O.gen_segment_lengths(SEED, m, 16384, n)
    vals = O.gen_normal_f32(SEED, n)
    d_lens, d_vals = A.DeviceArray.from_numpy(stream, lens), A.DeviceArray.from_numpy(stream, vals)
    d_off = A.DeviceArray(m + 1, np.int64)
    say("C2 offsets scan")
    A.scan(stream, ADD, L2R, np.int64(0), d_lens, 1, m, d_off)
    off = d_off.to_numpy(stream)
    assert np.array_equal(off, O.scan(ADD, L2R, np.int64(0), lens))
    d_out = A.DeviceArray(m, np.float32)
    say("C2 fold1Seg (+)")
    A.fold_seg(stream, ADD, None, d_vals, 1, n, d_off, m, d_out)
    truth, ab = O.fold_seg_f64(vals, off)
    err = np.abs(d_out.to_numpy(stream).astype(np.float64) - truth)
    bound = fp_bound(np.diff(off), ab)
    rel = err / np.maximum(ab, 1e-300)
    say(f"C2 fold1Seg (+) per-segment |y-s|/sum|x|: max {rel.max():.3e} p99 {np.quantile(rel, 0.99):.3e}; "
        f"max err/bound {(err / bound).max():.3f}")
    assert np.all(err <= bound)
    say("C2 fold1Seg max")
    A.fold_seg(stream, MAX, None, d_vals, 1, n, d_off, m, d_out)
    assert np.array_equal(d_out.to_numpy(stream), O.fold_seg(MAX, None, vals, off))


# ---------------------------------------------------------------------- C3
def test_c3_histogram_full_size(stream):
    n = 1 << 26
    idx = O.gen_zipf_i32(SEED, 1.1, 4096, n)
    d_idx = A.DeviceArray.from_numpy(stream, idx)
    zero = A.DeviceArray.from_numpy(stream, np.zeros(4096, np.int32))
    bins = A.DeviceArray(4096, np.int32)
    A.permute(stream, ADD, zero, bins, 4096, None, 1, d_idx, n)
    got = bins.to_numpy(stream)
    assert np.array_equal(got, np.bincount(idx, minlength=4096)) and int(got.sum()) == n


@pytest.mark.parametrize("case", [(2, 2, A.CLAMP, "5x5 clamp"), (1, 1, A.MIRROR, "3x3 mirror")])
def test_c3_stencils_full_size(stream, case):
    ry, rx, b, name = case
    H = W = 8192
    img = O.gen_normal_f32(SEED, H * W).reshape(H, W)
    taps = [1, 4, 6, 4, 1] if ry == 2 else [1, 2, 1]
    w = (np.outer(taps, taps) / float(sum(taps) ** 2)).astype(np.float32)
    d_img, d_out = A.DeviceArray.from_numpy(stream, img), A.DeviceArray(H * W, np.float32)
    say(f"C3 stencil {name}")
    A.stencil2d(stream, ry, rx, w, b, None, d_img, H, W, d_out)
    got = d_out.to_numpy(stream).reshape(H, W).astype(np.float64)
    truth, ab = O.stencil2d_f64(ry, rx, w, b, None, img)
    assert np.all(np.abs(got - truth) <= w.size * 2.0 ** -23 * ab + 1e-30)


# ---------------------------------------------------------------------- C4
def test_c4_all_batched_problems(stream):
    """10,000 problems: even i Float32 scanl1 (+) U(-1,1), odd i Int32 fold (+) 0
    U[-100,100], sizes floor(exp(U(ln 2^8, ln 2^20))), one persistent launch
    replayed from a captured graph; every problem against the oracle."""
    P = 10000
    rng = np.random.default_rng(SEED)
    sizes = np.floor(np.exp(rng.uniform(math.log(2 ** 8), math.log(2 ** 20), size=P))).astype(np.int64)
    pad = lambda k: (k + 63) // 64 * 64
    scan_idx, fold_idx = np.arange(0, P, 2), np.arange(1, P, 2)
    s_off = np.concatenate([[0], np.cumsum([pad(int(sizes[i])) for i in scan_idx])]).astype(np.int64)
    f_off = np.concatenate([[0], np.cumsum([pad(int(sizes[i])) for i in fold_idx])]).astype(np.int64)
    xs = O.gen_uniform_f32(SEED, -1.0, 1.0, int(s_off[-1]))
    xf = O.gen_uniform_i32(SEED + 1, -100, 100, int(f_off[-1]))
    d_xs, d_xf = A.DeviceArray.from_numpy(stream, xs), A.DeviceArray.from_numpy(stream, xf)
    d_ys, d_yf = A.DeviceArray(int(s_off[-1]), np.float32), A.DeviceArray(len(fold_idx) * 64, np.int32)
    probs = []
    for j, i in enumerate(scan_idx):
        k = int(sizes[i])
        probs.append((A.BATCH_SCAN, ADD, None, d_xs.view(int(s_off[j]), k), k, d_ys.view(int(s_off[j]), k)))
    for j, i in enumerate(fold_idx):
        k = int(sizes[i])
        probs.append((A.BATCH_FOLD, ADD, np.int32(0), d_xf.view(int(f_off[j]), k), k, d_yf.view(j * 64, 1)))
    say(f"C4 batch of {P} problems, {int(sizes.sum())} elements")
    b = A.Batch(0, probs, tile=0)
    say(f"C4 shape (block, vectors, run, tickets) = {b.shape()}")
    g = A.Graph.capture(stream, lambda: b.launch(stream))
    for rep in range(2):
        g.launch(stream)
        stream.synchronize()
        say(f"C4 replay {rep} done")
        got_f = d_yf.to_numpy(stream)[::64]
        for j, i in enumerate(fold_idx):
            k = int(sizes[i])
            assert got_f[j] == np.sum(xf[int(f_off[j]):int(f_off[j]) + k], dtype=np.int32), (rep, int(i))
        got_s = d_ys.to_numpy(stream)
        for j, i in enumerate(scan_idx):
            k, o = int(sizes[i]), int(s_off[j])
            x64 = xs[o:o + k].astype(np.float64)
            err = np.abs(got_s[o:o + k].astype(np.float64) - np.cumsum(x64))
            assert np.all(err <= fp_bound(np.arange(1, k + 1), np.cumsum(np.abs(x64))) + 1e-300), (rep, int(i))
        d_ys_clear = np.zeros(1, np.float32)
        A.check(A.lib().acc_hip_fill_async(stream.handle, A.c_void_p(d_ys.ptr), d_ys_clear.ctypes.data_as(A.c_void_p),
                                           4, d_ys.count))
        A.check(A.lib().acc_hip_fill_async(stream.handle, A.c_void_p(d_yf.ptr), d_ys_clear.ctypes.data_as(A.c_void_p),
                                           4, d_yf.count))
    g.destroy()
    b.destroy()


# ---------------------------------------------------------------------- C5
def test_c5_served_mix_up_to_2_24(stream):
    """Requests of the served mix (35% scanl1, 30% fold1Seg, 20% histogram,
    15% 3x3 mirror stencil; Float32/Int32/Int64) with sizes log-uniform up
    to 2^24, submitted at once (so they run concurrently), each checked."""
    N = 1 << 24
    host = {"f32": O.gen_normal_f32(1, N), "i32": O.gen_uniform_i32(2, -1000, 1000, N)}
    host["i64"] = O.gen_uniform_i32(3, -1000, 1000, N).astype(np.int64)
    pool = {k: A.DeviceArray.from_numpy(stream, v) for k, v in host.items()}
    lens = O.gen_segment_lengths(4, N // 1024, 16384, N)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    d_off = A.DeviceArray.from_numpy(stream, off)
    idx = O.gen_zipf_i32(5, 1.1, 4096, N)
    d_idx = A.DeviceArray.from_numpy(stream, idx)
    dts = {"f32": np.float32, "i32": np.int32, "i64": np.int64}
    rng = np.random.default_rng(SEED + 7)
    srv = A.Server(0, slots=8)
    reqs = []
    sizes = [N] + [int(math.exp(rng.uniform(math.log(2 ** 10), math.log(2 ** 24)))) for _ in range(119)]
    for i, n in enumerate(sizes):
        key = ["f32", "i32", "i64"][i % 3 if i < 12 else rng.integers(0, 3)]
        dt = dts[key]
        kind = i % 4 if i < 12 else rng.choice(4, p=[0.35, 0.30, 0.20, 0.15])
        if kind == 0:
            r = A.Server.request(A.REQ_SCAN, dt, ADD, src=pool[key], n=n)
        elif kind == 1:
            m = max(1, int(np.searchsorted(off, n, side="right")) - 1)
            n = int(off[m])
            r = A.Server.request(A.REQ_FOLD_SEG, dt, ADD, src=pool[key], n=n, m=m, aux=d_off)
        elif kind == 2:
            r = A.Server.request(A.REQ_PERMUTE, dt, ADD, seed=0, n=n, m=4096, aux=d_idx, fill=1)
        else:
            n = max(3, int(math.sqrt(n)))
            w = (np.outer([1, 2, 1], [1, 2, 1]) / (16 if key == "f32" else 1)).astype(dt)
            r = A.Server.request(A.REQ_STENCIL2D, dt, src=pool[key], n=n, m=n, ry=1, rx=1, boundary=A.MIRROR,
                                 weights=w)
        reqs.append((int(kind), key, dt, n, r))
    ids = [srv.submit(r, keep=True) for (_, _, _, _, r) in reqs]
    say(f"C5 {len(ids)} requests submitted")
    for rid, (kind, key, dt, n, r) in zip(ids, reqs):
        ptr, lat = srv.wait(rid)
        x = host[key]
        if kind == 0:
            got = A.copy_to_host(stream, ptr, n, dt)
            if key == "f32":
                truth, ab = O.scan_f64(x[:n])
                ok = np.all(np.abs(got.astype(np.float64) - truth) <= fp_bound(np.arange(1, n + 1), ab) + 1e-300)
            else:
                ok = np.array_equal(got, np.cumsum(x[:n], dtype=dt))
        elif kind == 1:
            m = int(r[0].m)
            got = A.copy_to_host(stream, ptr, m, dt)
            if key == "f32":
                truth, ab = O.fold_seg_f64(x[:n], off[:m + 1])
                ok = np.all(np.abs(got.astype(np.float64) - truth) <= fp_bound(np.diff(off[:m + 1]), ab) + 1e-300)
            else:
                ok = np.array_equal(got, O.fold_seg(ADD, None, x[:n], off[:m + 1]))
        elif kind == 2:
            got = A.copy_to_host(stream, ptr, 4096, dt)
            ok = np.array_equal(got, np.bincount(idx[:n], minlength=4096).astype(dt))
        else:
            got = A.copy_to_host(stream, ptr, n * n, dt).reshape(n, n)
            img = x[:n * n].reshape(n, n)
            if key == "f32":
                w = (np.outer([1, 2, 1], [1, 2, 1]) / 16).astype(np.float32)
                truth, ab = O.stencil2d_f64(1, 1, w, A.MIRROR, None, img)
                ok = np.all(np.abs(got.astype(np.float64) - truth) <= 9 * 2.0 ** -23 * ab + 1e-30)
            else:
                w = np.outer([1, 2, 1], [1, 2, 1]).astype(dt)
                ok = np.array_equal(got, O.stencil2d(1, 1, w, A.MIRROR, None, img))
        assert ok, (kind, key, n)
        srv.release(rid)
    st = srv.stats()
    assert st.completed == len(ids) and st.failed == 0
    srv.destroy()
