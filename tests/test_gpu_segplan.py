"""Segment tile plans: foldSeg over one offsets array with the plan built once
(acc_hip_seg_plan_build + acc_hip_fold_seg_planned).  The planned call runs the
same tile kernels on the same tiles as acc_hip_fold_seg, so its results must
be bit-identical to the unplanned call's (floating sums included), which the
oracle checks in turn (integers and max bit-exact, floating (+) within the FP32
bound of the FP64 truth)."""
import numpy as np
import pytest

import acc_hip as A
import oracle_py as O

pytestmark = pytest.mark.gpu

TYPES = [np.float32, np.int32, np.float64, np.int64]


def segments(rng, n, nseg, empties=False):
    if empties:
        lens = rng.multinomial(n, np.ones(nseg) / nseg)
    else:
        cut = np.sort(rng.choice(np.arange(1, n), size=nseg - 1, replace=False))
        lens = np.diff(np.concatenate([[0], cut, [n]]))
    return np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)


def data(rng, dt, n):
    if np.dtype(dt).kind == "f":
        return rng.standard_normal(n).astype(dt)
    return rng.integers(-1000, 1000, size=n).astype(dt)


def unplanned(stream, op, seed, d, outer, inner, d_off, nseg, dt):
    o = A.DeviceArray.from_numpy(stream, np.full(outer * nseg, 7, dt))
    A.fold_seg(stream, op, seed, d, outer, inner, d_off, nseg, o)
    return o.to_numpy(stream)


def planned(stream, plan, op, seed, d, outer, nseg, dt):
    o = A.DeviceArray.from_numpy(stream, np.full(outer * nseg, 7, dt))
    plan.fold_seg(stream, op, seed, d, outer, o)
    return o.to_numpy(stream)


def against_oracle(op, seed, x, off, got, outer=1):
    want = O.fold_seg(op, seed, x, off, outer)
    if seed is None:  # fold1Seg leaves empty segments alone
        empty = np.tile(np.diff(off) == 0, outer)
        want = np.where(empty, got, want)
    if np.dtype(x.dtype).kind != "f" or op == A.MAX:
        assert np.array_equal(got, want)
    else:
        lens = np.tile(np.diff(off), outer)
        ab = np.abs(O.fold_seg(A.ADD, None, np.abs(x.astype(np.float64)), off, outer))
        bound = 1e-5 * np.maximum(1, np.log2(np.maximum(lens, 2))) * ab + 1e-300
        assert np.all(np.abs(got.astype(np.float64) - want.astype(np.float64)) <= bound + np.abs(want) * 1e-6)


@pytest.mark.parametrize("dt", TYPES)
@pytest.mark.parametrize("offs", [np.int32, np.int64])
def test_planned_equals_unplanned(stream, dt, offs):
    rng = np.random.default_rng(TYPES.index(dt) * 2 + (offs == np.int64))
    n, nseg = 1_000_003, 3001
    x = data(rng, dt, n)
    off = segments(rng, n, nseg).astype(offs)
    d, d_off = A.DeviceArray.from_numpy(stream, x), A.DeviceArray.from_numpy(stream, off)
    plan = A.SegPlan(0)
    plan.build(stream, dt, d, n, d_off, nseg)
    for op in (A.ADD, A.MAX):  # two folds on one plan
        got = planned(stream, plan, op, None, d, 1, nseg, dt)
        assert np.array_equal(got, unplanned(stream, op, None, d, 1, n, d_off, nseg, dt))
        against_oracle(op, None, x, off.astype(np.int64), got)
    plan.destroy()


def test_one_plan_many_arrays_and_modes(stream):
    """Short segments (more than 512 starts in a tile), long ones (a segment
    over many tiles), empty ones with a seed, for two different arrays."""
    rng = np.random.default_rng(5)
    n = 1 << 21
    lens = np.concatenate([rng.integers(0, 4, size=60000), [300000], rng.integers(1, 2000, size=900)])
    lens[-1] += n - lens.sum()
    assert lens[-1] >= 0
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    nseg = lens.size
    d_off = A.DeviceArray.from_numpy(stream, off)
    plan = A.SegPlan(0)
    for k, dt in enumerate((np.float32, np.int32)):
        x = data(rng, dt, n)
        d = A.DeviceArray.from_numpy(stream, x)
        plan.build(stream, dt, d, n, d_off, nseg)
        for op, seed in ((A.ADD, dt(3)), (A.MAX, None), (A.ADD, None)):
            got = planned(stream, plan, op, seed, d, 1, nseg, dt)
            assert np.array_equal(got, unplanned(stream, op, seed, d, 1, n, d_off, nseg, dt)), (dt, op, seed)
            against_oracle(op, seed, x, off, got)
    plan.destroy()


def test_misaligned_views_and_rows(stream):
    rng = np.random.default_rng(9)
    inner, nseg, outer = 40000, 37, 3  # rows of 160000 bytes: 16-byte multiples
    x = rng.standard_normal(outer * inner + 5).astype(np.float32)
    off = segments(rng, inner, nseg)
    base = A.DeviceArray.from_numpy(stream, x)
    d_off = A.DeviceArray.from_numpy(stream, off)
    plan = A.SegPlan(0)
    v = base.view(1, outer * inner)  # 4 bytes past a 16-byte boundary
    plan.build(stream, np.float32, v, inner, d_off, nseg)
    got = planned(stream, plan, A.ADD, None, v, outer, nseg, np.float32)
    assert np.array_equal(got, unplanned(stream, A.ADD, None, v, outer, inner, d_off, nseg, np.float32))
    against_oracle(A.ADD, None, x[1:1 + outer * inner], off, got, outer)
    # another alignment, or rows that are not 16-byte multiples, are refused
    o = A.DeviceArray(outer * nseg, np.float32)
    with pytest.raises(A.AccError):
        plan.fold_seg(stream, A.ADD, None, base.view(2, inner), 1, o)
    odd = A.SegPlan(0)
    odd.build(stream, np.float32, base, inner + 1, A.DeviceArray.from_numpy(stream, segments(rng, inner + 1, 5)), 5)
    with pytest.raises(A.AccError):
        odd.fold_seg(stream, A.ADD, None, base, 2, o)
    # the element type is the plan's
    with pytest.raises(A.AccError):
        plan.fold_seg(stream, A.ADD, None, A.DeviceArray(inner, np.int32), 1, A.DeviceArray(nseg, np.int32))
    odd.destroy()
    plan.destroy()


def test_unbuilt_plan_and_rebuild_larger(stream):
    rng = np.random.default_rng(13)
    plan = A.SegPlan(0)
    with pytest.raises(A.AccError):
        plan.fold_seg(stream, A.ADD, None, A.DeviceArray(16, np.float32), 1, A.DeviceArray(1, np.float32))
    for n, nseg in ((5000, 7), (3_000_000, 2000), (123, 123)):  # the second grows the plan
        x = data(rng, np.int64, n)
        off = segments(rng, n, nseg) if nseg < n else np.arange(n + 1, dtype=np.int64)
        d, d_off = A.DeviceArray.from_numpy(stream, x), A.DeviceArray.from_numpy(stream, off)
        plan.build(stream, np.int64, d, n, d_off, nseg)
        got = planned(stream, plan, A.ADD, np.int64(-4), d, 1, nseg, np.int64)
        against_oracle(A.ADD, np.int64(-4), x, off, got)
    plan.destroy()


def test_c2_size_planned_equals_unplanned(stream):
    """The benchmark step's shape: 2^26 Float32 in 2^16 segments, (+) and max."""
    from acc_hip import workloads as W
    lens = W.segment_lengths(42, 1 << 16, 16384, 1 << 26)
    off = W.offsets(lens)
    x = W.normal_f32(42, 1 << 26)
    d, d_off = A.DeviceArray.from_numpy(stream, x), A.DeviceArray.from_numpy(stream, off)
    plan = A.SegPlan(0)
    plan.build(stream, np.float32, d, x.size, d_off, lens.size)
    for op in (A.ADD, A.MAX):
        got = planned(stream, plan, op, None, d, 1, lens.size, np.float32)
        assert np.array_equal(got, unplanned(stream, op, None, d, 1, x.size, d_off, lens.size, np.float32))
    plan.destroy()


def test_empty_inputs(stream):
    """No elements: every segment is empty (the seed, or untouched without one);
    no segments: nothing to do."""
    plan = A.SegPlan(0)
    d = A.DeviceArray(4, np.float32)
    off = A.DeviceArray.from_numpy(stream, np.zeros(6, np.int64))
    plan.build(stream, np.float32, d, 0, off, 5)
    got = planned(stream, plan, A.ADD, np.float32(2.5), d, 1, 5, np.float32)
    assert np.array_equal(got, np.full(5, 2.5, np.float32))
    got = planned(stream, plan, A.MAX, None, d, 1, 5, np.float32)
    assert np.array_equal(got, np.full(5, 7, np.float32))
    plan.build(stream, np.float32, d, 4, off, 0)
    o = A.DeviceArray.from_numpy(stream, np.full(1, 7, np.float32))
    plan.fold_seg(stream, A.ADD, np.float32(1), d, 1, o)
    assert o.to_numpy(stream)[0] == 7
    plan.destroy()


def lengths_case(case, rng):
    if case == "mixed":
        from acc_hip import workloads as W
        return W.segment_lengths(3, 3000, 16384, 2_000_003)
    if case == "long":  # segments over hundreds of fold tiles, empty ones between
        return np.array([5, 3_000_000, 0, 7, 1_000_000, 0, 0, 12], np.int64)
    if case == "tiny":  # more than 64 starts per fold tile, many empty segments
        return rng.integers(0, 3, size=50000).astype(np.int64)
    if case == "tile_edge":  # nseg a multiple of the scan's tile, and one more
        return rng.integers(1, 100, size=2048).astype(np.int64)
    return rng.integers(1, 100, size=2049).astype(np.int64)


@pytest.mark.parametrize("lt", [np.int32, np.int64])
@pytest.mark.parametrize("case", ["mixed", "long", "tiny", "tile_edge", "tile_edge_plus1"])
def test_build_from_lengths(stream, lt, case):
    """scanl (+) 0 of the lengths and the plan in one launch: the offsets equal
    numpy's, and folds on that plan equal the unplanned folds over them."""
    rng = np.random.default_rng(len(case) * 7 + (lt == np.int64))
    lens = lengths_case(case, rng)
    nseg, n = lens.size, int(lens.sum())
    x = data(rng, np.float32, n + 3)
    base = A.DeviceArray.from_numpy(stream, x)
    d = base.view(3, n) if case == "mixed" else base.view(0, n)  # a view 12 bytes past a 16-byte boundary
    xs = x[3:3 + n] if case == "mixed" else x[:n]
    d_len = A.DeviceArray.from_numpy(stream, lens.astype(lt))
    d_off = A.DeviceArray(nseg + 1, lt)
    plan = A.SegPlan(0)
    plan.build_lengths(stream, np.float32, d, n, d_len, nseg, d_off)
    off = d_off.to_numpy(stream)
    assert np.array_equal(off, np.concatenate([[0], np.cumsum(lens)]).astype(lt))
    for op, seed in ((A.ADD, None), (A.MAX, None), (A.ADD, np.float32(1.5))):
        got = planned(stream, plan, op, seed, d, 1, nseg, np.float32)
        assert np.array_equal(got, unplanned(stream, op, seed, d, 1, n, d_off, nseg, np.float32)), (op, seed)
        against_oracle(op, seed, xs, off.astype(np.int64), got)
    plan.destroy()


def test_build_from_lengths_offsets_wrap_like_the_scan(stream):
    """Int32 lengths whose sum passes 2^31: the offsets wrap as the scan's do."""
    lens = np.full(3000, 1_000_000, np.int32)
    d_len = A.DeviceArray.from_numpy(stream, lens)
    d_off = A.DeviceArray(lens.size + 1, np.int32)
    want = A.DeviceArray(lens.size + 1, np.int32)
    A.scan(stream, A.ADD, A.LEFT_TO_RIGHT, np.int32(0), d_len, 1, lens.size, want)
    plan = A.SegPlan(0)
    plan.build_lengths(stream, np.float32, A.DeviceArray(16, np.float32), 16, d_len, lens.size, d_off)
    assert np.array_equal(d_off.to_numpy(stream), want.to_numpy(stream))
    assert np.array_equal(d_off.to_numpy(stream), np.concatenate([[0], np.cumsum(lens, dtype=np.int32)]))
    plan.destroy()


def test_c2_step_from_lengths(stream):
    """The benchmark step's shape through the one-launch offsets and plan."""
    from acc_hip import workloads as W
    lens = W.segment_lengths(42, 1 << 16, 16384, 1 << 26)
    x = W.normal_f32(42, 1 << 26)
    d, d_len = A.DeviceArray.from_numpy(stream, x), A.DeviceArray.from_numpy(stream, lens)
    d_off = A.DeviceArray(lens.size + 1, np.int64)
    plan = A.SegPlan(0)
    for _ in range(2):  # a rebuild on the same plan
        plan.build_lengths(stream, np.float32, d, x.size, d_len, lens.size, d_off)
        assert np.array_equal(d_off.to_numpy(stream), W.offsets(lens))
        for op in (A.ADD, A.MAX):
            got = planned(stream, plan, op, None, d, 1, lens.size, np.float32)
            assert np.array_equal(got, unplanned(stream, op, None, d, 1, x.size, d_off, lens.size, np.float32))
    plan.destroy()


@pytest.mark.parametrize("dt", TYPES)
def test_several_folds_in_one_call(stream, dt):
    """acc_hip_fold_seg_planned_multi: each fold bit-identical to its own
    planned call (the tile kernels are the same; the fix-ups share a launch)."""
    rng = np.random.default_rng(17 + TYPES.index(dt))
    n, nseg = 700_001, 1999
    off = segments(rng, n, nseg, empties=True)
    d_off = A.DeviceArray.from_numpy(stream, off)
    xs = [data(rng, dt, n) for _ in range(4)]
    ds = [A.DeviceArray.from_numpy(stream, x) for x in xs]
    plan = A.SegPlan(0)
    plan.build(stream, dt, ds[0], n, d_off, nseg)
    mul = A.MUL if np.dtype(dt).kind != "f" else A.ADD
    spec = [(A.ADD, None), (A.MAX, dt(2)), (A.MIN, None), (mul, dt(3))]
    outs = [A.DeviceArray.from_numpy(stream, np.full(nseg, 7, dt)) for _ in spec]
    plan.fold_seg_multi(stream, [(op, sd, d, o) for (op, sd), d, o in zip(spec, ds, outs)])
    for (op, sd), d, o, x in zip(spec, ds, outs, xs):
        got = o.to_numpy(stream)
        assert np.array_equal(got, planned(stream, plan, op, sd, d, 1, nseg, dt)), (op, sd)
        against_oracle(op, sd, x, off, got) if op != A.MUL else None
    plan.destroy()


@pytest.mark.parametrize("ops", [(A.ADD, A.MAX), (A.MAX, A.ADD), (A.ADD, A.ADD), (A.MAX, A.MAX)])
@pytest.mark.parametrize("offs", [np.int32, np.int64])
def test_two_float32_folds_in_one_launch(stream, ops, offs):
    """Two Float32 folds of (+) or max in one multi call run both folds' tiles
    in one launch (k_fold_seg_pair); each result is bit-identical to its own
    planned call, with and without a seed, empty segments included."""
    rng = np.random.default_rng(31 + 2 * ops[0] + ops[1] + (offs == np.int64))
    n, nseg = 1_000_003, 3001
    off = segments(rng, n, nseg, empties=True).astype(offs)
    xs = [data(rng, np.float32, n) for _ in range(2)]
    ds = [A.DeviceArray.from_numpy(stream, x) for x in xs]
    d_off = A.DeviceArray.from_numpy(stream, off)
    plan = A.SegPlan(0)
    plan.build(stream, np.float32, ds[0], n, d_off, nseg)
    spec = [(ops[0], None), (ops[1], np.float32(2.5))]
    outs = [A.DeviceArray.from_numpy(stream, np.full(nseg, 7, np.float32)) for _ in spec]
    plan.fold_seg_multi(stream, [(op, sd, d, o) for (op, sd), d, o in zip(spec, ds, outs)])
    for (op, sd), d, o, x in zip(spec, ds, outs, xs):
        got = o.to_numpy(stream)
        assert np.array_equal(got, planned(stream, plan, op, sd, d, 1, nseg, np.float32)), (op, sd)
        against_oracle(op, sd, x, off.astype(np.int64), got)
    plan.destroy()


def test_two_float32_folds_in_one_launch_at_c2_size(stream):
    """The benchmark step's folds (+) and max over 2^26 Float32 values in 2^16
    segments through one multi call: each equals its own planned call."""
    from acc_hip import workloads as W
    lens = W.segment_lengths(42, 1 << 16, 16384, 1 << 26)
    x = W.normal_f32(42, 1 << 26)
    d, d_len = A.DeviceArray.from_numpy(stream, x), A.DeviceArray.from_numpy(stream, lens)
    d_off = A.DeviceArray(lens.size + 1, np.int64)
    plan = A.SegPlan(0)
    plan.build_lengths(stream, np.float32, d, x.size, d_len, lens.size, d_off)
    o_sum, o_max = A.DeviceArray(lens.size, np.float32), A.DeviceArray(lens.size, np.float32)
    plan.fold_seg_multi(stream, [(A.ADD, None, d, o_sum), (A.MAX, None, d, o_max)])
    assert np.array_equal(o_sum.to_numpy(stream), planned(stream, plan, A.ADD, None, d, 1, lens.size, np.float32))
    assert np.array_equal(o_max.to_numpy(stream), planned(stream, plan, A.MAX, None, d, 1, lens.size, np.float32))
    plan.destroy()


def test_several_folds_limits(stream):
    rng = np.random.default_rng(23)
    plan = A.SegPlan(0)
    # eight folds of one small tile (no fix-up launch), and nine refused
    x = data(rng, np.float32, 100)
    off = segments(rng, 100, 9)
    d, d_off = A.DeviceArray.from_numpy(stream, x), A.DeviceArray.from_numpy(stream, off)
    plan.build(stream, np.float32, d, 100, d_off, 9)
    outs = [A.DeviceArray(9, np.float32) for _ in range(9)]
    plan.fold_seg_multi(stream, [(A.ADD if k % 2 else A.MAX, None, d, outs[k]) for k in range(8)])
    for k in range(8):
        op = A.ADD if k % 2 else A.MAX
        assert np.array_equal(outs[k].to_numpy(stream), planned(stream, plan, op, None, d, 1, 9, np.float32))
    with pytest.raises(A.AccError):
        plan.fold_seg_multi(stream, [(A.ADD, None, d, outs[k]) for k in range(9)])
    # no elements: the seeds
    plan.build(stream, np.float32, d, 0, A.DeviceArray.from_numpy(stream, np.zeros(4, np.int64)), 3)
    o1, o2 = A.DeviceArray.from_numpy(stream, np.full(3, 7, np.float32)), A.DeviceArray(3, np.float32)
    plan.fold_seg_multi(stream, [(A.ADD, None, d, o1), (A.MAX, np.float32(-1), d, o2)])
    assert np.array_equal(o1.to_numpy(stream), np.full(3, 7, np.float32))
    assert np.array_equal(o2.to_numpy(stream), np.full(3, -1, np.float32))
    plan.destroy()
