"""The runtime half of the C ABI on the GPU: code objects loaded, inspected and
launched through the module API; the out-of-memory status under an allocation
limit; a cross-stream dependency through an event; graphs bound to their
capture stream.  All calls through ctypes -> libaccelerate_hip.so."""
import ctypes
import os

import numpy as np
import pytest

import acc_hip as A

pytestmark = pytest.mark.gpu

L = A.lib
HSACO = os.path.join(os.path.dirname(A.LIB_PATH), "foreign_axpy.hsaco")


def load_module():
    image = open(HSACO, "rb").read()
    buf = ctypes.create_string_buffer(image, len(image))
    m = ctypes.c_void_p()
    A.check(L().acc_hip_module_load_data(0, buf, len(image), ctypes.byref(m)))
    return m, buf


def get_function(m, name):
    f = ctypes.c_void_p()
    A.check(L().acc_hip_module_get_function(m, name.encode(), ctypes.byref(f)))
    return f


def test_module_load_inspect_launch(stream):
    """Link.link: load the image, look the kernel up, read its attributes
    (PTX/Link.hs:57-121); launch with an argument vector (PTX/Execute.hs:860-881)."""
    m, keep = load_module()
    f = get_function(m, "acc_foreign_axpy")
    regs, lds, mt = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    A.check(L().acc_hip_function_attributes(f, ctypes.byref(regs), ctypes.byref(lds), ctypes.byref(mt)))
    assert regs.value > 0 and lds.value == 0 and mt.value == 256
    n = 100003
    x = np.random.default_rng(1).standard_normal(n).astype(np.float32)
    y = np.random.default_rng(2).standard_normal(n).astype(np.float32)
    dx, dy = A.DeviceArray.from_numpy(stream, x), A.DeviceArray.from_numpy(stream, y)
    a = ctypes.c_float(2.5)
    px, py, cn = ctypes.c_void_p(dx.ptr), ctypes.c_void_p(dy.ptr), ctypes.c_int64(n)
    args = (ctypes.c_void_p * 4)(*[ctypes.cast(ctypes.byref(v), ctypes.c_void_p) for v in (a, px, py, cn)])
    A.check(L().acc_hip_launch(f, 64, 1, 1, 256, 1, 1, 0, stream.handle, args))
    got = dy.to_numpy(stream)
    # a fused multiply-add rounds once; the float64 restatement may round
    # twice, so allow one unit in the last place
    want = (np.float64(2.5) * x.astype(np.float64) + y.astype(np.float64)).astype(np.float32)
    assert np.all(np.abs(got - want) <= np.spacing(np.abs(want)))
    A.check(L().acc_hip_module_unload(m))


def test_zero_size_launch_is_skipped(stream):
    """executeOp skips launches with an empty grid (PTX/Execute.hs:847-854)."""
    m, keep = load_module()
    f = get_function(m, "acc_foreign_count")
    c = A.DeviceArray.from_numpy(stream, np.zeros(1, np.uint32))
    pc = ctypes.c_void_p(c.ptr)
    args = (ctypes.c_void_p * 1)(ctypes.cast(ctypes.byref(pc), ctypes.c_void_p))
    A.check(L().acc_hip_launch(f, 0, 1, 1, 64, 1, 1, 0, stream.handle, args))
    A.check(L().acc_hip_launch(f, 3, 0, 1, 64, 1, 1, 0, stream.handle, args))
    assert c.to_numpy(stream)[0] == 0
    A.check(L().acc_hip_launch(f, 3, 2, 1, 64, 1, 1, 0, stream.handle, args))
    assert c.to_numpy(stream)[0] == 6
    A.check(L().acc_hip_module_unload(m))


def test_missing_function_is_an_error(stream):
    m, keep = load_module()
    f = ctypes.c_void_p()
    assert L().acc_hip_module_get_function(m, b"no_such_kernel", ctypes.byref(f)) == A.HIP_ERROR
    A.check(L().acc_hip_module_unload(m))


def test_allocation_limit_gives_out_of_memory_then_recovers(stream):
    """Remote.malloc returns Nothing on OOM so the caller can evict and retry
    (PTX/Array/Remote.hs:63-74): the status is distinct, and after a free the
    same allocation succeeds."""
    used = ctypes.c_size_t()
    A.check(L().acc_hip_allocated_bytes(0, ctypes.byref(used)))
    limit = used.value + (8 << 20)
    A.check(L().acc_hip_set_alloc_limit(0, limit))
    try:
        a = A.DeviceArray(6 << 20, np.uint8)
        with pytest.raises(A.OutOfMemory):
            A.DeviceArray(4 << 20, np.uint8)
        free, total = ctypes.c_size_t(), ctypes.c_size_t()
        A.check(L().acc_hip_mem_get_info(0, ctypes.byref(free), ctypes.byref(total)))
        assert free.value <= 2 << 20 and total.value == limit
        a.free()
        b = A.DeviceArray(4 << 20, np.uint8)
        b.free()
    finally:
        A.check(L().acc_hip_set_alloc_limit(0, 0))


def test_stream_wait_event_orders_a_consumer_after_a_producer(stream):
    """Future get = cuStreamWaitEvent (PTX/Execute/Async.hs:134-149): stream B
    consumes what stream A produced, ordered only by the event."""
    a, b = A.Stream(0), A.Stream(0)
    n = 1 << 26
    x = A.DeviceArray(n, np.int64)
    out = A.DeviceArray(1, np.int64)
    ev = A.Event(0, timing=False)
    for rep in range(3):
        A.generate(a, A.GEN_IOTA, n, x, rep, 1)  # x[i] = rep + i: a long producer
        ev.record(a)
        A.check(L().acc_hip_stream_wait_event(b.handle, ev.handle))
        A.fold(b, A.ADD, np.int64(0), x, 1, n, out)
        assert out.to_numpy(b)[0] == n * rep + n * (n - 1) // 2
    a.destroy()
    b.destroy()


def test_event_query_and_elapsed(stream):
    e0, e1 = A.Event(0), A.Event(0)
    x = A.DeviceArray(1 << 24, np.int32)
    e0.record(stream)
    A.generate(stream, A.GEN_FILL, 1 << 24, x, 7)
    e1.record(stream)
    e1.synchronize()
    assert L().acc_hip_event_query(e1.handle) == A.SUCCESS
    assert e0.elapsed_ms(e1) >= 0.0


def test_graph_is_bound_to_its_capture_stream(stream):
    """A captured graph's skeletons use the capture stream's workspace: replay
    on another stream is refused; the graph keeps the workspace alive after
    the stream is gone."""
    s2 = A.Stream(0)
    n = 1 << 20
    x = np.arange(n, dtype=np.int32) % 7
    d = A.DeviceArray.from_numpy(s2, x)
    out = A.DeviceArray(n, np.int32)
    A.check(L().acc_hip_stream_reserve(s2.handle, 1 << 20))
    g = A.Graph.capture(s2, lambda: A.scan(s2, A.ADD, A.LEFT_TO_RIGHT, None, d, 1, n, out))
    g.launch(s2)
    assert np.array_equal(out.to_numpy(s2), np.cumsum(x, dtype=np.int32))
    with pytest.raises(A.AccError):
        g.launch(stream)
    before = ctypes.c_size_t()
    A.check(L().acc_hip_allocated_bytes(0, ctypes.byref(before)))
    s2.destroy()
    mid = ctypes.c_size_t()
    A.check(L().acc_hip_allocated_bytes(0, ctypes.byref(mid)))
    assert mid.value == before.value  # the graph still holds the workspace
    g.destroy()
    after = ctypes.c_size_t()
    A.check(L().acc_hip_allocated_bytes(0, ctypes.byref(after)))
    assert after.value < before.value


def test_pinned_host_memory_round_trip(stream):
    p = ctypes.c_void_p()
    n = 1 << 20
    A.check(L().acc_hip_host_alloc(n * 4, ctypes.byref(p)))
    host = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_int32)), shape=(n,))
    host[:] = np.arange(n, dtype=np.int32)
    d = A.DeviceArray(n, np.int32)
    A.check(L().acc_hip_memcpy_h2d_async(stream.handle, ctypes.c_void_p(d.ptr), p, n * 4))
    A.map_(stream, A.UNOP_NEG, d, n, d)
    A.check(L().acc_hip_memcpy_d2h_async(stream.handle, p, ctypes.c_void_p(d.ptr), n * 4))
    stream.synchronize()
    assert np.array_equal(host, -np.arange(n, dtype=np.int32))
    A.check(L().acc_hip_host_free(p))


def test_bandwidth_probe_moves_the_bytes_it_times(stream):
    """The ceiling probe: the copy moves src into dst, the write-only fill
    writes every 16-byte vector, the read leaves both untouched; times are
    positive; bad shapes are refused."""
    n = 1 << 22
    src = A.DeviceArray.from_numpy(stream, np.arange(n, dtype=np.uint32))
    dst = A.DeviceArray.from_numpy(stream, np.zeros(n, np.uint32))
    for vec in (4, 8, 16):
        for nt in (False, True):
            assert A.bandwidth_probe(stream, A.PROBE_READ, vec, 2, nt, src, None, 4 * n, reps=2) > 0
    assert np.array_equal(dst.to_numpy(stream), np.zeros(n, np.uint32))
    assert A.bandwidth_probe(stream, A.PROBE_COPY, 4, 2, True, src, dst, 4 * n, reps=2) > 0
    assert np.array_equal(dst.to_numpy(stream), np.arange(n, dtype=np.uint32))
    assert A.bandwidth_probe(stream, A.PROBE_WRITE, 8, 1, False, None, dst, 4 * n, reps=2) > 0
    got = dst.to_numpy(stream).reshape(-1, 4)
    assert np.all(got[:, :3] == np.array([1, 2, 3], np.uint32))
    with pytest.raises(A.AccError):
        A.bandwidth_probe(stream, A.PROBE_READ, 5, 2, False, src, None, 4 * n)
    with pytest.raises(A.AccError):
        A.bandwidth_probe(stream, A.PROBE_COPY, 4, 2, False, src, None, 4 * n)


def test_foreign_array_function_between_skeletons(stream):
    """aforeign (ForeignAcc, PTX/Foreign.hs:74-84): a foreign array function
    runs on the backend's stream between two skeletons -- here the stand-alone
    axpy kernel launched by raw HIP on the stream acc_hip_stream_native gives
    (the way a foreign function calls a library), then a fold -- and an event
    recorded after it is the result's future; the result equals the fallback
    computed by the skeletons alone."""
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipModuleLaunchKernel.restype = ctypes.c_int
    native = ctypes.c_void_p()
    A.check(L().acc_hip_stream_native(stream.handle, ctypes.byref(native)))
    assert native.value
    m, keep = load_module()
    f = get_function(m, "acc_foreign_axpy")
    n = (1 << 20) + 3
    x = np.random.default_rng(5).integers(-8, 8, size=n).astype(np.float32)
    dx = A.DeviceArray.from_numpy(stream, x)
    dy = A.DeviceArray(n, np.float32)
    A.scan(stream, A.ADD, A.LEFT_TO_RIGHT, None, dx, 1, n, dy)  # skeleton: y = scanl1 (+) x
    a, px, py, cn = ctypes.c_float(3.0), ctypes.c_void_p(dx.ptr), ctypes.c_void_p(dy.ptr), ctypes.c_int64(n)
    args = (ctypes.c_void_p * 4)(*[ctypes.cast(ctypes.byref(v), ctypes.c_void_p) for v in (a, px, py, cn)])
    # the foreign function: y = 3 x + y, by raw HIP on the backend's stream
    assert hip.hipModuleLaunchKernel(f, 1024, 1, 1, 256, 1, 1, 0, native, args, None) == 0
    out = A.DeviceArray(1, np.float32)
    A.fold(stream, A.ADD, np.float32(0), dy, 1, n, out)  # skeleton on the foreign result
    ev = A.Event(0)
    ev.record(stream)  # the future of the whole pipeline
    ev.synchronize()
    # the fallback's value: every scan prefix and every 3 x + y is an exact
    # small integer in Float32, so the device's y equals this one; the fold
    # is held to the FP32 bound
    y = np.cumsum(x, dtype=np.float64)
    assert np.array_equal(dy.to_numpy(stream).astype(np.float64), 3.0 * x + y)
    v = 3.0 * x.astype(np.float64) + y
    assert abs(float(out.to_numpy(stream)[0]) - v.sum()) <= 1e-5 * np.log2(n) * np.abs(v).sum()
    A.check(L().acc_hip_module_unload(m))
