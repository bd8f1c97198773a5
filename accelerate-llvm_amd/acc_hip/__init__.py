"""ctypes binding of libaccelerate_hip.so (the C ABI in include/accelerate_hip.h).

This is the binding a non-C host (here: the tests, the benchmark and the
serving loop) uses; a Haskell backend binds the same symbols with
`foreign import ccall`.  Every skeleton call goes to the gfx950 kernels in the
shared library; there is no CPU fallback: if the library or a GPU is missing,
calls raise.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_float, c_int, c_int64, c_size_t, c_uint, c_void_p

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "lib", "libaccelerate_hip.so")

# acc_status
SUCCESS, INVALID_VALUE, OUT_OF_MEMORY, HIP_ERROR, UNSUPPORTED, EMPTY_ARRAY, NOT_READY, NO_DEVICE = range(8)
# acc_type
INT32, INT64, WORD32, WORD64, FLOAT32, FLOAT64 = range(6)
# acc_op
ADD, MUL, MIN, MAX, AND, OR, XOR, SUB, EXCHANGE = range(9)
# acc_direction
LEFT_TO_RIGHT, RIGHT_TO_LEFT = 0, 1
# acc_boundary
CLAMP, MIRROR, WRAP, CONSTANT, FUNCTION = range(5)
# acc_unop (map / transform element functions)
UNOP_ID, UNOP_NEG, UNOP_ABS, UNOP_AFFINE, UNOP_CAST = range(5)
# acc_genop
GEN_FILL, GEN_IOTA = range(2)
# acc_ixmap (transform / backpermute index maps)
IX_GATHER, IX_REVERSE, IX_TRANSPOSE = range(3)

_NP_TO_ACC = {
    np.dtype(np.int32): INT32,
    np.dtype(np.int64): INT64,
    np.dtype(np.uint32): WORD32,
    np.dtype(np.uint64): WORD64,
    np.dtype(np.float32): FLOAT32,
    np.dtype(np.float64): FLOAT64,
}


class AccError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[acc status {status}] {message}")
        self.status = status


class OutOfMemory(AccError):
    pass


class EmptyArray(AccError):
    pass


_lib = None


class DeviceProps(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 64),
        ("arch", ctypes.c_char * 32),
        ("compute_units", c_int),
        ("wave_size", c_int),
        ("max_threads_per_block", c_int),
        ("lds_per_block", c_size_t),
        ("total_global_mem", c_size_t),
        ("clock_khz", c_int),
        ("memory_clock_khz", c_int),
        ("memory_bus_width", c_int),
        ("l2_cache_bytes", c_int),
    ]


class BatchProblem(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("type", ctypes.c_int32),
        ("op", ctypes.c_int32),
        ("has_seed", ctypes.c_int32),
        ("n", ctypes.c_int64),
        ("in_", c_void_p),
        ("out", c_void_p),
        ("seed", ctypes.c_uint64),
        ("aux", c_void_p),
        ("m", ctypes.c_int64),
        ("aux_type", ctypes.c_int32),
        ("params", ctypes.c_int32),
        ("fill", ctypes.c_uint64),
    ]


BATCH_SCAN, BATCH_FOLD, BATCH_FOLD_SEG, BATCH_HISTOGRAM, BATCH_STENCIL = 0, 1, 2, 3, 4

REQ_FOLD, REQ_SCAN, REQ_FOLD_SEG, REQ_PERMUTE, REQ_STENCIL2D = range(5)


class Request(ctypes.Structure):
    _fields_ = [
        ("skeleton", ctypes.c_int32),
        ("type", ctypes.c_int32),
        ("op", ctypes.c_int32),
        ("has_seed", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
        ("fill", ctypes.c_uint64),
        ("in_", c_void_p),
        ("n", ctypes.c_int64),
        ("m", ctypes.c_int64),
        ("aux", c_void_p),
        ("aux_type", ctypes.c_int32),
        ("direction", ctypes.c_int32),
        ("ry", ctypes.c_int32),
        ("rx", ctypes.c_int32),
        ("boundary", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("weights", c_void_p),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("completed", ctypes.c_uint64),
        ("failed", ctypes.c_uint64),
        ("p50_us", ctypes.c_double),
        ("p99_us", ctypes.c_double),
        ("mean_us", ctypes.c_double),
        ("max_us", ctypes.c_double),
        ("req_per_s", ctypes.c_double),
        ("bytes_per_s", ctypes.c_double),
        ("span_s", ctypes.c_double),
        ("arena_bytes", c_size_t),
        ("arena_used", c_size_t),
        ("rounds", ctypes.c_uint64),
        ("fused", ctypes.c_uint64),
    ]


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("fuse", ctypes.c_int32),
        ("round_requests", ctypes.c_int32),
        ("round_bytes", ctypes.c_int64),
        ("block", ctypes.c_int32),
        ("vectors", ctypes.c_int32),
        ("run", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]

class KernelInfo(ctypes.Structure):
    _fields_ = [
        ("origin", c_int),
        ("lock_free_permute", c_int),
        ("compile_ms", ctypes.c_double),
        ("key", ctypes.c_char * 40),
    ]


class KernelArgs(ctypes.Structure):
    _fields_ = [
        ("in_", c_void_p * 4),
        ("out", c_void_p * 4),
        ("defaults", c_void_p * 4),
        ("sums", c_void_p * 4),
        ("offsets", c_void_p),
        ("outer", ctypes.c_int64),
        ("inner", ctypes.c_int64),
        ("nseg", ctypes.c_int64),
        ("nfree", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("free_ptr", c_void_p * 8),
        ("free_len", ctypes.c_int64 * 8),
    ]


KERNEL_COMPILED, KERNEL_FROM_DISK_CACHE, KERNEL_FROM_IMAGE, KERNEL_LOADED = range(4)


class TableStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "uploads", "downloads", "evictions", "nursery_hits", "oom_retries", "bytes_up", "bytes_down",
        "bytes_evicted", "resident_bytes", "nursery_bytes", "entries")]

# exported symbol -> (restype, argtypes); every declaration in the header
SIGNATURES = {
    "acc_hip_last_error": (c_char_p, []),
    "acc_hip_abi_version": (c_int, []),
    "acc_hip_device_count": (c_int, [POINTER(c_int)]),
    "acc_hip_device_props": (c_int, [c_int, POINTER(DeviceProps)]),
    "acc_hip_device_synchronize": (c_int, [c_int]),
    "acc_hip_malloc": (c_int, [c_int, c_size_t, POINTER(c_void_p)]),
    "acc_hip_free": (c_int, [c_int, c_void_p]),
    "acc_hip_mem_get_info": (c_int, [c_int, POINTER(c_size_t), POINTER(c_size_t)]),
    "acc_hip_set_alloc_limit": (c_int, [c_int, c_size_t]),
    "acc_hip_allocated_bytes": (c_int, [c_int, POINTER(c_size_t)]),
    "acc_hip_host_alloc": (c_int, [c_size_t, POINTER(c_void_p)]),
    "acc_hip_host_free": (c_int, [c_void_p]),
    "acc_hip_memcpy_h2d_async": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    "acc_hip_memcpy_d2h_async": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    "acc_hip_memcpy_d2d_async": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    "acc_hip_fill_async": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t, c_size_t]),
    "acc_hip_stream_create": (c_int, [c_int, POINTER(c_void_p)]),
    "acc_hip_stream_destroy": (c_int, [c_void_p]),
    "acc_hip_stream_synchronize": (c_int, [c_void_p]),
    "acc_hip_stream_query": (c_int, [c_void_p]),
    "acc_hip_stream_device": (c_int, [c_void_p, POINTER(c_int)]),
    "acc_hip_stream_reserve": (c_int, [c_void_p, c_size_t]),
    "acc_hip_event_create": (c_int, [c_int, c_int, POINTER(c_void_p)]),
    "acc_hip_event_destroy": (c_int, [c_void_p]),
    "acc_hip_event_record": (c_int, [c_void_p, c_void_p]),
    "acc_hip_stream_wait_event": (c_int, [c_void_p, c_void_p]),
    "acc_hip_event_query": (c_int, [c_void_p]),
    "acc_hip_event_synchronize": (c_int, [c_void_p]),
    "acc_hip_event_elapsed_ms": (c_int, [c_void_p, c_void_p, POINTER(c_float)]),
    "acc_hip_module_load_data": (c_int, [c_int, c_void_p, c_size_t, POINTER(c_void_p)]),
    "acc_hip_module_unload": (c_int, [c_void_p]),
    "acc_hip_module_get_function": (c_int, [c_void_p, c_char_p, POINTER(c_void_p)]),
    "acc_hip_function_attributes": (c_int, [c_void_p, POINTER(c_int), POINTER(c_int), POINTER(c_int)]),
    "acc_hip_launch": (c_int, [c_void_p, c_uint, c_uint, c_uint, c_uint, c_uint, c_uint, c_uint, c_void_p,
                               POINTER(c_void_p)]),
    "acc_hip_graph_begin_capture": (c_int, [c_void_p]),
    "acc_hip_graph_end_capture": (c_int, [c_void_p, POINTER(c_void_p)]),
    "acc_hip_graph_launch": (c_int, [c_void_p, c_void_p]),
    "acc_hip_graph_destroy": (c_int, [c_void_p]),
    "acc_hip_fold": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_int64, c_int64, c_void_p]),
    "acc_hip_fold_seg": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_int64, c_int64, c_int, c_void_p,
                                 c_int64, c_void_p]),
    "acc_hip_fold_seg_noalloc": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_int64, c_int64, c_int,
                                         c_void_p, c_int64, c_void_p]),
    "acc_hip_seg_plan_create": (c_int, [c_int, POINTER(c_void_p)]),
    "acc_hip_seg_plan_build": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int64, c_int, c_void_p, c_int64]),
    "acc_hip_seg_plan_build_lengths": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int64, c_int, c_void_p, c_int64,
                                               c_void_p]),
    "acc_hip_fold_seg_planned": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_int64, c_void_p]),
    "acc_hip_fold_seg_planned_multi": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                               c_void_p]),
    "acc_hip_seg_plan_destroy": (c_int, [c_void_p]),
    "acc_hip_scan": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_int64, c_int64, c_void_p]),
    "acc_hip_scan_prime": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_int64, c_int64, c_void_p,
                                   c_void_p]),
    "acc_hip_permute": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_int,
                                c_void_p, c_int64]),
    "acc_hip_stencil2d": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_void_p, c_int64,
                                  c_int64, c_void_p]),
    "acc_hip_stencil3d": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_void_p,
                                  c_int64, c_int64, c_int64, c_void_p]),
    "acc_hip_stencil2_2d": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_void_p, c_int64,
                                    c_int64, c_void_p, c_int, c_void_p, c_void_p, c_int64, c_int64, c_void_p]),
    "acc_hip_map": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_int64, c_void_p]),
    "acc_hip_zipwith": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_int64, c_void_p]),
    "acc_hip_generate": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int64, c_void_p]),
    "acc_hip_transform": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_int64, c_int64, c_int,
                                  c_void_p, c_int64, c_void_p]),
    "acc_hip_node_source": (c_int, [c_char_p, c_char_p, c_size_t, POINTER(c_size_t)]),
    "acc_hip_node_precompile": (c_int, [c_char_p, POINTER(c_int)]),
    "acc_hip_kernel_compile": (c_int, [c_int, c_char_p, POINTER(c_void_p)]),
    "acc_hip_kernel_from_image": (c_int, [c_int, c_char_p, c_void_p, c_size_t, POINTER(c_void_p)]),
    "acc_hip_kernel_image": (c_int, [c_void_p, c_void_p, c_size_t, POINTER(c_size_t)]),
    "acc_hip_kernel_info": (c_int, [c_void_p, POINTER(KernelInfo)]),
    "acc_hip_kernel_run": (c_int, [c_void_p, c_void_p, POINTER(KernelArgs)]),
    "acc_hip_kernel_destroy": (c_int, [c_void_p]),
    "acc_hip_table_create": (c_int, [c_int, POINTER(c_void_p)]),
    "acc_hip_table_use": (c_int, [c_void_p, ctypes.c_uint64, c_void_p, c_size_t, c_void_p, POINTER(c_void_p)]),
    "acc_hip_table_alloc": (c_int, [c_void_p, ctypes.c_uint64, c_void_p, c_size_t, POINTER(c_void_p)]),
    "acc_hip_table_get": (c_int, [c_void_p, ctypes.c_uint64, c_void_p, POINTER(c_void_p)]),
    "acc_hip_table_mark": (c_int, [c_void_p, ctypes.c_uint64, c_void_p, c_int]),
    "acc_hip_table_peek": (c_int, [c_void_p, ctypes.c_uint64, POINTER(c_int)]),
    "acc_hip_table_free": (c_int, [c_void_p, ctypes.c_uint64]),
    "acc_hip_table_stats": (c_int, [c_void_p, POINTER(TableStats)]),
    "acc_hip_table_destroy": (c_int, [c_void_p]),
    "acc_hip_batch_create": (c_int, [c_int, POINTER(BatchProblem), c_int64, c_int, POINTER(c_void_p)]),
    "acc_hip_batch_launch": (c_int, [c_void_p, c_void_p]),
    "acc_hip_batch_launch_plain": (c_int, [c_void_p, c_void_p]),
    "acc_hip_batch_tile": (c_int, [c_void_p, POINTER(c_int), POINTER(ctypes.c_uint32)]),
    "acc_hip_batch_shape": (c_int, [c_void_p, POINTER(c_int), POINTER(c_int), POINTER(c_int),
                                    POINTER(ctypes.c_uint32)]),
    "acc_hip_batch_destroy": (c_int, [c_void_p]),
    "acc_hip_server_create": (c_int, [c_int, c_int, c_size_t, POINTER(c_void_p)]),
    "acc_hip_server_configure": (c_int, [c_void_p, POINTER(ServerConfig)]),
    "acc_hip_server_submit": (c_int, [c_void_p, POINTER(Request), c_int64, c_int, POINTER(ctypes.c_uint64)]),
    "acc_hip_server_wait": (c_int, [c_void_p, ctypes.c_uint64, POINTER(c_void_p), POINTER(ctypes.c_double)]),
    "acc_hip_server_release": (c_int, [c_void_p, ctypes.c_uint64]),
    "acc_hip_server_drain": (c_int, [c_void_p]),
    "acc_hip_server_stats": (c_int, [c_void_p, POINTER(ServerStats)]),
    "acc_hip_server_reset_stats": (c_int, [c_void_p]),
    "acc_hip_server_destroy": (c_int, [c_void_p]),
    "acc_hip_stream_native": (c_int, [c_void_p, POINTER(c_void_p)]),
    "acc_hip_bandwidth_probe": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_size_t, c_int,
                                        POINTER(ctypes.c_double)]),
}


def lib():
    """Load the shared library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C accelerate-llvm_amd`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    if status == SUCCESS:
        return
    msg = lib().acc_hip_last_error().decode(errors="replace")
    if status == OUT_OF_MEMORY:
        raise OutOfMemory(status, msg)
    if status == EMPTY_ARRAY:
        raise EmptyArray(status, msg)
    raise AccError(status, msg)


def acc_type(dtype) -> int:
    try:
        return _NP_TO_ACC[np.dtype(dtype)]
    except KeyError:
        raise TypeError(f"unsupported element type {dtype}") from None


def device_count() -> int:
    n = c_int(0)
    check(lib().acc_hip_device_count(ctypes.byref(n)))
    return n.value


def device_props(device: int = 0) -> DeviceProps:
    p = DeviceProps()
    check(lib().acc_hip_device_props(device, ctypes.byref(p)))
    return p


class Stream:
    def __init__(self, device: int = 0):
        self.device = device
        h = c_void_p()
        check(lib().acc_hip_stream_create(device, ctypes.byref(h)))
        self.handle = h

    def synchronize(self) -> None:
        check(lib().acc_hip_stream_synchronize(self.handle))

    def wait_event(self, event: "Event") -> None:
        """Later work on this stream waits for `event` (acc_hip_stream_wait_event)."""
        check(lib().acc_hip_stream_wait_event(self.handle, event.handle))

    def destroy(self) -> None:
        if self.handle:
            lib().acc_hip_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Event:
    def __init__(self, device: int = 0, timing: bool = True):
        h = c_void_p()
        check(lib().acc_hip_event_create(device, 1 if timing else 0, ctypes.byref(h)))
        self.handle = h

    def record(self, stream: Stream) -> None:
        check(lib().acc_hip_event_record(self.handle, stream.handle))

    def synchronize(self) -> None:
        check(lib().acc_hip_event_synchronize(self.handle))

    def elapsed_ms(self, later: "Event") -> float:
        ms = c_float(0)
        check(lib().acc_hip_event_elapsed_ms(self.handle, later.handle, ctypes.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            if self.handle:
                lib().acc_hip_event_destroy(self.handle)
        except Exception:
            pass


class DeviceArray:
    """A device buffer of `count` elements of `dtype` (one struct-of-arrays field)."""

    def __init__(self, count: int, dtype, device: int = 0):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.device = device
        self.nbytes = self.count * self.dtype.itemsize
        p = c_void_p()
        check(lib().acc_hip_malloc(device, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value or 0

    @classmethod
    def from_numpy(cls, stream: Stream, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype, stream.device)
        if a.size:
            check(lib().acc_hip_memcpy_h2d_async(stream.handle, c_void_p(d.ptr), a.ctypes.data_as(c_void_p), a.nbytes))
            stream.synchronize()
        return d

    def to_numpy(self, stream: Stream) -> np.ndarray:
        out = np.empty(self.count, self.dtype)
        if self.count:
            check(lib().acc_hip_memcpy_d2h_async(stream.handle, out.ctypes.data_as(c_void_p), c_void_p(self.ptr),
                                                 self.nbytes))
            stream.synchronize()
        return out

    def view(self, offset_elems: int, count: int) -> "DeviceView":
        return DeviceView(self, offset_elems, count)

    def free(self) -> None:
        if self.ptr:
            check(lib().acc_hip_free(self.device, c_void_p(self.ptr)))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceView:
    """A sub-range of a DeviceArray (not owning)."""

    def __init__(self, base: DeviceArray, offset: int, count: int):
        self.base = base
        self.dtype = base.dtype
        self.count = count
        self.device = base.device
        self.ptr = base.ptr + offset * base.dtype.itemsize
        self.nbytes = count * base.dtype.itemsize

    def to_numpy(self, stream: Stream) -> np.ndarray:
        out = np.empty(self.count, self.dtype)
        if self.count:
            check(lib().acc_hip_memcpy_d2h_async(stream.handle, out.ctypes.data_as(c_void_p), c_void_p(self.ptr),
                                                 self.nbytes))
            stream.synchronize()
        return out


def _scalar(dtype, v):
    """A scalar, or a sequence (the function boundary's c0, cy, cx), as a C array."""
    if v is None:
        return None, None
    a = np.ascontiguousarray(np.array(v, dtype=dtype).reshape(-1))
    return a, a.ctypes.data_as(c_void_p)


def _p(x):
    return c_void_p(x.ptr) if x is not None else None


# ----------------------------------------------------------------- skeletons
def fold(stream, op, seed, src, outer, inner, out):
    keep, sp = _scalar(src.dtype, seed)
    check(lib().acc_hip_fold(stream.handle, acc_type(src.dtype), op, sp, _p(src), outer, inner, _p(out)))


def fold_seg(stream, op, seed, src, outer, inner, offsets, nseg, out):
    keep, sp = _scalar(src.dtype, seed)
    check(lib().acc_hip_fold_seg(stream.handle, acc_type(src.dtype), op, sp, _p(src), outer, inner,
                                 acc_type(offsets.dtype), _p(offsets), nseg, _p(out)))


def fold_seg_noalloc(stream, op, seed, src, outer, inner, offsets, nseg, out):
    """fold_seg with the library's fixed device scratch instead of allocated workspace."""
    keep, sp = _scalar(src.dtype, seed)
    check(lib().acc_hip_fold_seg_noalloc(stream.handle, acc_type(src.dtype), op, sp, _p(src), outer, inner,
                                         acc_type(offsets.dtype), _p(offsets), nseg, _p(out)))


class SegPlan:
    """A segment descriptor's tile plan (acc_hip_seg_plan_*): build it once for
    an offsets array, then fold any number of arrays over those segments."""

    def __init__(self, device=0):
        h = c_void_p()
        check(lib().acc_hip_seg_plan_create(device, ctypes.byref(h)))
        self.handle = h

    def build(self, stream, dtype, src, n, offsets, nseg):
        """Plan nseg segments (offsets: DeviceArray/View) over rows of n elements
        of dtype whose first row starts at src."""
        check(lib().acc_hip_seg_plan_build(stream.handle, self.handle, acc_type(dtype), _p(src), n,
                                           acc_type(offsets.dtype), _p(offsets), nseg))

    def build_lengths(self, stream, dtype, src, n, lengths, nseg, offsets):
        """offsets[0..nseg] = scanl (+) 0 lengths, and the plan over them, in one launch."""
        check(lib().acc_hip_seg_plan_build_lengths(stream.handle, self.handle, acc_type(dtype), _p(src), n,
                                                   acc_type(lengths.dtype), _p(lengths), nseg, _p(offsets)))

    def fold_seg(self, stream, op, seed, src, outer, out):
        keep, sp = _scalar(src.dtype, seed)
        check(lib().acc_hip_fold_seg_planned(stream.handle, self.handle, acc_type(src.dtype), op, sp, _p(src), outer,
                                             _p(out)))

    def fold_seg_multi(self, stream, folds):
        """Several foldSegs on this plan in one call: folds = [(op, seed_or_None, src, out), ...]."""
        k = len(folds)
        ops = (c_int * k)(*[f[0] for f in folds])
        keep = [_scalar(f[2].dtype, f[1]) for f in folds]
        seeds = (c_void_p * k)(*[kp[1] for kp in keep])
        ins = (c_void_p * k)(*[_p(f[2]) for f in folds])
        outs = (c_void_p * k)(*[_p(f[3]) for f in folds])
        check(lib().acc_hip_fold_seg_planned_multi(stream.handle, self.handle, acc_type(folds[0][2].dtype), k, ops,
                                                   seeds, ins, outs))

    def destroy(self):
        if self.handle:
            check(lib().acc_hip_seg_plan_destroy(self.handle))
            self.handle = None


def scan(stream, op, direction, seed, src, outer, inner, out):
    keep, sp = _scalar(src.dtype, seed)
    check(lib().acc_hip_scan(stream.handle, acc_type(src.dtype), op, direction, sp, _p(src), outer, inner, _p(out)))


def scan_prime(stream, op, direction, seed, src, outer, inner, out, sums):
    keep, sp = _scalar(src.dtype, seed)
    check(lib().acc_hip_scan_prime(stream.handle, acc_type(src.dtype), op, direction, sp, _p(src), outer, inner,
                                   _p(out), _p(sums)))


def permute(stream, op, defaults, out, m, src, src_fill, index, n):
    dt = out.dtype
    keep, fp = _scalar(dt, src_fill)
    check(lib().acc_hip_permute(stream.handle, acc_type(dt), op, _p(defaults), _p(out), m, _p(src), fp,
                                acc_type(index.dtype), _p(index), n))


def stencil2d(stream, ry, rx, weights, boundary, constant, src, height, width, out):
    dt = src.dtype
    w = np.ascontiguousarray(np.asarray(weights, dtype=dt).reshape(-1))
    keep, cp = _scalar(dt, constant)
    check(lib().acc_hip_stencil2d(stream.handle, acc_type(dt), ry, rx, w.ctypes.data_as(c_void_p), boundary, cp,
                                  _p(src), height, width, _p(out)))


def stencil2_2d(stream, ry, rx, weights1, boundary1, constant1, src1, height1, width1,
                weights2, boundary2, constant2, src2, height2, width2, out):
    """stencil2 on DIM2 arrays: the weighted sums of both inputs' neighbourhoods
    over the intersection of their shapes, each input with its own boundary."""
    dt = src1.dtype
    w1 = np.ascontiguousarray(np.asarray(weights1, dtype=dt).reshape(-1))
    w2 = np.ascontiguousarray(np.asarray(weights2, dtype=dt).reshape(-1))
    k1, c1 = _scalar(dt, constant1)
    k2, c2 = _scalar(dt, constant2)
    check(lib().acc_hip_stencil2_2d(stream.handle, acc_type(dt), ry, rx, w1.ctypes.data_as(c_void_p), boundary1, c1,
                                    _p(src1), height1, width1, w2.ctypes.data_as(c_void_p), boundary2, c2, _p(src2),
                                    height2, width2, _p(out)))


def stencil3d(stream, rz, ry, rx, weights, boundary, constant, src, depth, height, width, out):
    dt = src.dtype
    w = np.ascontiguousarray(np.asarray(weights, dtype=dt).reshape(-1))
    keep, cp = _scalar(dt, constant)
    check(lib().acc_hip_stencil3d(stream.handle, acc_type(dt), rz, ry, rx, w.ctypes.data_as(c_void_p), boundary,
                                  cp, _p(src), depth, height, width, _p(out)))


def _pair(dtype, a, b=0):
    v = np.array([a, b], dtype=dtype)
    return v, v.ctypes.data_as(c_void_p)


def map_(stream, f, src, n, out, a=None, b=None):
    """out[i] = f(src[i]); f one of UNOP_*; UNOP_AFFINE takes a and b; UNOP_CAST converts
    src's element type to out's."""
    keep, pp = _pair(out.dtype, a, b) if f == UNOP_AFFINE else (None, None)
    check(lib().acc_hip_map(stream.handle, acc_type(src.dtype), acc_type(out.dtype), f, pp, _p(src), n, _p(out)))


def zipwith(stream, op, a, b, n, out):
    check(lib().acc_hip_zipwith(stream.handle, acc_type(out.dtype), op, _p(a), _p(b), n, _p(out)))


def generate(stream, g, n, out, a, b=0):
    keep, pp = _pair(out.dtype, a, b)
    check(lib().acc_hip_generate(stream.handle, acc_type(out.dtype), g, pp, n, _p(out)))


def transform(stream, p, f, src, in_outer, in_inner, index, n, out, a=None, b=None):
    """out[i] = f(src[p(i)]); p one of IX_*, f one of UNOP_ID/NEG/ABS/AFFINE (backpermute: UNOP_ID)."""
    keep, pp = _pair(out.dtype, a, b) if f == UNOP_AFFINE else (None, None)
    it = acc_type(index.dtype) if index is not None else INT32
    check(lib().acc_hip_transform(stream.handle, acc_type(out.dtype), p, f, pp, _p(src), in_outer, in_inner, it,
                                  _p(index), n, _p(out)))


def _bits(dtype, v):
    a = np.zeros(1, np.uint64)
    b = np.array([v], dtype=dtype)
    a.view(np.uint8)[: b.nbytes] = b.view(np.uint8)
    return int(a[0])


class Batch:
    """Many scans/folds in one persistent launch (acc_hip_batch_*)."""

    def __init__(self, device, problems, tile=0):
        """problems: list of (kind, op, seed_or_None, src, n, out[, aux, m, fill, params]):
        src/out DeviceArray/View; foldSeg: aux = offsets, m = segments; histogram: src =
        indices, out = m bins of the result type, seed = their initial value, fill = the
        source value; stencil: an n x m image, aux = the weights on the device, fill = the
        constant boundary's value, params = ry | rx << 8 | boundary << 16."""
        arr = (BatchProblem * len(problems))()
        for i, prob in enumerate(problems):
            kind, op, seed, src, n, out = prob[:6]
            aux, m, fill, params = list(prob[6:10]) + [None, 0, 0, 0][len(prob[6:10]):]
            dt = out.dtype if kind == BATCH_HISTOGRAM else src.dtype
            arr[i].kind = kind
            arr[i].type = acc_type(dt)
            arr[i].op = op
            arr[i].has_seed = 0 if seed is None else 1
            arr[i].n = n
            arr[i].in_ = src.ptr if src is not None else None
            arr[i].out = out.ptr
            arr[i].seed = 0 if seed is None else _bits(dt, seed)
            if aux is not None:
                arr[i].aux = aux.ptr
                arr[i].aux_type = acc_type(aux.dtype)
            elif kind == BATCH_HISTOGRAM:
                arr[i].aux_type = acc_type(src.dtype)
            arr[i].m = m
            arr[i].fill = _bits(dt, fill) if kind in (BATCH_HISTOGRAM, BATCH_STENCIL) else 0
            arr[i].params = params
        h = c_void_p()
        check(lib().acc_hip_batch_create(device, arr, len(problems), tile, ctypes.byref(h)))
        self.handle = h

    def launch(self, stream):
        check(lib().acc_hip_batch_launch(self.handle, stream.handle))

    def launch_plain(self, stream):
        """The same problems, one plain C call each (the per-call path)."""
        check(lib().acc_hip_batch_launch_plain(self.handle, stream.handle))

    def tile(self):
        k, t = c_int(), ctypes.c_uint32()
        check(lib().acc_hip_batch_tile(self.handle, ctypes.byref(k), ctypes.byref(t)))
        return k.value, t.value

    def shape(self):
        """(workgroup size, vectors per thread, tiles per ticket, tickets)"""
        b, k, r, t = c_int(), c_int(), c_int(), ctypes.c_uint32()
        check(lib().acc_hip_batch_shape(self.handle, ctypes.byref(b), ctypes.byref(k), ctypes.byref(r),
                                        ctypes.byref(t)))
        return b.value, k.value, r.value, t.value

    def destroy(self):
        if self.handle:
            lib().acc_hip_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Graph:
    """A captured hipGraph (acc_hip_graph_*)."""

    def __init__(self, handle):
        self.handle = handle

    @classmethod
    def capture(cls, stream, fn):
        check(lib().acc_hip_graph_begin_capture(stream.handle))
        try:
            fn()
        finally:
            h = c_void_p()
            rc = lib().acc_hip_graph_end_capture(stream.handle, ctypes.byref(h))
        check(rc)
        return cls(h)

    def launch(self, stream):
        check(lib().acc_hip_graph_launch(self.handle, stream.handle))

    def destroy(self):
        if self.handle:
            lib().acc_hip_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Server:
    """The serving loop (acc_hip_server_*)."""

    def __init__(self, device=0, slots=4, hbm_budget=0, fuse=True, round_requests=0, round_bytes=0,
                 shape=(0, 0, 0)):
        h = c_void_p()
        check(lib().acc_hip_server_create(device, slots, hbm_budget, ctypes.byref(h)))
        self.handle = h
        self._keep = []  # host buffers referenced by queued requests (stencil weights)
        self.configure(fuse, round_requests, round_bytes, shape)

    def configure(self, fuse=True, round_requests=0, round_bytes=0, shape=(0, 0, 0)):
        """fuse: run requests with a batched form in fused rounds; shape:
        (workgroup size, vectors per thread, tiles per run) of the fused launch."""
        c = ServerConfig(1 if fuse else 0, round_requests, round_bytes, shape[0], shape[1], shape[2], 0)
        check(lib().acc_hip_server_configure(self.handle, ctypes.byref(c)))

    @staticmethod
    def request(skeleton, dtype, op=ADD, seed=None, src=None, n=0, m=0, aux=None, fill=0, direction=0,
                ry=0, rx=0, boundary=0, weights=None):
        r = Request()
        r.skeleton = skeleton
        r.type = acc_type(dtype)
        r.op = op
        r.has_seed = 0 if seed is None else 1
        r.seed = 0 if seed is None else _bits(dtype, seed)
        r.fill = _bits(dtype, fill)
        r.in_ = src.ptr if src is not None else None
        r.n = n
        r.m = m
        r.aux = aux.ptr if aux is not None else None
        r.aux_type = acc_type(aux.dtype) if aux is not None else INT32
        r.direction = direction
        r.ry, r.rx, r.boundary = ry, rx, boundary
        w = None
        if weights is not None:
            w = np.ascontiguousarray(np.asarray(weights, dtype=dtype).reshape(-1))
            r.weights = w.ctypes.data_as(c_void_p)
        return r, w

    def submit(self, req, arrival_offset_ns=0, keep=False):
        r, w = req
        rid = ctypes.c_uint64()
        check(lib().acc_hip_server_submit(self.handle, ctypes.byref(r), int(arrival_offset_ns), 1 if keep else 0,
                                          ctypes.byref(rid)))
        return rid.value

    def wait(self, rid):
        p, lat = c_void_p(), ctypes.c_double()
        check(lib().acc_hip_server_wait(self.handle, rid, ctypes.byref(p), ctypes.byref(lat)))
        return p.value, lat.value

    def release(self, rid):
        check(lib().acc_hip_server_release(self.handle, rid))

    def drain(self):
        check(lib().acc_hip_server_drain(self.handle))

    def stats(self):
        st = ServerStats()
        check(lib().acc_hip_server_stats(self.handle, ctypes.byref(st)))
        return st

    def reset_stats(self):
        check(lib().acc_hip_server_reset_stats(self.handle))

    def destroy(self):
        if self.handle:
            lib().acc_hip_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


PROBE_READ, PROBE_COPY, PROBE_WRITE = 0, 1, 2


def bandwidth_probe(stream, kind, vectors, blocks_per_cu, nontemporal, src, dst, nbytes, reps=10):
    """Mean microseconds per launch of the read / copy / write probe kernel."""
    us = ctypes.c_double()
    check(lib().acc_hip_bandwidth_probe(stream.handle, kind, vectors, blocks_per_cu, 1 if nontemporal else 0,
                                        c_void_p(src.ptr) if src is not None else None,
                                        c_void_p(dst.ptr) if dst is not None else None, nbytes, reps,
                                        ctypes.byref(us)))
    return us.value


def copy_to_host(stream, ptr, count, dtype):
    out = np.empty(count, dtype)
    if count:
        check(lib().acc_hip_memcpy_d2h_async(stream.handle, out.ctypes.data_as(c_void_p), c_void_p(ptr), out.nbytes))
        stream.synchronize()
    return out


# ------------------------------------------------------ run-time compiled skeletons
def node_source(node: str) -> str:
    """The HIP source a skeleton node translates to."""
    n = c_size_t()
    check(lib().acc_hip_node_source(node.encode(), None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value + 1)
    check(lib().acc_hip_node_source(node.encode(), buf, n.value + 1, ctypes.byref(n)))
    return buf.value.decode()


def precompile(node: str) -> int:
    """Compile a node into the on-disk code-object cache (no device needed)."""
    o = c_int()
    check(lib().acc_hip_node_precompile(node.encode(), ctypes.byref(o)))
    return o.value


def _ptrs(xs):
    if xs is None:
        return []
    if not isinstance(xs, (list, tuple)):
        xs = [xs]
    return [x if isinstance(x, int) else (x.ptr if x is not None else 0) for x in xs]


class Kernel:
    """A skeleton node compiled at run time (acc_hip_kernel_*)."""

    def __init__(self, node: str, device: int = 0, image: bytes = None):
        h = c_void_p()
        if image is None:
            check(lib().acc_hip_kernel_compile(device, node.encode(), ctypes.byref(h)))
        else:
            buf = ctypes.create_string_buffer(image, len(image))
            check(lib().acc_hip_kernel_from_image(device, node.encode(), buf, len(image), ctypes.byref(h)))
        self.handle = h
        self.node = node

    def info(self) -> KernelInfo:
        i = KernelInfo()
        check(lib().acc_hip_kernel_info(self.handle, ctypes.byref(i)))
        return i

    def image(self) -> bytes:
        n = c_size_t()
        check(lib().acc_hip_kernel_image(self.handle, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        check(lib().acc_hip_kernel_image(self.handle, buf, n.value, ctypes.byref(n)))
        return buf.raw[: n.value]

    def run(self, stream, inputs=None, outputs=None, outer=1, inner=0, nseg=0, offsets=None, defaults=None,
            sums=None, free=()):
        """inputs/outputs/defaults/sums: a DeviceArray (one field) or a list (one per field);
        free: sequence of DeviceArray (free arrays read by (read k i))."""
        a = KernelArgs()
        for k, p in enumerate(_ptrs(inputs)):
            a.in_[k] = p
        for k, p in enumerate(_ptrs(outputs)):
            a.out[k] = p
        for k, p in enumerate(_ptrs(defaults)):
            a.defaults[k] = p
        for k, p in enumerate(_ptrs(sums)):
            a.sums[k] = p
        a.offsets = offsets.ptr if offsets is not None else None
        a.outer, a.inner, a.nseg = outer, inner, nseg
        a.nfree = len(free)
        for k, f in enumerate(free):
            a.free_ptr[k] = f.ptr
            a.free_len[k] = f.count
        check(lib().acc_hip_kernel_run(self.handle, stream.handle, ctypes.byref(a)))

    def destroy(self):
        if self.handle:
            lib().acc_hip_kernel_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


# ------------------------------------------------------------ memory table
class DevRef:
    """A raw device pointer with an element type and count (not owning)."""

    def __init__(self, ptr, count, dtype, device=0):
        self.ptr, self.count, self.dtype, self.device = ptr, count, np.dtype(dtype), device
        self.nbytes = self.count * self.dtype.itemsize


class Table:
    """The remote memory table (acc_hip_table_*): numpy arrays by key."""

    def __init__(self, device=0):
        h = c_void_p()
        check(lib().acc_hip_table_create(device, ctypes.byref(h)))
        self.handle, self.device = h, device
        self.host = {}  # key -> numpy array kept alive while registered

    def use(self, stream, key, a: np.ndarray) -> DevRef:
        a = np.ascontiguousarray(a)
        self.host[key] = a
        p = c_void_p()
        check(lib().acc_hip_table_use(self.handle, key, a.ctypes.data_as(c_void_p), a.nbytes, stream.handle,
                                      ctypes.byref(p)))
        return DevRef(p.value, a.size, a.dtype, self.device)

    def alloc(self, key, count, dtype, host_backing=True) -> DevRef:
        h = np.empty(count, dtype) if host_backing else None
        self.host[key] = h
        p = c_void_p()
        check(lib().acc_hip_table_alloc(self.handle, key, h.ctypes.data_as(c_void_p) if h is not None else None,
                                        count * np.dtype(dtype).itemsize, ctypes.byref(p)))
        return DevRef(p.value, count, dtype, self.device)

    def get(self, stream, key, count, dtype) -> DevRef:
        p = c_void_p()
        check(lib().acc_hip_table_get(self.handle, key, stream.handle, ctypes.byref(p)))
        return DevRef(p.value, count, dtype, self.device)

    def mark(self, stream, key, writes=False):
        check(lib().acc_hip_table_mark(self.handle, key, stream.handle, 1 if writes else 0))

    def peek(self, key) -> tuple:
        c = c_int()
        check(lib().acc_hip_table_peek(self.handle, key, ctypes.byref(c)))
        return self.host[key], bool(c.value)

    def free(self, key):
        check(lib().acc_hip_table_free(self.handle, key))
        self.host.pop(key, None)

    def stats(self) -> TableStats:
        st = TableStats()
        check(lib().acc_hip_table_stats(self.handle, ctypes.byref(st)))
        return st

    def destroy(self):
        if self.handle:
            lib().acc_hip_table_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
