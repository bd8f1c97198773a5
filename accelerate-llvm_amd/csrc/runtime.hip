// Runtime half of the C ABI: devices, memory, transfers, streams, events,
// code objects, launches and graphs.  Each function replaces one group of
// `Foreign.CUDA.*` call sites of accelerate-llvm-ptx (cited in
// include/accelerate_hip.h).
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <unordered_map>
#include <vector>
#include <memory>
#include <atomic>
#include "host.hpp"
#include "lookback.hpp"
#include "vecmem.hpp"
#include "lookback.hpp"

namespace acc {

static thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  // the runtime keeps the error as the thread's "last error": clear it, so
  // that the launch check of a later call does not report this failure again
  (void)hipGetLastError();
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    return fail(ACC_ERROR_OUT_OF_MEMORY, "%s: %s", what, hipGetErrorString(e));
  if (e == hipErrorNoDevice) return fail(ACC_ERROR_NO_DEVICE, "%s: %s", what, hipGetErrorString(e));
  return fail(ACC_ERROR_HIP, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}

size_t type_size(acc_type t) {
  switch (t) {
    case ACC_INT32: case ACC_WORD32: case ACC_FLOAT32: return 4;
    case ACC_INT64: case ACC_WORD64: case ACC_FLOAT64: return 8;
  }
  return 0;
}

// ------------------------------------------------------------------ devices
static std::mutex g_dev_mu;
static std::vector<std::unique_ptr<DeviceInfo>> g_devs;

int device_info(int device, const DeviceInfo** out) {
  std::lock_guard<std::mutex> lk(g_dev_mu);
  if (g_devs.empty()) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
    g_devs.resize(n);
  }
  if (device < 0 || device >= (int)g_devs.size())
    return fail(ACC_ERROR_INVALID_VALUE, "device %d out of range (%zu devices)", device, g_devs.size());
  if (!g_devs[device]) {
    hipDeviceProp_t p;
    hipError_t e = hipGetDeviceProperties(&p, device);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
    auto d = std::make_unique<DeviceInfo>();
    d->compute_units = p.multiProcessorCount;
    d->max_threads = p.maxThreadsPerBlock;
    d->lds_per_block = p.sharedMemPerBlock;
    d->total_mem = p.totalGlobalMem;
    d->arch = p.gcnArchName;
    g_devs[device] = std::move(d);
  }
  *out = g_devs[device].get();
  return ACC_SUCCESS;
}

int blocks_per_cu(const void* kernel, int block) {
  static std::mutex mu;
  static std::unordered_map<const void*, int> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(kernel);
  if (it != cache.end()) return it->second;
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, block, 0) != hipSuccess || nb < 1) {
    (void)hipGetLastError();
    nb = 1;
  }
  cache[kernel] = nb;
  return nb;
}

// ------------------------------------------------------------------- memory
struct Accounting {
  std::mutex mu;
  std::unordered_map<void*, size_t> live;
  size_t bytes = 0;
  size_t limit = 0;
};
static Accounting g_acct[64];

int dev_malloc(int device, size_t bytes, void** p) {
  if (device < 0 || device >= 64) return fail(ACC_ERROR_INVALID_VALUE, "bad device %d", device);
  Accounting& a = g_acct[device];
  {
    std::lock_guard<std::mutex> lk(a.mu);
    if (a.limit && a.bytes + bytes > a.limit)
      return fail(ACC_ERROR_OUT_OF_MEMORY, "allocation of %zu bytes exceeds limit (%zu of %zu in use)",
                  bytes, a.bytes, a.limit);
  }
  ACC_CHECK(set_device(device));
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, bytes ? bytes : 1);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return hip_fail(e, "hipMalloc");
  }
  std::lock_guard<std::mutex> lk(a.mu);
  a.live[q] = bytes;
  a.bytes += bytes;
  *p = q;
  return ACC_SUCCESS;
}

int dev_free(int device, void* p) {
  if (!p) return ACC_SUCCESS;
  if (device < 0 || device >= 64) return fail(ACC_ERROR_INVALID_VALUE, "bad device %d", device);
  Accounting& a = g_acct[device];
  {
    std::lock_guard<std::mutex> lk(a.mu);
    auto it = a.live.find(p);
    if (it == a.live.end()) return fail(ACC_ERROR_INVALID_VALUE, "free of unknown pointer %p", p);
    a.bytes -= it->second;
    a.live.erase(it);
  }
  ACC_CHECK(set_device(device));
  ACC_HIP_CHECK(hipFree(p));
  return ACC_SUCCESS;
}

// ------------------------------------------------------------------ streams
static constexpr size_t kCtlBytes = kSyncCtlBytes;

Workspace::~Workspace() {
  if (ptr) (void)dev_free(device, ptr);
}

struct ScratchRegion {
  int device = -1;
  void* base = nullptr;  // workspace
  size_t bytes = 0;
  void* scr = nullptr;   // untagged scratch
  size_t scr_bytes = 0;
};
static thread_local ScratchRegion tl_scratch;

ScratchOverride::ScratchOverride(int device, void* ws, size_t ws_bytes, void* scratch, size_t scratch_bytes) {
  tl_scratch = {device, ws, ws_bytes, scratch, scratch_bytes};
}
ScratchOverride::~ScratchOverride() { tl_scratch = ScratchRegion{}; }

int stream_workspace(acc_stream_t s, size_t status_bytes, SyncCtl** ctl, uint64_t** status) {
  if (!s) return fail(ACC_ERROR_INVALID_VALUE, "null stream");
  ACC_CHECK(set_device(s->device));
  if (tl_scratch.base) {
    if (tl_scratch.device != s->device)
      return fail(ACC_ERROR_INVALID_VALUE, "scratch of device %d used on device %d", tl_scratch.device, s->device);
    if (kCtlBytes + status_bytes > tl_scratch.bytes)
      return fail(ACC_ERROR_INVALID_VALUE,
                  "allocation-free call needs %zu bytes of scratch; the maximum is %zu bytes",
                  kCtlBytes + status_bytes, tl_scratch.bytes);
    *ctl = (SyncCtl*)tl_scratch.base;
    *status = (uint64_t*)((char*)tl_scratch.base + kCtlBytes);
    return ACC_SUCCESS;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  if (!s->ws || status_bytes > s->ws->status_bytes) {
    if (s->capturing)
      return fail(ACC_ERROR_INVALID_VALUE,
                  "stream workspace must grow to %zu bytes during capture; call "
                  "acc_hip_stream_reserve before capturing", status_bytes);
    size_t cap = s->ws ? s->ws->status_bytes : (size_t)1 << 20;
    while (cap < status_bytes) cap *= 2;
    auto ws = std::make_shared<Workspace>();
    ws->device = s->device;
    ACC_CHECK(dev_malloc(s->device, kCtlBytes + cap, &ws->ptr));
    ws->status_bytes = cap;
    // not under capture (refused above), so the clear is an ordinary
    // stream-ordered memset
    ACC_HIP_CHECK(hipMemsetAsync(ws->ptr, 0, kCtlBytes + cap, s->stream));
    // the old workspace may still be in use by queued work: it is retired,
    // not freed (no synchronisation of the stream or the device here)
    if (s->ws) s->retired.push_back(s->ws);
    s->ws = ws;
  }
  *ctl = (SyncCtl*)s->ws->ptr;
  *status = (uint64_t*)((char*)s->ws->ptr + kCtlBytes);
  return ACC_SUCCESS;
}

int stream_scratch(acc_stream_t s, size_t bytes, void** p) {
  if (!s) return fail(ACC_ERROR_INVALID_VALUE, "null stream");
  ACC_CHECK(set_device(s->device));
  if (tl_scratch.base) {
    if (tl_scratch.device != s->device)
      return fail(ACC_ERROR_INVALID_VALUE, "scratch of device %d used on device %d", tl_scratch.device, s->device);
    if (bytes > tl_scratch.scr_bytes)
      return fail(ACC_ERROR_INVALID_VALUE,
                  "allocation-free call needs %zu bytes of untagged scratch; the maximum is %zu bytes", bytes,
                  tl_scratch.scr_bytes);
    *p = tl_scratch.scr;
    return ACC_SUCCESS;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  if (!s->scr || bytes > s->scr->status_bytes) {
    if (s->capturing)
      return fail(ACC_ERROR_INVALID_VALUE,
                  "stream scratch must grow to %zu bytes during capture; call acc_hip_stream_reserve before "
                  "capturing", bytes);
    size_t cap = s->scr ? s->scr->status_bytes : (size_t)1 << 20;
    while (cap < bytes) cap *= 2;
    auto w = std::make_shared<Workspace>();
    w->device = s->device;
    ACC_CHECK(dev_malloc(s->device, cap, &w->ptr));
    w->status_bytes = cap;
    // (queued work may still use the old one: retired, not freed)
    if (s->scr) s->retired.push_back(s->scr);
    s->scr = w;
  }
  *p = s->scr->ptr;
  return ACC_SUCCESS;
}

// ---------------------------------------------------------------- fill kernel
// vec (dst aligned to the element): a head of elements up to the first
// 16-byte boundary, whole 16-byte stores of the replicated value (ordinary
// stores: for write-only streams they measured faster than non-temporal
// ones), then a tail of elements; otherwise element by element.
template <typename W>
__global__ __launch_bounds__(256) void k_fill(W* __restrict__ dst, W value, size_t count, int vec) {
  constexpr int E = 16 / (int)sizeof(W);
  const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  if (!vec) {
    for (size_t i = tid; i < count; i += stride) dst[i] = value;
    return;
  }
  const size_t mis = ((uintptr_t)dst % 16) / sizeof(W);
  const size_t head = mis ? min((size_t)E - mis, count) : 0;
  if (tid < head) dst[tid] = value;
  W rep[E];
#pragma unroll
  for (int j = 0; j < E; ++j) rep[j] = value;
  u32x4 v;
  __builtin_memcpy(&v, rep, 16);
  u32x4* vd = (u32x4*)(dst + head);
  const size_t nv = (count - head) / E;
  for (size_t i = tid; i < nv; i += stride) vd[i] = v;
  const size_t done = head + nv * E;  // count - done < E
  if (tid < count - done) dst[done + tid] = value;
}

// Device-to-device copy of nv 16-byte vectors: four non-temporal dwordx4
// loads in flight per thread, then four non-temporal stores (measured 6.1
// TB/s on 1 GiB, against 5.2 TB/s for hipMemcpyAsync's blit).
__global__ __launch_bounds__(256) void k_copy16(const u32x4* __restrict__ in, u32x4* __restrict__ out, size_t nv) {
  constexpr int U = 4;
  const size_t stride = (size_t)gridDim.x * 256 * U;
  for (size_t i0 = (size_t)blockIdx.x * 256 * U + threadIdx.x; i0 < nv; i0 += stride) {
    u32x4 r[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const size_t i = i0 + (size_t)u * 256;
      if (i < nv) r[u] = __builtin_nontemporal_load(in + i);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const size_t i = i0 + (size_t)u * 256;
      if (i < nv) __builtin_nontemporal_store(r[u], out + i);
    }
  }
}

}  // namespace acc

using namespace acc;

extern "C" {

const char* acc_hip_last_error(void) { return g_last_error.c_str(); }
int acc_hip_abi_version(void) { return ACC_HIP_ABI_VERSION; }

int acc_hip_device_count(int* count) {
  if (!count) return fail(ACC_ERROR_INVALID_VALUE, "null count");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e == hipErrorNoDevice) { *count = 0; (void)hipGetLastError(); return ACC_SUCCESS; }
  if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
  *count = n;
  return ACC_SUCCESS;
}

int acc_hip_device_props(int device, acc_device_props* props) {
  if (!props) return fail(ACC_ERROR_INVALID_VALUE, "null props");
  hipDeviceProp_t p;
  ACC_HIP_CHECK(hipGetDeviceProperties(&p, device));
  memset(props, 0, sizeof *props);
  snprintf(props->name, sizeof props->name, "%s", p.name);
  snprintf(props->arch, sizeof props->arch, "%s", p.gcnArchName);
  props->compute_units = p.multiProcessorCount;
  props->wave_size = p.warpSize;
  props->max_threads_per_block = p.maxThreadsPerBlock;
  props->lds_per_block = p.sharedMemPerBlock;
  props->total_global_mem = p.totalGlobalMem;
  props->clock_khz = p.clockRate;
  props->memory_clock_khz = p.memoryClockRate;
  props->memory_bus_width = p.memoryBusWidth;
  props->l2_cache_bytes = p.l2CacheSize;
  return ACC_SUCCESS;
}

int acc_hip_device_synchronize(int device) {
  ACC_CHECK(set_device(device));
  ACC_HIP_CHECK(hipDeviceSynchronize());
  return ACC_SUCCESS;
}

int acc_hip_malloc(int device, size_t bytes, void** ptr) {
  if (!ptr) return fail(ACC_ERROR_INVALID_VALUE, "null ptr");
  return dev_malloc(device, bytes, ptr);
}

int acc_hip_free(int device, void* ptr) { return dev_free(device, ptr); }

int acc_hip_mem_get_info(int device, size_t* free_bytes, size_t* total_bytes) {
  ACC_CHECK(set_device(device));
  size_t f = 0, t = 0;
  ACC_HIP_CHECK(hipMemGetInfo(&f, &t));
  if (device >= 0 && device < 64) {
    Accounting& a = g_acct[device];
    std::lock_guard<std::mutex> lk(a.mu);
    if (a.limit) {
      size_t left = a.limit > a.bytes ? a.limit - a.bytes : 0;
      if (left < f) f = left;
      if (a.limit < t) t = a.limit;
    }
  }
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return ACC_SUCCESS;
}

int acc_hip_set_alloc_limit(int device, size_t limit_bytes) {
  if (device < 0 || device >= 64) return fail(ACC_ERROR_INVALID_VALUE, "bad device %d", device);
  std::lock_guard<std::mutex> lk(g_acct[device].mu);
  g_acct[device].limit = limit_bytes;
  return ACC_SUCCESS;
}

int acc_hip_allocated_bytes(int device, size_t* bytes) {
  if (device < 0 || device >= 64 || !bytes) return fail(ACC_ERROR_INVALID_VALUE, "bad argument");
  std::lock_guard<std::mutex> lk(g_acct[device].mu);
  *bytes = g_acct[device].bytes;
  return ACC_SUCCESS;
}

int acc_hip_host_alloc(size_t bytes, void** ptr) {
  if (!ptr) return fail(ACC_ERROR_INVALID_VALUE, "null ptr");
  ACC_HIP_CHECK(hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault));
  return ACC_SUCCESS;
}

int acc_hip_host_free(void* ptr) {
  if (!ptr) return ACC_SUCCESS;
  ACC_HIP_CHECK(hipHostFree(ptr));
  return ACC_SUCCESS;
}

static int copy_async(acc_stream_t s, void* dst, const void* src, size_t bytes, hipMemcpyKind k) {
  if (!s) return fail(ACC_ERROR_INVALID_VALUE, "null stream");
  if (!bytes) return ACC_SUCCESS;
  ACC_CHECK(set_device(s->device));
  ACC_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, k, s->stream));
  return ACC_SUCCESS;
}

int acc_hip_memcpy_h2d_async(acc_stream_t s, void* dst, const void* src, size_t bytes) {
  return copy_async(s, dst, src, bytes, hipMemcpyHostToDevice);
}
int acc_hip_memcpy_d2h_async(acc_stream_t s, void* dst, const void* src, size_t bytes) {
  return copy_async(s, dst, src, bytes, hipMemcpyDeviceToHost);
}
int acc_hip_memcpy_d2d_async(acc_stream_t s, void* dst, const void* src, size_t bytes) {
  // 16-byte aligned buffers of a whole number of vectors: the streaming
  // copy kernel; anything else: the runtime's copy
  if (s && dst && src && bytes >= 16 && bytes % 16 == 0 && (uintptr_t)dst % 16 == 0 && (uintptr_t)src % 16 == 0) {
    ACC_CHECK(set_device(s->device));
    const size_t nv = bytes / 16;
    const DeviceInfo* di;
    ACC_CHECK(device_info(s->device, &di));
    size_t grid = (nv + 1023) / 1024;
    const size_t cap = (size_t)di->compute_units * 32;  // 8192 workgroups on 256 CUs
    if (grid > cap) grid = cap;
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)grid), dim3(256), 0, s->stream, (const u32x4*)src, (u32x4*)dst, nv);
    ACC_HIP_CHECK(hipGetLastError());
    return ACC_SUCCESS;
  }
  return copy_async(s, dst, src, bytes, hipMemcpyDeviceToDevice);
}

int acc_hip_fill_async(acc_stream_t s, void* dst, const void* value, size_t elem_bytes, size_t count) {
  if (!s || !value) return fail(ACC_ERROR_INVALID_VALUE, "null argument");
  if (!count) return ACC_SUCCESS;
  if (!dst) return fail(ACC_ERROR_INVALID_VALUE, "null destination");
  ACC_CHECK(set_device(s->device));
  if (elem_bytes != 1 && elem_bytes != 2 && elem_bytes != 4 && elem_bytes != 8)
    return fail(ACC_ERROR_INVALID_VALUE, "fill element size %zu", elem_bytes);
  const int vec = (uintptr_t)dst % elem_bytes == 0 ? 1 : 0;
  // one thread per 16 bytes (per element without vec), at most 8192 workgroups
  const size_t work = vec ? count * elem_bytes / 16 + 1 : count;
  const unsigned grid = (unsigned)std::min<size_t>((work + 255) / 256, 8192);
  switch (elem_bytes) {
    case 1: { uint8_t v; memcpy(&v, value, 1); hipLaunchKernelGGL(k_fill<uint8_t>, grid, 256, 0, s->stream, (uint8_t*)dst, v, count, vec); break; }
    case 2: { uint16_t v; memcpy(&v, value, 2); hipLaunchKernelGGL(k_fill<uint16_t>, grid, 256, 0, s->stream, (uint16_t*)dst, v, count, vec); break; }
    case 4: { uint32_t v; memcpy(&v, value, 4); hipLaunchKernelGGL(k_fill<uint32_t>, grid, 256, 0, s->stream, (uint32_t*)dst, v, count, vec); break; }
    case 8: { uint64_t v; memcpy(&v, value, 8); hipLaunchKernelGGL(k_fill<uint64_t>, grid, 256, 0, s->stream, (uint64_t*)dst, v, count, vec); break; }
  }
  ACC_HIP_CHECK(hipGetLastError());
  return ACC_SUCCESS;
}

int acc_hip_stream_create(int device, acc_stream_t* out) {
  if (!out) return fail(ACC_ERROR_INVALID_VALUE, "null stream out");
  const DeviceInfo* di;
  ACC_CHECK(device_info(device, &di));
  ACC_CHECK(set_device(device));
  static std::atomic<uint64_t> next_id{1};
  auto s = std::make_unique<acc_stream_s>();
  s->device = device;
  s->id = next_id++;
  ACC_HIP_CHECK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  *out = s.release();
  return ACC_SUCCESS;
}

int acc_hip_stream_destroy(acc_stream_t s) {
  if (!s) return ACC_SUCCESS;
  int rc = set_device(s->device);
  if (rc == ACC_SUCCESS) {
    (void)hipStreamSynchronize(s->stream);
    // a workspace that a live graph still references outlives the stream
    s->ws.reset();
    s->scr.reset();
    s->retired.clear();
    (void)hipStreamDestroy(s->stream);
  }
  delete s;
  return rc;
}

int acc_hip_stream_synchronize(acc_stream_t s) {
  if (!s) return fail(ACC_ERROR_INVALID_VALUE, "null stream");
  ACC_CHECK(set_device(s->device));
  ACC_HIP_CHECK(hipStreamSynchronize(s->stream));
  return ACC_SUCCESS;
}

int acc_hip_stream_query(acc_stream_t s) {
  if (!s) return fail(ACC_ERROR_INVALID_VALUE, "null stream");
  hipError_t e = hipStreamQuery(s->stream);
  if (e == hipSuccess) return ACC_SUCCESS;
  if (e == hipErrorNotReady) { (void)hipGetLastError(); return ACC_ERROR_NOT_READY; }
  return hip_fail(e, "hipStreamQuery");
}

int acc_hip_stream_device(acc_stream_t s, int* device) {
  if (!s || !device) return fail(ACC_ERROR_INVALID_VALUE, "null argument");
  *device = s->device;
  return ACC_SUCCESS;
}

int acc_hip_stream_native(acc_stream_t s, void** hip_stream) {
  if (!s || !hip_stream) return fail(ACC_ERROR_INVALID_VALUE, "null argument");
  *hip_stream = (void*)s->stream;
  return ACC_SUCCESS;
}

int acc_hip_stream_reserve(acc_stream_t s, size_t status_bytes) {
  SyncCtl* c;
  uint64_t* st;
  void* p;
  ACC_CHECK(stream_workspace(s, status_bytes, &c, &st));
  return stream_scratch(s, status_bytes, &p);
}

int acc_hip_event_create(int device, int timing, acc_event_t* out) {
  if (!out) return fail(ACC_ERROR_INVALID_VALUE, "null event out");
  ACC_CHECK(set_device(device));
  auto e = std::make_unique<acc_event_s>();
  e->device = device;
  ACC_HIP_CHECK(hipEventCreateWithFlags(&e->event, timing ? hipEventDefault : hipEventDisableTiming));
  *out = e.release();
  return ACC_SUCCESS;
}

int acc_hip_event_destroy(acc_event_t e) {
  if (!e) return ACC_SUCCESS;
  (void)hipEventDestroy(e->event);
  delete e;
  return ACC_SUCCESS;
}

int acc_hip_event_record(acc_event_t e, acc_stream_t s) {
  if (!e || !s) return fail(ACC_ERROR_INVALID_VALUE, "null argument");
  ACC_CHECK(set_device(s->device));
  ACC_HIP_CHECK(hipEventRecord(e->event, s->stream));
  return ACC_SUCCESS;
}

int acc_hip_stream_wait_event(acc_stream_t s, acc_event_t e) {
  if (!e || !s) return fail(ACC_ERROR_INVALID_VALUE, "null argument");
  ACC_CHECK(set_device(s->device));
  ACC_HIP_CHECK(hipStreamWaitEvent(s->stream, e->event, 0));
  return ACC_SUCCESS;
}

int acc_hip_event_query(acc_event_t e) {
  if (!e) return fail(ACC_ERROR_INVALID_VALUE, "null event");
  hipError_t r = hipEventQuery(e->event);
  if (r == hipSuccess) return ACC_SUCCESS;
  if (r == hipErrorNotReady) { (void)hipGetLastError(); return ACC_ERROR_NOT_READY; }
  return hip_fail(r, "hipEventQuery");
}

int acc_hip_event_synchronize(acc_event_t e) {
  if (!e) return fail(ACC_ERROR_INVALID_VALUE, "null event");
  ACC_HIP_CHECK(hipEventSynchronize(e->event));
  return ACC_SUCCESS;
}

int acc_hip_event_elapsed_ms(acc_event_t a, acc_event_t b, float* ms) {
  if (!a || !b || !ms) return fail(ACC_ERROR_INVALID_VALUE, "null argument");
  ACC_HIP_CHECK(hipEventElapsedTime(ms, a->event, b->event));
  return ACC_SUCCESS;
}

int acc_hip_module_load_data(int device, const void* image, size_t bytes, acc_module_t* out) {
  (void)bytes;
  if (!image || !out) return fail(ACC_ERROR_INVALID_VALUE, "null argument");
  ACC_CHECK(set_device(device));
  auto m = std::make_unique<acc_module_s>();
  m->device = device;
  ACC_HIP_CHECK(hipModuleLoadData(&m->module, image));
  *out = m.release();
  return ACC_SUCCESS;
}

int acc_hip_module_unload(acc_module_t m) {
  if (!m) return ACC_SUCCESS;
  (void)set_device(m->device);
  hipError_t e = hipModuleUnload(m->module);
  delete m;
  if (e != hipSuccess) return hip_fail(e, "hipModuleUnload");
  return ACC_SUCCESS;
}

int acc_hip_module_get_function(acc_module_t m, const char* name, acc_function_t* f) {
  if (!m || !name || !f) return fail(ACC_ERROR_INVALID_VALUE, "null argument");
  hipFunction_t fn;
  ACC_HIP_CHECK(hipModuleGetFunction(&fn, m->module, name));
  *f = (acc_function_t)fn;
  return ACC_SUCCESS;
}

int acc_hip_function_attributes(acc_function_t f, int* regs, int* lds, int* max_threads) {
  if (!f) return fail(ACC_ERROR_INVALID_VALUE, "null function");
  hipFunction_t fn = (hipFunction_t)f;
  int v = 0;
  if (regs) { ACC_HIP_CHECK(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_NUM_REGS, fn)); *regs = v; }
  if (lds) { ACC_HIP_CHECK(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_SHARED_SIZE_BYTES, fn)); *lds = v; }
  if (max_threads) { ACC_HIP_CHECK(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_MAX_THREADS_PER_BLOCK, fn)); *max_threads = v; }
  return ACC_SUCCESS;
}

int acc_hip_launch(acc_function_t f, unsigned gx, unsigned gy, unsigned gz, unsigned bx, unsigned by,
                   unsigned bz, unsigned lds, acc_stream_t s, void** args) {
  if (!f || !s) return fail(ACC_ERROR_INVALID_VALUE, "null argument");
  if ((uint64_t)gx * gy * gz == 0) return ACC_SUCCESS;  // zero-size launches are skipped (PTX/Execute.hs:847-854)
  ACC_CHECK(set_device(s->device));
  ACC_HIP_CHECK(hipModuleLaunchKernel((hipFunction_t)f, gx, gy, gz, bx, by, bz, lds, s->stream, args, nullptr));
  return ACC_SUCCESS;
}

int acc_hip_graph_begin_capture(acc_stream_t s) {
  if (!s) return fail(ACC_ERROR_INVALID_VALUE, "null stream");
  ACC_CHECK(set_device(s->device));
  ACC_HIP_CHECK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  std::lock_guard<std::mutex> lk(s->mu);
  s->capturing = true;
  return ACC_SUCCESS;
}

int acc_hip_graph_end_capture(acc_stream_t s, acc_graph_t* out) {
  if (!s || !out) return fail(ACC_ERROR_INVALID_VALUE, "null argument");
  ACC_CHECK(set_device(s->device));
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->capturing = false;
  }
  auto g = std::make_unique<acc_graph_s>();
  g->device = s->device;
  g->stream_id = s->id;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    g->ws = s->ws;  // growth is refused during capture: these are the ones the kernels captured
    g->scr = s->scr;
  }
  ACC_HIP_CHECK(hipStreamEndCapture(s->stream, &g->graph));
  hipError_t e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g->graph);
    return hip_fail(e, "hipGraphInstantiate");
  }
  *out = g.release();
  return ACC_SUCCESS;
}

int acc_hip_graph_launch(acc_graph_t g, acc_stream_t s) {
  if (!g || !s) return fail(ACC_ERROR_INVALID_VALUE, "null argument");
  // the graph's skeleton kernels use the capture stream's tickets and status
  // words: a replay on another stream could interleave with that stream's own
  // skeletons on the same counters
  if (s->id != g->stream_id)
    return fail(ACC_ERROR_INVALID_VALUE, "graph: launch on a stream other than the one it was captured on");
  ACC_CHECK(set_device(s->device));
  ACC_HIP_CHECK(hipGraphLaunch(g->exec, s->stream));
  return ACC_SUCCESS;
}

int acc_hip_graph_destroy(acc_graph_t g) {
  if (!g) return ACC_SUCCESS;
  (void)set_device(g->device);
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
  return ACC_SUCCESS;
}

}  // extern "C"
