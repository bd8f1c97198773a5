// HBM ceilings of the device, measured with the simplest kernels that can
// reach them: a read-only reduction, a copy and a write-only fill over a
// buffer, each a grid of `blocks_per_cu` 256-thread workgroups per CU
// striding over 16-byte vectors with `vectors` loads (or stores) in flight
// per lane, with or without the non-temporal hint.  No counterpart in the
// reference: it calibrates the roofline fractions the skeletons are
// measured against (a skeleton cannot move its bytes faster than these
// kernels move theirs).
#include "host.hpp"
#include "vecmem.hpp"

namespace acc {

template <int U, bool NT>
__global__ __launch_bounds__(256) void k_probe_read(const u32x4* in, size_t nvec, uint32_t* sink) {
  size_t i = (size_t)blockIdx.x * 256 * U + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * 256 * U;
  uint32_t acc = 0;
  for (; i + (U - 1) * 256 < nvec; i += stride) {
    u32x4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      v[u] = NT ? __builtin_nontemporal_load((const g_u32x4*)(in + i + u * 256)) : *(const g_u32x4*)(in + i + u * 256);
#pragma unroll
    for (int u = 0; u < U; ++u) acc += v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
  }
  if (acc == 0x9e3779b9u) sink[0] = acc;  // keeps the loads; practically never true
}

template <int U, bool NT>
__global__ __launch_bounds__(256) void k_probe_copy(const u32x4* in, u32x4* out, size_t nvec) {
  size_t i = (size_t)blockIdx.x * 256 * U + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * 256 * U;
  for (; i + (U - 1) * 256 < nvec; i += stride) {
    u32x4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      v[u] = NT ? __builtin_nontemporal_load((const g_u32x4*)(in + i + u * 256)) : *(const g_u32x4*)(in + i + u * 256);
#pragma unroll
    for (int u = 0; u < U; ++u) *(g_u32x4*)(out + i + u * 256) = v[u];
  }
}

template <int U, bool NT>
__global__ __launch_bounds__(256) void k_probe_write(u32x4* out, size_t nvec) {
  size_t i = (size_t)blockIdx.x * 256 * U + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * 256 * U;
  const u32x4 v = {1u, 2u, 3u, (uint32_t)threadIdx.x};
  for (; i + (U - 1) * 256 < nvec; i += stride) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (NT) __builtin_nontemporal_store(v, (g_u32x4*)(out + i + u * 256));
      else *(g_u32x4*)(out + i + u * 256) = v;
    }
  }
}

template <int U, bool NT>
static void probe_launch(int kind, int grid, hipStream_t st, void* a, void* b, size_t nvec, uint32_t* sink) {
  if (kind == ACC_PROBE_READ)
    hipLaunchKernelGGL((k_probe_read<U, NT>), dim3(grid), dim3(256), 0, st, (const u32x4*)a, nvec, sink);
  else if (kind == ACC_PROBE_COPY)
    hipLaunchKernelGGL((k_probe_copy<U, NT>), dim3(grid), dim3(256), 0, st, (const u32x4*)a, (u32x4*)b, nvec);
  else
    hipLaunchKernelGGL((k_probe_write<U, NT>), dim3(grid), dim3(256), 0, st, (u32x4*)b, nvec);
}

}  // namespace acc

using namespace acc;

extern "C" int acc_hip_bandwidth_probe(acc_stream_t s, int kind, int vectors, int blocks_per_cu, int nontemporal,
                                       const void* src, void* dst, size_t bytes, int reps, double* us) {
  if (!s || !us || reps < 1) return fail(ACC_ERROR_INVALID_VALUE, "probe: bad arguments");
  if (kind < ACC_PROBE_READ || kind > ACC_PROBE_WRITE) return fail(ACC_ERROR_INVALID_VALUE, "probe: kind %d", kind);
  if (vectors != 4 && vectors != 8 && vectors != 16)
    return fail(ACC_ERROR_INVALID_VALUE, "probe: vectors must be 4, 8 or 16");
  if (blocks_per_cu < 1 || blocks_per_cu > 32) return fail(ACC_ERROR_INVALID_VALUE, "probe: 1..32 blocks per CU");
  if ((kind != ACC_PROBE_WRITE && !src) || (kind != ACC_PROBE_READ && !dst) || bytes < 16 ||
      (src && (uintptr_t)src % 16) || (dst && (uintptr_t)dst % 16))
    return fail(ACC_ERROR_INVALID_VALUE, "probe: 16-byte aligned buffers required");
  ACC_CHECK(set_device(s->device));
  const DeviceInfo* di;
  ACC_CHECK(device_info(s->device, &di));
  const int grid = di->compute_units * blocks_per_cu;
  const size_t nvec = bytes / 16;
  void* sink = nullptr;
  ACC_CHECK(dev_malloc(s->device, 64, &sink));
  auto launch = [&] {
    void* a = (void*)src;
    uint32_t* k = (uint32_t*)sink;
    if (nontemporal) {
      if (vectors == 4) probe_launch<4, true>(kind, grid, s->stream, a, dst, nvec, k);
      else if (vectors == 8) probe_launch<8, true>(kind, grid, s->stream, a, dst, nvec, k);
      else probe_launch<16, true>(kind, grid, s->stream, a, dst, nvec, k);
    } else {
      if (vectors == 4) probe_launch<4, false>(kind, grid, s->stream, a, dst, nvec, k);
      else if (vectors == 8) probe_launch<8, false>(kind, grid, s->stream, a, dst, nvec, k);
      else probe_launch<16, false>(kind, grid, s->stream, a, dst, nvec, k);
    }
  };
  hipEvent_t e0, e1;
  int rc = ACC_SUCCESS;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = fail(ACC_ERROR_HIP, "probe: events");
  if (!rc) {
    launch();  // warm-up
    (void)hipEventRecord(e0, s->stream);
    for (int r = 0; r < reps; ++r) launch();
    (void)hipEventRecord(e1, s->stream);
    hipError_t e = hipEventSynchronize(e1);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) rc = hip_fail(e, "probe");
    *us = (double)ms * 1e3 / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  (void)dev_free(s->device, sink);
  return rc;
}
