// Allocation-free foldSeg / fold1Seg: the same kernels as acc_hip_fold_seg,
// with their workspace (the launch epoch and tickets, the fix-up's tagged
// status words) and untagged scratch (per-tile carry records, per-tile
// segment ranges and starts) in a fixed device array of this file instead of
// the stream's allocations.  No device
// memory is allocated or freed by a call, so callers that cannot allocate
// (code under graph capture, tight serving loops, harnesses that forbid
// hipMalloc) can use it.
//
// `scratch` is zero at load.  Nothing in it needs clearing between calls:
// the status words carry the launch epoch, which the last workgroup of every
// launch advances, and every per-tile record and table is written by one
// kernel of a call before the next kernel of the same call reads it.  Every call on a device
// shares that device's copy, so two calls that overlap on one device (on
// different streams, or from different threads) race.
#include "host.hpp"

namespace acc {

// 16 MiB: the first 64 KiB are the workspace (launch epoch, tickets, the
// fix-up's status words), the rest the records and tables of a fold1Seg over
// about 2^28.8 Float32 (or 2^27.8 Float64) elements in any number of
// segments; larger calls are refused before any launch.
constexpr size_t kNoAllocScratchBytes = (size_t)16 << 20;
constexpr size_t kNoAllocWorkspaceBytes = (size_t)64 << 10;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// Device address of this device's copy of `scratch`, looked up once per
// device with that device current.  The array's host-side name is only a
// stand-in for the symbol and is never passed to a kernel or a copy.
static int scratch_address(int device, void** out) {
  static std::mutex mu;
  static void* addr[64] = {};
  if (device < 0 || device >= 64) return fail(ACC_ERROR_INVALID_VALUE, "bad device %d", device);
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[device]) {
    ACC_CHECK(set_device(device));
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(ACC_ERROR_HIP, "hipGetSymbolAddress(scratch) returned no address");
    addr[device] = p;
  }
  *out = addr[device];
  return ACC_SUCCESS;
}

}  // namespace acc

using namespace acc;

extern "C" int acc_hip_fold_seg_noalloc(acc_stream_t s, acc_type t, acc_op op, const void* seed, const void* in,
                                        int64_t outer, int64_t inner, acc_type offset_type, const void* offsets,
                                        int64_t nseg, void* out) {
  if (!s) return fail(ACC_ERROR_INVALID_VALUE, "foldSeg (no-alloc): null stream");
  void* base = nullptr;
  ACC_CHECK(scratch_address(s->device, &base));
  ScratchOverride guard(s->device, base, kNoAllocWorkspaceBytes, (char*)base + kNoAllocWorkspaceBytes,
                        kNoAllocScratchBytes - kNoAllocWorkspaceBytes);
  return acc_hip_fold_seg(s, t, op, seed, in, outer, inner, offset_type, offsets, nseg, out);
}
