// Host-side internals shared by the runtime and the skeleton drivers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <mutex>
#include <vector>
#include <memory>
#include "../../include/accelerate_hip.h"

namespace acc {

struct SyncCtl;

// Sets the thread-local error message and returns `code`.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int hip_fail(hipError_t e, const char* what);

#define ACC_HIP_CHECK(expr)                                  \
  do {                                                       \
    hipError_t _e = (expr);                                  \
    if (_e != hipSuccess) return ::acc::hip_fail(_e, #expr); \
  } while (0)

#define ACC_CHECK(expr)                 \
  do {                                  \
    int _s = (expr);                    \
    if (_s != ACC_SUCCESS) return _s;   \
  } while (0)

struct DeviceInfo {
  int compute_units = 0;
  int max_threads = 0;
  size_t lds_per_block = 0;
  size_t total_mem = 0;
  std::string arch;
};

int device_info(int device, const DeviceInfo** out);

// Device allocation with the library's accounting (allocation limit knob).
int dev_malloc(int device, size_t bytes, void** p);
int dev_free(int device, void* p);

}  // namespace acc

namespace acc {
// One device allocation of skeleton workspace: [SyncCtl (256 B) | status
// words].  Shared ownership: the stream that allocated it and every graph
// captured on that stream while it was current keep it alive, since a graph's
// kernels hold its address as frozen arguments.
struct Workspace {
  int device = 0;
  void* ptr = nullptr;
  size_t status_bytes = 0;  // capacity of the status region
  ~Workspace();
};
}  // namespace acc

// A stream and the device workspace that its skeletons use.  The workspace
// only grows; growth allocates a new one and retires the old one (which
// queued work may still use) without synchronising.
struct acc_stream_s {
  int device = 0;
  uint64_t id = 0;  // unique per stream created by the library (never reused)
  hipStream_t stream = nullptr;
  std::mutex mu;
  std::shared_ptr<acc::Workspace> ws;
  std::shared_ptr<acc::Workspace> scr;  // untagged scratch (status_bytes = its capacity)
  bool capturing = false;
  // workspaces outgrown by this stream: released when the stream is destroyed
  // (hipFree would synchronise the whole device at the moment of growth)
  std::vector<std::shared_ptr<acc::Workspace>> retired;
};

struct acc_event_s {
  int device = 0;
  hipEvent_t event = nullptr;
};

struct acc_module_s {
  int device = 0;
  hipModule_t module = nullptr;
};

struct acc_graph_s {
  int device = 0;
  uint64_t stream_id = 0;  // the capture stream: the only stream it may be launched on
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::shared_ptr<acc::Workspace> ws;   // the capture stream's workspace, kept alive
  std::shared_ptr<acc::Workspace> scr;  // and its untagged scratch
};

namespace acc {

// Workspace of `s` with at least `status_bytes` of status words; sets the
// device.  Returns the SyncCtl pointer and the status region.  While a
// ScratchOverride is alive on the calling thread, the override's memory is
// returned instead (nothing is allocated) or, if it is too small, an error.
//
// The status region holds tagged words only (look-back status, published
// partials): every skeleton on the stream reads it by tag, and a word of
// arbitrary data there can carry the current tag by chance.  Data without
// tags (per-tile records, index tables, locks) goes to stream_scratch.
int stream_workspace(acc_stream_t s, size_t status_bytes, SyncCtl** ctl, uint64_t** status);

// Untagged scratch of `s` (at least `bytes`, 256-byte aligned, contents
// undefined); sets the device.  Same override and capture rules as the
// workspace.
int stream_scratch(acc_stream_t s, size_t bytes, void** p);

// Routes the skeleton workspace ([SyncCtl (256 B) | status words]) and the
// untagged scratch of the calling thread's launches to caller-provided device
// regions for the lifetime of the object (the allocation-free entry points).
struct ScratchOverride {
  ScratchOverride(int device, void* ws, size_t ws_bytes, void* scratch, size_t scratch_bytes);
  ~ScratchOverride();
  ScratchOverride(const ScratchOverride&) = delete;
  ScratchOverride& operator=(const ScratchOverride&) = delete;
};

inline int set_device(int device) {
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  return ACC_SUCCESS;
}

size_t type_size(acc_type t);

// Resident workgroups per CU of a kernel at a block size (cached).
int blocks_per_cu(const void* kernel, int block);

}  // namespace acc

// All rows of a rank >= 2 scan in one launch (scan_rows.hip).
int acc_scan_rows(acc_stream_t s, acc_type t, acc_op op, bool reverse, bool exclusive, const void* seed,
                  const void* in, int64_t outer, int64_t inner, void* out, int64_t out_row, void* sums,
                  bool seed_at_end);
