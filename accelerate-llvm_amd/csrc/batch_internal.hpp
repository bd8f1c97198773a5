// Batch planning shared by the batched form (batch.hip) and the server's
// fused rounds (server.hip): validation, the per-problem descriptors and the
// run→problem maps for one (workgroup size, vectors per thread, run length),
// and the launch of the stencil kernel and the persistent kernel over a plan
// whose descriptors, maps and workspace are already on the device.
#pragma once
#include <vector>
#include "host.hpp"

namespace acc {

struct BatchDescBytes {
  uint64_t w[12];  // one device descriptor (its layout is private to batch.hip)
};

struct BatchPlan {
  std::vector<BatchDescBytes> desc;
  std::vector<uint32_t> run_prob;  // ticket -> problem
  std::vector<uint32_t> stencil_prob;  // stencil kernel workgroup -> problem
  int block = 256, k = 4, R = 4;
  uint32_t total_runs = 0;
  uint64_t status_words = 0;       // 8-byte status words after the SyncCtl (kSyncCtlBytes)
  size_t lds_bytes = 0;            // dynamic LDS (histogram bins)
};

int batch_validate(const acc_batch_problem* problems, int64_t count);
bool batch_shape_supported(int block, int k);
int batch_plan(const acc_batch_problem* problems, int64_t count, int block, int k, int R, BatchPlan* plan);
int batch_launch_plan(int device, const BatchPlan& plan, const void* d_desc, const void* d_map,
                      const void* d_smap, void* d_ws, hipStream_t stream);

}  // namespace acc
