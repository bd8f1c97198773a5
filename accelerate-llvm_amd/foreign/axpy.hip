// A stand-alone kernel compiled to its own gfx950 code object
// (lib/foreign_axpy.hsaco), the way a program's own kernel reaches the
// device through acc_hip_module_load_data / _module_get_function /
// _launch (the reference's Link.link and launch: PTX/Link.hs:57-121,
// PTX/Execute.hs:860-881).  y[i] = a * x[i] + y[i] for i < n, grid-stride.
#include <hip/hip_runtime.h>
#include <stdint.h>

extern "C" __global__ __launch_bounds__(256) void acc_foreign_axpy(float a, const float* __restrict__ x,
                                                                   float* __restrict__ y, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    y[i] = __fmaf_rn(a, x[i], y[i]);
}

// Counts its own invocations: a launch with a zero-sized grid must not run.
extern "C" __global__ void acc_foreign_count(uint32_t* counter) {
  if (threadIdx.x == 0) atomicAdd(counter, 1u);
}
