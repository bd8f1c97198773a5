// The stencil boundary index map, shared by the 2-D and 3-D stencils.
#pragma once
#ifndef __HIPCC_RTC__  // hiprtc provides the HIP runtime itself
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>
#include <type_traits>
#include "../../include/accelerate_hip.h"

namespace acc {

// Coordinate i of an extent n mapped into [0, n) for clamp, mirror and wrap
// (LLVM/CodeGen/Stencil.hs:224-253: one step of the map, as the reference
// applies it); a constant boundary's coordinates are mapped too (to any
// index inside: the caller substitutes the constant).
__device__ __forceinline__ int64_t bound_index(int b, int64_t i, int64_t n) {
  if (i < 0) {
    if (b == ACC_BOUNDARY_CLAMP) i = 0;
    else if (b == ACC_BOUNDARY_MIRROR) i = -i;
    else if (b == ACC_BOUNDARY_WRAP) i = n + i;
  } else if (i >= n) {
    if (b == ACC_BOUNDARY_CLAMP) i = n - 1;
    else if (b == ACC_BOUNDARY_MIRROR) i = n - (i - n + 2);
    else if (b == ACC_BOUNDARY_WRAP) i = i - n;
  }
  // arrays smaller than the radius: keep the single-step map in bounds
  return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

// The Function boundary's value at a 2-D index (y, x) outside the array
// (LLVM/CodeGen/Stencil.hs:215-218) for the affine function set:
// (c0 + cy*y) + cx*x in the element type (integers wrap, floating operations
// round one by one).  The callers keep extents below 2^31 - 8 for it, so the
// halo's coordinates are 32-bit (one conversion instruction each).
template <typename T> __device__ __forceinline__ T bound_function(const T (&c)[3], int32_t y, int32_t x) {
#pragma clang fp contract(off)
  if constexpr (std::is_integral<T>::value) {
    using U = typename std::make_unsigned<T>::type;
    return (T)((U)((U)c[0] + (U)c[1] * (U)(T)y) + (U)c[2] * (U)(T)x);
  } else {
    const T t = c[0] + c[1] * (T)y;
    return t + c[2] * (T)x;
  }
}

}  // namespace acc
