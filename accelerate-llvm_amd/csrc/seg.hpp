// Segmented combination: a value with a "a segment starts at or after the
// first element" flag.  Combining (a, b) restarts at b when b holds a
// segment start, so a scan with SegOp never carries a value across a start.
// The identity is (identity of op, 0).
#pragma once
#ifndef __HIPCC_RTC__  // hiprtc provides the HIP runtime itself
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

namespace acc {

template <typename T> struct SegV {
  T v;
  uint32_t f;
};

template <typename Op> struct SegOp {
  Op op;
  template <typename T>
  __device__ __forceinline__ SegV<T> operator()(SegV<T> a, SegV<T> b) const {
    SegV<T> r;
    r.v = b.f ? b.v : op(a.v, b.v);
    r.f = a.f | b.f;
    return r;
  }
};

}  // namespace acc
