// 16-byte (dwordx4) global loads and stores of VEC = 16 / sizeof(T) elements.
#pragma once
#ifndef __HIPCC_RTC__  // hiprtc provides the HIP runtime itself
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

namespace acc {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
// Every access here goes through a pointer in the global address space: a
// pointer the compiler cannot place (one loaded from memory, as the batched
// kernel's descriptors hold) would otherwise give flat accesses, which also
// count against the LDS counter and make LDS operations wait for them.
typedef __attribute__((address_space(1))) u32x4 g_u32x4;

template <typename T> struct Vec {
  static constexpr int N = 16 / (int)sizeof(T);
  T e[N];
};

template <typename T> __device__ __forceinline__ Vec<T> load16(const T* p) {
  u32x4 raw = *(const g_u32x4*)p;
  Vec<T> v;
  __builtin_memcpy(&v, &raw, 16);
  return v;
}

// streaming load: the data is read once
template <typename T> __device__ __forceinline__ Vec<T> load16_stream(const T* p) {
  u32x4 raw = __builtin_nontemporal_load((const g_u32x4*)p);
  Vec<T> v;
  __builtin_memcpy(&v, &raw, 16);
  return v;
}

template <typename T> __device__ __forceinline__ void store16(T* p, const Vec<T>& v) {
  u32x4 raw;
  __builtin_memcpy(&raw, &v, 16);
  *(g_u32x4*)p = raw;
}

template <typename T> __device__ __forceinline__ void store16_stream(T* p, const Vec<T>& v) {
  u32x4 raw;
  __builtin_memcpy(&raw, &v, 16);
  __builtin_nontemporal_store(raw, (g_u32x4*)p);
}

// a scalar global load / store
template <typename T> __device__ __forceinline__ T gload(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}
template <typename T> __device__ __forceinline__ void gstore(T* p, T v) {
  *(__attribute__((address_space(1))) T*)p = v;
}

}  // namespace acc
