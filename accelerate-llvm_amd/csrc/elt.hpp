// Element types and combine functions spliced into the skeleton templates.
//
// The scalar semantics follow the reference's expression code generator:
//   * integer (+), (-), (*) wrap (two's complement): the reference emits
//     `add`/`sub`/`mul` without `nsw`/`nuw` flags
//     (accelerate-llvm/src/LLVM/AST/Type/Instruction.hs:425-434);
//   * floating `min`/`max` are C `fminf`/`fmaxf` (`fmin`/`fmax` for double):
//     accelerate-llvm/src/Data/Array/Accelerate/LLVM/CodeGen/Arithmetic.hs:456-466;
//   * integer `min`/`max` are plain comparisons (same file, :716-741).
// Every combine function is a binary functor `T op(T a, T b)` where `a` is the
// element on the left in the reduction order (the accumulator for a left-to-
// right traversal).
#pragma once
#ifndef __HIPCC_RTC__  // hiprtc provides the HIP runtime itself
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>
#include <type_traits>
#ifndef __HIPCC_RTC__  // nothing of the C interface is needed in device code
#include "../../include/accelerate_hip.h"
#endif

namespace acc {

template <typename T> struct unsigned_of { using type = T; };
template <> struct unsigned_of<int32_t> { using type = uint32_t; };
template <> struct unsigned_of<int64_t> { using type = uint64_t; };

struct OpAdd {
  template <typename T> __host__ __device__ __forceinline__ T operator()(T a, T b) const {
    using U = typename unsigned_of<T>::type;
    return (T)((U)a + (U)b);
  }
};
template <> __host__ __device__ __forceinline__ float OpAdd::operator()(float a, float b) const { return a + b; }
template <> __host__ __device__ __forceinline__ double OpAdd::operator()(double a, double b) const { return a + b; }

struct OpMul {
  template <typename T> __host__ __device__ __forceinline__ T operator()(T a, T b) const {
    using U = typename unsigned_of<T>::type;
    return (T)((U)a * (U)b);
  }
};
template <> __host__ __device__ __forceinline__ float OpMul::operator()(float a, float b) const { return a * b; }
template <> __host__ __device__ __forceinline__ double OpMul::operator()(double a, double b) const { return a * b; }

struct OpMin {
  template <typename T> __host__ __device__ __forceinline__ T operator()(T a, T b) const { return a < b ? a : b; }
};
template <> __host__ __device__ __forceinline__ float OpMin::operator()(float a, float b) const { return fminf(a, b); }
template <> __host__ __device__ __forceinline__ double OpMin::operator()(double a, double b) const { return fmin(a, b); }

struct OpMax {
  template <typename T> __host__ __device__ __forceinline__ T operator()(T a, T b) const { return a < b ? b : a; }
};
template <> __host__ __device__ __forceinline__ float OpMax::operator()(float a, float b) const { return fmaxf(a, b); }
template <> __host__ __device__ __forceinline__ double OpMax::operator()(double a, double b) const { return fmax(a, b); }

struct OpAnd {
  template <typename T> __host__ __device__ __forceinline__ T operator()(T a, T b) const { return a & b; }
};
struct OpOr {
  template <typename T> __host__ __device__ __forceinline__ T operator()(T a, T b) const { return a | b; }
};
struct OpXor {
  template <typename T> __host__ __device__ __forceinline__ T operator()(T a, T b) const { return a ^ b; }
};

// Right-to-left traversals combine `x (+) acc` (native CodeGen/Scan.hs:204-206):
// the reversed sequence is scanned left-to-right with the arguments swapped.
template <typename Op> struct Flip {
  Op op;
  template <typename T> __host__ __device__ __forceinline__ T operator()(T a, T b) const { return op(b, a); }
};

// Two-sided identity of each combine: op(e, x) == x == op(x, e) exactly, for
// every x including NaNs and signed zeros.  Used to pad the ragged edge of a
// tile.  Floating (+) uses -0.0 (-0.0 + x == x also for x == -0.0); fminf and
// fmaxf return the other operand when one is a quiet NaN (C99 7.12.12), so a
// quiet NaN is their identity.
template <typename T> struct Limits;
template <> struct Limits<int32_t> {
  static __device__ __forceinline__ int32_t lo() { return INT32_MIN; }
  static __device__ __forceinline__ int32_t hi() { return INT32_MAX; }
};
template <> struct Limits<int64_t> {
  static __device__ __forceinline__ int64_t lo() { return INT64_MIN; }
  static __device__ __forceinline__ int64_t hi() { return INT64_MAX; }
};
template <> struct Limits<uint32_t> {
  static __device__ __forceinline__ uint32_t lo() { return 0u; }
  static __device__ __forceinline__ uint32_t hi() { return UINT32_MAX; }
};
template <> struct Limits<uint64_t> {
  static __device__ __forceinline__ uint64_t lo() { return 0ull; }
  static __device__ __forceinline__ uint64_t hi() { return UINT64_MAX; }
};

template <typename T> __device__ __forceinline__ T neg_zero() { return T(0); }
template <> __device__ __forceinline__ float neg_zero<float>() { return __int_as_float(0x80000000); }
template <> __device__ __forceinline__ double neg_zero<double>() {
  return __longlong_as_double((long long)0x8000000000000000ull);
}
template <typename T> __device__ __forceinline__ T quiet_nan() { return T(0); }
template <> __device__ __forceinline__ float quiet_nan<float>() { return __int_as_float(0x7fc00000); }
template <> __device__ __forceinline__ double quiet_nan<double>() {
  return __longlong_as_double((long long)0x7ff8000000000000ull);
}

template <typename T, typename Op> struct Ident;
template <typename T> struct Ident<T, OpAdd> {
  static __device__ __forceinline__ T value() { return neg_zero<T>(); }
};
template <typename T> struct Ident<T, OpMul> {
  static __device__ __forceinline__ T value() { return T(1); }
};
template <typename T> struct Ident<T, OpMin> {
  static __device__ __forceinline__ T value() {
    if constexpr (std::is_floating_point<T>::value) return quiet_nan<T>();
    else return Limits<T>::hi();
  }
};
template <typename T> struct Ident<T, OpMax> {
  static __device__ __forceinline__ T value() {
    if constexpr (std::is_floating_point<T>::value) return quiet_nan<T>();
    else return Limits<T>::lo();
  }
};
template <typename T> struct Ident<T, OpAnd> {
  static __device__ __forceinline__ T value() { return (T)~(T)0; }
};
template <typename T> struct Ident<T, OpOr> {
  static __device__ __forceinline__ T value() { return T(0); }
};
template <typename T> struct Ident<T, OpXor> {
  static __device__ __forceinline__ T value() { return T(0); }
};
template <typename T, typename Op> struct Ident<T, Flip<Op>> {
  static __device__ __forceinline__ T value() { return Ident<T, Op>::value(); }
};
template <typename T, typename Op> __device__ __forceinline__ T identity(Op) { return Ident<T, Op>::value(); }

// A value that may be absent: the "have a value" bit travels through wave and
// block stages where a lane or a tile may hold no element (fold1 of ragged
// tiles, the segmented look-back).
template <typename T> struct Maybe {
  T v;
  uint32_t ok;
};

template <typename T, typename Op>
__device__ __forceinline__ Maybe<T> mcombine(Op op, Maybe<T> a, Maybe<T> b) {
  Maybe<T> r;
  r.v = a.ok ? (b.ok ? op(a.v, b.v) : a.v) : b.v;
  r.ok = a.ok | b.ok;
  return r;
}

}  // namespace acc
