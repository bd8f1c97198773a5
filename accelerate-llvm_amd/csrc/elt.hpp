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
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/accelerate_hip.h"

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

// A value that may be absent. Reductions without an identity element carry
// the "have a value" bit through the wave and block stages so that the
// ragged edges of a tile need no neutral element (fold1, scanl1 and the
// NaN-propagation rules of fminf/fmaxf all stay exact).
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
