// Compile-time splicing of element type and combine function into the
// skeleton templates: one instantiation per (type, function) pair that the
// C ABI offers without run-time compilation.
#pragma once
#include <type_traits>
#include "elt.hpp"
#include "host.hpp"

namespace acc {

template <typename F>
int with_type(acc_type t, F&& f) {
  switch (t) {
    case ACC_INT32: return f((int32_t)0);
    case ACC_INT64: return f((int64_t)0);
    case ACC_WORD32: return f((uint32_t)0);
    case ACC_WORD64: return f((uint64_t)0);
    case ACC_FLOAT32: return f((float)0);
    case ACC_FLOAT64: return f((double)0);
  }
  return fail(ACC_ERROR_UNSUPPORTED, "element type %d", (int)t);
}

// fold/scan combine functions
template <typename T, typename F>
int with_op(acc_op op, F&& f) {
  switch (op) {
    case ACC_OP_ADD: return f(OpAdd{});
    case ACC_OP_MUL: return f(OpMul{});
    case ACC_OP_MIN: return f(OpMin{});
    case ACC_OP_MAX: return f(OpMax{});
    default: break;
  }
  if constexpr (std::is_integral<T>::value) {
    switch (op) {
      case ACC_OP_AND: return f(OpAnd{});
      case ACC_OP_OR: return f(OpOr{});
      case ACC_OP_XOR: return f(OpXor{});
      default: break;
    }
  }
  return fail(ACC_ERROR_UNSUPPORTED, "combine function %d for this element type", (int)op);
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace acc
