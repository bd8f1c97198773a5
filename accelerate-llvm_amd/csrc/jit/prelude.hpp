// Scalar semantics of the program's own functions, for kernels compiled at
// run time: one device function per primitive of the expression language,
// each following the reference's code generator
// (accelerate-llvm/src/Data/Array/Accelerate/LLVM/CodeGen/Arithmetic.hs,
// dispatched from CodeGen/Exp.hs:244-310):
//   * integer add/sub/mul wrap (no nsw/nuw: LLVM/AST/Type/Instruction.hs:425-434);
//   * negate x = x * (-1)                                   (Arithmetic.hs:69-73)
//   * abs: fabs; identity for unsigned; abs minBound = minBound (:75-86)
//   * signum: (x /= 0) .|. (x `shiftRA` (bits-1)) for signed, (x /= 0) for
//     unsigned, (x > 0) - (x < 0) in the float type for floating (:88-115)
//   * quot/rem truncate; div/mod floor as the reference computes them (:121-225)
//   * shift amounts are converted to the value's type (:239-258); an amount
//     outside [0, bits) is masked to its low bits here (LLVM leaves it poison)
//   * rotate = funnel shift, amount modulo bits (:261-279)
//   * popCount/countLeadingZeros/countTrailingZeros of 0 count every bit, as Int (:281-315)
//   * floating min/max are fmin/fmax (:456-466); integer min/max select (:456-472)
//   * float comparisons are ordered: (/=) is "one", false when either is NaN (Instruction.hs:518-524)
//   * round = C round (half away from zero) then truncate (:416-419); floor/ceiling likewise
//   * logBase b x = log x / log b, log2/log10 for a constant base 2/10 (:375-392)
//   * isInfinite x = |x| == inf (:398-410)
//   * fromIntegral: sign/zero extension by the source's signedness, truncation, or
//     int->float conversion; toFloating: int->float, fptrunc/fpext (:501-535)
// Kernels are compiled with -ffp-contract=off: every floating operation rounds
// on its own (the reference's fast-math flags are off by default,
// Instruction.hs:654-662).
#pragma once
#include <stdint.h>
#include <type_traits>

namespace ap {

template <typename T> struct uns { using type = T; };
template <> struct uns<int8_t> { using type = uint8_t; };
template <> struct uns<int16_t> { using type = uint16_t; };
template <> struct uns<int32_t> { using type = uint32_t; };
template <> struct uns<int64_t> { using type = uint64_t; };
template <typename T> constexpr bool is_f = std::is_floating_point<T>::value;
template <typename T> constexpr bool is_s = std::is_signed<T>::value && !std::is_floating_point<T>::value;
template <typename T> constexpr int bits = (int)sizeof(T) * 8;

template <typename T> __device__ __forceinline__ T add(T a, T b) {
  if constexpr (is_f<T>) return a + b;
  else return (T)((typename uns<T>::type)a + (typename uns<T>::type)b);
}
template <typename T> __device__ __forceinline__ T sub(T a, T b) {
  if constexpr (is_f<T>) return a - b;
  else return (T)((typename uns<T>::type)a - (typename uns<T>::type)b);
}
template <typename T> __device__ __forceinline__ T mul(T a, T b) {
  if constexpr (is_f<T>) return a * b;
  else {
    using U = typename uns<T>::type;
    // promote narrow unsigned products past int so the wrap is defined
    using W = typename std::conditional<(sizeof(U) < 4), uint32_t, U>::type;
    return (T)(U)((W)(U)a * (W)(U)b);
  }
}
template <typename T> __device__ __forceinline__ T neg(T a) { return mul<T>(a, (T)-1); }
template <typename T> __device__ __forceinline__ T abs_(T a) {
  if constexpr (is_f<T>) {
    if constexpr (sizeof(T) == 4) return fabsf(a);
    else return fabs(a);
  } else if constexpr (!is_s<T>) {
    return a;
  } else {
    return a < 0 ? neg<T>(a) : a;  // minBound * (-1) wraps to minBound
  }
}
template <typename T> __device__ __forceinline__ T signum(T a) {
  if constexpr (is_f<T>) return (T)(a > (T)0 ? 1 : 0) - (T)(a < (T)0 ? 1 : 0);
  else if constexpr (!is_s<T>) return (T)(a != 0 ? 1 : 0);
  else return (T)((a != 0 ? 1 : 0) | (a >> (bits<T> - 1)));
}
template <typename T> __device__ __forceinline__ T quot(T a, T b) { return (T)(a / b); }
template <typename T> __device__ __forceinline__ T rem(T a, T b) { return (T)(a % b); }
template <typename T> __device__ __forceinline__ T idiv(T a, T b) {
  if constexpr (!is_s<T>) return (T)(a / b);
  else {
    if (a > 0 && b < 0) return sub<T>(quot<T>(sub<T>(a, 1), b), 1);
    if (a < 0 && b > 0) return sub<T>(quot<T>(add<T>(a, 1), b), 1);
    return quot<T>(a, b);
  }
}
template <typename T> __device__ __forceinline__ T mod(T a, T b) {
  if constexpr (!is_s<T>) return (T)(a % b);
  else {
    const T r = rem<T>(a, b);
    if ((a > 0 && b < 0) || (a < 0 && b > 0)) return r != 0 ? add<T>(r, b) : (T)0;
    return r;
  }
}
template <typename T> __device__ __forceinline__ T band(T a, T b) { return (T)(a & b); }
template <typename T> __device__ __forceinline__ T bor(T a, T b) { return (T)(a | b); }
template <typename T> __device__ __forceinline__ T bxor(T a, T b) { return (T)(a ^ b); }
template <typename T> __device__ __forceinline__ T bnot(T a) { return (T)(a ^ (T)-1); }
template <typename T> __device__ __forceinline__ int shamt(int64_t i) { return (int)((T)i) & (bits<T> - 1); }
template <typename T> __device__ __forceinline__ T shiftl(T a, int64_t i) {
  using U = typename uns<T>::type;
  return (T)(U)((U)a << shamt<T>(i));
}
template <typename T> __device__ __forceinline__ T shiftr(T a, int64_t i) {
  return (T)(a >> shamt<T>(i));  // arithmetic for signed, logical for unsigned
}
template <typename T> __device__ __forceinline__ T rotatel(T a, int64_t i) {
  using U = typename uns<T>::type;
  const int s = (int)((uint64_t)(U)(T)i % (uint64_t)bits<T>);
  const U u = (U)a;
  return s ? (T)(U)((U)(u << s) | (U)(u >> (bits<T> - s))) : a;
}
template <typename T> __device__ __forceinline__ T rotater(T a, int64_t i) {
  using U = typename uns<T>::type;
  const int s = (int)((uint64_t)(U)(T)i % (uint64_t)bits<T>);
  const U u = (U)a;
  return s ? (T)(U)((U)(u >> s) | (U)(u << (bits<T> - s))) : a;
}
template <typename T> __device__ __forceinline__ int64_t popcount(T a) {
  using U = typename uns<T>::type;
  if constexpr (sizeof(T) == 8) return (int64_t)__builtin_popcountll((uint64_t)(U)a);
  else return (int64_t)__builtin_popcount((uint32_t)(U)a);
}
template <typename T> __device__ __forceinline__ int64_t clz(T a) {
  using U = typename uns<T>::type;
  const U u = (U)a;
  if (u == 0) return bits<T>;
  if constexpr (sizeof(T) == 8) return (int64_t)__builtin_clzll((uint64_t)u);
  else return (int64_t)__builtin_clz((uint32_t)u) - (32 - bits<T>);
}
template <typename T> __device__ __forceinline__ int64_t ctz(T a) {
  using U = typename uns<T>::type;
  const U u = (U)a;
  if (u == 0) return bits<T>;
  if constexpr (sizeof(T) == 8) return (int64_t)__builtin_ctzll((uint64_t)u);
  else return (int64_t)__builtin_ctz((uint32_t)u);
}

// ---- floating
template <typename T> __device__ __forceinline__ T fdiv(T a, T b) { return a / b; }
template <typename T> __device__ __forceinline__ T recip(T a) { return (T)1 / a; }
#define AP_MATH1(name, f32, f64) \
  template <typename T> __device__ __forceinline__ T name(T a) { \
    if constexpr (sizeof(T) == 4) return f32(a); else return f64(a); }
AP_MATH1(sin_, sinf, sin)
AP_MATH1(cos_, cosf, cos)
AP_MATH1(tan_, tanf, tan)
AP_MATH1(sinh_, sinhf, sinh)
AP_MATH1(cosh_, coshf, cosh)
AP_MATH1(tanh_, tanhf, tanh)
AP_MATH1(asin_, asinf, asin)
AP_MATH1(acos_, acosf, acos)
AP_MATH1(atan_, atanf, atan)
AP_MATH1(asinh_, asinhf, asinh)
AP_MATH1(acosh_, acoshf, acosh)
AP_MATH1(atanh_, atanhf, atanh)
AP_MATH1(exp_, expf, exp)
AP_MATH1(sqrt_, sqrtf, sqrt)
AP_MATH1(log_, logf, log)
AP_MATH1(log2_, log2f, log2)
AP_MATH1(log10_, log10f, log10)
#undef AP_MATH1
template <typename T> __device__ __forceinline__ T atan2_(T a, T b) {
  if constexpr (sizeof(T) == 4) return atan2f(a, b); else return atan2(a, b);
}
template <typename T> __device__ __forceinline__ T fpow(T a, T b) {
  if constexpr (sizeof(T) == 4) return powf(a, b); else return pow(a, b);
}
template <typename T> __device__ __forceinline__ T logbase(T b, T x) { return log_<T>(x) / log_<T>(b); }
template <typename T> __device__ __forceinline__ bool isnan_(T a) { return a != a; }
template <typename T> __device__ __forceinline__ bool isinfinite(T a) {
  return abs_<T>(a) == (T)__builtin_huge_valf();
}
template <typename R, typename T> __device__ __forceinline__ R truncate_(T a) { return (R)a; }
template <typename R, typename T> __device__ __forceinline__ R round_(T a) {
  if constexpr (sizeof(T) == 4) return (R)roundf(a); else return (R)round(a);
}
template <typename R, typename T> __device__ __forceinline__ R floor_(T a) {
  if constexpr (sizeof(T) == 4) return (R)floorf(a); else return (R)floor(a);
}
template <typename R, typename T> __device__ __forceinline__ R ceiling_(T a) {
  if constexpr (sizeof(T) == 4) return (R)ceilf(a); else return (R)ceil(a);
}

// ---- relational
template <typename T> __device__ __forceinline__ bool lt(T a, T b) { return a < b; }
template <typename T> __device__ __forceinline__ bool gt(T a, T b) { return a > b; }
template <typename T> __device__ __forceinline__ bool lte(T a, T b) { return a <= b; }
template <typename T> __device__ __forceinline__ bool gte(T a, T b) { return a >= b; }
template <typename T> __device__ __forceinline__ bool eq(T a, T b) { return a == b; }
template <typename T> __device__ __forceinline__ bool neq(T a, T b) {
  if constexpr (is_f<T>) return a < b || a > b;  // ordered: false on NaN
  else return a != b;
}
template <typename T> __device__ __forceinline__ T max_(T a, T b) {
  if constexpr (is_f<T>) {
    if constexpr (sizeof(T) == 4) return fmaxf(a, b); else return fmax(a, b);
  } else {
    return a >= b ? a : b;
  }
}
template <typename T> __device__ __forceinline__ T min_(T a, T b) {
  if constexpr (is_f<T>) {
    if constexpr (sizeof(T) == 4) return fminf(a, b); else return fmin(a, b);
  } else {
    return a <= b ? a : b;
  }
}

// ---- conversions (C conversions are the LLVM sext/zext/trunc/sitofp/uitofp/fptrunc/fpext)
template <typename R, typename T> __device__ __forceinline__ R fromintegral(T a) { return (R)a; }
template <typename R, typename T> __device__ __forceinline__ R tofloating(T a) { return (R)a; }
template <typename R, typename T> __device__ __forceinline__ R bitcast(T a) {
  static_assert(sizeof(R) == sizeof(T), "bitcast between types of different sizes");
  R r;
  __builtin_memcpy(&r, &a, sizeof(R));
  return r;
}

// A free array read at a computed index; the index is clamped into the array
// (the reference leaves an out-of-range read undefined).
template <typename T> __device__ __forceinline__ T read(const void* p, int64_t n, int64_t i) {
  i = i < 0 ? 0 : (i >= n ? n - 1 : i);
  return n > 0 ? ((const T*)p)[i] : (T)0;
}

}  // namespace ap
