// Wave64 primitives for gfx950: DPP lane moves, in-order wave scans and
// reductions, lane broadcast.  All functions must be called by the whole
// wavefront in converged control flow (DPP reads from other lanes).
//
// The scans preserve the left-to-right order of lanes, so they are exact for
// associative but non-commutative combine functions, which is the contract of
// Accelerate's fold/scan (accelerate-llvm-ptx/.../PTX/CodeGen/Fold.hs:425-431
// reduces in order for the same reason).
#pragma once
#ifndef __HIPCC_RTC__  // hiprtc provides the HIP runtime itself
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>
#include "elt.hpp"

namespace acc {

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// DPP control codes (GCN3+ ISA; gfx9 keeps row broadcasts).
enum : int {
  DPP_ROW_SHR1 = 0x111,
  DPP_ROW_SHR2 = 0x112,
  DPP_ROW_SHR4 = 0x114,
  DPP_ROW_SHR8 = 0x118,
  DPP_ROW_BCAST15 = 0x142,
  DPP_ROW_BCAST31 = 0x143,
  DPP_WAVE_SHR1 = 0x138,
};

template <typename T> struct Words {
  static constexpr int N = (int)((sizeof(T) + 3) / 4);
  uint32_t w[N];
};

template <typename T> __device__ __forceinline__ Words<T> to_words(const T& v) {
  Words<T> r;
#pragma unroll
  for (int i = 0; i < Words<T>::N; ++i) r.w[i] = 0;
  __builtin_memcpy(&r, &v, sizeof(T));
  return r;
}
template <typename T> __device__ __forceinline__ T from_words(const Words<T>& r) {
  T v;
  __builtin_memcpy(&v, &r, sizeof(T));
  return v;
}

template <int CTRL, int ROW_MASK = 0xF, int BANK_MASK = 0xF, typename T>
__device__ __forceinline__ T dpp_move(const T& v) {
  Words<T> a = to_words(v), r;
#pragma unroll
  for (int i = 0; i < Words<T>::N; ++i)
    r.w[i] = (uint32_t)__builtin_amdgcn_update_dpp((int)a.w[i], (int)a.w[i], CTRL, ROW_MASK,
                                                   BANK_MASK, false);
  return from_words<T>(r);
}

template <typename T> __device__ __forceinline__ T read_lane(const T& v, int lane) {
  Words<T> a = to_words(v), r;
#pragma unroll
  for (int i = 0; i < Words<T>::N; ++i) r.w[i] = (uint32_t)__builtin_amdgcn_readlane((int)a.w[i], lane);
  return from_words<T>(r);
}

// value of lane (l - 1); lane 0 receives its own value (callers mask it)
template <typename T> __device__ __forceinline__ T shift_up1(const T& v) {
  Words<T> a = to_words(v), r;
  const int src = (lane_id() == 0 ? 0 : lane_id() - 1) << 2;
#pragma unroll
  for (int i = 0; i < Words<T>::N; ++i) r.w[i] = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)a.w[i]);
  return from_words<T>(r);
}

// Inclusive scan over the 64 lanes, lane 0 first.
template <typename T, typename Op>
__device__ __forceinline__ T wave_inclusive_scan(T x, Op op) {
  const int lane = lane_id();
  const int r = lane & 15;
  { T y = dpp_move<DPP_ROW_SHR1>(x); if (r >= 1) x = op(y, x); }
  { T y = dpp_move<DPP_ROW_SHR2>(x); if (r >= 2) x = op(y, x); }
  { T y = dpp_move<DPP_ROW_SHR4>(x); if (r >= 4) x = op(y, x); }
  { T y = dpp_move<DPP_ROW_SHR8>(x); if (r >= 8) x = op(y, x); }
  { T y = dpp_move<DPP_ROW_BCAST15, 0xA>(x); if (lane & 16) x = op(y, x); }
  { T y = dpp_move<DPP_ROW_BCAST31, 0xC>(x); if (lane & 32) x = op(y, x); }
  return x;
}

// A DPP lane move whose result is used only in lanes with an in-range
// source (a row shift within its row, a row broadcast into its row mask):
// the other lanes' values are unspecified, so no copy of the source is kept.
template <int CTRL, int ROW_MASK = 0xF, int BANK_MASK = 0xF, typename T>
__device__ __forceinline__ T dpp_fetch(const T& v) {
  Words<T> a = to_words(v), r;
#pragma unroll
  for (int i = 0; i < Words<T>::N; ++i)
    r.w[i] = (uint32_t)__builtin_amdgcn_mov_dpp((int)a.w[i], CTRL, ROW_MASK, BANK_MASK, true);
  return from_words<T>(r);
}

// Value of lane (l - 1) by a wave-wide DPP shift; lane 0 receives fill.
template <typename T> __device__ __forceinline__ T wave_shr1(const T& v, const T& fill) {
  Words<T> a = to_words(v), f = to_words(fill), r;
#pragma unroll
  for (int i = 0; i < Words<T>::N; ++i)
    r.w[i] = (uint32_t)__builtin_amdgcn_update_dpp((int)f.w[i], (int)a.w[i], DPP_WAVE_SHR1, 0xF, 0xF, false);
  return from_words<T>(r);
}

// Segmented inclusive scan over the 64 lanes, lane 0 first.  A segment
// starts at each lane whose bit is set in heads (a wave-uniform ballot);
// lane i combines lanes [max(L, 0), i] where L is the last head at or below
// i.  Each lane first finds its distance s to L (lane + 1 when there is
// none); the lane-shift step over distance d then applies exactly when
// d <= s, so the steps move only values, never flags.  below: a head lies
// strictly below this lane.
template <typename T, typename Op>
__device__ __forceinline__ T wave_seg_inclusive_scan(T x, uint64_t heads, Op op, bool* below) {
  const int lane = lane_id();
  const int r = lane & 15;
  const uint64_t lt = heads & ((1ull << lane) - 1ull);  // heads strictly below
  const bool me = (heads >> lane) & 1ull;
  *below = lt != 0;
  const int s = me ? 0 : (lt ? lane - (63 - __builtin_clzll(lt)) : lane + 1);
  const int rs = min(r, s);
  { T y = dpp_fetch<DPP_ROW_SHR1>(x); if (rs >= 1) x = op(y, x); }
  { T y = dpp_fetch<DPP_ROW_SHR2>(x); if (rs >= 2) x = op(y, x); }
  { T y = dpp_fetch<DPP_ROW_SHR4>(x); if (rs >= 4) x = op(y, x); }
  { T y = dpp_fetch<DPP_ROW_SHR8>(x); if (rs >= 8) x = op(y, x); }
  { T y = dpp_fetch<DPP_ROW_BCAST15, 0xA>(x); if ((lane & 16) && s > r) x = op(y, x); }
  { T y = dpp_fetch<DPP_ROW_BCAST31, 0xC>(x); if ((lane & 32) && s > lane - 32) x = op(y, x); }
  return x;
}

// Inclusive scan of optional values (absent lanes are skipped, order kept).
template <typename T, typename Op>
__device__ __forceinline__ Maybe<T> wave_inclusive_scan_m(Maybe<T> x, Op op) {
  auto mop = [op](Maybe<T> a, Maybe<T> b) { return mcombine(op, a, b); };
  return wave_inclusive_scan(x, mop);
}

// Reduction of all 64 lanes in lane order; the result is uniform.
template <typename T, typename Op>
__device__ __forceinline__ T wave_reduce(T x, Op op) {
  return read_lane(wave_inclusive_scan(x, op), 63);
}

template <typename T, typename Op>
__device__ __forceinline__ Maybe<T> wave_reduce_m(Maybe<T> x, Op op) {
  return read_lane(wave_inclusive_scan_m(x, op), 63);
}

}  // namespace acc
