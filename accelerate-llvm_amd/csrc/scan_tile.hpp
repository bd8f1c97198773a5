// One tile of a single-pass decoupled look-back scan, shared by the plain
// scan kernel and the batched kernel.  See scan.hip for the element
// semantics of each scan variant.
#pragma once
#include "dispatch.hpp"
#include "lookback.hpp"
#include "vecmem.hpp"

namespace acc {

// The scan of one problem.  Tiles first..last of the shared status array
// belong to it; `base` passed to scan_tile is the element offset of the tile
// within the problem.
template <typename T> struct TileIO {
  const T* in;
  T* out;
  int64_t n;
  uint32_t first;      // status index of the problem's first tile
  uint32_t last;       // status index of the problem's last tile
  uint32_t exclusive;  // emit exclusive values
  uint32_t has_seed;
  T seed;
  T* total_out;  // nullable: where the total goes
  T* seed_out;   // nullable: where the seed goes (scanr z: out[n])
  uint64_t* status;
};

template <typename T, typename Op, bool MASKED> struct Lane;

// Element storage without masks: plain values.
template <typename T, typename Op> struct Lane<T, Op, false> {
  using E = T;
  static __device__ __forceinline__ E make(T v, bool) { return v; }
  static __device__ __forceinline__ E cmb(Op op, E a, E b) { return op(a, b); }
  static __device__ __forceinline__ Maybe<T> lift(E e) { return Maybe<T>{e, 1u}; }
  static __device__ __forceinline__ E wscan(E x, Op op) { return wave_inclusive_scan(x, op); }
};

// Element storage in the ragged tile: optional values.
template <typename T, typename Op> struct Lane<T, Op, true> {
  using E = Maybe<T>;
  static __device__ __forceinline__ E make(T v, bool ok) { return Maybe<T>{v, ok ? 1u : 0u}; }
  static __device__ __forceinline__ E cmb(Op op, E a, E b) { return mcombine(op, a, b); }
  static __device__ __forceinline__ Maybe<T> lift(E e) { return e; }
  static __device__ __forceinline__ E wscan(E x, Op op) { return wave_inclusive_scan_m(x, op); }
};

template <typename T> __device__ __forceinline__ Maybe<T> none() {
  Maybe<T> m;
  m.v = T();
  m.ok = 0;
  return m;
}

template <typename T, typename Op, int BLOCK, int K, bool REVERSE, bool ALIGNED, bool MASKED>
__device__ __forceinline__ void scan_tile(const TileIO<T>& a, Op op, uint32_t t, int64_t base,
                                          uint32_t epoch, Maybe<T>* s_wave, Maybe<T>* s_prefix) {
  constexpr int VEC = Vec<T>::N;
  constexpr int W = BLOCK / 64;
  constexpr int NW = Tagged<T>::NW;
  using L = Lane<T, Op, MASKED>;
  using E = typename L::E;
  const int lane = lane_id();
  const int w = threadIdx.x / 64;

  // physical vector index (within the tile) of logical chunk k of this lane
  auto vidx = [&](int k) -> int64_t {
    return REVERSE ? ((int64_t)((W - 1 - w) * K + (K - 1 - k)) * 64 + (63 - lane))
                   : ((int64_t)(w * K + k) * 64 + lane);
  };

  // ---- load, thread-level inclusive scan of each vector
  E x[K][VEC];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int64_t e0 = base + vidx(k) * VEC;
    T v[VEC];
    bool ok[VEC];
    if (ALIGNED && !MASKED) {
      Vec<T> r = load16_stream(a.in + e0);
#pragma unroll
      for (int j = 0; j < VEC; ++j) { v[j] = r.e[j]; ok[j] = true; }
    } else {
#pragma unroll
      for (int j = 0; j < VEC; ++j) {
        const int64_t i = e0 + j;
        ok[j] = !MASKED || i < a.n;
        v[j] = ok[j] ? a.in[i] : T();
      }
    }
#pragma unroll
    for (int j = 0; j < VEC; ++j) {
      const int jj = REVERSE ? VEC - 1 - j : j;
      x[k][j] = L::make(v[jj], ok[jj]);
    }
#pragma unroll
    for (int j = 1; j < VEC; ++j) x[k][j] = L::cmb(op, x[k][j - 1], x[k][j]);
  }

  // ---- wave scan of lane totals, chunk by chunk
  Maybe<T> pre[K];
  Maybe<T> run = none<T>();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    E inc = L::wscan(x[k][VEC - 1], op);
    E up = shift_up1(inc);
    Maybe<T> exl = lane > 0 ? L::lift(up) : none<T>();
    pre[k] = mcombine(op, run, exl);
    run = mcombine(op, run, L::lift(read_lane(inc, 63)));
  }
  if (lane == 0) s_wave[w] = run;
  __syncthreads();
  Maybe<T> wp = none<T>(), agg = none<T>();
#pragma unroll
  for (int i = 0; i < W; ++i) {
    const Maybe<T> v = s_wave[i];
    if (i < w) wp = mcombine(op, wp, v);
    agg = mcombine(op, agg, v);
  }

  // ---- tile prefix: seed for the first tile, look-back otherwise
  if (w == 0) {
    Maybe<T> X = none<T>();
    uint64_t* slot = a.status + (size_t)t * NW;
    if (t == a.first) {
      if (a.has_seed) X = Maybe<T>{a.seed, 1u};
    } else {
      if (lane == 0) Tagged<T>::publish(slot, agg.v, make_tag(epoch, ST_AGGREGATE));
      X = Maybe<T>{wave_lookback<T>(a.status, t, epoch, op, a.first), 1u};
    }
    const Maybe<T> incl = mcombine(op, X, agg);
    if (lane == 0) {
      if (t != a.last) Tagged<T>::publish(slot, incl.v, make_tag(epoch, ST_PREFIX));
      *s_prefix = X;
      if (t == a.last && a.total_out) *a.total_out = incl.v;
      if (t == a.first && a.seed_out) *a.seed_out = a.seed;
    }
  }
  __syncthreads();
  const Maybe<T> X = *s_prefix;
  const Maybe<T> XW = mcombine(op, X, wp);

  // ---- outputs
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const Maybe<T> P = mcombine(op, XW, pre[k]);
    T y[VEC];
    bool ok[VEC];
#pragma unroll
    for (int j = 0; j < VEC; ++j) {
      const Maybe<T> xi = L::lift(x[k][j]);
      if (a.exclusive) {
        // exclusive value: everything before this element (always defined: a seed is given)
        const Maybe<T> prev = j == 0 ? P : mcombine(op, P, L::lift(x[k][j - 1]));
        y[j] = prev.v;
      } else {
        y[j] = mcombine(op, P, xi).v;
      }
      ok[j] = true;
    }
    if constexpr (MASKED) {
      // validity of each element in its own (physical) position
#pragma unroll
      for (int j = 0; j < VEC; ++j) {
        const int64_t i = base + vidx(k) * VEC + (REVERSE ? VEC - 1 - j : j);
        ok[j] = i < a.n;
      }
    }
    const int64_t e0 = base + vidx(k) * VEC;
    if (ALIGNED && !MASKED) {
      Vec<T> r;
#pragma unroll
      for (int j = 0; j < VEC; ++j) r.e[j] = y[REVERSE ? VEC - 1 - j : j];
      store16(a.out + e0, r);
    } else {
#pragma unroll
      for (int j = 0; j < VEC; ++j) {
        const int jl = REVERSE ? VEC - 1 - j : j;  // logical slot of physical element j
        if (ok[jl]) a.out[e0 + j] = y[jl];
      }
    }
  }
}

}  // namespace acc
