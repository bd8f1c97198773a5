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

template <typename T> __device__ __forceinline__ Maybe<T> none() {
  Maybe<T> m;
  m.v = T();
  m.ok = 0;
  return m;
}

// Scan of one tile of BLOCK*K*VEC elements.  Each lane holds K dwordx4
// vectors; vector k of wave w is vector (w*K + k)*64 + lane of the tile, so
// every load instruction of a wave reads 1 KiB of consecutive bytes.  The
// ragged tile (MASKED) is padded with the combine's identity, which leaves
// every valid prefix, the aggregate and the total exactly as without padding;
// only its one partial vector uses element accesses.
template <typename T, typename Op, int BLOCK, int K, bool REVERSE, bool ALIGNED, bool MASKED>
__device__ __forceinline__ void scan_tile(const TileIO<T>& a, Op op, uint32_t t, int64_t base,
                                          uint32_t epoch, T* s_wave, T* s_prefix) {
  constexpr int VEC = Vec<T>::N;
  constexpr int W = BLOCK / 64;
  constexpr int NW = Tagged<T>::NW;
  const T id = identity<T>(op);
  const int lane = lane_id();
  const int w = threadIdx.x / 64;

  // physical vector index (within the tile) of logical chunk k of this lane
  auto vidx = [&](int k) -> int64_t {
    return REVERSE ? ((int64_t)((W - 1 - w) * K + (K - 1 - k)) * 64 + (63 - lane))
                   : ((int64_t)(w * K + k) * 64 + lane);
  };

  // ---- ragged tile: bit k*VEC + j says physical element j of vector k is
  // inside the array (one 64-bit mask per lane instead of a lane mask per
  // element)
  static_assert(K * VEC <= 64, "validity mask");
  constexpr uint32_t VMASK = (1u << VEC) - 1u;
  uint64_t valid = ~0ull;
  if constexpr (MASKED) {
    valid = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int64_t r = a.n - (base + vidx(k) * VEC);
      const int nv = r <= 0 ? 0 : (r >= VEC ? VEC : (int)r);
      valid |= (uint64_t)((1u << nv) - 1u) << (k * VEC);
    }
  }
  auto ok = [&](int k, int j) -> bool { return !MASKED || ((valid >> (k * VEC + j)) & 1u); };

  // ---- load every vector first (so all of the tile's loads are in flight
  // together), then the thread-level inclusive scan of each vector
  // (traversal order)
  T x[K][VEC];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int64_t e0 = base + vidx(k) * VEC;
    if (ALIGNED) {
      // In the ragged tile a vector past the end re-reads the vector holding
      // the last element (an aligned 16-byte block with a valid element never
      // crosses a page) and every element past the end becomes the identity.
      const int64_t ea = MASKED ? min(e0, ((a.n - 1) / VEC) * VEC) : e0;
      const Vec<T> r = load16_stream(a.in + ea);
#pragma unroll
      for (int j = 0; j < VEC; ++j) x[k][j] = r.e[j];
    } else {
#pragma unroll
      for (int j = 0; j < VEC; ++j) x[k][j] = ok(k, j) ? gload(a.in + e0 + j) : id;
    }
  }
  // keep the loads above ahead of the arithmetic below
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    T v[VEC];
#pragma unroll
    for (int j = 0; j < VEC; ++j) v[j] = ALIGNED && !ok(k, j) ? id : x[k][j];
#pragma unroll
    for (int j = 0; j < VEC; ++j) x[k][j] = v[REVERSE ? VEC - 1 - j : j];
#pragma unroll
    for (int j = 1; j < VEC; ++j) x[k][j] = op(x[k][j - 1], x[k][j]);
  }

  // ---- wave scan of lane totals, chunk by chunk
  T pre[K];
  T run = id;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const T inc = wave_inclusive_scan(x[k][VEC - 1], op);
    const T up = shift_up1(inc);
    pre[k] = op(run, lane > 0 ? up : id);
    run = op(run, read_lane(inc, 63));
  }
  if (lane == 0) s_wave[w] = run;
  __syncthreads();
  T wp = id, agg = id;
#pragma unroll
  for (int i = 0; i < W; ++i) {
    const T v = s_wave[i];
    if (i < w) wp = op(wp, v);
    agg = op(agg, v);
  }

  // ---- tile prefix: seed for the first tile, look-back otherwise
  if (w == 0) {
    T X = id;
    uint64_t* slot = a.status + (size_t)t * NW;
    if (t == a.first) {
      if (a.has_seed) X = a.seed;
    } else {
      if (lane == 0) Tagged<T>::publish(slot, agg, make_tag(epoch, ST_AGGREGATE));
      // a first window of 2 x 64 predecessors, then windows of 6 x 64: the
      // published-prefix front of a long scan advances up to 384 tiles per
      // round trip once the nearest ones are found not to carry it (measured
      // on 2^28 Int32 / Float32 and 2^27 Int64 / Float64 scanl1, us: first
      // and later windows of 3 x 64 406 / 404 / 459 / 460; 2 then 6 394 /
      // 399 / 422 / 413; 2 then 8 394 / 401 / 425 / 423; 2 then 3 or 4 410
      // (Int32); 1 then 3..8 453-460; 4 then 4 425; 6 then 6 491; 8 then 8
      // 500: the larger windows cost registers and occupancy; at the same
      // registers a wider first window is still slower: 3 then 6 402 / 402 /
      // 473 / 467, 3 then 8 403 / 400 / 469 / 463, 4 then 6 or 8 422 / 420 /
      // 497 / 494)
#ifndef ACC_LB_C0
#define ACC_LB_C0 2
#define ACC_LB_C1 6
#endif
      X = wave_lookback<T, Op, ACC_LB_C0, ACC_LB_C1>(a.status, t, epoch, op, a.first);
    }
    const T incl = op(X, agg);
    if (lane == 0) {
      if (t != a.last) Tagged<T>::publish(slot, incl, make_tag(epoch, ST_PREFIX));
      *s_prefix = X;
      if (t == a.last && a.total_out) gstore(a.total_out, incl);
      if (t == a.first && a.seed_out) gstore(a.seed_out, a.seed);
    }
  }
  __syncthreads();
  const T XW = op(*s_prefix, wp);

  // ---- outputs
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const T P = op(XW, pre[k]);
    T y[VEC];
#pragma unroll
    for (int j = 0; j < VEC; ++j)
      y[j] = a.exclusive ? (j == 0 ? P : op(P, x[k][j - 1])) : op(P, x[k][j]);
    const int64_t e0 = base + vidx(k) * VEC;
    const uint32_t vk = MASKED ? (uint32_t)(valid >> (k * VEC)) & VMASK : VMASK;
    if (ALIGNED && vk == VMASK) {
      Vec<T> r;
#pragma unroll
      for (int j = 0; j < VEC; ++j) r.e[j] = y[REVERSE ? VEC - 1 - j : j];
      store16_stream(a.out + e0, r);  // written once: non-temporal
    } else if (vk) {
#pragma unroll
      for (int j = 0; j < VEC; ++j)
        if (vk >> j & 1u) gstore(a.out + e0 + j, y[REVERSE ? VEC - 1 - j : j]);
    }
  }
}

}  // namespace acc
