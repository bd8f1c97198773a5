// Skeleton templates for kernels compiled at run time around the program's
// own scalar functions (the reference compiles every skeleton node with its
// functions inlined: PTX/Compile.hs:84-165, the templates PTX/CodeGen/*.hs).
//
// The generated source supplies, per node:
//   E     the element: a struct of 1..4 scalar fields (struct of arrays in memory)
//   Fn    the combine function, E operator()(const E& a, const E& b); `a` is the
//         left operand in traversal order.  It need only be associative: every
//         reduction here combines in element order (no commutativity, no
//         identity element is assumed; PTX/CodeGen/Fold.hs:421-423)
//   Seed  has (constexpr bool) and E operator()() - the seed, combined on the left
//   Ld    E operator()(int64_t i) and four(i, E*) - element i of the input, from
//         its field arrays or computed by a delayed producer function (the
//         reference's IRDelayed fusion, LLVM/CodeGen/Sugar.hs:70-76)
//   St    operator()(int64_t i, const E&) and four(i, const E*) - output stores
//
// Cross-workgroup publication reuses the tagged status words and decoupled
// look-back of the ahead-of-time skeletons (lookback.hpp), which work for any
// trivially copyable element and any associative combine.
#pragma once
#include "args.hpp"
#include "../lookback.hpp"

namespace accjit {

using acc::lane_id;
using acc::make_tag;
using acc::Maybe;
using acc::SyncCtl;
using acc::Tagged;

constexpr int BLOCK = 256;
constexpr int WAVES = BLOCK / 64;

template <typename E> __device__ __forceinline__ Maybe<E> none() {
  Maybe<E> m;
  m.v = E{};
  m.ok = 0;
  return m;
}
template <typename E> __device__ __forceinline__ Maybe<E> just(const E& v) {
  Maybe<E> m;
  m.v = v;
  m.ok = 1;
  return m;
}
template <typename E, typename Fn>
__device__ __forceinline__ Maybe<E> mc(const Fn& fn, const Maybe<E>& a, const Maybe<E>& b) {
  return acc::mcombine(fn, a, b);
}
template <typename E, typename Fn>
__device__ __forceinline__ E lapp(const Fn& fn, const Maybe<E>& p, const E& x) {
  return p.ok ? fn(p.v, x) : x;
}

// ------------------------------------------------------------------ fold
// fold / fold1 of a vector in one launch.  Workgroup b owns elements
// [b*chunk, (b+1)*chunk); wave w of it the w-th contiguous quarter, which it
// reduces 256 elements at a time (4 consecutive elements per lane, an in-lane
// fold, then an in-order wave reduction).  Workgroup partials are published
// as tagged words; the last workgroup to finish combines them in workgroup
// order and puts the seed on the left.
template <typename E, typename Fn, typename Seed, typename Ld, typename St>
__device__ void fold_all(const Args& a, const Fn& fn, const Seed& seed, const Ld& ld, const St& st) {
  constexpr int U = 2;
  constexpr int NW = Tagged<Maybe<E>>::NW;
  __shared__ uint32_t s_epoch, s_last;
  __shared__ Maybe<E> s_w[WAVES];
  SyncCtl* ctl = (SyncCtl*)a.ctl;
  const uint32_t epoch = acc::block_epoch(ctl, &s_epoch);
  const uint32_t tag = make_tag(epoch, acc::ST_PREFIX);
  const int lane = lane_id(), w = threadIdx.x / 64;
  const int64_t n = a.inner;
  const int64_t q = a.chunk / WAVES;  // a multiple of 256
  const int64_t w0 = (int64_t)blockIdx.x * a.chunk + w * q;
  const int64_t w1 = w0 + q < n ? w0 + q : n;
  Maybe<E> acc = none<E>();
  int64_t i = w0;
  for (; i + 256 * U <= w1; i += 256 * U) {
    E x[U][4];
#pragma unroll
    for (int u = 0; u < U; ++u) ld.four(i + u * 256 + lane * 4, x[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      E r = x[u][0];
#pragma unroll
      for (int j = 1; j < 4; ++j) r = fn(r, x[u][j]);
      acc = mc(fn, acc, just(acc::wave_reduce(r, fn)));
    }
  }
  for (; i < w1; i += 256) {
    const int64_t e = i + lane * 4;
    Maybe<E> r = none<E>();
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (e + j < w1) r = mc(fn, r, just(ld(e + j)));
    acc = mc(fn, acc, acc::wave_reduce_m(r, fn));
  }
  if (lane == 0) s_w[w] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    Maybe<E> b = s_w[0];
#pragma unroll
    for (int k = 1; k < WAVES; ++k) b = mc(fn, b, s_w[k]);
    Tagged<Maybe<E>>::publish(a.status + (size_t)blockIdx.x * NW, b, tag);
    s_last = (acc::add_relaxed(&ctl->done, 1u) == gridDim.x - 1) ? 1u : 0u;
  }
  __syncthreads();
  if (!s_last) return;
  // the last workgroup: thread t combines a contiguous run of partials, then
  // the runs are combined in thread order
  const uint32_t nb = gridDim.x;
  const uint32_t per = (nb + BLOCK - 1) / BLOCK;
  Maybe<E> g = none<E>();
  for (uint32_t k = 0; k < per; ++k) {
    const uint32_t b = threadIdx.x * per + k;
    if (b < nb) {
      Maybe<E> p;
      while (Tagged<Maybe<E>>::peek(a.status + (size_t)b * NW, p) != tag) __builtin_amdgcn_s_sleep(1);
      g = mc(fn, g, p);
    }
  }
  g = acc::wave_reduce_m(g, fn);
  __syncthreads();
  if (lane == 0) s_w[w] = g;
  __syncthreads();
  if (threadIdx.x == 0) {
    Maybe<E> r = s_w[0];
#pragma unroll
    for (int k = 1; k < WAVES; ++k) r = mc(fn, r, s_w[k]);
    if constexpr (Seed::has) r = mc(fn, just(seed()), r);
    st(0, r.v);
    acc::st_relaxed(&ctl->done, 0u);
    acc::advance_epoch(ctl, epoch);
  }
}

// ------------------------------------------------------- segmented folds
// Segment s of row r is elements [off(s), off(s+1)) of the row.  A wave
// (WAVE_PER_SEG) or a whole workgroup reduces one segment in order; the
// reference makes the same choice between a warp and a block per segment by
// the mean segment length (PTX/Execute.hs:361-366).
struct OffUniform {
  int64_t inner;
  __device__ __forceinline__ int64_t operator()(int64_t s) const { return s * inner; }
};
template <typename I> struct OffArray {
  const I* p;
  __device__ __forceinline__ int64_t operator()(int64_t s) const { return (int64_t)p[s]; }
};

template <typename E, typename Fn, typename Ld>
__device__ __forceinline__ Maybe<E> wave_fold_range(const Fn& fn, const Ld& ld, int64_t lo, int64_t hi) {
  const int lane = lane_id();
  Maybe<E> acc = none<E>();
  for (int64_t i = lo; i < hi; i += 256) {
    const int64_t e = i + lane * 4;
    Maybe<E> r = none<E>();
    if (i + 256 <= hi) {
      E x[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) x[j] = ld(e + j);
      E t = x[0];
#pragma unroll
      for (int j = 1; j < 4; ++j) t = fn(t, x[j]);
      r = just(t);
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (e + j < hi) r = mc(fn, r, just(ld(e + j)));
    }
    acc = mc(fn, acc, acc::wave_reduce_m(r, fn));
  }
  return acc;
}

template <typename E, typename Fn, typename Seed, typename Ld, typename St, typename Off, bool WAVE_PER_SEG>
__device__ void fold_seg(const Args& a, const Fn& fn, const Seed& seed, const Ld& ld, const St& st, const Off& off) {
  __shared__ Maybe<E> s_w[WAVES];
  const int lane = lane_id(), w = threadIdx.x / 64;
  const int64_t total = a.outer * a.nseg;
  if (WAVE_PER_SEG) {
    const int64_t waves = (int64_t)gridDim.x * WAVES;
    for (int64_t g = (int64_t)blockIdx.x * WAVES + w; g < total; g += waves) {
      const int64_t r = g / a.nseg, s = g - r * a.nseg;
      const int64_t base = r * a.inner;
      const int64_t lo = off(s), hi = off(s + 1);
      Maybe<E> acc = wave_fold_range<E>(fn, ld, base + lo, base + hi);
      if constexpr (Seed::has) acc = mc(fn, just(seed()), acc);
      if (lane == 0 && acc.ok) st(g, acc.v);
    }
  } else {
    for (int64_t g = blockIdx.x; g < total; g += gridDim.x) {
      const int64_t r = g / a.nseg, s = g - r * a.nseg;
      const int64_t base = r * a.inner;
      const int64_t lo = off(s), hi = off(s + 1);
      // wave w: the w-th quarter of the segment (a multiple of 256 elements)
      const int64_t q = ((hi - lo + WAVES - 1) / WAVES + 255) / 256 * 256;
      const int64_t a0 = lo + w * q, a1 = a0 + q < hi ? a0 + q : hi;
      const Maybe<E> acc = wave_fold_range<E>(fn, ld, base + a0, base + (a1 > a0 ? a1 : a0));
      if (lane == 0) s_w[w] = acc;
      __syncthreads();
      if (threadIdx.x == 0) {
        Maybe<E> b = s_w[0];
#pragma unroll
        for (int k = 1; k < WAVES; ++k) b = mc(fn, b, s_w[k]);
        if constexpr (Seed::has) b = mc(fn, just(seed()), b);
        if (b.ok) st(g, b.v);
      }
      __syncthreads();
    }
  }
}

// ------------------------------------------------------------------ scan
// Single-pass decoupled look-back over tiles of BLOCK*K*4 elements taken by
// ticket (as the ahead-of-time scan: each workgroup takes its next ticket when
// it is ready to start the tile).  In traversal order, lane l of wave w holds
// positions ((w*K + k)*64 + l)*4 + j of the tile for chunk k; REV maps
// traversal position p to array index n-1-p and flips the combine
// (`x (+) acc`, NAT/CodeGen/Scan.hs:204-206).  Without an identity, a missing
// prefix travels as an absent Maybe; the ragged last tile reads clamped
// positions past the end, whose values only flow into positions that are
// never stored (the scan runs forward), and its total is taken at the last
// valid position.
//   exclusive == 0 : out[i] = inclusive value (scanl1/scanr1, scanr z)
//   exclusive == 1 : out[i] = exclusive value (scanl z, scan')
//   has_total      : the total goes to tot(...) (scanl z: out[n]; scan': sums)
template <typename E, typename Fn, typename Seed, typename Ld, typename St, typename Tot, int K, bool REV>
__device__ __forceinline__ void scan_tile(const Args& a, const Fn& fn0, const Seed& seed, const Ld& ld, const St& st,
                                          const Tot& tot, uint32_t t, uint32_t epoch, bool exclusive, E* s_wave,
                                          Maybe<E>* s_prefix) {
  constexpr int NW = Tagged<E>::NW;
  const acc::Flip<Fn> flip{fn0};
  auto fn = [&](const E& x, const E& y) -> E { if constexpr (REV) return flip(x, y); else return fn0(x, y); };
  const int lane = lane_id(), w = threadIdx.x / 64;
  const int64_t n = a.inner;
  constexpr int64_t TILE = (int64_t)BLOCK * K * 4;
  const int64_t base = (int64_t)t * TILE;
  const bool masked = base + TILE > n;
  E x[K][4];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int64_t p0 = base + ((int64_t)(w * K + k) * 64 + lane) * 4;
    if (!REV && !masked) {
      ld.four(p0, x[k]);
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        int64_t p = p0 + j;
        p = p < n ? p : n - 1;
        x[k][j] = ld(REV ? n - 1 - p : p);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int j = 1; j < 4; ++j) x[k][j] = fn(x[k][j - 1], x[k][j]);
  Maybe<E> pre[K];
  Maybe<E> run = none<E>();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const E inc = acc::wave_inclusive_scan(x[k][3], fn);
    const E up = acc::shift_up1(inc);
    pre[k] = mc(fn, run, lane > 0 ? just(up) : none<E>());
    run = mc(fn, run, just(acc::read_lane(inc, 63)));
  }
  if (lane == 0) s_wave[w] = run.v;
  __syncthreads();
  Maybe<E> wp = none<E>();
  E agg = s_wave[0];
#pragma unroll
  for (int i = 0; i < WAVES; ++i) {
    if (i < w) wp = mc(fn, wp, just(s_wave[i]));
    if (i > 0) agg = fn(agg, s_wave[i]);
  }
  if (w == 0) {
    Maybe<E> X = none<E>();
    uint64_t* slot = a.status + (size_t)t * NW;
    if (t == 0) {
      if constexpr (Seed::has) {
        X = just(seed());
        if (a.has_total == 2 && lane == 0) tot(seed());  // scanr z: out[n] = z
      }
    } else {
      if (lane == 0) Tagged<E>::publish(slot, agg, make_tag(epoch, acc::ST_AGGREGATE));
      X = just(acc::wave_lookback<E>(a.status, t, epoch, fn, 0u));
    }
    if (lane == 0) {
      if (t + 1 != a.ntiles) Tagged<E>::publish(slot, lapp(fn, X, agg), make_tag(epoch, acc::ST_PREFIX));
      *s_prefix = X;
    }
  }
  __syncthreads();
  const Maybe<E> XW = mc(fn, *s_prefix, wp);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const Maybe<E> P = mc(fn, XW, pre[k]);
    E y[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) y[j] = exclusive ? (j == 0 ? P.v : lapp(fn, P, x[k][j - 1])) : lapp(fn, P, x[k][j]);
    const int64_t p0 = base + ((int64_t)(w * K + k) * 64 + lane) * 4;
    if (!REV && !masked) {
      st.four(p0, y);
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (p0 + j < n) st(REV ? n - 1 - (p0 + j) : p0 + j, y[j]);
    }
    if (a.has_total == 1 && p0 <= n - 1 && n - 1 < p0 + 4) {
      const int j = (int)(n - 1 - p0);
#pragma unroll
      for (int jj = 0; jj < 4; ++jj)
        if (jj == j) tot(lapp(fn, P, x[k][jj]));
    }
  }
}

template <typename E, typename Fn, typename Seed, typename Ld, typename St, typename Tot, int K, bool REV>
__device__ void scan_all(const Args& a, const Fn& fn, const Seed& seed, const Ld& ld, const St& st, const Tot& tot,
                         bool exclusive) {
  __shared__ uint64_t s_tick;
  __shared__ E s_wave[WAVES];
  __shared__ Maybe<E> s_prefix;
  SyncCtl* ctl = (SyncCtl*)a.ctl;
  if (threadIdx.x == 0) s_tick = acc::take_ticket(ctl);
  __syncthreads();
  const uint64_t tk0 = s_tick;
  const uint32_t epoch = (uint32_t)(tk0 >> 32);
  uint32_t t = (uint32_t)tk0;
  while (t < a.ntiles) {
    scan_tile<E, Fn, Seed, Ld, St, Tot, K, REV>(a, fn, seed, ld, st, tot, t, epoch, exclusive, s_wave, &s_prefix);
    __syncthreads();
    if (threadIdx.x == 0) s_tick = acc::take_ticket(ctl);
    __syncthreads();
    t = (uint32_t)s_tick;
  }
  acc::block_retire(ctl, gridDim.x, epoch);
}

// Rows of a rank >= 2 scan: one workgroup per row (grid-stride over rows),
// 1024 elements at a time with the carry held in LDS (the reference's scanDim
// also gives each row one block, PTX/CodeGen/Scan.hs:746-939).  The row's
// outputs start at r*out_row (out_row = inner, or inner + 1 with a total or
// the seed closing the row); Tot receives (row, total).
template <typename E, typename Fn, typename Seed, typename Ld, typename St, typename Tot, bool REV>
__device__ void scan_rows(const Args& a, const Fn& fn0, const Seed& seed, const Ld& ld, const St& st, const Tot& tot,
                          bool exclusive, int64_t out_row) {
  __shared__ E s_wave[WAVES];
  __shared__ Maybe<E> s_carry;
  const acc::Flip<Fn> flip{fn0};
  auto fn = [&](const E& x, const E& y) -> E { if constexpr (REV) return flip(x, y); else return fn0(x, y); };
  const int lane = lane_id(), w = threadIdx.x / 64;
  const int64_t n = a.inner;
  for (int64_t r = blockIdx.x; r < a.outer; r += gridDim.x) {
    if (threadIdx.x == 0) {
      Maybe<E> c = none<E>();
      if constexpr (Seed::has) {
        c = just(seed());
        if (a.has_total == 2) tot(r, seed());  // scanr z: the row closes with z
      }
      s_carry = c;
    }
    __syncthreads();
    for (int64_t b = 0; b < n; b += 1024) {
      const int64_t p0 = b + (int64_t)threadIdx.x * 4;
      E x[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        int64_t p = p0 + j;
        p = p < n ? p : n - 1;
        x[j] = ld(r * n + (REV ? n - 1 - p : p));
      }
#pragma unroll
      for (int j = 1; j < 4; ++j) x[j] = fn(x[j - 1], x[j]);
      const E inc = acc::wave_inclusive_scan(x[3], fn);
      const E up = acc::shift_up1(inc);
      if (lane == 63) s_wave[w] = inc;
      __syncthreads();
      Maybe<E> P = s_carry;
#pragma unroll
      for (int i = 0; i < WAVES; ++i)
        if (i < w) P = mc(fn, P, just(s_wave[i]));
      P = mc(fn, P, lane > 0 ? just(up) : none<E>());
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const E y = exclusive ? (j == 0 ? P.v : lapp(fn, P, x[j - 1])) : lapp(fn, P, x[j]);
        if (p0 + j < n) st(r * out_row + (REV ? n - 1 - (p0 + j) : p0 + j), y);
        if (p0 + j == n - 1 && a.has_total == 1) tot(r, lapp(fn, P, x[j]));
      }
      __syncthreads();
      if (threadIdx.x == BLOCK - 1) {
        Maybe<E> c = s_carry;
#pragma unroll
        for (int i = 0; i < WAVES; ++i) c = mc(fn, c, just(s_wave[i]));
        s_carry = c;
      }
      __syncthreads();
    }
  }
}

// Compare-and-swap update of a one-field element of 4 or 8 bytes (W): the
// combine is applied to the current value until no other update came between
// the read and the swap.
template <typename E, typename W, typename Fn>
__device__ __forceinline__ void cas_update(W* q, const E& v, const Fn& fn) {
  static_assert(sizeof(E) == sizeof(W), "one-field element of the word's size");
  W old = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (true) {
    E o;
    __builtin_memcpy(&o, &old, sizeof(E));
    const E nv = fn(v, o);
    W nw;
    __builtin_memcpy(&nw, &nv, sizeof(E));
    if (__hip_atomic_compare_exchange_strong(q, &old, nw, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT))
      break;
  }
}

// --------------------------------------------------------------- permute
// out[project(i)] = combine(src(i), out[project(i)]) for every source index i
// whose projection lies inside the output (anything else is Nothing, the
// element is dropped: PTX/CodeGen/Permute.hs:151-157).  The combine must be
// associative and commutative (:67-68), so the lanes of a wave that update
// the same slot first combine their values (in lane order, a masked wave
// reduction) and one lane applies the result: a slot is updated at most once
// per wave and grid step, however skewed the indices.  A one-field element of
// 4 or 8 bytes is updated with a compare-and-swap loop on its bit pattern;
// any other element under a per-slot spin lock (the reference's
// mkPermute_mutex, PTX/CodeGen/Permute.hs:219-258): the lock word is polled
// with plain loads and taken with one swap at agent scope (acquire), the
// fields are read, combined and written, and the lock is released.  Only one
// lane of a wave waits at a time, so no lane waits on a lane of its own wave.
// A wait is bounded (kLockTries polls with back-off, seconds): a lane that
// cannot take a lock in that time drops its update rather than hang.
constexpr uint32_t kLockTries = 1u << 22;

template <typename E, typename Fn, typename Ld, typename Proj, typename Out, bool CAS>
__device__ void permute(const Args& a, const Fn& fn, const Ld& ld, const Proj& proj, const Out& out) {
  const int lane = lane_id();
  const int64_t nsrc = a.outer * a.inner;
  const int64_t stride = (int64_t)gridDim.x * BLOCK;
  // wave-uniform trip count: every lane of the wave runs every step (the wave
  // reductions below need the whole wave)
  for (int64_t base = (int64_t)blockIdx.x * BLOCK + (threadIdx.x & ~63); base < nsrc; base += stride) {
    const int64_t i = base + lane;
    int64_t j = -1;
    E v{};
    if (i < nsrc) {
      j = proj(i);
      if (j >= 0 && j < a.nseg) v = ld(i);
      else j = -1;
    }
    bool pending = j >= 0;
    while (true) {
      const uint64_t act = __ballot(pending);
      if (act == 0) break;
      const int leader = __builtin_ctzll(act);
      const int64_t jl = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)j >> 32), leader) << 32) |
                                   (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(uint64_t)j, leader));
      const bool mine = pending && j == jl;
      const Maybe<E> r = acc::wave_reduce_m(mine ? just(v) : none<E>(), fn);
      if (lane == leader) {
        if constexpr (CAS) {
          out.cas_update(jl, r.v, fn);
        } else {
          uint32_t* lk = a.locks + jl;
          for (uint32_t tries = 0; tries < kLockTries; ++tries) {
            if (acc::ld_relaxed(lk) == 0u &&
                __hip_atomic_exchange(lk, 1u, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0u) {
              out.store(jl, fn(r.v, out.load(jl)));
              __hip_atomic_store(lk, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
              break;
            }
            if (tries < 4) __builtin_amdgcn_s_sleep(1);
            else if (tries < 16) __builtin_amdgcn_s_sleep(4);
            else __builtin_amdgcn_s_sleep(16);
          }
        }
      }
      pending = pending && !mine;
    }
  }
}

// --------------------------------------------------------------- stencil
// One DIM2 stencil application with the program's own stencil function: a
// workgroup stages a (16 + 2ry) x (64 + 2rx) halo tile in LDS, mapping
// coordinates outside the array through the boundary (Bnd: a value for an
// index outside), then each thread applies the function at 4 outputs of one
// column (the neighbourhood read from LDS as `nb[dy*STRIDE + dx]` around the
// centre).
template <typename Ti, typename To, typename Fs, typename Bnd, int RY, int RX>
__device__ void stencil2d(const Args& a, const Fs& f, const Bnd& bnd) {
  constexpr int TH = 16, TW = 64;
  constexpr int LH = TH + 2 * RY, LW = TW + 2 * RX;
  __shared__ Ti tile[LH * LW];
  const int64_t H = a.outer, Wd = a.inner;
  const Ti* in = (const Ti*)a.in[0];
  To* out = (To*)a.out[0];
  const int64_t tiles_x = (Wd + TW - 1) / TW;
  const int64_t ty0 = (int64_t)(blockIdx.x / tiles_x) * TH, tx0 = (int64_t)(blockIdx.x % tiles_x) * TW;
  for (int c = threadIdx.x; c < LH * LW; c += BLOCK) {
    const int ly = c / LW, lx = c - ly * LW;
    const int64_t y = ty0 + ly - RY, x = tx0 + lx - RX;
    tile[c] = (y >= 0 && y < H && x >= 0 && x < Wd) ? in[y * Wd + x] : bnd(in, y, x, H, Wd);
  }
  __syncthreads();
  const int tx = threadIdx.x % 64, tyq = threadIdx.x / 64;
  const int64_t x = tx0 + tx;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int ly = tyq * 4 + k;
    const int64_t y = ty0 + ly;
    if (y < H && x < Wd) out[y * Wd + x] = f(&tile[(ly + RY) * LW + tx + RX]);
  }
}

// The boundary maps of LLVM/CodeGen/Stencil.hs:224-253 (one step each, then
// kept inside for extents smaller than the radius).
__device__ __forceinline__ int64_t map_clamp(int64_t i, int64_t n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); }
__device__ __forceinline__ int64_t map_mirror(int64_t i, int64_t n) {
  return map_clamp(i < 0 ? -i : (i >= n ? n - (i - n + 2) : i), n);
}
__device__ __forceinline__ int64_t map_wrap(int64_t i, int64_t n) {
  return map_clamp(i < 0 ? n + i : (i >= n ? i - n : i), n);
}

// ------------------------------------------------------------- generate
// out[i] = f(i) for i < outer*inner, four consecutive elements per thread
// (map, zipWith, generate, backpermute and transform are all this kernel
// with the program's function reading its arrays through `read`,
// PTX/CodeGen/{Map,Generate,Transform}.hs).
template <typename E, typename Gen, typename St>
__device__ void generate(const Args& a, const Gen& gen, const St& st) {
  const int64_t n = a.outer * a.inner;
  const int64_t stride = (int64_t)gridDim.x * BLOCK * 4;
  for (int64_t i = ((int64_t)blockIdx.x * BLOCK + threadIdx.x) * 4; i < n; i += stride) {
    if (i + 4 <= n) {
      E y[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) y[j] = gen(i + j, a.inner);
      st.four(i, y);
    } else {
      for (int64_t k = i; k < n; ++k) st(k, gen(k, a.inner));
    }
  }
}

}  // namespace accjit
