// Cross-workgroup publication for single-pass skeletons on gfx950.
//
// gfx950 has eight XCDs, each with its own L2, so values shared between
// workgroups are published with agent-scope *relaxed* atomics (which the
// compiler routes around the non-coherent L2) instead of release/acquire
// fences (which would write back the whole L2).  Ordering between a value and
// its status is obtained by packing them into the same 64-bit word:
//
//   word = (tag << 32) | payload32,   tag = ((epoch + 1) << 2) | status
//
// A value of sizeof(T) bytes occupies ceil(sizeof(T)/4) such words.  A reader
// accepts a value only when every word carries the same tag, so a torn read
// (some words from an older publication) is seen as "not ready" and retried.
// `epoch` is a per-stream launch counter kept in device memory (SyncCtl),
// bumped by the last workgroup of each launch, so status words left over from
// earlier launches never need clearing (and the same holds under hipGraph
// replay, where kernel arguments are frozen).
//
// Dynamic tile tickets (SyncCtl::ticket) give every workgroup a tile index in
// dispatch order, so a workgroup only ever waits on tiles already owned by
// running or finished workgroups: no forward-progress assumption about
// workgroups that have not been dispatched.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "wave64.hpp"

namespace acc {

struct alignas(64) SyncCtl {
  uint64_t tick;  // high 32 bits: launch epoch; low 32 bits: next tile ticket
  uint32_t done;  // workgroups finished in this launch
  uint32_t pad[13];
};

enum : uint32_t { ST_INVALID = 0, ST_AGGREGATE = 1, ST_PREFIX = 2 };

__device__ __forceinline__ uint32_t ld_relaxed(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t add_relaxed(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t make_tag(uint32_t epoch, uint32_t status) {
  return ((epoch + 1u) << 2) | status;
}

template <typename T> struct Tagged {
  static constexpr int NW = Words<T>::N;

  __device__ static void publish(uint64_t* slot, const T& v, uint32_t tag) {
    Words<T> w = to_words(v);
#pragma unroll
    for (int i = 0; i < NW; ++i)
      __hip_atomic_store(slot + i, ((uint64_t)tag << 32) | w.w[i], __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
  }

  // Returns the common tag of all words, or 0 when the words disagree.
  __device__ static uint32_t peek(const uint64_t* slot, T& v) {
    Words<T> w;
    uint32_t tag = 0;
    bool same = true;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
      uint64_t x = __hip_atomic_load(slot + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      uint32_t t = (uint32_t)(x >> 32);
      if (i == 0) tag = t; else same = same && (t == tag);
      w.w[i] = (uint32_t)x;
    }
    v = from_words<T>(w);
    return same ? tag : 0u;
  }
};

// Take a ticket (thread 0) and broadcast it: high 32 bits epoch, low 32 bits
// the ticket.  One atomic gives both, so a tile costs one round trip here.
__device__ __forceinline__ uint64_t take_ticket(SyncCtl* ctl) {
  return __hip_atomic_fetch_add(&ctl->tick, (uint64_t)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t block_epoch(SyncCtl* ctl, uint32_t* lds_slot) {
  if (threadIdx.x == 0)
    *lds_slot = (uint32_t)(__hip_atomic_load(&ctl->tick, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> 32);
  __syncthreads();
  uint32_t e = *lds_slot;
  __syncthreads();
  return e;
}

// Called once per workgroup after its last use of the status words.  The last
// workgroup of the launch resets the ticket and advances the epoch.  Every
// workgroup has taken all its tickets by then.
__device__ __forceinline__ void block_retire(SyncCtl* ctl, uint32_t nblocks, uint32_t epoch) {
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t prev = add_relaxed(&ctl->done, 1u);
    if (prev == nblocks - 1) {
      st_relaxed(&ctl->done, 0u);
      __hip_atomic_store(&ctl->tick, (uint64_t)(epoch + 1u) << 32, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// Decoupled look-back, executed by one full wave of the block that owns tile
// `tile` (> first).  Returns the in-order combination of tiles [first, tile).
// Several independent problems can share one status array: tile `first`
// (the problem's first tile) always publishes ST_PREFIX and everything below
// it is cut off.
// `status` holds NW words per tile.  Predecessors publish ST_AGGREGATE (their
// own reduction) early and ST_PREFIX (inclusive of everything before them)
// once known; tile 0 publishes ST_PREFIX directly.
//
// The first window covers the 64 nearest predecessors (one per lane); later
// windows give each lane S consecutive tiles (64*S per round trip), which
// bounds the latency of a deep look-back without reading far on the common,
// shallow one.  Lanes whose tile is not yet published re-poll with an
// exponential back-off.
template <typename T, typename Op, int S = 4>
__device__ T wave_lookback(const uint64_t* status, uint32_t tile, uint32_t epoch, Op op, uint32_t first = 0) {
  constexpr int NW = Tagged<T>::NW;
  const int lane = lane_id();
  const uint32_t tag_a = make_tag(epoch, ST_AGGREGATE);
  const uint32_t tag_p = make_tag(epoch, ST_PREFIX);
  // segmented in-order combine: an element flagged PREFIX starts a new run
  auto seg = [op](Maybe<T> a, Maybe<T> b) {
    Maybe<T> r;
    r.v = b.ok ? b.v : op(a.v, b.v);
    r.ok = a.ok | b.ok;
    return r;
  };
  bool have_acc = false;
  T acc{};
  int64_t hi = (int64_t)tile - 1;  // highest predecessor not yet combined
  int per = 1;                      // tiles per lane in this window
  while (true) {
    const int64_t lo = hi - 64 * per + 1;
    const int64_t idx0 = lo + (int64_t)lane * per;
    T v[S];
    bool is_p[S];
    uint32_t ready = 0;
#pragma unroll
    for (int k = 0; k < S; ++k) {
      v[k] = T{};
      is_p[k] = false;
      if (k >= per || idx0 + k < (int64_t)first) {
        ready |= 1u << k;
        is_p[k] = idx0 + k < (int64_t)first;  // before the first tile: cut off
      }
    }
    int backoff = 1;
    while (true) {
#pragma unroll
      for (int k = 0; k < S; ++k) {
        if (!(ready >> k & 1u)) {
          const uint32_t tag = Tagged<T>::peek(status + (size_t)(idx0 + k) * NW, v[k]);
          if (tag == tag_a || tag == tag_p) {
            ready |= 1u << k;
            is_p[k] = tag == tag_p;
          }
        }
      }
      const bool done = ready == (1u << S) - 1u;
      if (__all(done)) break;
      if (backoff == 1) __builtin_amdgcn_s_sleep(1);
      else if (backoff == 2) __builtin_amdgcn_s_sleep(2);
      else if (backoff == 4) __builtin_amdgcn_s_sleep(4);
      else __builtin_amdgcn_s_sleep(8);
      backoff = backoff < 8 ? backoff * 2 : 8;
    }
    // lane-level segmented fold of its tiles, in order
    Maybe<T> m;
    m.v = v[0];
    m.ok = is_p[0] ? 1u : 0u;
#pragma unroll
    for (int k = 1; k < S; ++k) {
      if (k < per) {
        Maybe<T> x;
        x.v = v[k];
        x.ok = is_p[k] ? 1u : 0u;
        m = seg(m, x);
      }
    }
    const Maybe<T> s = read_lane(wave_inclusive_scan(m, seg), 63);
    if (have_acc) acc = op(s.v, acc); else { acc = s.v; have_acc = true; }
    if (s.ok) break;
    hi = lo - 1;
    per = S;
  }
  return acc;
}

}  // namespace acc
