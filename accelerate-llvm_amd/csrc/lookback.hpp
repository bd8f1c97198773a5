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
  uint32_t ticket;  // next tile ticket
  uint32_t done;    // workgroups finished in this launch
  uint32_t epoch;   // launch counter (tags)
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

// Take the next tile ticket (thread 0 of the block) and broadcast it.
__device__ __forceinline__ uint32_t block_ticket(SyncCtl* ctl, uint32_t* lds_slot) {
  if (threadIdx.x == 0) *lds_slot = add_relaxed(&ctl->ticket, 1u);
  __syncthreads();
  uint32_t t = *lds_slot;
  __syncthreads();
  return t;
}

__device__ __forceinline__ uint32_t block_epoch(SyncCtl* ctl, uint32_t* lds_slot) {
  if (threadIdx.x == 0) *lds_slot = ld_relaxed(&ctl->epoch);
  __syncthreads();
  uint32_t e = *lds_slot;
  __syncthreads();
  return e;
}

// Called once per workgroup after its last use of the status words.  The last
// workgroup of the launch resets the ticket and advances the epoch.
__device__ __forceinline__ void block_retire(SyncCtl* ctl, uint32_t nblocks) {
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t prev = add_relaxed(&ctl->done, 1u);
    if (prev == nblocks - 1) {
      st_relaxed(&ctl->ticket, 0u);
      st_relaxed(&ctl->done, 0u);
      add_relaxed(&ctl->epoch, 1u);
    }
  }
}

// Decoupled look-back, executed by one full wave of the block that owns tile
// `tile` (> 0).  Returns the in-order combination of all tiles before `tile`.
// `status` holds NW words per tile.  Predecessors publish ST_AGGREGATE (their
// own reduction) early and ST_PREFIX (inclusive of everything before them)
// once known; tile 0 publishes ST_PREFIX directly.
template <typename T, typename Op>
__device__ T wave_lookback(const uint64_t* status, uint32_t tile, uint32_t epoch, Op op) {
  constexpr int NW = Tagged<T>::NW;
  const int lane = lane_id();
  const uint32_t tag_a = make_tag(epoch, ST_AGGREGATE);
  const uint32_t tag_p = make_tag(epoch, ST_PREFIX);
  // segmented in-order combine: a lane flagged PREFIX starts a new run
  auto seg = [op](Maybe<T> a, Maybe<T> b) {
    Maybe<T> r;
    r.v = b.ok ? b.v : op(a.v, b.v);
    r.ok = a.ok | b.ok;
    return r;
  };
  bool have_acc = false;
  T acc;
  int64_t hi = (int64_t)tile - 1;  // highest predecessor covered by this window
  while (true) {
    int64_t idx = hi - 63 + lane;
    T v{};
    uint32_t tag = 0;
    if (idx >= 0) {
      while (true) {
        tag = Tagged<T>::peek(status + (size_t)idx * NW, v);
        if (tag == tag_a || tag == tag_p) break;
        __builtin_amdgcn_s_sleep(1);
      }
    } else {
      tag = tag_p;  // lanes before tile 0 are cut off by tile 0's PREFIX
    }
    const bool is_p = (tag == tag_p);
    Maybe<T> m;
    m.v = v;
    m.ok = is_p ? 1u : 0u;
    Maybe<T> s = read_lane(wave_inclusive_scan(m, seg), 63);
    const uint64_t any_p = __ballot(is_p);
    // s.v = combination of lanes from the highest PREFIX lane (or lane 0) to 63
    if (have_acc) acc = op(s.v, acc); else { acc = s.v; have_acc = true; }
    if (any_p) break;
    hi -= 64;
  }
  return acc;
}

}  // namespace acc
