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
// Dynamic tile tickets give every workgroup a tile index in dispatch order,
// so a workgroup only ever waits on tiles already owned by running or
// finished workgroups: no forward-progress assumption about workgroups that
// have not been dispatched.
//
// One counter hands out the tickets (about one per 12 ns, measured; atomics
// to one 128-byte line serialise).  Striping them over several counters
// would multiply that, but a stripe's next ticket can then only be taken by
// one of the stripe's own workgroups: two persistent launches that are each
// partly resident (concurrent streams) could then hold the slots each
// other's missing stripe needs.  With one counter every dispatched workgroup
// takes the next ticket, so the running ones always hold a prefix.
#pragma once
#ifndef __HIPCC_RTC__  // hiprtc provides the HIP runtime itself
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>
#include "wave64.hpp"

namespace acc {

struct alignas(64) SyncCtl {
  uint64_t tick;  // high 32 bits: launch epoch; low 32 bits: next tile ticket
  uint32_t done;  // workgroups finished in this launch
  uint32_t pad[13];
};

// bytes reserved for the SyncCtl at the start of a workspace (the status
// words follow)
constexpr size_t kSyncCtlBytes = 256;
static_assert(sizeof(SyncCtl) <= kSyncCtlBytes, "SyncCtl size");

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

// The next launch's epoch, with the ticket back at 0.
__device__ __forceinline__ void advance_epoch(SyncCtl* ctl, uint32_t epoch) {
  __hip_atomic_store(&ctl->tick, (uint64_t)(epoch + 1u) << 32, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
      advance_epoch(ctl, epoch);
    }
  }
}

// Decoupled look-back, executed by one full wave of the block that owns tile
// `tile` (> first).  Returns the in-order combination of tiles [first, tile).
// Several independent problems can share one status array: tile `first`
// (the problem's first tile) always publishes ST_PREFIX and everything below
// it is cut off.
//
// A window is C chunks of 64 consecutive predecessors, lane l of chunk k
// holding tile lo + 64k + l, so every load instruction reads 64 adjacent
// status words.  The look-back needs only the tiles from the nearest
// ST_PREFIX up to tile-1: it finishes as soon as that suffix is published,
// without waiting for older tiles.  The first window has C0 chunks, later
// ones C1: the ST_PREFIX frontier of a long scan advances by up to 64*C
// tiles per memory round trip.  Unpublished tiles are re-polled with an
// exponential back-off.
template <typename T, typename Op, int C0 = 1, int C1 = 1>
__device__ T wave_lookback(const uint64_t* status, uint32_t tile, uint32_t epoch, Op op, uint32_t first = 0) {
  constexpr int NW = Tagged<T>::NW;
  constexpr int CM = C0 > C1 ? C0 : C1;
  const int lane = lane_id();
  const uint32_t tag_a = make_tag(epoch, ST_AGGREGATE);
  const uint32_t tag_p = make_tag(epoch, ST_PREFIX);
  // segmented in-order combine: an element flagged ok starts a new run
  auto seg = [op](Maybe<T> a, Maybe<T> b) {
    Maybe<T> r;
    r.v = b.ok ? b.v : op(a.v, b.v);
    r.ok = a.ok | b.ok;
    return r;
  };
  bool have_acc = false;
  T acc{};
  int64_t hi = (int64_t)tile - 1;  // highest predecessor not yet combined
  int nc = C0;
  while (true) {
    const int64_t lo = hi - 64 * nc + 1;
    T v[CM];
    uint32_t ready = 0, isp = 0;  // bit k: this lane's tile of chunk k is published / a PREFIX
#pragma unroll
    for (int k = 0; k < CM; ++k) {
      v[k] = T{};
      if (k < nc && lo + 64 * k + lane < (int64_t)first) {  // before the first tile: cut off
        ready |= 1u << k;
        isp |= 1u << k;
      }
    }
    int backoff = 1, kf = -1, L = -1;
    while (true) {
#pragma unroll
      for (int k = 0; k < CM; ++k) {
        if (k < nc && !(ready >> k & 1u)) {
          const uint32_t tag = Tagged<T>::peek(status + (size_t)(lo + 64 * k + lane) * NW, v[k]);
          if (tag == tag_a || tag == tag_p) {
            ready |= 1u << k;
            isp |= (tag == tag_p ? 1u : 0u) << k;
          }
        }
      }
      // walk chunks from the nearest: find the highest PREFIX above the
      // highest unpublished tile
      bool wait = false, found = false;
#pragma unroll
      for (int k = CM - 1; k >= 0; --k) {
        if (k < nc && !found && !wait) {
          const uint64_t R = __ballot(ready >> k & 1u), P = __ballot(isp >> k & 1u);
          const uint64_t nr = ~R;
          const uint64_t above = nr ? ~((2ull << (63 - __builtin_clzll(nr))) - 1ull) : ~0ull;
          const uint64_t Pa = P & (nr && (63 - __builtin_clzll(nr)) == 63 ? 0ull : above);
          if (Pa) {
            found = true;
            kf = k;
            L = 63 - __builtin_clzll(Pa);
          } else if (nr) {
            wait = true;
          }
        }
      }
      if (found || !wait) break;
      if (backoff == 1) __builtin_amdgcn_s_sleep(1);
      else if (backoff == 2) __builtin_amdgcn_s_sleep(2);
      else if (backoff == 4) __builtin_amdgcn_s_sleep(4);
      else __builtin_amdgcn_s_sleep(8);
      backoff = backoff < 8 ? backoff * 2 : 8;
    }
    // combine chunks kf..nc-1 (all chunks when no PREFIX was found) in order;
    // in chunk kf, lane L's PREFIX cuts off the lanes below it
    Maybe<T> w;
    w.v = T{};
    w.ok = 0;
    bool wset = false;
#pragma unroll
    for (int k = 0; k < CM; ++k) {
      if (k < nc && k >= (kf < 0 ? 0 : kf)) {
        Maybe<T> m;
        m.v = v[k];
        m.ok = (k == kf && lane == L) ? 1u : 0u;
        const Maybe<T> c = read_lane(wave_inclusive_scan(m, seg), 63);
        if (wset) w.v = op(w.v, c.v); else { w = c; wset = true; }
      }
    }
    if (have_acc) acc = op(w.v, acc); else { acc = w.v; have_acc = true; }
    if (kf >= 0) break;
    hi = lo - 1;
    nc = C1;
  }
  return acc;
}

}  // namespace acc
