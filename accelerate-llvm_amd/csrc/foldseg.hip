// foldSeg / fold1Seg, and fold along the innermost dimension of a rank >= 2
// array (every row one segment), load-balanced over equal tiles.
//
// The reference assigns one block or one warp per segment
// (PTX/CodeGen/FoldSeg.hs:79-264 and :274-458, selected at
// PTX/Execute.hs:361-366) and one block per row for foldDim
// (PTX/CodeGen/Fold.hs:296-405), so long segments serialise and short ones
// idle most lanes.  Here the elements are cut into equal tiles regardless of
// segment boundaries and the fold becomes a segmented scan over
// (head-flag, value) pairs, evaluated in registers; the value at the last
// element of each segment is its result.
//
// Three launches, none of which waits on another workgroup:
//   k_tile_seg  (offsets only) per tile boundary: the first segment of the
//               tile starting there and the positions of its first SLOT
//               segment starts, so a tile stages its segment heads with one
//               load issued ahead of its data; also writes the seed into
//               empty segments;
//   k_fold_seg  one workgroup per tile: every segment end whose segment
//               starts inside the tile is final at once and written; the one
//               end that continues a segment from earlier tiles ("open") and
//               the tile's segmented aggregate go to a per-tile record
//               (k_fold_seg_pair: the tiles of two folds over one plan in
//               one launch, the same tile body);
//   k_seg_fix   a segmented scan over the tile aggregates (256 tiles per
//               workgroup, chained by a look-back), which completes the open
//               ends (at most one per tile).  (One workgroup over all the
//               records took 16 us for 2^13 tiles, against 5: one CU cannot
//               issue the scattered record loads fast enough, measured.)
// No kernel waits on another workgroup.  (Completing the open ends inside
// k_fold_seg -- tagged records, an arrival count per group of tiles, the
// group's last tile scanning the records -- was built and measured: the
// per-tile atomic's round trip kept every tile resident longer and the tile
// kernel lost as much time as the fix-up launch takes; searching the offsets
// inside the tiles instead of k_tile_seg put dependent round trips ahead of
// the data loads and cost more.)
//
// Segments are given as offsets (the `scanl (+) 0` of the lengths; the
// frontend has already applied it, PTX/Execute.hs:359,
// NAT/CodeGen/FoldSeg.hs:99-101) or, for foldDim, as the uniform row length.
// Empty segments yield the seed (PTX/CodeGen/FoldSeg.hs:170-178); fold1Seg
// leaves them untouched (it requires non-empty segments,
// NAT/CodeGen/FoldSeg.hs:96-97).
//
// HBM roofline: n*sizeof(T) + (m+1)*sizeof(I) + m*sizeof(T) bytes.
#include "dispatch.hpp"
#include "host.hpp"
#include "lookback.hpp"
#include "seg.hpp"
#include "vecmem.hpp"
#include "wave64.hpp"

namespace acc {

// Element positions inside the kernels are relative to the 16-byte boundary
// at or below the row's first element: position p is row element p - shift,
// and positions below `shift` are padding (dead elements).  So every tile and
// every dwordx4 load is aligned, also for sub-array views and for the rows
// of a rank-2 array.

constexpr int SLOT = 64;  // segment starts per tile handed over by k_tile_seg

// k_tile_seg's result per tile boundary b = 0 .. ntiles: first[b] is the
// last segment starting at or before position b * TILE (0 if none).  Tile t
// is covered by the starts first[t] .. first[t+1] + 1 (cnt = first[t+1] -
// first[t] + 2 of them; the last starts after the tile or is past the last
// segment).

// Per-tile record for the fix-up.
template <typename T> struct TileCarry {
  SegV<T> agg;     // segmented combination of the whole tile
  T open_v;        // the tile's first elements up to the end of their segment
  int64_t open_g;  // that segment, or -1 when the tile has no open end
};

// The records as four arrays (every array 256-byte aligned), so that the
// fix-up's reads of consecutive tiles' records coalesce.
template <typename T> struct CarryArrays {
  T* agg_v;
  uint8_t* agg_f;
  T* open_v;
  int64_t* open_g;
  __device__ __forceinline__ TileCarry<T> get(uint32_t t) const {
    return TileCarry<T>{SegV<T>{agg_v[t], (uint32_t)agg_f[t]}, open_v[t], open_g[t]};
  }
  __device__ __forceinline__ void put(uint32_t t, const TileCarry<T>& c) const {
    agg_v[t] = c.agg.v;
    agg_f[t] = c.agg.f ? 1 : 0;
    open_v[t] = c.open_v;
    open_g[t] = c.open_g;
  }
};

// Segment starts given by an offsets array: start(g) = off[g] + shift, g in [0, S].
template <typename I> struct OffsetStarts {
  const I* off;
  const int64_t* first;  // per tile boundary (ntiles + 1)
  const int32_t* slots;  // SLOT tile-relative starts per tile (valid when cnt <= SLOT)
  int64_t shift;
  __device__ __forceinline__ int64_t start(int64_t g) const { return (int64_t)off[g] + shift; }
};

// foldDim: start(g) = g * len + shift.
struct UniformStarts {
  int64_t len;
  int64_t shift;
  __device__ __forceinline__ int64_t start(int64_t g) const { return g * len + shift; }
};

template <typename T, typename Starts> struct SegArgs {
  const T* in;  // 16-byte aligned: position 0
  T* out;
  Starts st;
  int64_t nseg;   // S
  int64_t e_lo;   // position of the row's first element (the shift)
  int64_t e_end;  // one past the position of the row's last element
  uint32_t ntiles;
  uint32_t has_seed;
  T seed;
  CarryArrays<T> carry;
};

// A tile-relative position clamped to [-1, tile + 1]: order against the
// tile's own positions is kept.
__device__ __forceinline__ int32_t rel_pos(int64_t p, int64_t base, int64_t tile) {
  const int64_t r = p - base;
  return r < 0 ? -1 : (r > tile ? (int32_t)tile + 1 : (int32_t)r);
}

// One wave per tile boundary: the boundary's first segment by one 64-ary
// search (three round trips for 2^16 segments), and the positions of the
// next tile's first SLOT starts relative to it.  (A search per boundary,
// not two per tile: each probe is its own cache line, so the pre-pass is
// bound by the probes' address processing, and the end of tile t is the
// start of tile t+1.)  Then, when a seed is given, every empty segment gets
// the seed (grid-stride over the segments).
template <typename T, typename I>
__global__ __launch_bounds__(256) void k_tile_seg(const I* off, int64_t nseg, int64_t tile, uint32_t ntiles,
                                                  int64_t shift, int64_t* first, int32_t* slots, T* out,
                                                  uint32_t fill_empty, T seed) {
  const uint32_t t = blockIdx.x * 4u + threadIdx.x / 64u;  // boundary
  const int lane = lane_id();
  if (t <= ntiles) {  // whole waves
    const int64_t base = (int64_t)t * tile;  // position
    // last g in [0, nseg] with off[g] <= e for e = the row element at the
    // boundary (0 if none)
    const int64_t e = base - shift;
    int64_t lo = 0, hi = (int64_t)off[0] > e ? 0 : nseg;
    while (hi > lo) {
      const int64_t st = (hi - lo + 64) / 64;
      const int64_t g = lo + (int64_t)lane * st;
      const bool p = g <= hi;
      const int64_t v = p ? (int64_t)off[g] : 0;
      const int c = __popcll(__ballot(p && v <= e));  // >= 1: off[lo] <= e
      lo = lo + (int64_t)(c - 1) * st;
      hi = min(hi, lo + st - 1);
    }
    if (lane == 0) first[t] = lo;
    if (t < ntiles) {
      const int64_t g = lo + lane;
      slots[(size_t)t * SLOT + lane] = g > nseg ? (int32_t)tile + 1 : rel_pos((int64_t)off[g] + shift, base, tile);
    }
  }
  if (fill_empty) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < nseg; g += stride)
      if (off[g] == off[g + 1]) out[g] = seed;
  }
}

// The offsets (`scanl (+) 0` of the segment lengths) and their tile plan in
// one launch.  A single-pass scan of the lengths: each workgroup takes a tile
// of SP_LT lengths by ticket, scans it, and gets its prefix by the decoupled
// look-back; it then holds the offsets of its segments g0 .. g0+SP_LT-1 and,
// from the SLOT lengths after its tile, of the SLOT segments after them, in
// LDS.  The tile boundaries whose first segment (k_tile_seg's first[]) lies
// in its segments are exactly those at positions from its first offset up to
// (not including) the next tile's first offset, so it writes their first[]
// (a binary search in LDS) and each one's starts up to the first past that
// fold tile (all a fold tile reads of them).  So every boundary is written
// once, and a segment spanning many fold tiles costs two slot words per tile.
constexpr int SP_BLOCK = 256, SP_PER = 4, SP_LT = SP_BLOCK * SP_PER;

template <typename I> struct ScanPlanArgs {
  const I* len;
  I* off;  // nseg + 1
  int64_t nseg;
  int64_t tile, ntiles, shift;  // the fold's tiles
  int64_t* first;
  int32_t* slots;
  SyncCtl* ctl;
  uint64_t* status;
};

template <typename I>
__global__ __launch_bounds__(SP_BLOCK) void k_scan_plan(ScanPlanArgs<I> a) {
  constexpr int W = SP_BLOCK / 64;
  constexpr int NW = Tagged<I>::NW;
  __shared__ int64_t s_off[SP_LT + SLOT];
  __shared__ I s_wave[W];
  __shared__ I s_prefix;
  __shared__ uint64_t s_tick;
  OpAdd add{};
  auto sub = [](I u, I v) {  // wrapping, as the scan's additions
    using U = typename unsigned_of<I>::type;
    return (I)((U)u - (U)v);
  };
  const int lane = lane_id(), w = threadIdx.x / 64;
  if (threadIdx.x == 0) s_tick = take_ticket(a.ctl);
  __syncthreads();
  const uint32_t epoch = (uint32_t)(s_tick >> 32);
  const uint32_t j = (uint32_t)s_tick;  // one ticket per workgroup
  const int64_t g0 = (int64_t)j * SP_LT;
  // this thread's lengths (none past the last segment) and one of the SLOT after the tile
  I x[SP_PER];
#pragma unroll
  for (int k = 0; k < SP_PER; ++k) {
    const int64_t g = g0 + threadIdx.x * SP_PER + k;
    x[k] = g < a.nseg ? a.len[g] : (I)0;
  }
  I xe = 0;
  if (threadIdx.x < SLOT && g0 + SP_LT + threadIdx.x < a.nseg) xe = a.len[g0 + SP_LT + threadIdx.x];
  I tot = x[0];
#pragma unroll
  for (int k = 1; k < SP_PER; ++k) tot = add(tot, x[k]);
  const I inc = wave_inclusive_scan(tot, add);
  if (lane == 63) s_wave[w] = inc;
  const I e_inc = wave_inclusive_scan(xe, add);  // the lengths after the tile (wave 0's lanes)
  __syncthreads();
  I wp = 0, agg = 0;
#pragma unroll
  for (int i = 0; i < W; ++i) {
    if (i < w) wp = add(wp, s_wave[i]);
    agg = add(agg, s_wave[i]);
  }
  if (w == 0) {
    I X = 0;
    uint64_t* slot = a.status + (size_t)j * NW;
    const bool last = j + 1 == gridDim.x;
    if (j == 0) {
      if (lane == 0 && !last) Tagged<I>::publish(slot, agg, make_tag(epoch, ST_PREFIX));
    } else {
      if (lane == 0) Tagged<I>::publish(slot, agg, make_tag(epoch, ST_AGGREGATE));
      X = wave_lookback<I, OpAdd, 2, 2>(a.status, j, epoch, add);  // 128 predecessors per round trip
      if (lane == 0 && !last) Tagged<I>::publish(slot, add(X, agg), make_tag(epoch, ST_PREFIX));
    }
    if (lane == 0) s_prefix = X;
  }
  __syncthreads();
  const I P = s_prefix;
  I o = add(P, add(wp, sub(inc, tot)));  // offset of this thread's first segment
#pragma unroll
  for (int k = 0; k < SP_PER; ++k) {
    const int idx = threadIdx.x * SP_PER + k;
    if (g0 + idx <= a.nseg) a.off[g0 + idx] = o;
    s_off[idx] = (int64_t)o;
    o = add(o, x[k]);
  }
  if (threadIdx.x < SLOT) s_off[SP_LT + threadIdx.x] = (int64_t)add(add(P, agg), sub(e_inc, xe));
  __syncthreads();
  // the plan entries of the boundaries whose first segment is one of g0 .. g1-1
  const int64_t g1 = min(g0 + SP_LT, a.nseg + 1);
  const int64_t lo_pos = s_off[0] + a.shift, hi_pos = s_off[g1 - g0] + a.shift;
  const int64_t b_lo = g0 == 0 ? 0 : (lo_pos <= 0 ? 0 : (lo_pos + a.tile - 1) / a.tile);
  int64_t b_hi = g1 > a.nseg ? a.ntiles : (hi_pos <= 0 ? -1 : (hi_pos + a.tile - 1) / a.tile - 1);
  b_hi = min(b_hi, a.ntiles);
  for (int64_t b = b_lo + threadIdx.x; b <= b_hi; b += SP_BLOCK) {
    const int64_t e = b * a.tile;
    int L = 0, R = (int)(g1 - g0) - 1;  // last segment starting at or before e (g0 when none)
    while (L < R) {
      const int m = (L + R + 1) >> 1;
      if (s_off[m] + a.shift <= e) L = m;
      else R = m - 1;
    }
    const int64_t g = g0 + L;
    a.first[b] = g;
    if (b < a.ntiles) {
      // four starts per 16-byte store (rows are 256-byte aligned), up to the
      // first past the fold tile
      for (int i = 0; i < SLOT; i += 4) {
        int32_t r[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
          r[k] = g + i + k > a.nseg ? (int32_t)a.tile + 1 : rel_pos(s_off[L + i + k] + a.shift, e, a.tile);
        typedef int32_t i32x4 __attribute__((ext_vector_type(4)));
        const i32x4 v = {r[0], r[1], r[2], r[3]};
        *(__attribute__((address_space(1))) i32x4*)(a.slots + (size_t)b * SLOT + i) = v;
        if (r[3] > a.tile) break;
      }
    }
  }
  block_retire(a.ctl, gridDim.x, epoch);
}

// The seed into every empty segment (a call on a tile plan built ahead, whose
// pre-pass did not see the output).
template <typename T, typename I>
__global__ __launch_bounds__(256) void k_fill_empty(const I* off, int64_t nseg, T* out, T seed) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < nseg; g += stride)
    if (off[g] == off[g + 1]) out[g] = seed;
}

// One workgroup per tile.  The tile is read with coalesced dwordx4 loads
// (all issued first), then processed in NH halves: a half's vectors are
// written to LDS in a transposed layout, and each thread reads back one
// contiguous run of R elements and folds it sequentially (head bits from the
// LDS bitmap, one word per run), so the only cross-lane work is one
// segmented scan of the run aggregates per half.  Row stride of the LDS
// layout: R elements plus 16 bytes, which keeps the 16-byte reads of a wave
// free of bank conflicts.
template <typename T, typename Op, typename Starts, int BLOCK, int K, int CAP>
__global__ __launch_bounds__(BLOCK) void k_fold_seg(SegArgs<T, Starts> a) {
  constexpr bool UNIFORM = std::is_same<Starts, UniformStarts>::value;
  constexpr int VEC = Vec<T>::N;
  constexpr int W = BLOCK / 64;
  constexpr int TILE = BLOCK * K * VEC;
  constexpr int NBW = TILE / 32 + 1;  // head bitmap: tile positions 0..TILE
#ifndef ACC_SEG_KH
#define ACC_SEG_KH 4
#endif
  constexpr int KH = ACC_SEG_KH;      // vectors per thread per part
  constexpr int NH = K / KH;          // halves per tile
  constexpr int R = KH * VEC;         // run: elements per thread per half
  constexpr int HALF = BLOCK * R;
  constexpr int RS = R + 16 / (int)sizeof(T);  // LDS row stride of a run
  static_assert(K % KH == 0 && R <= 16, "run layout");
  static_assert(TILE + 1 < 32767, "tile-relative starts are 16-bit");
  using S = SegV<T>;
  __shared__ uint32_t s_heads[NBW];
  __shared__ int16_t s_rel[UNIFORM ? 1 : CAP];  // tile-relative starts (-1 .. TILE+1)
  __shared__ int16_t s_pre[UNIFORM ? 1 : NBW];  // heads in the bitmap words before each word
  __shared__ int32_t s_nin, s_neg, s_mode;      // starts in [0, TILE]; one before 0; search mode
  __shared__ S s_wave[W];
  __shared__ int64_t s_open_g;
  __shared__ T s_open_v;
  __shared__ __align__(16) T s_data[BLOCK * RS];
  const uint32_t t = blockIdx.x;
#include "foldseg_tile.inc"
}

// The tiles of two foldSegs over one tile plan in one launch: workgroups
// [0, n0) take the first fold's tiles, the rest the second's, so the second
// fold's first tiles fill the CUs that the first fold's last ones leave idle
// (one launch tail instead of two).  Both share the LDS declared here.
template <typename T, typename Op0, typename Op1, typename Starts, int BLOCK, int K, int CAP>
__global__ __launch_bounds__(BLOCK) void k_fold_seg_pair(SegArgs<T, Starts> a0, SegArgs<T, Starts> a1, uint32_t n0) {
  constexpr bool UNIFORM = std::is_same<Starts, UniformStarts>::value;
  constexpr int VEC = Vec<T>::N;
  constexpr int W = BLOCK / 64;
  constexpr int TILE = BLOCK * K * VEC;
  constexpr int NBW = TILE / 32 + 1;  // head bitmap: tile positions 0..TILE
#ifndef ACC_SEG_KH
#define ACC_SEG_KH 4
#endif
  constexpr int KH = ACC_SEG_KH;      // vectors per thread per part
  constexpr int NH = K / KH;          // halves per tile
  constexpr int R = KH * VEC;         // run: elements per thread per half
  constexpr int HALF = BLOCK * R;
  constexpr int RS = R + 16 / (int)sizeof(T);  // LDS row stride of a run
  static_assert(K % KH == 0 && R <= 16, "run layout");
  static_assert(TILE + 1 < 32767, "tile-relative starts are 16-bit");
  using S = SegV<T>;
  __shared__ uint32_t s_heads[NBW];
  __shared__ int16_t s_rel[UNIFORM ? 1 : CAP];  // tile-relative starts (-1 .. TILE+1)
  __shared__ int16_t s_pre[UNIFORM ? 1 : NBW];  // heads in the bitmap words before each word
  __shared__ int32_t s_nin, s_neg, s_mode;      // starts in [0, TILE]; one before 0; search mode
  __shared__ S s_wave[W];
  __shared__ int64_t s_open_g;
  __shared__ T s_open_v;
  __shared__ __align__(16) T s_data[BLOCK * RS];
  if (blockIdx.x < n0) {
    using Op = Op0;
    const SegArgs<T, Starts>& a = a0;
    const uint32_t t = blockIdx.x;
#include "foldseg_tile.inc"
  } else {
    using Op = Op1;
    const SegArgs<T, Starts>& a = a1;
    const uint32_t t = blockIdx.x - n0;
#include "foldseg_tile.inc"
  }
}

// Segmented exclusive scan over the tile aggregates: it gives each tile the
// carry of the segment it starts in, which completes the open ends.  Each
// workgroup (numbered by ticket) takes BLOCK consecutive tiles, one per
// thread (coalesced), scans them, publishes its aggregate and takes the
// carry from the workgroups before it with a decoupled look-back (with at
// most 64 workgroups one window covers them all).
// One workgroup's part of the fix-up: tiles grp*BLOCK .. +BLOCK-1 of a
// foldSeg whose look-back status words start at `first` (status slot
// first + grp).
template <typename T, typename Op, int BLOCK>
__device__ __forceinline__ void seg_fix_group(const CarryArrays<T>& carry, uint32_t ntiles, T* out, int64_t nseg,
                                              uint32_t has_seed, T seed, uint64_t* status, uint32_t first,
                                              uint32_t grp, uint32_t ngrp, uint32_t epoch) {
  constexpr int W = BLOCK / 64;
  using S = SegV<T>;
  constexpr int NW = Tagged<S>::NW;
  __shared__ S s_part[W];
  __shared__ S s_carry;
  SegOp<Op> sop{Op{}};
  Op op{};
  const S sid{identity<T>(op), 0u};
  const int lane = lane_id(), w = threadIdx.x / 64;
  const uint32_t t = grp * BLOCK + threadIdx.x;
  TileCarry<T> c;
  if (t < ntiles) c = carry.get(t);
  else c = TileCarry<T>{sid, sid.v, -1};
  const S inc = wave_inclusive_scan(c.agg, sop);
  const S up = shift_up1(inc);  // all lanes take part in the cross-lane move
  if (lane == 63) s_part[w] = inc;
  __syncthreads();
  S wp = sid, agg = sid;
#pragma unroll
  for (int i = 0; i < W; ++i) {
    const S p = s_part[i];
    if (i < w) wp = sop(wp, p);
    agg = sop(agg, p);
  }
  if (w == 0) {
    S X = sid;
    const uint32_t b = first + grp;
    uint64_t* slot = status + (size_t)b * NW;
    const bool last = grp + 1 == ngrp;
    if (grp > 0) {
      if (lane == 0) Tagged<S>::publish(slot, agg, make_tag(epoch, agg.f ? ST_PREFIX : ST_AGGREGATE));
      X = wave_lookback<S>(status, b, epoch, sop, first);
      if (lane == 0 && !agg.f && !last) Tagged<S>::publish(slot, sop(X, agg), make_tag(epoch, ST_PREFIX));
    } else if (lane == 0 && !last) {
      Tagged<S>::publish(slot, agg, make_tag(epoch, ST_PREFIX));
    }
    if (lane == 0) s_carry = X;
  }
  __syncthreads();
  const S X = sop(sop(s_carry, wp), lane > 0 ? up : sid);  // carry into this thread's tile
  if (c.open_g >= 0 && c.open_g < nseg) {
    T v = op(X.v, c.open_v);
    if (has_seed) v = op(seed, v);
    out[c.open_g] = v;
  }
}

template <typename T, typename Op, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_seg_fix(CarryArrays<T> carry, uint32_t ntiles, T* out,
                                                   int64_t nseg, uint32_t has_seed, T seed, SyncCtl* ctl,
                                                   uint64_t* status) {
  __shared__ uint64_t s_tick;
  if (threadIdx.x == 0) s_tick = take_ticket(ctl);
  __syncthreads();
  const uint32_t epoch = (uint32_t)(s_tick >> 32);
  const uint32_t b = (uint32_t)s_tick;  // < gridDim.x: one ticket per workgroup
  seg_fix_group<T, Op, BLOCK>(carry, ntiles, out, nseg, has_seed, seed, status, 0u, b, gridDim.x, epoch);
  block_retire(ctl, gridDim.x, epoch);
}

// The fix-ups of several foldSegs over one tile plan in one launch: ticket t
// is group t % ngrp of fold t / ngrp, each fold with its own records, output,
// combine and status words (a look-back never crosses into another fold's).
constexpr int kMultiFolds = 8;
template <typename T> struct MultiFix {
  CarryArrays<T> carry[kMultiFolds];
  T* out[kMultiFolds];
  T seed[kMultiFolds];
  int32_t op[kMultiFolds];
  uint32_t has_seed[kMultiFolds];
};

template <typename T, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_seg_fix_multi(MultiFix<T> m, uint32_t ntiles, uint32_t ngrp, int64_t nseg,
                                                         SyncCtl* ctl, uint64_t* status) {
  __shared__ uint64_t s_tick;
  if (threadIdx.x == 0) s_tick = take_ticket(ctl);
  __syncthreads();
  const uint32_t epoch = (uint32_t)(s_tick >> 32);
  const uint32_t tk = (uint32_t)s_tick;  // < gridDim.x
  const uint32_t f = tk / ngrp, grp = tk - f * ngrp;
  const uint32_t first = f * ngrp;
  auto go = [&](auto opv) {
    using Op = decltype(opv);
    seg_fix_group<T, Op, BLOCK>(m.carry[f], ntiles, m.out[f], nseg, m.has_seed[f], m.seed[f], status, first, grp,
                                ngrp, epoch);
  };
  const int32_t opc = m.op[f];  // uniform
  if (opc == ACC_OP_ADD) go(OpAdd{});
  else if (opc == ACC_OP_MUL) go(OpMul{});
  else if (opc == ACC_OP_MIN) go(OpMin{});
  else if (opc == ACC_OP_MAX) go(OpMax{});
  else if constexpr (std::is_integral<T>::value) {
    if (opc == ACC_OP_AND) go(OpAnd{});
    else if (opc == ACC_OP_OR) go(OpOr{});
    else go(OpXor{});
  }
  block_retire(ctl, gridDim.x, epoch);
}

#ifndef ACC_SEG_K
#define ACC_SEG_K 8
#endif
template <typename T> struct SegTune {
  static constexpr int BLOCK = 256;
  static constexpr int K = ACC_SEG_K;  // dwordx4 vectors per thread
#ifndef ACC_SEG_CAP
#define ACC_SEG_CAP 512
#endif
  static constexpr int CAP = ACC_SEG_CAP;  // starts staged in LDS (beyond: searched in HBM)
};

// k_tile_seg over offsets of 4 or 8 bytes (out / fill_empty: the seed into
// empty segments as well).
template <typename T>
int launch_tile_plan(acc_stream_t s, const void* offsets, size_t offset_bytes, int64_t nseg, int64_t tile,
                     int64_t ntiles, int64_t shift, int64_t* first, int32_t* slots, T* out, bool fill_empty, T seed) {
  const unsigned pg = (unsigned)((ntiles + 1 + 3) / 4);  // a wave per tile boundary
  if (offset_bytes == 4)
    hipLaunchKernelGGL((k_tile_seg<T, int32_t>), dim3(pg), dim3(256), 0, s->stream, (const int32_t*)offsets, nseg,
                       tile, (uint32_t)ntiles, shift, first, slots, out, fill_empty ? 1u : 0u, seed);
  else
    hipLaunchKernelGGL((k_tile_seg<T, int64_t>), dim3(pg), dim3(256), 0, s->stream, (const int64_t*)offsets, nseg,
                       tile, (uint32_t)ntiles, shift, first, slots, out, fill_empty ? 1u : 0u, seed);
  ACC_HIP_CHECK(hipGetLastError());
  return ACC_SUCCESS;
}

template <typename T>
int launch_fill_empty(acc_stream_t s, const void* offsets, size_t offset_bytes, int64_t nseg, T* out, T seed) {
  const DeviceInfo* di;
  ACC_CHECK(device_info(s->device, &di));
  const int64_t blocks = std::min<int64_t>(ceil_div(nseg, (int64_t)256), (int64_t)di->compute_units * 4);
  const unsigned g = (unsigned)std::max<int64_t>(1, blocks);
  if (offset_bytes == 4)
    hipLaunchKernelGGL((k_fill_empty<T, int32_t>), dim3(g), dim3(256), 0, s->stream, (const int32_t*)offsets, nseg,
                       out, seed);
  else
    hipLaunchKernelGGL((k_fill_empty<T, int64_t>), dim3(g), dim3(256), 0, s->stream, (const int64_t*)offsets, nseg,
                       out, seed);
  ACC_HIP_CHECK(hipGetLastError());
  return ACC_SUCCESS;
}

// A tile plan built ahead (acc_hip_seg_plan_build) for the same offsets,
// element type, element count and alignment: k_tile_seg's results.
struct TilePlan {
  const int64_t* first;
  const int32_t* slots;
};

template <typename T, typename Op, typename Starts>
int fold_seg_launch(acc_stream_t s, const T* in, T* out, Starts st, int64_t nseg, int64_t e_end,
                    const T* seed, bool check_empty, const void* offsets, size_t offset_bytes,
                    const TilePlan* plan = nullptr) {
  constexpr int BLOCK = SegTune<T>::BLOCK, K = SegTune<T>::K, CAP = SegTune<T>::CAP;
  constexpr int64_t TILE = (int64_t)BLOCK * K * Vec<T>::N;
  if ((uintptr_t)in % sizeof(T)) return fail(ACC_ERROR_INVALID_VALUE, "foldSeg: input not element-aligned");
  // positions count from the 16-byte boundary at or below `in`
  const int64_t shift = (int64_t)(((uintptr_t)in % 16) / sizeof(T));
  const T* in0 = in - shift;
  st.shift = shift;
  const int64_t ntiles = e_end == 0 ? 0 : ceil_div(e_end + shift, TILE);
  if (ntiles == 0) {
    // no elements: every segment is empty
    if (!seed || !check_empty) return ACC_SUCCESS;
    return acc_hip_fill_async(s, out, seed, sizeof(T), (size_t)nseg);
  }
  if (ntiles >= 0x7fffffffLL) return fail(ACC_ERROR_UNSUPPORTED, "foldSeg: too many elements");
  // untagged scratch: per-tile carry records, then (offsets) segment ranges
  // and starts; the workspace holds only the fix-up's tagged status words
  // (a scan on the same stream reads that region by tag)
  SyncCtl* ctl;
  uint64_t* fix_status;
  void* scr;
  auto al256 = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t cv_bytes = al256((size_t)ntiles * sizeof(T)), cf_bytes = al256((size_t)ntiles);
  const size_t cg_bytes = (size_t)ntiles * 8;
  const size_t carry_bytes = al256(2 * cv_bytes + cf_bytes + cg_bytes);
  const bool own_plan = offsets && !plan;
  const size_t segs_bytes = own_plan ? ((size_t)(ntiles + 1) * sizeof(int64_t) + 255) / 256 * 256 : 0;
  const size_t slot_bytes = own_plan ? (size_t)ntiles * SLOT * sizeof(int32_t) : 0;
  constexpr int FIX = 256;  // tiles per k_seg_fix workgroup
  const int64_t nfix = ceil_div(ntiles, (int64_t)FIX);
  const size_t fix_bytes = (size_t)nfix * Tagged<SegV<T>>::NW * 8;
  ACC_CHECK(stream_workspace(s, fix_bytes, &ctl, &fix_status));
  ACC_CHECK(stream_scratch(s, carry_bytes + segs_bytes + slot_bytes, &scr));
  char* ws = (char*)scr;
  char* cb = ws;
  const CarryArrays<T> carry{(T*)cb, (uint8_t*)(cb + cv_bytes), (T*)(cb + cv_bytes + cf_bytes),
                             (int64_t*)(cb + 2 * cv_bytes + cf_bytes)};
  const bool fill_empty = seed && check_empty;
  if constexpr (!std::is_same<Starts, UniformStarts>::value) {
    if (plan) {
      st.first = plan->first;
      st.slots = plan->slots;
      if (fill_empty) ACC_CHECK(launch_fill_empty(s, offsets, offset_bytes, nseg, out, *seed));
    } else {
      int64_t* first = (int64_t*)((char*)ws + carry_bytes);
      int32_t* slots = (int32_t*)((char*)ws + carry_bytes + segs_bytes);
      ACC_CHECK(launch_tile_plan<T>(s, offsets, offset_bytes, nseg, TILE, ntiles, shift, first, slots, out,
                                    fill_empty, seed ? *seed : T()));
      st.first = first;
      st.slots = slots;
    }
  }
  SegArgs<T, Starts> a{in0, out, st, nseg, shift, e_end + shift, (uint32_t)ntiles, seed ? 1u : 0u,
                       seed ? *seed : T(), carry};
  hipLaunchKernelGGL((k_fold_seg<T, Op, Starts, BLOCK, K, CAP>), dim3((unsigned)ntiles), dim3(BLOCK), 0, s->stream,
                     a);
  ACC_HIP_CHECK(hipGetLastError());
  if (ntiles > 1) {
    hipLaunchKernelGGL((k_seg_fix<T, Op, FIX>), dim3((unsigned)nfix), dim3(FIX), 0, s->stream, carry,
                       (uint32_t)ntiles, out, nseg, seed ? 1u : 0u, seed ? *seed : T(), ctl, fix_status);
    ACC_HIP_CHECK(hipGetLastError());
  }
  return ACC_SUCCESS;
}

int fold_dim(acc_stream_t s, acc_type t, acc_op op, const void* seed, const void* in, int64_t outer,
             int64_t inner, void* out) {
  return with_type(t, [&](auto z) {
    using T = decltype(z);
    return with_op<T>(op, [&](auto f) {
      using Op = decltype(f);
      return fold_seg_launch<T, Op, UniformStarts>(s, (const T*)in, (T*)out, UniformStarts{inner, 0}, outer,
                                                   outer * inner, (const T*)seed, false, nullptr, 0);
    });
  });
}

}  // namespace acc

using namespace acc;

extern "C" int acc_hip_fold_seg(acc_stream_t s, acc_type t, acc_op op, const void* seed, const void* in,
                                int64_t outer, int64_t inner, acc_type offset_type, const void* offsets,
                                int64_t nseg, void* out) {
  if (!s) return fail(ACC_ERROR_INVALID_VALUE, "foldSeg: null stream");
  if (outer < 0 || inner < 0 || nseg < 0) return fail(ACC_ERROR_INVALID_VALUE, "foldSeg: negative extent");
  if (offset_type != ACC_INT32 && offset_type != ACC_INT64)
    return fail(ACC_ERROR_UNSUPPORTED, "foldSeg: offsets must be Int32 or Int64");
  if (outer == 0 || nseg == 0) return ACC_SUCCESS;
  if (!in || !out || !offsets) return fail(ACC_ERROR_INVALID_VALUE, "foldSeg: null array");
  const size_t B = type_size(t);
  if (!B) return fail(ACC_ERROR_UNSUPPORTED, "foldSeg: element type %d", (int)t);
  ACC_CHECK(set_device(s->device));
  const size_t I = offset_type == ACC_INT32 ? 4 : 8;
  return with_type(t, [&](auto z) {
    using T = decltype(z);
    return with_op<T>(op, [&](auto f) {
      using Op = decltype(f);
      for (int64_t r = 0; r < outer; ++r) {
        const T* rin = (const T*)in + r * inner;
        T* rout = (T*)out + r * nseg;
        int rc;
        if (I == 4)
          rc = fold_seg_launch<T, Op, OffsetStarts<int32_t>>(
              s, rin, rout, OffsetStarts<int32_t>{(const int32_t*)offsets, nullptr, nullptr, 0}, nseg, inner,
              (const T*)seed, true, offsets, 4);
        else
          rc = fold_seg_launch<T, Op, OffsetStarts<int64_t>>(
              s, rin, rout, OffsetStarts<int64_t>{(const int64_t*)offsets, nullptr, nullptr, 0}, nseg, inner,
              (const T*)seed, true, offsets, 8);
        if (rc) return rc;
      }
      return (int)ACC_SUCCESS;
    });
  });
}

// ------------------------------------------------------------ tile plans
// A segment descriptor's tile plan, built once and used by every foldSeg over
// the same offsets (one k_tile_seg for all of them).  The plan holds the
// results of k_tile_seg for one element type, element count and alignment.
struct acc_segplan_s {
  int device = 0;
  void* buf = nullptr;  // first[ntiles + 1] (256-byte aligned), then slots[ntiles * SLOT]
  size_t cap = 0;
  bool built = false;
  acc_type type = ACC_INT32;
  int64_t n = 0, nseg = 0, shift = 0, ntiles = 0;
  acc_type offset_type = ACC_INT64;
  const void* offsets = nullptr;
};

extern "C" int acc_hip_seg_plan_create(int device, acc_segplan_t* plan) {
  if (!plan) return fail(ACC_ERROR_INVALID_VALUE, "segment plan: null argument");
  ACC_CHECK(set_device(device));
  auto p = new acc_segplan_s();
  p->device = device;
  *plan = p;
  return ACC_SUCCESS;
}

extern "C" int acc_hip_seg_plan_destroy(acc_segplan_t plan) {
  if (!plan) return ACC_SUCCESS;
  int rc = ACC_SUCCESS;
  if (plan->buf) {
    ACC_CHECK(set_device(plan->device));
    const hipError_t e = hipDeviceSynchronize();  // folds still using it
    if (e != hipSuccess) rc = hip_fail(e, "segment plan: synchronising before the free");
    const int rf = dev_free(plan->device, plan->buf);
    if (!rc) rc = rf;
  }
  delete plan;
  return rc;
}

static int plan_reserve(acc_segplan_t plan, size_t bytes) {
  if (bytes <= plan->cap) return ACC_SUCCESS;
  if (plan->buf) {
    // folds on any stream may still read the old plan
    ACC_HIP_CHECK(hipDeviceSynchronize());
    ACC_CHECK(dev_free(plan->device, plan->buf));
    plan->buf = nullptr;
    plan->cap = 0;
  }
  ACC_CHECK(dev_malloc(plan->device, bytes, &plan->buf));
  plan->cap = bytes;
  return ACC_SUCCESS;
}

extern "C" int acc_hip_seg_plan_build(acc_stream_t s, acc_segplan_t plan, acc_type t, const void* in, int64_t n,
                                      acc_type offset_type, const void* offsets, int64_t nseg) {
  if (!s || !plan) return fail(ACC_ERROR_INVALID_VALUE, "segment plan: null argument");
  if (s->device != plan->device) return fail(ACC_ERROR_INVALID_VALUE, "segment plan: stream on another device");
  if (n < 0 || nseg < 0) return fail(ACC_ERROR_INVALID_VALUE, "segment plan: negative extent");
  if (offset_type != ACC_INT32 && offset_type != ACC_INT64)
    return fail(ACC_ERROR_UNSUPPORTED, "segment plan: offsets must be Int32 or Int64");
  const size_t B = type_size(t);
  if (!B) return fail(ACC_ERROR_UNSUPPORTED, "segment plan: element type %d", (int)t);
  if (nseg > 0 && !offsets) return fail(ACC_ERROR_INVALID_VALUE, "segment plan: null offsets");
  if ((uintptr_t)in % B) return fail(ACC_ERROR_INVALID_VALUE, "segment plan: input not element-aligned");
  ACC_CHECK(set_device(s->device));
  plan->built = false;
  return with_type(t, [&](auto z) {
    using T = decltype(z);
    constexpr int64_t TILE = (int64_t)SegTune<T>::BLOCK * SegTune<T>::K * Vec<T>::N;
    const int64_t shift = (int64_t)(((uintptr_t)in % 16) / sizeof(T));
    const int64_t ntiles = n == 0 ? 0 : ceil_div(n + shift, TILE);
    if (ntiles >= 0x7fffffffLL) return fail(ACC_ERROR_UNSUPPORTED, "segment plan: too many elements");
    const size_t first_bytes = ((size_t)(ntiles + 1) * sizeof(int64_t) + 255) / 256 * 256;
    const size_t bytes = first_bytes + (size_t)ntiles * SLOT * sizeof(int32_t);
    ACC_CHECK(plan_reserve(plan, bytes));
    if (ntiles > 0 && nseg > 0)
      ACC_CHECK(launch_tile_plan<T>(s, offsets, offset_type == ACC_INT32 ? 4 : 8, nseg, TILE, ntiles, shift,
                                    (int64_t*)plan->buf, (int32_t*)((char*)plan->buf + first_bytes), nullptr,
                                    false, T()));
    plan->type = t;
    plan->n = n;
    plan->nseg = nseg;
    plan->shift = shift;
    plan->ntiles = ntiles;
    plan->offset_type = offset_type;
    plan->offsets = offsets;
    plan->built = true;
    return (int)ACC_SUCCESS;
  });
}

extern "C" int acc_hip_seg_plan_build_lengths(acc_stream_t s, acc_segplan_t plan, acc_type t, const void* in,
                                              int64_t n, acc_type len_type, const void* lengths, int64_t nseg,
                                              void* offsets) {
  if (!s || !plan) return fail(ACC_ERROR_INVALID_VALUE, "segment plan: null argument");
  if (s->device != plan->device) return fail(ACC_ERROR_INVALID_VALUE, "segment plan: stream on another device");
  if (n < 0 || nseg < 0) return fail(ACC_ERROR_INVALID_VALUE, "segment plan: negative extent");
  if (len_type != ACC_INT32 && len_type != ACC_INT64)
    return fail(ACC_ERROR_UNSUPPORTED, "segment plan: lengths must be Int32 or Int64");
  const size_t B = type_size(t);
  if (!B) return fail(ACC_ERROR_UNSUPPORTED, "segment plan: element type %d", (int)t);
  if (!offsets || (nseg > 0 && !lengths)) return fail(ACC_ERROR_INVALID_VALUE, "segment plan: null array");
  if ((uintptr_t)in % B) return fail(ACC_ERROR_INVALID_VALUE, "segment plan: input not element-aligned");
  if (nseg >= ((int64_t)1 << 40)) return fail(ACC_ERROR_UNSUPPORTED, "segment plan: too many segments");
  ACC_CHECK(set_device(s->device));
  plan->built = false;
  return with_type(t, [&](auto z) {
    using T = decltype(z);
    constexpr int64_t TILE = (int64_t)SegTune<T>::BLOCK * SegTune<T>::K * Vec<T>::N;
    const int64_t shift = (int64_t)(((uintptr_t)in % 16) / sizeof(T));
    const int64_t ntiles = n == 0 ? 0 : ceil_div(n + shift, TILE);
    if (ntiles >= 0x7fffffffLL) return fail(ACC_ERROR_UNSUPPORTED, "segment plan: too many elements");
    const size_t first_bytes = ((size_t)(ntiles + 1) * sizeof(int64_t) + 255) / 256 * 256;
    ACC_CHECK(plan_reserve(plan, first_bytes + (size_t)ntiles * SLOT * sizeof(int32_t)));
    const int64_t ltiles = ceil_div(nseg + 1, (int64_t)SP_LT);
    const size_t I = len_type == ACC_INT32 ? 4 : 8;
    SyncCtl* ctl;
    uint64_t* status;
    ACC_CHECK(stream_workspace(s, (size_t)ltiles * (I == 4 ? 1 : 2) * 8, &ctl, &status));
    int64_t* first = (int64_t*)plan->buf;
    int32_t* slots = (int32_t*)((char*)plan->buf + first_bytes);
    if (I == 4) {
      ScanPlanArgs<int32_t> a{(const int32_t*)lengths, (int32_t*)offsets, nseg, TILE, ntiles, shift, first, slots,
                              ctl, status};
      hipLaunchKernelGGL((k_scan_plan<int32_t>), dim3((unsigned)ltiles), dim3(SP_BLOCK), 0, s->stream, a);
    } else {
      ScanPlanArgs<int64_t> a{(const int64_t*)lengths, (int64_t*)offsets, nseg, TILE, ntiles, shift, first, slots,
                              ctl, status};
      hipLaunchKernelGGL((k_scan_plan<int64_t>), dim3((unsigned)ltiles), dim3(SP_BLOCK), 0, s->stream, a);
    }
    ACC_HIP_CHECK(hipGetLastError());
    plan->type = t;
    plan->n = n;
    plan->nseg = nseg;
    plan->shift = shift;
    plan->ntiles = ntiles;
    plan->offset_type = len_type;
    plan->offsets = offsets;
    plan->built = true;
    return (int)ACC_SUCCESS;
  });
}

extern "C" int acc_hip_fold_seg_planned(acc_stream_t s, acc_segplan_t plan, acc_type t, acc_op op, const void* seed,
                                        const void* in, int64_t outer, void* out) {
  if (!s || !plan) return fail(ACC_ERROR_INVALID_VALUE, "foldSeg (planned): null argument");
  if (!plan->built) return fail(ACC_ERROR_INVALID_VALUE, "foldSeg (planned): the plan has not been built");
  if (s->device != plan->device) return fail(ACC_ERROR_INVALID_VALUE, "foldSeg (planned): stream on another device");
  if (t != plan->type) return fail(ACC_ERROR_INVALID_VALUE, "foldSeg (planned): element type differs from the plan's");
  if (outer < 0) return fail(ACC_ERROR_INVALID_VALUE, "foldSeg (planned): negative extent");
  const int64_t inner = plan->n, nseg = plan->nseg;
  if (outer == 0 || nseg == 0) return ACC_SUCCESS;
  if (!in || !out) return fail(ACC_ERROR_INVALID_VALUE, "foldSeg (planned): null array");
  const size_t B = type_size(t);
  if ((uintptr_t)in % B) return fail(ACC_ERROR_INVALID_VALUE, "foldSeg (planned): input not element-aligned");
  // every row must sit at the plan's position relative to a 16-byte boundary
  if ((int64_t)(((uintptr_t)in % 16) / B) != plan->shift || (outer > 1 && (inner * (int64_t)B) % 16 != 0))
    return fail(ACC_ERROR_INVALID_VALUE, "foldSeg (planned): input alignment differs from the plan's");
  ACC_CHECK(set_device(s->device));
  const size_t I = plan->offset_type == ACC_INT32 ? 4 : 8;
  const size_t first_bytes = ((size_t)(plan->ntiles + 1) * sizeof(int64_t) + 255) / 256 * 256;
  const TilePlan tp{(const int64_t*)plan->buf, (const int32_t*)((const char*)plan->buf + first_bytes)};
  const void* offsets = plan->offsets;
  return with_type(t, [&](auto z) {
    using T = decltype(z);
    return with_op<T>(op, [&](auto f) {
      using Op = decltype(f);
      for (int64_t r = 0; r < outer; ++r) {
        const T* rin = (const T*)in + r * inner;
        T* rout = (T*)out + r * nseg;
        int rc;
        if (I == 4)
          rc = fold_seg_launch<T, Op, OffsetStarts<int32_t>>(
              s, rin, rout, OffsetStarts<int32_t>{(const int32_t*)offsets, nullptr, nullptr, 0}, nseg, inner,
              (const T*)seed, true, offsets, 4, &tp);
        else
          rc = fold_seg_launch<T, Op, OffsetStarts<int64_t>>(
              s, rin, rout, OffsetStarts<int64_t>{(const int64_t*)offsets, nullptr, nullptr, 0}, nseg, inner,
              (const T*)seed, true, offsets, 8, &tp);
        if (rc) return rc;
      }
      return (int)ACC_SUCCESS;
    });
  });
}

extern "C" int acc_hip_fold_seg_planned_multi(acc_stream_t s, acc_segplan_t plan, acc_type t, int count,
                                              const acc_op* ops, const void* const* seeds, const void* const* ins,
                                              void* const* outs) {
  if (!s || !plan || !ops || !ins || !outs) return fail(ACC_ERROR_INVALID_VALUE, "foldSeg (planned, several): null argument");
  if (!plan->built) return fail(ACC_ERROR_INVALID_VALUE, "foldSeg (planned, several): the plan has not been built");
  if (s->device != plan->device)
    return fail(ACC_ERROR_INVALID_VALUE, "foldSeg (planned, several): stream on another device");
  if (t != plan->type)
    return fail(ACC_ERROR_INVALID_VALUE, "foldSeg (planned, several): element type differs from the plan's");
  if (count < 1 || count > kMultiFolds)
    return fail(ACC_ERROR_INVALID_VALUE, "foldSeg (planned, several): 1 to %d folds", kMultiFolds);
  const size_t B = type_size(t);
  for (int i = 0; i < count; ++i) {
    if (!ins[i] || !outs[i]) return fail(ACC_ERROR_INVALID_VALUE, "foldSeg (planned, several): null array %d", i);
    if ((uintptr_t)ins[i] % B || (int64_t)(((uintptr_t)ins[i] % 16) / B) != plan->shift)
      return fail(ACC_ERROR_INVALID_VALUE, "foldSeg (planned, several): input %d alignment differs from the plan's", i);
  }
  const int64_t n = plan->n, nseg = plan->nseg, ntiles = plan->ntiles;
  if (nseg == 0) return ACC_SUCCESS;
  ACC_CHECK(set_device(s->device));
  const size_t I = plan->offset_type == ACC_INT32 ? 4 : 8;
  const size_t first_bytes = ((size_t)(ntiles + 1) * sizeof(int64_t) + 255) / 256 * 256;
  const TilePlan tp{(const int64_t*)plan->buf, (const int32_t*)((const char*)plan->buf + first_bytes)};
  const void* offsets = plan->offsets;
  return with_type(t, [&](auto z) {
    using T = decltype(z);
    constexpr int BLOCK = SegTune<T>::BLOCK, K = SegTune<T>::K, CAP = SegTune<T>::CAP;
    constexpr int FIX = 256;
    if (ntiles == 0) {  // no elements: every segment is empty
      for (int i = 0; i < count; ++i)
        if (seeds && seeds[i]) ACC_CHECK(acc_hip_fill_async(s, outs[i], seeds[i], sizeof(T), (size_t)nseg));
      return (int)ACC_SUCCESS;
    }
    // per fold: its own tile records (untagged scratch); one block of status
    // words per fold for the shared fix-up
    auto al256 = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t cv_bytes = al256((size_t)ntiles * sizeof(T)), cf_bytes = al256((size_t)ntiles);
    const size_t carry_bytes = al256(2 * cv_bytes + cf_bytes + (size_t)ntiles * 8);
    const int64_t nfix = ceil_div(ntiles, (int64_t)FIX);
    SyncCtl* ctl;
    uint64_t* status;
    void* scr;
    ACC_CHECK(stream_workspace(s, (size_t)count * nfix * Tagged<SegV<T>>::NW * 8, &ctl, &status));
    ACC_CHECK(stream_scratch(s, (size_t)count * carry_bytes, &scr));
    MultiFix<T> m{};
    // two Float32 folds, each (+) or max (the common pair: a segment's sum and
    // maximum): their tiles in one k_fold_seg_pair launch
    const bool pair = std::is_same<T, float>::value && count == 2 && 2 * ntiles <= (int64_t)UINT32_MAX &&
                      (ops[0] == ACC_OP_ADD || ops[0] == ACC_OP_MAX) && (ops[1] == ACC_OP_ADD || ops[1] == ACC_OP_MAX);
    auto fold_args = [&](int i, auto st) {
      using Starts = decltype(st);
      st.first = tp.first;
      st.slots = tp.slots;
      st.shift = plan->shift;
      const T* in0 = (const T*)ins[i] - plan->shift;
      return SegArgs<T, Starts>{in0, m.out[i], st, nseg, plan->shift, n + plan->shift, (uint32_t)ntiles,
                                m.has_seed[i], m.seed[i], m.carry[i]};
    };
    for (int i = 0; i < count; ++i) {
      char* cb = (char*)scr + (size_t)i * carry_bytes;
      m.carry[i] = CarryArrays<T>{(T*)cb, (uint8_t*)(cb + cv_bytes), (T*)(cb + cv_bytes + cf_bytes),
                                  (int64_t*)(cb + 2 * cv_bytes + cf_bytes)};
      m.out[i] = (T*)outs[i];
      m.has_seed[i] = seeds && seeds[i] ? 1u : 0u;
      m.seed[i] = m.has_seed[i] ? *(const T*)seeds[i] : T();
      m.op[i] = (int32_t)ops[i];
      if (m.has_seed[i]) ACC_CHECK(launch_fill_empty(s, offsets, I, nseg, m.out[i], m.seed[i]));
      if (pair) continue;  // both folds' tiles in one launch below
      const int rc = with_op<T>(ops[i], [&](auto f) {
        using Op = decltype(f);
        auto launch = [&](auto st) {
          using Starts = decltype(st);
          const SegArgs<T, Starts> a = fold_args(i, st);
          hipLaunchKernelGGL((k_fold_seg<T, Op, Starts, BLOCK, K, CAP>), dim3((unsigned)ntiles), dim3(BLOCK), 0,
                             s->stream, a);
          ACC_HIP_CHECK(hipGetLastError());
          return (int)ACC_SUCCESS;
        };
        if (I == 4) return launch(OffsetStarts<int32_t>{(const int32_t*)offsets, nullptr, nullptr, 0});
        return launch(OffsetStarts<int64_t>{(const int64_t*)offsets, nullptr, nullptr, 0});
      });
      if (rc) return rc;
    }
    if constexpr (std::is_same<T, float>::value) {
      if (pair) {
        auto launch = [&](auto st) {
          using Starts = decltype(st);
          const SegArgs<T, Starts> a0 = fold_args(0, st), a1 = fold_args(1, st);
          auto go = [&](auto o0, auto o1) {
            hipLaunchKernelGGL((k_fold_seg_pair<T, decltype(o0), decltype(o1), Starts, BLOCK, K, CAP>),
                               dim3((unsigned)(2 * ntiles)), dim3(BLOCK), 0, s->stream, a0, a1, (uint32_t)ntiles);
            ACC_HIP_CHECK(hipGetLastError());
            return (int)ACC_SUCCESS;
          };
          if (ops[0] == ACC_OP_ADD) return ops[1] == ACC_OP_ADD ? go(OpAdd{}, OpAdd{}) : go(OpAdd{}, OpMax{});
          return ops[1] == ACC_OP_ADD ? go(OpMax{}, OpAdd{}) : go(OpMax{}, OpMax{});
        };
        const int rc = I == 4 ? launch(OffsetStarts<int32_t>{(const int32_t*)offsets, nullptr, nullptr, 0})
                              : launch(OffsetStarts<int64_t>{(const int64_t*)offsets, nullptr, nullptr, 0});
        if (rc) return rc;
      }
    }
    if (ntiles > 1) {
      hipLaunchKernelGGL((k_seg_fix_multi<T, FIX>), dim3((unsigned)(count * nfix)), dim3(FIX), 0, s->stream, m,
                         (uint32_t)ntiles, (uint32_t)nfix, nseg, ctl, status);
      ACC_HIP_CHECK(hipGetLastError());
    }
    return (int)ACC_SUCCESS;
  });
}
