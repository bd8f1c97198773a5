// Batched execution: many independent scans, folds, segmented folds and
// histograms of uneven sizes in ONE persistent launch, capturable as a
// hipGraph.
//
// The plain path costs one or more kernel launches per problem (and the
// reference's path 1-4 launches plus host marshalling, PTX/Execute.hs:835-881),
// so small problems are dominated by launch latency.  Here every problem is
// cut into runs of work; the runs of all problems form one ticket space;
// persistent workgroups take runs in ticket order (a run→problem map gives
// the problem in one load) and dispatch on the problem's class (kind x
// element type x combine function):
//   fold         a run is R consecutive tiles of BLOCK*K*VEC elements,
//                reduced in registers as one piece; intermediate runs publish
//                their aggregates and the last run looks back and writes the
//                result with the seed on the left
//   scan         a problem of at most R tiles is one run; a larger one has
//                one tile per run, each a tile of the single-pass look-back
//                publishing its inclusive prefix
//                Each problem owns a slice of the status words, so no
//                look-back crosses into another.
//   foldSeg      a run is 32 consecutive segments, their elements cut into
//                one-vector-per-lane chunks dealt to the waves (the reference
//                uses a warp per segment, PTX/CodeGen/FoldSeg.hs:274-458,
//                which leaves a wave one load in flight on short segments)
//   histogram    a run is 128 Ki index positions folded into a private LDS
//                copy of the bins, then flushed with one global atomic per
//                non-zero bin;
//                the problem's first run writes the initial bins and
//                publishes a tagged flag, which every run waits for before
//                it flushes (the first run is ticketed before the others, so
//                the wait ends).
// One ticket per run instead of per tile: the ticket counter, which costs
// ~12 ns per ticket at the device (measured), bounded the first version of
// this kernel (309 k tiles = 3.8 ms); runs divide the tickets by R.
//
// The workgroup size, the tile size (vectors per thread) and the run length
// are chosen by an autotuner that times the candidates on the batch itself
// and remembers the winner per batch signature in a small text file.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include <vector>
#include <algorithm>
#include <chrono>
#include "scan_tile.hpp"
#include "boundary.hpp"
#include "batch_internal.hpp"

namespace acc {

enum : int32_t { BK_SCAN = 0, BK_FOLD = 1, BK_FOLD_SEG = 2, BK_HIST = 3, BK_STENCIL = 4 };

struct BatchDesc {  // one device descriptor per problem (96 bytes, asserted below)
  int32_t cls;          // kind * 8 + type index * 2 + op index
  uint32_t has_seed;
  int64_t n;
  const void* in;
  void* out;
  uint64_t seed;        // bit pattern
  uint64_t status_off;  // first status word of this problem
  uint32_t ntiles;      // scan/fold: tiles; foldSeg: segments; histogram: indices / HIST_RUN;
                        // stencil: column strips
  uint32_t run0;        // first ticket of this problem
  uint32_t R;           // scan/fold: tiles per run; stencil: boundary
  uint32_t aux64;       // foldSeg offsets / histogram indices are Int64
  const void* aux;      // foldSeg: offsets; stencil: weights
  int64_t m;            // foldSeg: segments; histogram: bins; stencil: columns (n: rows)
  uint64_t fill;        // histogram: constant source value; stencil: the constant boundary's value
  uint64_t reserved;    // pads to the host's descriptor stride (96 bytes)
};

struct BatchArgs {
  const BatchDesc* desc;
  const uint32_t* run_prob;  // ticket -> problem
  uint32_t total_runs;
  SyncCtl* ctl;
  uint64_t* status;
};

constexpr int SEG_RUN = 32;          // segments per foldSeg run
constexpr int64_t HIST_RUN = 131072; // index positions per histogram run
constexpr int HIST_MAX_BINS = 4096;  // bins held in LDS (8-byte bins: 32 KiB)

template <typename T> __device__ __forceinline__ T bits_to(uint64_t b) {
  T v;
  __builtin_memcpy(&v, &b, sizeof(T));
  return v;
}

// Fold of run r: tiles [r*R, min(ntiles, r*R + R)) reduced in registers as
// one piece, then one block reduction and one status slot per run; the
// problem's last run looks back over the runs before it and writes the
// result with the seed on the left.  Every run depends only on runs with
// lower tickets, so the look-back ends, and no run waits inside its tiles.
template <typename T, typename Op, int BLOCK, int K>
__device__ __forceinline__ void fold_run(const BatchDesc& d, uint64_t* status, uint32_t r, uint32_t epoch,
                                         T* s_wave) {
  constexpr int VEC = Vec<T>::N;
  constexpr int W = BLOCK / 64;
  constexpr int64_t TILE = (int64_t)BLOCK * K * VEC;
  constexpr int NW = Tagged<T>::NW;
  Op op{};
  const int lane = lane_id(), w = threadIdx.x / 64;
  const T* in = (const T*)d.in;
  const T id = identity<T>(op);
  const uint32_t nruns = (d.ntiles + d.R - 1) / d.R;
  const uint32_t t0 = r * d.R, t1 = min(d.ntiles, t0 + d.R);
  T acc = id;
  for (uint32_t t = t0; t < t1; ++t) {
    const int64_t base = (int64_t)t * TILE;
    Vec<T> v[K];
    if (base + TILE <= d.n) {
#pragma unroll
      for (int k = 0; k < K; ++k) v[k] = load16_stream(in + base + ((int64_t)(w * K + k) * 64 + lane) * VEC);
#pragma unroll
      for (int k = 0; k < K; ++k)
#pragma unroll
        for (int j = 0; j < VEC; ++j) acc = op(acc, v[k].e[j]);
    } else {
      // the problem's last tile: a vector past the end re-reads the vector
      // holding the last element (an aligned 16-byte block with a valid
      // element never crosses a page), so every load is one unconditional
      // dwordx4, and the elements past the end are replaced by the identity
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int64_t e0 = base + ((int64_t)(w * K + k) * 64 + lane) * VEC;
        v[k] = load16_stream(in + min(e0, (d.n - 1) / VEC * VEC));
      }
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int64_t e0 = base + ((int64_t)(w * K + k) * 64 + lane) * VEC;
#pragma unroll
        for (int j = 0; j < VEC; ++j) acc = op(acc, e0 + j < d.n ? v[k].e[j] : id);
      }
    }
  }
  const T wr = wave_reduce(acc, op);
  if (lane == 0) s_wave[w] = wr;
  __syncthreads();
  if (w == 0) {
    T agg = id;
#pragma unroll
    for (int i = 0; i < W; ++i) agg = op(agg, s_wave[i]);
    uint64_t* slot = status + (size_t)r * NW;
    const bool first = r == 0, last = r + 1 == nruns;
    if (!last) {
      if (lane == 0) Tagged<T>::publish(slot, agg, make_tag(epoch, first ? ST_PREFIX : ST_AGGREGATE));
    } else {
      T res = agg;
      if (!first) res = op(wave_lookback<T>(status, r, epoch, op), res);
      if (d.has_seed) res = op(bits_to<T>(d.seed), res);  // seed on the left
      if (lane == 0) gstore((T*)d.out, res);
    }
  }
}

// Scan of run r.  A problem of at most R tiles is one run, scanned tile
// after tile by one workgroup (each tile's look-back finds the inclusive
// prefix the same workgroup published for the tile before); a larger problem
// has one tile per run.  (Runs of several consecutive tiles of a large scan
// chain every run behind the previous one -- a run's first tile needs the
// aggregate of the previous run's last tile -- which makes the problem's
// scan serial, one tile at a time, with every later run's workgroup holding
// its slot while it waits.)
template <typename T, typename Op, int BLOCK, int K>
__device__ __forceinline__ void scan_run(const BatchDesc& d, uint64_t* status, uint32_t r, uint32_t epoch,
                                         unsigned char* s_raw) {
  constexpr int64_t TILE = (int64_t)BLOCK * K * Vec<T>::N;
  T* s_wave = (T*)s_raw;
  T* s_prefix = s_wave + BLOCK / 64;
  const uint32_t t0 = r * d.R, t1 = min(d.ntiles, t0 + d.R);
  for (uint32_t t = t0; t < t1; ++t) {
    const int64_t base = (int64_t)t * TILE;
    // scanl1 without a seed; with a seed the exclusive scanl (total at out[n])
    TileIO<T> io{(const T*)d.in, (T*)d.out, d.n, 0u, d.ntiles - 1, d.has_seed, d.has_seed, bits_to<T>(d.seed),
                 d.has_seed ? (T*)d.out + d.n : nullptr, nullptr, status};
    if (base + TILE <= d.n) scan_tile<T, Op, BLOCK, K, false, true, false>(io, Op{}, t, base, epoch, s_wave, s_prefix);
    else scan_tile<T, Op, BLOCK, K, false, true, true>(io, Op{}, t, base, epoch, s_wave, s_prefix);
    __syncthreads();
  }
}

template <typename T, typename Op, int BLOCK, int K>
__device__ __forceinline__ void scan_fold_run(const BatchDesc& d, int kind, uint64_t* status, uint32_t r,
                                              uint32_t epoch, unsigned char* s_raw) {
  if (kind == BK_SCAN) scan_run<T, Op, BLOCK, K>(d, status, r, epoch, s_raw);
  else fold_run<T, Op, BLOCK, K>(d, status, r, epoch, (T*)s_raw);
}

// foldSeg / fold1Seg.  A run is SEG_RUN consecutive segments, whose elements
// are contiguous.  The run's elements are cut into chunks of one 16-byte
// vector per lane (64 * VEC elements, aligned; a chunk never spans two
// segments, so a segment's first and last chunks are partial) and the chunks
// are dealt to the waves in blocks of SEG_U: each wave keeps SEG_U vector
// loads in flight, whatever the segments' lengths, and a long segment is
// shared by all the waves of the workgroup.  Per chunk: an in-lane fold of
// the elements inside its segment (identity elsewhere), the wave reduction,
// and a fold into the wave's partial result for that segment (LDS); at the
// end the waves' partials are folded in wave order and written, one
// coalesced store per run.  Empty segments get the seed (fold1Seg: left as
// they are).  The combination order depends only on the segment and the
// workgroup shape, so a result does not depend on the other problems.
constexpr int SEG_U = 8;  // chunks in flight per wave

struct SegTable {  // LDS: the run's segments and the waves' partial results
  int64_t lo[SEG_RUN], hi[SEG_RUN];  // positions (relative to the aligned base)
  int32_t pre[SEG_RUN], end[SEG_RUN];  // first chunk, one past the last chunk
  int32_t chunks;
  int32_t pad;
  uint64_t part[8 * SEG_RUN];  // [wave][segment], T in the low bytes
};

template <typename T, typename Op, int BLOCK>
__device__ __forceinline__ void fold_seg_run(const BatchDesc& d, uint32_t r, SegTable* tb) {
  constexpr int W = BLOCK / 64;
  static_assert(W <= 8, "partials for at most 8 waves");
  constexpr int VEC = Vec<T>::N;
  constexpr int64_t CH = 64 * VEC;
  Op op{};
  const int lane = lane_id(), w = threadIdx.x / 64;
  const T id = identity<T>(op);
  // positions are relative to the 16-byte boundary at or below `in`; the
  // valid ones are [shift, shift + n)
  const int shift = (int)(((uintptr_t)d.in % 16) / sizeof(T));
  const T* base = (const T*)d.in - shift;
  const int64_t last_vec = d.n > 0 ? (shift + d.n - 1) / VEC * VEC : 0;
  const int64_t s0 = (int64_t)r * SEG_RUN;
  const int ns = (int)min((int64_t)SEG_RUN, d.m - s0);
  T* part = (T*)tb->part;
  if (w == 0) {
    int64_t lo = 0, hi = 0;
    int nch = 0;
    if (lane < ns) {
      if (d.aux64) {
        lo = gload((const int64_t*)d.aux + s0 + lane);
        hi = gload((const int64_t*)d.aux + s0 + lane + 1);
      } else {
        lo = gload((const int32_t*)d.aux + s0 + lane);
        hi = gload((const int32_t*)d.aux + s0 + lane + 1);
      }
      lo += shift;
      hi += shift;
      if (hi > lo && d.n > 0) nch = (int)((hi - lo / VEC * VEC + CH - 1) / CH);
    }
    const int incl = wave_inclusive_scan(nch, OpAdd{});
    if (lane < ns) {
      tb->lo[lane] = lo;
      tb->hi[lane] = hi;
      tb->pre[lane] = incl - nch;
      tb->end[lane] = incl;
    }
    if (lane == 63) tb->chunks = incl;
  }
  for (int i = threadIdx.x; i < W * SEG_RUN; i += BLOCK) part[i] = id;
  __syncthreads();
  const int C = tb->chunks;
  // lane j < ns holds the end of segment j: the segment of chunk c is the
  // number of segments that end at or before c
  const int end_l = lane < ns ? tb->end[lane] : 0x7fffffff;
  for (int cb = w * SEG_U; cb < C; cb += W * SEG_U) {
    Vec<T> v[SEG_U];
    int64_t pos[SEG_U];
    int js[SEG_U];
#pragma unroll
    for (int u = 0; u < SEG_U; ++u) {
      const int c = cb + u;
      js[u] = -1;
      pos[u] = 0;
      if (c < C) {  // wave-uniform
        const int j = (int)__popcll(__ballot(end_l <= c));
        js[u] = j;
        pos[u] = tb->lo[j] / VEC * VEC + (int64_t)(c - tb->pre[j]) * CH + lane * VEC;
        // a vector past the segment's end reads the last vector holding an
        // element of the array (then every element of it is masked)
        v[u] = load16_stream(base + min(max(pos[u], (int64_t)0), last_vec));
      }
    }
#pragma unroll
    for (int u = 0; u < SEG_U; ++u) {
      const int j = js[u];
      if (j < 0) break;  // wave-uniform
      const int64_t lo = tb->lo[j], hi = tb->hi[j];
      T x = id;
#pragma unroll
      for (int e = 0; e < VEC; ++e) {
        const int64_t q = pos[u] + e;
        x = op(x, q >= lo && q < hi ? v[u].e[e] : id);
      }
      const T rr = wave_reduce(x, op);
      if (lane == 0) part[w * SEG_RUN + j] = op(part[w * SEG_RUN + j], rr);
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < ns) {
    const int j = threadIdx.x;
    T* out = (T*)d.out + s0;
    if (tb->end[j] == tb->pre[j]) {
      if (d.has_seed) gstore(out + j, bits_to<T>(d.seed));
    } else {
      T res = part[j];
#pragma unroll
      for (int k = 1; k < W; ++k) res = op(res, part[k * SEG_RUN + j]);
      gstore(out + j, d.has_seed ? op(bits_to<T>(d.seed), res) : res);  // seed on the left
    }
  }
  __syncthreads();  // the table is rebuilt by the next run
}

// The histogram: out[idx[i]] = out[idx[i]] (+) fill for 0 <= idx[i] < m, with
// out initially `seed` in every bin (permute (+) of a constant source).
// LDS float32 atomic adds run ~20x slower than integer or float64 ones on
// gfx950 (measured), so float32 bins accumulate in float64 and are rounded
// once, at the flush.
template <typename T> struct HistAcc { using type = T; };
template <> struct HistAcc<float> { using type = double; };

template <typename T>
__device__ __forceinline__ void lds_add(T* p, T v) {
  if constexpr (std::is_floating_point<T>::value) atomicAdd(p, v);
  else __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <typename T>
__device__ __forceinline__ void global_add(T* p, T v) {
  typedef __attribute__((address_space(1))) T gT;
  if constexpr (std::is_same<T, float>::value) (void)__builtin_amdgcn_global_atomic_fadd_f32((gT*)p, v);
  else if constexpr (std::is_same<T, double>::value) (void)__builtin_amdgcn_global_atomic_fadd_f64((gT*)p, v);
  else __hip_atomic_fetch_add((gT*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A run is HIST_RUN index positions (relative to the 16-byte boundary at or
// below the indices; the valid ones are [shift, shift + n)), read as 16-byte
// vectors, HIST_U per thread in flight, and folded into a private LDS copy of
// the bins.
constexpr int HIST_U = 8;

template <typename T, typename I, int BLOCK>
__device__ __forceinline__ void hist_run(const BatchDesc& d, uint64_t* status, uint32_t r, uint32_t epoch,
                                         unsigned char* s_bins) {
  using L = typename HistAcc<T>::type;
  L* bins = (L*)s_bins;
  constexpr int IV = Vec<I>::N;
  T* out = (T*)d.out;
  const int64_t m = d.m;
  const L fill = (L)bits_to<T>(d.fill);
  const uint32_t ready = make_tag(epoch, ST_PREFIX);
  if (r == 0) {
    // the first run initialises the bins, makes the stores visible to every
    // XCD, then publishes the flag
    const T z = bits_to<T>(d.seed);
    for (int64_t b = threadIdx.x; b < m; b += BLOCK) gstore(out + b, z);
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(status, (uint64_t)ready << 32, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  for (int64_t b = threadIdx.x; b < m; b += BLOCK) bins[b] = L(0);
  __syncthreads();
  const int shift = (int)(((uintptr_t)d.in % 16) / sizeof(I));
  const I* base = (const I*)d.in - shift;
  const int64_t v_lo = shift, v_hi = shift + d.n;  // valid positions
  const int64_t last_vec = d.n > 0 ? (v_hi - 1) / IV * IV : 0;
  const int64_t p0 = (int64_t)r * HIST_RUN, p1 = d.n > 0 ? min(v_hi, p0 + HIST_RUN) : 0;
  constexpr int64_t STEP = (int64_t)BLOCK * IV * HIST_U;
  for (int64_t q0 = p0 + (int64_t)threadIdx.x * IV; q0 < p1; q0 += STEP) {
    Vec<I> v[HIST_U];
#pragma unroll
    for (int u = 0; u < HIST_U; ++u) v[u] = load16_stream(base + min(q0 + (int64_t)u * BLOCK * IV, last_vec));
#pragma unroll
    for (int u = 0; u < HIST_U; ++u) {
      const int64_t q = q0 + (int64_t)u * BLOCK * IV;
#pragma unroll
      for (int e = 0; e < IV; ++e) {
        const int64_t j = (int64_t)v[u].e[e];
        if (q + e >= v_lo && q + e < p1 && j >= 0 && j < m) lds_add(&bins[j], fill);
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0)
    while ((uint32_t)(__hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> 32) != ready)
      __builtin_amdgcn_s_sleep(2);
  __syncthreads();
  for (int64_t b = threadIdx.x; b < m; b += BLOCK) {
    const L v = bins[b];
    if (v != L(0)) global_add(&out[b], (T)v);
  }
  __syncthreads();
}

// Stencil (weighted sum over a (2RY+1) x (2RX+1) neighbourhood, taps in
// tuple order: rows outer, columns inner, LLVM/CodeGen/Stencil.hs:126-180;
// boundaries as in LLVM/CodeGen/Stencil.hs:224-253).  A run is a strip of
// ST_STRIP columns by BLOCK/64 * ST_ROWS rows; each wave slides down ST_ROWS
// output rows.  Lane l holds columns l + 64 j (j < ST_NJ) of the strip, so
// every row is read with ST_NJ fully coalesced loads per lane and each input
// row once per wave; horizontal neighbours come from the neighbouring lanes
// (ds_bpermute), the two halo columns by a broadcast load.  Each input row is
// added into the 2RY+1 output rows that use it (accumulators rotate in
// registers), which keeps every output's taps in tuple order.  Row and column
// indices outside the image go through the boundary map (the constant
// boundary substitutes its value), so no element is special-cased.
constexpr int ST_NJ = 2;             // columns per lane
constexpr int ST_STRIP = 64 * ST_NJ;  // columns per run
constexpr int ST_ROWS = 16;    // output rows per wave per run

template <typename T> __device__ __forceinline__ T lane_fetch(const T& v, int src) {
  Words<T> a = to_words(v), r;
#pragma unroll
  for (int i = 0; i < Words<T>::N; ++i) r.w[i] = (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)a.w[i]);
  return from_words<T>(r);
}

template <typename T, int RY, int RX, int BLOCK>
__device__ __forceinline__ void stencil_run(const BatchDesc& d, uint32_t r) {
  constexpr int KH = 2 * RY + 1, KW = 2 * RX + 1;
  OpAdd add{};
  OpMul mul{};
  const int lane = lane_id(), w = threadIdx.x / 64;
  const int64_t H = d.n, Wd = d.m;
  const int64_t strip = r % d.ntiles, rb = r / d.ntiles;
  const int64_t c0 = strip * ST_STRIP;
  const int64_t y0 = (rb * (BLOCK / 64) + w) * ST_ROWS, y1 = min(H, y0 + ST_ROWS);
  if (y0 >= y1) return;  // wave-uniform; no barrier below
  const int bnd = (int)d.R;
  const bool cst = bnd == ACC_BOUNDARY_CONSTANT;
  const T cval = bits_to<T>(d.fill);
  const T* in = (const T*)d.in;
  T* out = (T*)d.out;
  T wt[KH * KW];
#pragma unroll
  for (int i = 0; i < KH * KW; ++i) wt[i] = gload((const T*)d.aux + i);
  // the lane's columns (mapped) and whether each is an output column
  // (column indices fit 32 bits: the batch takes images of < 2^30 columns)
  int32_t col[ST_NJ];
  bool live[ST_NJ];
#pragma unroll
  for (int j = 0; j < ST_NJ; ++j) {
    const int64_t c = c0 + lane + 64 * j;
    live[j] = c < Wd;
    col[j] = (int32_t)(live[j] ? c : bound_index(bnd, c, Wd));
  }
  // halo columns: c0 - RX .. c0 - 1 and c0 + ST_STRIP .. + RX - 1
  int32_t hcol[2 * RX];
  bool hout[2 * RX];
#pragma unroll
  for (int k = 0; k < 2 * RX; ++k) {
    const int64_t c = k < RX ? c0 - RX + k : c0 + ST_STRIP + (k - RX);
    hout[k] = c < 0 || c >= Wd;
    hcol[k] = (int32_t)bound_index(bnd, c, Wd);
  }
  T acc[KH][ST_NJ];  // acc[k]: output row yi - RY + k while input row yi is added
#pragma unroll
  for (int k = 0; k < KH; ++k)
#pragma unroll
    for (int j = 0; j < ST_NJ; ++j) acc[k][j] = T();
  for (int64_t yi = y0 - RY; yi < y1 + RY; ++yi) {
    const bool rout = yi < 0 || yi >= H;
    const T* row = in + bound_index(bnd, yi, H) * Wd;
    T x[ST_NJ], h[2 * RX];
#pragma unroll
    for (int j = 0; j < ST_NJ; ++j) x[j] = (cst && (rout || !live[j])) ? cval : gload(row + col[j]);
#pragma unroll
    for (int k = 0; k < 2 * RX; ++k) h[k] = (cst && (rout || hout[k])) ? cval : gload(row + hcol[k]);
#pragma unroll
    for (int dx = -RX; dx <= RX; ++dx) {
      // e[j]: the value at column (lane + 64 j) + dx of the strip
      T e[ST_NJ];
      if (dx == 0) {
#pragma unroll
        for (int j = 0; j < ST_NJ; ++j) e[j] = x[j];
      } else {
        const int src = (lane + dx) & 63;
        T sh[ST_NJ];
#pragma unroll
        for (int j = 0; j < ST_NJ; ++j) sh[j] = lane_fetch(x[j], src);
        const bool wrap = dx < 0 ? lane + dx < 0 : lane + dx > 63;
        // the halo value of a wrapping lane at the strip's edge
        T he = h[0];
#pragma unroll
        for (int k = 0; k < 2 * RX; ++k) {
          const int hk = dx < 0 ? RX + lane + dx : RX + lane + dx - 64;  // index into h
          if (hk == k) he = h[k];
        }
#pragma unroll
        for (int j = 0; j < ST_NJ; ++j) {
          const T nb = dx < 0 ? (j > 0 ? sh[j - 1] : he) : (j < ST_NJ - 1 ? sh[j + 1] : he);
          e[j] = wrap ? nb : sh[j];
        }
      }
#pragma unroll
      for (int k = 0; k < KH; ++k) {
        const int dy = RY - k;  // this input row is tap row dy of output row yi - dy
        const int64_t yo = yi - dy;
        if (yo >= y0 && yo < y1) {
#pragma unroll
          for (int j = 0; j < ST_NJ; ++j) {
            const T p = mul(wt[(dy + RY) * KW + dx + RX], e[j]);
            acc[k][j] = (dy == -RY && dx == -RX) ? p : add(acc[k][j], p);
          }
        }
      }
    }
    // output row yi - RY has all its taps
    const int64_t yo = yi - RY;
    if (yo >= y0 && yo < y1) {
#pragma unroll
      for (int j = 0; j < ST_NJ; ++j)
        if (live[j]) gstore(out + yo * Wd + c0 + lane + 64 * j, acc[0][j]);
    }
#pragma unroll
    for (int k = 0; k + 1 < KH; ++k)
#pragma unroll
      for (int j = 0; j < ST_NJ; ++j) acc[k][j] = acc[k + 1][j];
  }
}



template <int BLOCK, int K>
__device__ __forceinline__ void batch_run(const BatchDesc& d, uint64_t* st, uint32_t r, uint32_t epoch,
                                          unsigned char* s_raw, unsigned char* s_dyn, SegTable* s_seg) {
  const int kind = d.cls >> 3;
  const int c = d.cls & 7;
  if (kind <= BK_FOLD) {
    switch (c) {
      case 0: scan_fold_run<int32_t, OpAdd, BLOCK, K>(d, kind, st, r, epoch, s_raw); break;
      case 1: scan_fold_run<int32_t, OpMax, BLOCK, K>(d, kind, st, r, epoch, s_raw); break;
      case 2: scan_fold_run<int64_t, OpAdd, BLOCK, K>(d, kind, st, r, epoch, s_raw); break;
      case 3: scan_fold_run<int64_t, OpMax, BLOCK, K>(d, kind, st, r, epoch, s_raw); break;
      case 4: scan_fold_run<float, OpAdd, BLOCK, K>(d, kind, st, r, epoch, s_raw); break;
      case 5: scan_fold_run<float, OpMax, BLOCK, K>(d, kind, st, r, epoch, s_raw); break;
      case 6: scan_fold_run<double, OpAdd, BLOCK, K>(d, kind, st, r, epoch, s_raw); break;
      default: scan_fold_run<double, OpMax, BLOCK, K>(d, kind, st, r, epoch, s_raw); break;
    }
  } else if (kind == BK_FOLD_SEG) {
    switch (c) {
      case 0: fold_seg_run<int32_t, OpAdd, BLOCK>(d, r, s_seg); break;
      case 1: fold_seg_run<int32_t, OpMax, BLOCK>(d, r, s_seg); break;
      case 2: fold_seg_run<int64_t, OpAdd, BLOCK>(d, r, s_seg); break;
      case 3: fold_seg_run<int64_t, OpMax, BLOCK>(d, r, s_seg); break;
      case 4: fold_seg_run<float, OpAdd, BLOCK>(d, r, s_seg); break;
      case 5: fold_seg_run<float, OpMax, BLOCK>(d, r, s_seg); break;
      case 6: fold_seg_run<double, OpAdd, BLOCK>(d, r, s_seg); break;
      default: fold_seg_run<double, OpMax, BLOCK>(d, r, s_seg); break;
    }
  } else {
    const bool i64 = d.aux64 != 0;
    switch (c >> 1) {
      case 0: i64 ? hist_run<int32_t, int64_t, BLOCK>(d, st, r, epoch, s_dyn)
                  : hist_run<int32_t, int32_t, BLOCK>(d, st, r, epoch, s_dyn); break;
      case 1: i64 ? hist_run<int64_t, int64_t, BLOCK>(d, st, r, epoch, s_dyn)
                  : hist_run<int64_t, int32_t, BLOCK>(d, st, r, epoch, s_dyn); break;
      case 2: i64 ? hist_run<float, int64_t, BLOCK>(d, st, r, epoch, s_dyn)
                  : hist_run<float, int32_t, BLOCK>(d, st, r, epoch, s_dyn); break;
      default: i64 ? hist_run<double, int64_t, BLOCK>(d, st, r, epoch, s_dyn)
                   : hist_run<double, int32_t, BLOCK>(d, st, r, epoch, s_dyn); break;
    }
  }
}

template <int BLOCK, int K>
__global__ __launch_bounds__(BLOCK) void k_batch(BatchArgs a) {
  extern __shared__ __align__(16) unsigned char s_dyn[];  // histogram bins (dynamic LDS)
  __shared__ uint64_t s_tick;
  __shared__ __align__(16) unsigned char s_raw[(BLOCK / 64 + 1) * sizeof(double)];
  __shared__ SegTable s_seg;
  if (threadIdx.x == 0) s_tick = take_ticket(a.ctl);
  __syncthreads();
  const uint64_t tk0 = s_tick;
  const uint32_t epoch = (uint32_t)(tk0 >> 32);
  uint32_t g = (uint32_t)tk0;
  while (g < a.total_runs) {
    // (the descriptors' pointers are loaded from memory, so every access
    // through them names the global address space explicitly: flat accesses
    // would also count against the LDS counter)
    const BatchDesc d = a.desc[a.run_prob[g]];
    batch_run<BLOCK, K>(d, a.status + d.status_off, g - d.run0, epoch, s_raw, s_dyn, &s_seg);
    __syncthreads();
    if (threadIdx.x == 0) s_tick = take_ticket(a.ctl);
    __syncthreads();
    g = (uint32_t)s_tick;
  }
  block_retire(a.ctl, gridDim.x, epoch);
}

// The stencil problems of a batch: one workgroup per run (stencil runs never
// wait on one another, so there are no tickets), launched after the
// persistent kernel on the same stream.  A kernel of its own keeps the
// stencil's registers out of the persistent kernel's budget (in it they cost
// a wave per SIMD of occupancy for the smaller tiles, measured).
constexpr int ST_BLOCK = 256;

__global__ __launch_bounds__(ST_BLOCK) void k_batch_stencil(const BatchDesc* desc, const uint32_t* run_prob) {
  const uint32_t g = blockIdx.x;
  const BatchDesc d = desc[run_prob[g]];
  switch ((d.cls & 7) >> 1) {
    case 0: stencil_run<int32_t, 1, 1, ST_BLOCK>(d, g - d.run0); break;
    case 1: stencil_run<int64_t, 1, 1, ST_BLOCK>(d, g - d.run0); break;
    case 2: stencil_run<float, 1, 1, ST_BLOCK>(d, g - d.run0); break;
    default: stencil_run<double, 1, 1, ST_BLOCK>(d, g - d.run0); break;
  }
}

// (workgroup size, vectors per thread) instantiations offered to the tuner
struct BatchShape {
  int block, k;
};
static const BatchShape kShapes[] = {{256, 2}, {256, 4}, {256, 8}, {512, 2}, {512, 4}};
static const int kRuns[] = {1, 4, 16};

static const void* shape_kernel(int block, int k) {
  if (block == 256 && k == 2) return (const void*)k_batch<256, 2>;
  if (block == 256 && k == 4) return (const void*)k_batch<256, 4>;
  if (block == 256 && k == 8) return (const void*)k_batch<256, 8>;
  if (block == 512 && k == 2) return (const void*)k_batch<512, 2>;
  if (block == 512 && k == 4) return (const void*)k_batch<512, 4>;
  return nullptr;
}

static int type_index(acc_type t) {
  switch (t) {
    case ACC_INT32: return 0;
    case ACC_INT64: return 1;
    case ACC_FLOAT32: return 2;
    case ACC_FLOAT64: return 3;
    default: return -1;
  }
}

bool batch_shape_supported(int block, int k) { return shape_kernel(block, k) != nullptr; }

int batch_validate(const acc_batch_problem* problems, int64_t count) {
  for (int64_t i = 0; i < count; ++i) {
    const acc_batch_problem& p = problems[i];
    const long long li = (long long)i;
    if (p.kind < ACC_BATCH_SCAN || p.kind > ACC_BATCH_STENCIL)
      return fail(ACC_ERROR_UNSUPPORTED, "batch: problem %lld: kind %d", li, p.kind);
    if (type_index((acc_type)p.type) < 0)
      return fail(ACC_ERROR_UNSUPPORTED, "batch: problem %lld: element type %d", li, p.type);
    if (p.kind == ACC_BATCH_STENCIL) {
      const int ry = p.params & 0xff, rx = (p.params >> 8) & 0xff, bd = (p.params >> 16) & 0xff;
      if (ry != 1 || rx != 1) return fail(ACC_ERROR_UNSUPPORTED, "batch: problem %lld: stencils of 3 x 3", li);
      if (bd != ACC_BOUNDARY_CLAMP && bd != ACC_BOUNDARY_MIRROR && bd != ACC_BOUNDARY_WRAP &&
          bd != ACC_BOUNDARY_CONSTANT)
        return fail(ACC_ERROR_UNSUPPORTED, "batch: problem %lld: boundary %d", li, bd);
      if (p.n <= 0 || p.m <= 0) return fail(ACC_ERROR_UNSUPPORTED, "batch: problem %lld: empty image", li);
      if (p.n > ((int64_t)1 << 31) / ST_ROWS || p.m >= ((int64_t)1 << 30))
        return fail(ACC_ERROR_UNSUPPORTED, "batch: problem %lld: image too large", li);
      if (!p.in || !p.out || !p.aux) return fail(ACC_ERROR_INVALID_VALUE, "batch: problem %lld: null array", li);
      if ((uintptr_t)p.in % type_size((acc_type)p.type) || (uintptr_t)p.out % type_size((acc_type)p.type) ||
          (uintptr_t)p.aux % type_size((acc_type)p.type))
        return fail(ACC_ERROR_INVALID_VALUE, "batch: problem %lld: arrays not element-aligned", li);
      continue;
    }
    if (p.kind == ACC_BATCH_HISTOGRAM ? p.op != ACC_OP_ADD : (p.op != ACC_OP_ADD && p.op != ACC_OP_MAX))
      return fail(ACC_ERROR_UNSUPPORTED, "batch: problem %lld: combine %d", li, p.op);
    if (p.n < 0 || p.m < 0) return fail(ACC_ERROR_INVALID_VALUE, "batch: problem %lld: negative size", li);
    if ((p.kind == ACC_BATCH_FOLD_SEG || p.kind == ACC_BATCH_HISTOGRAM) && p.aux_type != ACC_INT32 &&
        p.aux_type != ACC_INT64)
      return fail(ACC_ERROR_INVALID_VALUE, "batch: problem %lld: offsets / indices must be Int32 or Int64", li);
    if (p.kind == ACC_BATCH_SCAN || p.kind == ACC_BATCH_FOLD) {
      if (p.kind == ACC_BATCH_FOLD && p.n == 0 && !p.has_seed)
        return fail(ACC_ERROR_EMPTY_ARRAY, "batch: problem %lld: fold1 of an empty array", li);
      if (p.n == 0)
        return fail(ACC_ERROR_UNSUPPORTED, "batch: problem %lld: empty problems are served by the plain calls", li);
      if (!p.in || !p.out || (uintptr_t)p.in % 16 || (p.kind == ACC_BATCH_SCAN && (uintptr_t)p.out % 16))
        return fail(ACC_ERROR_INVALID_VALUE, "batch: problem %lld: arrays must be non-null and 16-byte aligned", li);
    } else if (p.kind == ACC_BATCH_FOLD_SEG) {
      if (p.m == 0) return fail(ACC_ERROR_UNSUPPORTED, "batch: problem %lld: no segments", li);
      if (!p.out || !p.aux || (p.n && !p.in)) return fail(ACC_ERROR_INVALID_VALUE, "batch: problem %lld: null array", li);
    } else {
      if (p.m == 0 || p.m > HIST_MAX_BINS)
        return fail(ACC_ERROR_UNSUPPORTED, "batch: problem %lld: histograms of 1..%d bins", li, HIST_MAX_BINS);
      if (!p.out || (p.n && !p.in)) return fail(ACC_ERROR_INVALID_VALUE, "batch: problem %lld: null array", li);
    }
  }
  return ACC_SUCCESS;
}

// The layout of one batch for a (shape, run length): descriptors, the
// run→problem map, status words.  Shared by acc_hip_batch_* and the server's
// fused rounds.
int batch_plan(const acc_batch_problem* problems, int64_t count, int block, int k, int R, BatchPlan* plan) {
  plan->desc.assign(count, BatchDescBytes{});
  plan->run_prob.clear();
  plan->stencil_prob.clear();
  uint64_t words = 0, runs = 0;
  size_t lds = 0;
  for (int64_t i = 0; i < count; ++i) {
    const acc_batch_problem& p = problems[i];
    const size_t B = type_size((acc_type)p.type);
    BatchDesc d;
    memset(&d, 0, sizeof d);
    d.cls = p.kind * 8 + type_index((acc_type)p.type) * 2 + (p.op == ACC_OP_MAX ? 1 : 0);
    d.has_seed = p.has_seed ? 1u : 0u;
    d.n = p.n;
    d.in = p.in;
    d.out = p.out;
    d.seed = p.seed;
    d.status_off = words;
    d.aux = p.aux;
    d.aux64 = p.aux_type == ACC_INT64 ? 1u : 0u;
    d.m = p.m;
    d.fill = p.fill;
    uint64_t nr = 0;
    if (p.kind == ACC_BATCH_SCAN || p.kind == ACC_BATCH_FOLD) {
      const int64_t tile = (int64_t)block * k * (int64_t)(16 / B);
      const int64_t nt = (p.n + tile - 1) / tile;
      d.ntiles = (uint32_t)nt;
      // a fold reduces R consecutive tiles per run with one status slot per
      // run; a scan is one run when it has at most R tiles, else one tile per
      // run (one status slot per tile)
      const int64_t r_eff = p.kind == ACC_BATCH_FOLD ? R : (nt <= R ? std::max<int64_t>(nt, 1) : 1);
      d.R = (uint32_t)r_eff;
      nr = (uint64_t)((nt + r_eff - 1) / r_eff);
      words += (p.kind == ACC_BATCH_FOLD ? nr : (uint64_t)nt) * (B == 8 ? 2 : 1);
    } else if (p.kind == ACC_BATCH_FOLD_SEG) {
      d.ntiles = (uint32_t)p.m;
      nr = (uint64_t)((p.m + SEG_RUN - 1) / SEG_RUN);
    } else if (p.kind == ACC_BATCH_STENCIL) {
      // runs of the stencil kernel: column strips x row blocks
      const int64_t strips = (p.m + ST_STRIP - 1) / ST_STRIP;
      const int64_t rows = (int64_t)(ST_BLOCK / 64) * ST_ROWS;
      const uint64_t ns = (uint64_t)(strips * ((p.n + rows - 1) / rows));
      d.ntiles = (uint32_t)strips;
      d.R = (uint32_t)((p.params >> 16) & 0xff);  // boundary
      d.cls = p.kind * 8 + type_index((acc_type)p.type) * 2;
      d.run0 = (uint32_t)plan->stencil_prob.size();
      if (d.run0 + ns > 0x7fffffffULL) return fail(ACC_ERROR_UNSUPPORTED, "batch: too many stencil runs");
      for (uint64_t r = 0; r < ns; ++r) plan->stencil_prob.push_back((uint32_t)i);
      memcpy(&plan->desc[i], &d, sizeof d);
      continue;
    } else {
      // runs over index positions relative to the 16-byte boundary below the
      // indices; run 0 also writes the initial bins (so there is one even
      // without indices)
      const int64_t ib = p.aux_type == ACC_INT64 ? 8 : 4;
      const int64_t shift = (int64_t)(((uintptr_t)p.in % 16) / ib);
      nr = p.n > 0 ? (uint64_t)((shift + p.n + HIST_RUN - 1) / HIST_RUN) : 1;
      d.ntiles = (uint32_t)nr;
      words += 1;
      lds = std::max(lds, (size_t)p.m * (p.type == ACC_FLOAT32 ? 8 : B));  // float32 bins accumulate in float64
    }
    d.run0 = (uint32_t)runs;
    runs += nr;
    if (runs > 0xffffffffULL) return fail(ACC_ERROR_UNSUPPORTED, "batch: too many runs");
    for (uint64_t r = 0; r < nr; ++r) plan->run_prob.push_back((uint32_t)i);
    memcpy(&plan->desc[i], &d, sizeof d);
  }
  plan->block = block;
  plan->k = k;
  plan->R = R;
  plan->total_runs = (uint32_t)runs;
  plan->status_words = words;
  plan->lds_bytes = lds;
  return ACC_SUCCESS;
}

int batch_launch_plan(int device, const BatchPlan& plan, const void* d_desc, const void* d_map,
                      const void* d_smap, void* d_ws, hipStream_t stream) {
  if (!plan.stencil_prob.empty()) {
    hipLaunchKernelGGL(k_batch_stencil, dim3((unsigned)plan.stencil_prob.size()), dim3(ST_BLOCK), 0, stream,
                       (const BatchDesc*)d_desc, (const uint32_t*)d_smap);
    ACC_HIP_CHECK(hipGetLastError());
  }
  if (!plan.total_runs) return ACC_SUCCESS;
  const DeviceInfo* di;
  ACC_CHECK(device_info(device, &di));
  BatchArgs a{(const BatchDesc*)d_desc, (const uint32_t*)d_map, plan.total_runs, (SyncCtl*)d_ws,
              (uint64_t*)((char*)d_ws + kSyncCtlBytes)};
  const void* kern = shape_kernel(plan.block, plan.k);
  if (!kern) return fail(ACC_ERROR_INVALID_VALUE, "batch: shape %d x %d", plan.block, plan.k);
  int per_cu = blocks_per_cu(kern, plan.block);
  if (plan.lds_bytes) {  // the histogram bins' dynamic LDS can lower the occupancy
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, plan.block, plan.lds_bytes) == hipSuccess && nb >= 1)
      per_cu = std::min(per_cu, nb);
    else
      (void)hipGetLastError();
  }
  int64_t grid = (int64_t)di->compute_units * per_cu;
  if (grid > (int64_t)plan.total_runs) grid = plan.total_runs;
  void* params[] = {&a};
  ACC_HIP_CHECK(hipLaunchKernel(kern, dim3((unsigned)grid), dim3(plan.block), params, plan.lds_bytes, stream));
  return ACC_SUCCESS;
}

// the host fills an array of BatchDescBytes that the kernel reads as BatchDesc:
// the strides must be equal
static_assert(sizeof(BatchDesc) == sizeof(BatchDescBytes), "descriptor size");

}  // namespace acc

// ----------------------------------------------------------------- host side
using namespace acc;

struct acc_batch_s {
  int device = 0;
  std::vector<acc_batch_problem> problems;
  std::vector<std::vector<unsigned char>> weights;  // host copies of stencil weights (the plain path's argument)
  BatchPlan plan;
  void* d_desc = nullptr;
  void* d_map = nullptr;
  void* d_smap = nullptr;  // stencil runs -> problem
  void* d_ws = nullptr;  // SyncCtl + status words
  size_t ws_bytes = 0;
};

namespace acc {

static void batch_free(acc_batch_t b) {
  if (b->d_desc) dev_free(b->device, b->d_desc);
  if (b->d_map) dev_free(b->device, b->d_map);
  if (b->d_smap) dev_free(b->device, b->d_smap);
  if (b->d_ws) dev_free(b->device, b->d_ws);
  b->d_desc = b->d_map = b->d_smap = b->d_ws = nullptr;
}

// Lay out runs for a shape and run length; upload descriptors, map and a
// zeroed workspace.
static int batch_layout(acc_batch_t b, int block, int k, int R) {
  batch_free(b);
  ACC_CHECK(batch_plan(b->problems.data(), (int64_t)b->problems.size(), block, k, R, &b->plan));
  const size_t n = b->problems.size();
  ACC_CHECK(dev_malloc(b->device, std::max<size_t>(n, 1) * sizeof(BatchDescBytes), &b->d_desc));
  ACC_CHECK(dev_malloc(b->device, std::max<size_t>(b->plan.run_prob.size(), 1) * sizeof(uint32_t), &b->d_map));
  ACC_CHECK(dev_malloc(b->device, std::max<size_t>(b->plan.stencil_prob.size(), 1) * sizeof(uint32_t), &b->d_smap));
  if (!b->plan.stencil_prob.empty())
    ACC_HIP_CHECK(hipMemcpy(b->d_smap, b->plan.stencil_prob.data(), b->plan.stencil_prob.size() * sizeof(uint32_t),
                            hipMemcpyHostToDevice));
  b->ws_bytes = kSyncCtlBytes + b->plan.status_words * 8;
  ACC_CHECK(dev_malloc(b->device, b->ws_bytes, &b->d_ws));
  ACC_HIP_CHECK(hipMemcpy(b->d_desc, b->plan.desc.data(), n * sizeof(BatchDescBytes), hipMemcpyHostToDevice));
  if (!b->plan.run_prob.empty())
    ACC_HIP_CHECK(hipMemcpy(b->d_map, b->plan.run_prob.data(), b->plan.run_prob.size() * sizeof(uint32_t),
                            hipMemcpyHostToDevice));
  ACC_HIP_CHECK(hipMemset(b->d_ws, 0, b->ws_bytes));
  // The uploads and the zeroing above go to the null stream, which does not
  // order the launches that follow on non-blocking streams (the tuner's own,
  // every library stream): without this wait a zeroing still in flight can
  // land inside the first launch, reset its ticket counter and erase status
  // words already published, and a look-back then waits for ever.
  ACC_HIP_CHECK(hipStreamSynchronize(nullptr));
  // warm the occupancy cache so a later launch inside a stream capture only
  // enqueues the kernel
  blocks_per_cu(shape_kernel(block, k), block);
  return ACC_SUCCESS;
}

static int batch_enqueue(acc_batch_t b, hipStream_t stream) {
  return batch_launch_plan(b->device, b->plan, b->d_desc, b->d_map, b->d_smap, b->d_ws, stream);
}

// Signature of a batch for the tuning cache: problem count, size histogram
// (log2 buckets) and class mix.
static std::string batch_signature(const acc_batch_t b) {
  int hist[64] = {0};
  int cls[32] = {0};
  for (auto& p : b->problems) {
    int lg = 0;
    while (lg < 62 && ((int64_t)1 << (lg + 1)) <= p.n) ++lg;
    hist[lg / 2]++;
    cls[(p.kind * 8 + type_index((acc_type)p.type) * 2 + (p.op == ACC_OP_MAX)) & 31]++;
  }
  std::string s = "n" + std::to_string(b->problems.size());
  const int np = std::max<int>(1, (int)b->problems.size());
  for (int i = 0; i < 32; ++i) if (hist[i]) s += "_h" + std::to_string(i) + ":" + std::to_string(hist[i] * 20 / np);
  for (int i = 0; i < 32; ++i) if (cls[i]) s += "_c" + std::to_string(i) + ":" + std::to_string(cls[i] * 20 / np);
  return s;
}

static std::string tune_path() {
  if (const char* p = getenv("ACC_HIP_TUNE_CACHE")) return p;
  std::string dir;
  if (const char* x = getenv("XDG_CACHE_HOME")) dir = x;
  else if (const char* h = getenv("HOME")) dir = std::string(h) + "/.cache";
  else dir = "/tmp";
  return dir + "/accelerate-hip-tune.txt";
}

static bool tune_lookup(const std::string& sig, int* block, int* k, int* R) {
  FILE* f = fopen(tune_path().c_str(), "r");
  if (!f) return false;
  char line[1024];
  bool found = false;
  while (fgets(line, sizeof line, f)) {
    char key[900];
    int b, kk, r;
    if (sscanf(line, "batch2 %899s %d %d %d", key, &b, &kk, &r) == 4 && sig == key && shape_kernel(b, kk) && r >= 1) {
      *block = b, *k = kk, *R = r;
      found = true;
    }
  }
  fclose(f);
  return found;
}

static void tune_store(const std::string& sig, int block, int k, int R) {
  FILE* f = fopen(tune_path().c_str(), "a");
  if (!f) return;
  fprintf(f, "batch2 %s %d %d %d\n", sig.c_str(), block, k, R);
  fclose(f);
}

// Time each candidate (workgroup size, vectors per thread, run length) on the
// batch itself; keep the fastest.
static int batch_autotune(acc_batch_t b) {
  const std::string sig = batch_signature(b);
  int bb, kk, rr;
  if (tune_lookup(sig, &bb, &kk, &rr)) return batch_layout(b, bb, kk, rr);
  // the timing runs read the inputs and write the outputs: order them after
  // everything already queued on the device (inputs produced on any stream)
  ACC_HIP_CHECK(hipDeviceSynchronize());
  hipStream_t st;
  ACC_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  hipEvent_t e0, e1;
  ACC_HIP_CHECK(hipEventCreate(&e0));
  ACC_HIP_CHECK(hipEventCreate(&e1));
  float best = 1e30f;
  BatchShape best_s = {256, 4};
  int best_r = 4;
  int rc = ACC_SUCCESS;
  for (const BatchShape& sh : kShapes) {
    for (int R : kRuns) {
      rc = batch_layout(b, sh.block, sh.k, R);
      if (!rc) rc = batch_enqueue(b, st);  // warm-up
      if (!rc) {
        (void)hipEventRecord(e0, st);
        for (int r = 0; r < 3 && !rc; ++r) rc = batch_enqueue(b, st);
        (void)hipEventRecord(e1, st);
        (void)hipEventSynchronize(e1);
      }
      float ms = 0;
      if (!rc && hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms < best) {
        best = ms;
        best_s = sh;
        best_r = R;
      }
      if (rc) break;
    }
    if (rc) break;
  }
  (void)hipStreamSynchronize(st);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(st);
  if (rc) return rc;
  tune_store(sig, best_s.block, best_s.k, best_r);
  return batch_layout(b, best_s.block, best_s.k, best_r);
}

}  // namespace acc

extern "C" {

int acc_hip_batch_create(int device, const acc_batch_problem* problems, int64_t count, int tile, acc_batch_t* out) {
  if (!out || (count > 0 && !problems) || count < 0) return fail(ACC_ERROR_INVALID_VALUE, "batch: bad arguments");
  if (tile != 0 && tile != 2 && tile != 4 && tile != 8)
    return fail(ACC_ERROR_INVALID_VALUE, "batch: tile must be 0 (autotune), 2, 4 or 8");
  ACC_CHECK(batch_validate(problems, count));
  ACC_CHECK(set_device(device));
  auto b = new acc_batch_s();
  b->device = device;
  b->problems.assign(problems, problems + count);
  b->weights.resize(count);
  int rc = ACC_SUCCESS;
  // the stencil weights are read here: wait for whatever is still producing
  // them on other streams (the null stream's copy does not order after those)
  bool stencils = false;
  for (int64_t i = 0; i < count; ++i) stencils |= problems[i].kind == ACC_BATCH_STENCIL;
  if (stencils) {
    const hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) rc = hip_fail(e, "batch: synchronising before reading stencil weights");
  }
  for (int64_t i = 0; i < count && !rc; ++i) {
    const acc_batch_problem& p = problems[i];
    if (p.kind != ACC_BATCH_STENCIL) continue;
    const int ry = p.params & 0xff, rx = (p.params >> 8) & 0xff;
    b->weights[i].resize((size_t)(2 * ry + 1) * (2 * rx + 1) * type_size((acc_type)p.type));
    const hipError_t e = hipMemcpy(b->weights[i].data(), p.aux, b->weights[i].size(), hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = hip_fail(e, "batch: reading stencil weights");
  }
  if (!rc) rc = tile ? batch_layout(b, 256, tile, 4) : batch_autotune(b);
  if (rc) {
    batch_free(b);
    delete b;
    return rc;
  }
  *out = b;
  return ACC_SUCCESS;
}

int acc_hip_batch_launch(acc_batch_t b, acc_stream_t s) {
  if (!b || !s) return fail(ACC_ERROR_INVALID_VALUE, "batch: null argument");
  if (s->device != b->device) return fail(ACC_ERROR_INVALID_VALUE, "batch: stream on another device");
  ACC_CHECK(set_device(s->device));
  return batch_enqueue(b, s->stream);
}

int acc_hip_batch_launch_plain(acc_batch_t b, acc_stream_t s) {
  if (!b || !s) return fail(ACC_ERROR_INVALID_VALUE, "batch: null argument");
  if (s->device != b->device) return fail(ACC_ERROR_INVALID_VALUE, "batch: stream on another device");
  for (const acc_batch_problem& p : b->problems) {
    const void* seed = p.has_seed ? (const void*)&p.seed : nullptr;
    const acc_type t = (acc_type)p.type;
    switch (p.kind) {
      case ACC_BATCH_SCAN:
        ACC_CHECK(acc_hip_scan(s, t, (acc_op)p.op, ACC_LEFT_TO_RIGHT, seed, p.in, 1, p.n, p.out));
        break;
      case ACC_BATCH_FOLD:
        ACC_CHECK(acc_hip_fold(s, t, (acc_op)p.op, seed, p.in, 1, p.n, p.out));
        break;
      case ACC_BATCH_FOLD_SEG:
        ACC_CHECK(acc_hip_fold_seg(s, t, (acc_op)p.op, seed, p.in, 1, p.n, (acc_type)p.aux_type, p.aux, p.m, p.out));
        break;
      case ACC_BATCH_STENCIL: {
        const uint64_t fill = p.fill;
        ACC_CHECK(acc_hip_stencil2d(s, t, p.params & 0xff, (p.params >> 8) & 0xff, b->weights[&p - &b->problems[0]].data(),
                                    (acc_boundary)((p.params >> 16) & 0xff), &fill, p.in, p.n, p.m, p.out));
        break;
      }
      default:
        ACC_CHECK(acc_hip_fill_async(s, p.out, &p.seed, type_size(t), (size_t)p.m));
        ACC_CHECK(acc_hip_permute(s, t, ACC_OP_ADD, p.out, p.out, p.m, nullptr, &p.fill, (acc_type)p.aux_type, p.in,
                                  p.n));
        break;
    }
  }
  return ACC_SUCCESS;
}

int acc_hip_batch_tile(acc_batch_t b, int* tile, uint32_t* total_tiles) {
  if (!b) return fail(ACC_ERROR_INVALID_VALUE, "batch: null");
  if (tile) *tile = b->plan.k;
  if (total_tiles) *total_tiles = b->plan.total_runs;
  return ACC_SUCCESS;
}

int acc_hip_batch_shape(acc_batch_t b, int* block, int* vectors, int* run, uint32_t* tickets) {
  if (!b) return fail(ACC_ERROR_INVALID_VALUE, "batch: null");
  if (block) *block = b->plan.block;
  if (vectors) *vectors = b->plan.k;
  if (run) *run = b->plan.R;
  if (tickets) *tickets = b->plan.total_runs;
  return ACC_SUCCESS;
}

int acc_hip_batch_destroy(acc_batch_t b) {
  if (!b) return ACC_SUCCESS;
  (void)set_device(b->device);
  (void)hipDeviceSynchronize();
  batch_free(b);
  delete b;
  return ACC_SUCCESS;
}

}  // extern "C"
