// foldSeg / fold1Seg, and fold along the innermost dimension of a rank >= 2
// array (every row one segment), in ONE load-balanced launch.
//
// The reference assigns one block or one warp per segment
// (PTX/CodeGen/FoldSeg.hs:79-264 and :274-458, selected at
// PTX/Execute.hs:361-366) and one block per row for foldDim
// (PTX/CodeGen/Fold.hs:296-405), so long segments serialise and short ones
// idle most lanes.  Here the elements are cut into equal tiles regardless of
// segment boundaries and the fold becomes a segmented scan over
// (head-flag, value) pairs, evaluated in registers; the value at the last
// element of each segment is its result.  A segment that crosses tiles takes
// its carry-in from a decoupled look-back in which every tile that contains a
// segment head publishes its final prefix immediately, so look-back depth is
// at most the number of tiles a segment spans.
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
#include "lookback.hpp"
#include "vecmem.hpp"

namespace acc {

template <typename T> struct SegV {
  T v;
  uint32_t f;  // a segment starts at (or after) the first element
};

template <typename Op> struct SegOp {
  Op op;
  template <typename T>
  __device__ __forceinline__ SegV<T> operator()(SegV<T> a, SegV<T> b) const {
    SegV<T> r;
    r.v = b.f ? b.v : op(a.v, b.v);
    r.f = a.f | b.f;
    return r;
  }
};

// Segment starts given by an offsets array: start(g) = off[g], g in [0, S].
// The segment range of each tile comes from a pre-pass (k_tile_seg).
template <typename I> struct OffsetStarts {
  const I* off;
  const int64_t* tseg;  // tseg[T] = last g with start(g) <= T * TILE (>= 0)
  __device__ __forceinline__ int64_t start(int64_t g) const { return (int64_t)off[g]; }
  __device__ __forceinline__ void tile_range(uint32_t t, int64_t, int64_t, int64_t, int64_t* lo, int64_t* hi) const {
    *lo = tseg[t];
    *hi = tseg[t + 1];  // an upper bound of the segment of the tile's last element
  }
};

// foldDim: start(g) = g * len.
struct UniformStarts {
  int64_t len;
  __device__ __forceinline__ int64_t start(int64_t g) const { return g * len; }
  __device__ __forceinline__ void tile_range(uint32_t, int64_t first, int64_t last, int64_t nseg, int64_t* lo,
                                             int64_t* hi) const {
    *lo = first / len;
    *hi = min(last / len, nseg);
  }
};

template <typename T, typename Starts> struct SegArgs {
  const T* in;
  T* out;
  Starts st;
  int64_t nseg;   // S
  int64_t e_end;  // elements per row
  uint32_t ntiles;
  uint32_t has_seed;
  T seed;
  uint32_t check_empty;  // look for empty segments (offsets only)
  SyncCtl* ctl;
  uint64_t* status;
};

// One wave per tile boundary T in [0, ntiles]: tseg[T] = last g in [0, S]
// with off[g] <= T * tile (0 if none), by a 64-ary search (one round trip
// narrows the range 64-fold).
template <typename I>
__global__ __launch_bounds__(256) void k_tile_seg(const I* off, int64_t nseg, int64_t tile, uint32_t ntiles,
                                                  int64_t* tseg) {
  const uint32_t T = blockIdx.x * 4u + threadIdx.x / 64u;
  if (T > ntiles) return;  // whole waves
  const int lane = lane_id();
  const int64_t e = (int64_t)T * tile;
  int64_t lo = 0, hi = nseg;
  if ((int64_t)off[0] > e) {
    hi = 0;  // no segment starts at or before e
  } else {
    while (hi > lo) {
      const int64_t step = (hi - lo + 64) / 64;
      const int64_t g = lo + (int64_t)lane * step;
      const bool le = g <= hi && (int64_t)off[g] <= e;
      const int c = __popcll(__ballot(le));  // >= 1: off[lo] <= e
      lo = lo + (int64_t)(c - 1) * step;
      hi = min(hi, lo + step - 1);
    }
  }
  if (lane == 0) tseg[T] = lo;
}

template <typename T, typename Op, typename Starts, int BLOCK, int K, int CAP>
__device__ __forceinline__ void seg_tile(const SegArgs<T, Starts>& a, uint32_t t, uint32_t epoch,
                                         int64_t* s_range, int64_t* s_starts, uint32_t* s_heads,
                                         Maybe<SegV<T>>* s_wave, Maybe<SegV<T>>* s_prefix) {
  constexpr int VEC = Vec<T>::N;
  constexpr int W = BLOCK / 64;
  constexpr int TILE = BLOCK * K * VEC;
  constexpr int NBW = TILE / 32 + 1;  // head bitmap: tile positions 0..TILE
  using S = SegV<T>;
  constexpr int NW = Tagged<S>::NW;
  SegOp<Op> sop{Op{}};
  Op op{};
  const int lane = lane_id();
  const int w = threadIdx.x / 64;
  const int64_t base = (int64_t)t * TILE;

  // ---- segment heads of the tile.  Segment g in [0, S) covers
  // [start(g), start(g+1)); "segment" S collects the elements after the last
  // one and elements before start(0) belong to "segment" -1 (no results).
  // Every start in [base, base + TILE] sets its bit in an LDS bitmap; the
  // starts themselves are staged (when they fit) for the rare end lookups.
  auto sv = [&](int64_t g) -> int64_t { return g > a.nseg ? INT64_MAX : a.st.start(g); };
  if (threadIdx.x == 0) a.st.tile_range(t, base, min(base + TILE, a.e_end) - 1, a.nseg, &s_range[0], &s_range[1]);
  for (int i = threadIdx.x; i < NBW; i += BLOCK) s_heads[i] = 0u;
  __syncthreads();
  const int64_t g_lo = s_range[0], g_hi = s_range[1];
  const int64_t cnt = g_hi - g_lo + 2;  // starts g_lo .. g_hi+1
  const bool staged = cnt <= CAP;
  for (int64_t i = threadIdx.x; i < cnt; i += BLOCK) {
    const int64_t st = sv(g_lo + i);
    if (staged) s_starts[i] = st;
    const int64_t p = st - base;
    if (p >= 0 && p <= TILE) atomicOr(&s_heads[p >> 5], 1u << (p & 31));
  }
  __syncthreads();
  // smallest segment whose start is at least `e`
  auto first_seg_at = [&](int64_t e) -> int64_t {
    int64_t x = g_lo, y = g_hi + 1;
    while (x < y) {
      const int64_t m = (x + y) >> 1;
      if ((staged ? s_starts[m - g_lo] : sv(m)) < e) x = m + 1; else y = m;
    }
    return x;
  };

  // ---- load; segmented scan inside each thread's dwordx4 vector
  T xv[K][VEC];
  uint32_t hb[K], eb[K];  // per vector: head bits of its elements, "next is a head" bits
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int p0 = ((w * K + k) * 64 + lane) * VEC;
    const int64_t e0 = base + p0;
    T v[VEC];
    uint32_t invalid = 0;
    if (e0 + VEC <= a.e_end) {
      Vec<T> r = load16_stream(a.in + e0);
#pragma unroll
      for (int j = 0; j < VEC; ++j) v[j] = r.e[j];
    } else {
#pragma unroll
      for (int j = 0; j < VEC; ++j) {
        const bool ok = e0 + j < a.e_end;
        v[j] = ok ? a.in[e0 + j] : T();
        invalid |= ok ? 0u : 1u << j;
      }
    }
    const int sh = p0 & 31;
    uint64_t word = s_heads[p0 >> 5];
    if (sh + VEC + 1 > 32) word |= (uint64_t)s_heads[(p0 >> 5) + 1] << 32;
    const uint32_t b = (uint32_t)(word >> sh) & ((1u << (VEC + 1)) - 1u);
    hb[k] = (b & ((1u << VEC) - 1u)) | invalid;  // elements past the row start a dead segment
    eb[k] = (b >> 1) & ~invalid;
    xv[k][0] = v[0];
#pragma unroll
    for (int j = 1; j < VEC; ++j) xv[k][j] = (hb[k] >> j & 1u) ? v[j] : op(xv[k][j - 1], v[j]);
  }

  // ---- segmented scan across lanes, chunks and waves
  auto none = []() { Maybe<S> m; m.v.v = T(); m.v.f = 0; m.ok = 0; return m; };
  Maybe<S> pre[K];
  Maybe<S> run = none();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    S tot_l;
    tot_l.v = xv[k][VEC - 1];
    tot_l.f = hb[k] ? 1u : 0u;
    S inc = wave_inclusive_scan(tot_l, sop);
    S up = shift_up1(inc);
    Maybe<S> exl = none();
    if (lane > 0) { exl.v = up; exl.ok = 1; }
    pre[k] = mcombine(sop, run, exl);
    Maybe<S> tot{read_lane(inc, 63), 1u};
    run = mcombine(sop, run, tot);
  }
  if (lane == 0) s_wave[w] = run;
  __syncthreads();
  Maybe<S> wp = none(), agg = none();
#pragma unroll
  for (int i = 0; i < W; ++i) {
    const Maybe<S> v = s_wave[i];
    if (i < w) wp = mcombine(sop, wp, v);
    agg = mcombine(sop, agg, v);
  }

  // ---- carry-in from predecessor tiles
  if (w == 0) {
    Maybe<S> X = none();
    uint64_t* slot = a.status + (size_t)t * NW;
    const bool last = t + 1 == a.ntiles;
    if (t == 0 || agg.v.f) {
      // a head inside the tile makes its inclusive prefix final at once
      if (lane == 0 && !last) Tagged<S>::publish(slot, agg.v, make_tag(epoch, ST_PREFIX));
    } else if (lane == 0 && !last) {
      Tagged<S>::publish(slot, agg.v, make_tag(epoch, ST_AGGREGATE));
    }
    // the first element needs a carry unless it starts a segment
    if (t > 0 && !(s_heads[0] & 1u)) {
      X.v = wave_lookback<S>(a.status, t, epoch, sop);
      X.ok = 1;
      if (!agg.v.f && lane == 0 && !last)
        Tagged<S>::publish(slot, sop(X.v, agg.v), make_tag(epoch, ST_PREFIX));
    }
    if (lane == 0) *s_prefix = X;
  }
  __syncthreads();
  const Maybe<S> XW = mcombine(sop, *s_prefix, wp);

  // ---- results at segment ends (the element before a head)
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (eb[k]) {
      // segmented prefix of everything before this vector
      const Maybe<S> P = mcombine(sop, XW, pre[k]);
      const int64_t e0 = base + ((int64_t)(w * K + k) * 64 + lane) * VEC;
#pragma unroll
      for (int j = 0; j < VEC; ++j) {
        if (eb[k] >> j & 1u) {
          const bool headed = (hb[k] & ((2u << j) - 1u)) != 0;  // a head at or before j in this vector
          T r = (headed || !P.ok) ? xv[k][j] : op(P.v.v, xv[k][j]);
          const int64_t g = first_seg_at(e0 + j + 1) - 1;
          if (g >= 0 && g < a.nseg) {
            if (a.has_seed) r = op(a.seed, r);  // seed on the left
            a.out[g] = r;
          }
        }
      }
    }
  }

  // ---- empty segments (their value is the seed)
  if (a.check_empty && a.has_seed) {
    const int64_t per = (a.nseg + a.ntiles - 1) / a.ntiles;
    const int64_t s0 = (int64_t)t * per, s1 = min(s0 + per, a.nseg);
    for (int64_t g = s0 + threadIdx.x; g < s1; g += BLOCK)
      if (a.st.start(g) == a.st.start(g + 1)) a.out[g] = a.seed;
  }
}

// Persistent workgroups taking tiles by ticket (next ticket prefetched).
template <typename T, typename Op, typename Starts, int BLOCK, int K, int CAP>
__global__ __launch_bounds__(BLOCK) void k_fold_seg(SegArgs<T, Starts> a) {
  constexpr int TILE = BLOCK * K * Vec<T>::N;
  __shared__ uint64_t s_tick[2];
  __shared__ int64_t s_range[2];
  __shared__ int64_t s_starts[CAP];
  __shared__ uint32_t s_heads[TILE / 32 + 1];
  __shared__ Maybe<SegV<T>> s_wave[BLOCK / 64];
  __shared__ Maybe<SegV<T>> s_prefix;
  if (threadIdx.x == 0) s_tick[0] = take_ticket(a.ctl);
  __syncthreads();
  const uint64_t tk0 = s_tick[0];
  const uint32_t epoch = (uint32_t)(tk0 >> 32);
  uint32_t t = (uint32_t)tk0;
  int cur = 0;
  while (t < a.ntiles) {
    if (threadIdx.x == 0) s_tick[cur ^ 1] = take_ticket(a.ctl);
    seg_tile<T, Op, Starts, BLOCK, K, CAP>(a, t, epoch, s_range, s_starts, s_heads, s_wave, &s_prefix);
    __syncthreads();
    cur ^= 1;
    t = (uint32_t)s_tick[cur];
  }
  block_retire(a.ctl, gridDim.x, epoch);
}

#ifndef ACC_SEG_K
#define ACC_SEG_K 8
#endif
template <typename T> struct SegTune {
  static constexpr int BLOCK = 256;
  static constexpr int K = ACC_SEG_K;  // dwordx4 vectors per thread
  static constexpr int CAP = 1024;
};

template <typename T, typename Op, typename Starts>
int fold_seg_launch(acc_stream_t s, const T* in, T* out, Starts st, int64_t nseg, int64_t e_end,
                    const T* seed, bool check_empty, const void* offsets, size_t offset_bytes) {
  constexpr int BLOCK = SegTune<T>::BLOCK, K = SegTune<T>::K, CAP = SegTune<T>::CAP;
  constexpr int64_t TILE = (int64_t)BLOCK * K * Vec<T>::N;
  constexpr int NW = Tagged<SegV<T>>::NW;
  if ((uintptr_t)in % 16) return fail(ACC_ERROR_UNSUPPORTED, "foldSeg: input must be 16-byte aligned");
  const int64_t ntiles = ceil_div(e_end, TILE);
  if (ntiles == 0) {
    // no elements: every segment is empty
    if (!seed || !check_empty) return ACC_SUCCESS;
    return acc_hip_fill_async(s, out, seed, sizeof(T), (size_t)nseg);
  }
  if (ntiles >= 0x7fffffffLL) return fail(ACC_ERROR_UNSUPPORTED, "foldSeg: too many elements");
  SyncCtl* ctl;
  uint64_t* status;
  const size_t status_bytes = ((size_t)ntiles * NW * 8 + 255) / 256 * 256;
  const size_t tseg_bytes = offsets ? (size_t)(ntiles + 1) * 8 : 0;
  ACC_CHECK(stream_workspace(s, status_bytes + tseg_bytes, &ctl, &status));
  if constexpr (!std::is_same<Starts, UniformStarts>::value) {
    int64_t* tseg = (int64_t*)((char*)status + status_bytes);
    const unsigned pg = (unsigned)((ntiles + 1 + 3) / 4);
    if (offset_bytes == 4)
      hipLaunchKernelGGL(k_tile_seg<int32_t>, dim3(pg), dim3(256), 0, s->stream, (const int32_t*)offsets, nseg,
                         TILE, (uint32_t)ntiles, tseg);
    else
      hipLaunchKernelGGL(k_tile_seg<int64_t>, dim3(pg), dim3(256), 0, s->stream, (const int64_t*)offsets, nseg,
                         TILE, (uint32_t)ntiles, tseg);
    ACC_HIP_CHECK(hipGetLastError());
    st.tseg = tseg;
  }
  SegArgs<T, Starts> a{in, out, st, nseg, e_end, (uint32_t)ntiles, seed ? 1u : 0u, seed ? *seed : T(),
                       check_empty ? 1u : 0u, ctl, status};
  auto kern = k_fold_seg<T, Op, Starts, BLOCK, K, CAP>;
  const DeviceInfo* di;
  ACC_CHECK(device_info(s->device, &di));
  int64_t grid = (int64_t)di->compute_units * blocks_per_cu((const void*)kern, BLOCK);
  if (grid > ntiles) grid = ntiles;
  hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(BLOCK), 0, s->stream, a);
  ACC_HIP_CHECK(hipGetLastError());
  return ACC_SUCCESS;
}

int fold_dim(acc_stream_t s, acc_type t, acc_op op, const void* seed, const void* in, int64_t outer,
             int64_t inner, void* out) {
  return with_type(t, [&](auto z) {
    using T = decltype(z);
    return with_op<T>(op, [&](auto f) {
      using Op = decltype(f);
      return fold_seg_launch<T, Op, UniformStarts>(s, (const T*)in, (T*)out, UniformStarts{inner}, outer,
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
              s, rin, rout, OffsetStarts<int32_t>{(const int32_t*)offsets, nullptr}, nseg, inner,
              (const T*)seed, true, offsets, 4);
        else
          rc = fold_seg_launch<T, Op, OffsetStarts<int64_t>>(
              s, rin, rout, OffsetStarts<int64_t>{(const int64_t*)offsets, nullptr}, nseg, inner,
              (const T*)seed, true, offsets, 8);
        if (rc) return rc;
      }
      return (int)ACC_SUCCESS;
    });
  });
}
