// Scans along the innermost dimension of a rank >= 2 array (scanDim,
// scan'Dim: PTX/Execute.hs:470-488,577-594) in ONE launch.
//
// The reference launches one block per row or a multi-kernel scan per row
// (PTX/CodeGen/Scan.hs:746-1126).  Here the rows are segments of one
// flattened scan: tiles of the flattened array are scanned with the
// segmented combine (SegOp) whose heads are the row starts (in traversal
// order), and a decoupled look-back carries a row across tiles.  A tile that
// contains a row start knows its inclusive prefix at once (the start cuts
// off everything before it) and publishes it before looking back, so a
// look-back is at most as deep as a row spans tiles; short rows need none.
//
// Output layout per row r (in traversal order, seed z on the left):
//   scanl1/scanr1 : out[r*inner + c]     = inclusive value at column c
//   scanl z       : out[r*(inner+1) + c] = exclusive value, out[r*(inner+1) + inner] = total
//   scanr z       : out[r*(inner+1) + c] = inclusive value, out[r*(inner+1) + inner] = z
//   scanl'/scanr' : out[r*inner + c]     = exclusive value, sums[r] = total
// (NAT/CodeGen/Scan.hs:136-270 for the element order of each variant.)
#include "dispatch.hpp"
#include "host.hpp"
#include "lookback.hpp"
#include "seg.hpp"
#include "vecmem.hpp"
#include "wave64.hpp"

namespace acc {

template <typename T> struct RowScanArgs {
  const T* in;
  T* out;
  int64_t n;        // outer * inner
  int64_t inner;
  int64_t out_row;  // inner, or inner + 1
  uint32_t exclusive;
  uint32_t has_seed;
  T seed;
  T* total_out;  // row r's total at total_out[r * total_stride] (nullable)
  int64_t total_stride;
  T* seed_out;   // row r's seed at seed_out[r * seed_stride] (nullable)
  int64_t seed_stride;
  uint32_t ntiles;
  uint32_t step_r;  // (64 * VEC) % inner: column step to one lane's next vector
  SyncCtl* ctl;
  uint64_t* status;
};

// Persistent workgroups taking tiles by ticket when ready (as the plain scan).
// Tile t covers traversal positions [t*TILE, (t+1)*TILE); right to left, it
// is physical tile ntiles-1-t read backwards.  Positions past the array are
// padding: identity values that start a dead segment and write nothing.
template <typename T, typename Op, int BLOCK, int K, bool REVERSE, bool ALIGNED>
__global__ __launch_bounds__(BLOCK) void k_scan_rows(RowScanArgs<T> a) {
  constexpr int VEC = Vec<T>::N;
  constexpr int W = BLOCK / 64;
  constexpr int TILE = BLOCK * K * VEC;
  using OpT = typename std::conditional<REVERSE, Flip<Op>, Op>::type;
  using S = SegV<T>;
  constexpr int NW = Tagged<S>::NW;
  __shared__ uint64_t s_tick;
  __shared__ S s_wave[W];
  __shared__ S s_carry;
  OpT op{};
  SegOp<OpT> sop{op};
  const T id = identity<T>(op);
  const S sid{id, 0u};
  const int lane = lane_id(), w = threadIdx.x / 64;
  // the column at which a row starts / ends in traversal order
  const uint32_t c_start = REVERSE ? (uint32_t)a.inner - 1 : 0u;
  const uint32_t c_end = REVERSE ? 0u : (uint32_t)a.inner - 1;

  if (threadIdx.x == 0) s_tick = take_ticket(a.ctl);
  __syncthreads();
  const uint32_t epoch = (uint32_t)(s_tick >> 32);
  uint32_t t = (uint32_t)s_tick;
  while (t < a.ntiles) {
    const int64_t base = (REVERSE ? (int64_t)(a.ntiles - 1 - t) : (int64_t)t) * TILE;
    auto vidx = [&](int k) -> int64_t {
      return REVERSE ? ((int64_t)((W - 1 - w) * K + (K - 1 - k)) * 64 + (63 - lane))
                     : ((int64_t)(w * K + k) * 64 + lane);
    };

    // row and column of the tile's first physical element (one 64-bit
    // division per tile; rows are shorter than 2^20, so a vector's column
    // offset from it fits 32 bits)
    const int64_t brow = base / a.inner;
    const uint32_t bcol = (uint32_t)(base - brow * a.inner);
    const uint32_t inner32 = (uint32_t)a.inner;

    // ---- load; per vector: traversal-ordered values, head bits (row starts
    // and padding), row and column of the first physical element
    // column and row of physical slot j of a vector starting at (col, row):
    // at most VEC - 1 wraps
    auto slot_pos = [&](uint32_t col, int64_t row, int j, uint32_t* c, int64_t* r) {
      uint32_t cj = col + (uint32_t)j;
      int64_t rj = row;
#pragma unroll
      for (int q = 0; q < VEC - 1; ++q)
        if (cj >= inner32) {
          cj -= inner32;
          ++rj;
        }
      *c = cj;
      *r = rj;
    };
    T x[K][VEC];
    uint32_t hb[K];
    // every load of the tile first, so they are all in flight together; a
    // vector past the end re-reads the vector holding the last element (an
    // aligned 16-byte block with a valid element never crosses a page)
    T raw[K][VEC];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int64_t e0 = base + vidx(k) * VEC;
      if (ALIGNED) {
        const Vec<T> r = load16_stream(a.in + min(e0, (a.n - 1) / VEC * VEC));
#pragma unroll
        for (int j = 0; j < VEC; ++j) raw[k][j] = r.e[j];
      } else {
#pragma unroll
        for (int j = 0; j < VEC; ++j) raw[k][j] = e0 + j < a.n ? a.in[e0 + j] : id;
      }
    }
    // keep the loads above ahead of the arithmetic below
    __builtin_amdgcn_sched_barrier(0);

    // row (relative to brow) and column of vector k's first physical element
    auto vec_pos = [&](int k, uint32_t* col, uint32_t* dr) {
      const uint32_t cc = bcol + (uint32_t)(vidx(k) * VEC);
      *dr = cc / inner32;
      *col = cc - *dr * inner32;
    };
    // the ragged tile (the physically last) has padding: identity values,
    // each a head of its own dead segment
    const bool ragged = base + TILE > a.n;
    // vector columns: one division per lane, then steps of 64 vectors
    // (forward) or back (right to left) by the step's remainder
    uint32_t col;
    {
      uint32_t dr0;
      vec_pos(0, &col, &dr0);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int64_t e0 = base + vidx(k) * VEC;
      uint32_t h = 0;
      if (inner32 >= (uint32_t)VEC) {
        // at most one row start in the vector: physical slot ps
        const uint32_t ps = REVERSE ? inner32 - 1u - col : (col == 0 ? 0u : inner32 - col);
        if (ps < (uint32_t)VEC) h = 1u << (REVERSE ? VEC - 1 - (int)ps : (int)ps);
      } else {
#pragma unroll
        for (int j = 0; j < VEC; ++j) {
          const int pj = REVERSE ? VEC - 1 - j : j;  // physical slot of traversal slot j
          uint32_t c;
          int64_t r;
          slot_pos(col, 0, pj, &c, &r);
          h |= (c == c_start ? 1u : 0u) << j;
        }
      }
      if (!REVERSE) {
        col += a.step_r;
        if (col >= inner32) col -= inner32;
      } else {
        const int c = (int)col - (int)a.step_r;
        col = (uint32_t)(c < 0 ? c + (int)inner32 : c);
      }
      if (ragged) {
#pragma unroll
        for (int j = 0; j < VEC; ++j) {
          const int pj = REVERSE ? VEC - 1 - j : j;
          const bool pad = e0 + pj >= a.n;
          h |= (pad ? 1u : 0u) << j;
          if (pad) raw[k][pj] = id;
        }
      }
#pragma unroll
      for (int j = 0; j < VEC; ++j) x[k][j] = raw[k][REVERSE ? VEC - 1 - j : j];
      hb[k] = h;
#pragma unroll
      for (int j = 1; j < VEC; ++j) x[k][j] = (h >> j & 1u) ? x[k][j] : op(x[k][j - 1], x[k][j]);
    }

    // ---- segmented scan of the vector totals across the wave and the block
    S pre[K];
    S run = sid;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const S inc = wave_inclusive_scan(S{x[k][VEC - 1], hb[k] ? 1u : 0u}, sop);
      const S up = shift_up1(inc);  // all lanes take part in the cross-lane move
      pre[k] = sop(run, lane > 0 ? up : sid);
      run = sop(run, read_lane(inc, 63));
    }
    if (lane == 0) s_wave[w] = run;
    __syncthreads();
    S wp = sid, agg = sid;
#pragma unroll
    for (int i = 0; i < W; ++i) {
      const S v = s_wave[i];
      if (i < w) wp = sop(wp, v);
      agg = sop(agg, v);
    }

    // ---- carry from predecessor tiles (only a row that enters the tile
    // needs it; a tile with a row start knows its prefix at once)
    if (w == 0) {
      S X = sid;
      uint64_t* slot = a.status + (size_t)t * NW;
      const bool last = t + 1 == a.ntiles;
      if (t == 0 || agg.f) {
        if (lane == 0 && !last) Tagged<S>::publish(slot, agg, make_tag(epoch, ST_PREFIX));
      } else if (lane == 0 && !last) {
        Tagged<S>::publish(slot, agg, make_tag(epoch, ST_AGGREGATE));
      }
      // the tile's first traversal element continues a row
      const int64_t first_p = REVERSE ? base + TILE - 1 : base;
      const bool first_is_start = first_p >= a.n || (uint32_t)(first_p % a.inner) == c_start;
      if (t > 0 && !first_is_start) {
        X = wave_lookback<S>(a.status, t, epoch, sop);
        if (lane == 0 && !agg.f && !last) Tagged<S>::publish(slot, sop(X, agg), make_tag(epoch, ST_PREFIX));
      }
      if (lane == 0) s_carry = X;
    }
    __syncthreads();
    const S XW = sop(s_carry, wp);

    // ---- outputs
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const S P = sop(XW, pre[k]);  // everything before this vector, segmented
      const int64_t e0 = base + vidx(k) * VEC;
      T y[VEC];  // by physical slot
      bool headed = false;  // a row start at or before traversal slot j in this vector
#pragma unroll
      for (int j = 0; j < VEC; ++j) {
        const bool h = hb[k] >> j & 1u;
        headed = headed || h;
        const T incl = headed ? x[k][j] : op(P.v, x[k][j]);  // row value through j
        T val;
        if (a.exclusive) {
          // row value before j: nothing at a row start
          const bool before = !h;
          const T prev = j == 0 ? P.v : (((hb[k] & ((1u << j) - 1u)) != 0) ? x[k][j - 1] : op(P.v, x[k][j - 1]));
          val = before ? (a.has_seed ? op(a.seed, prev) : prev) : a.seed;
        } else {
          val = a.has_seed ? op(a.seed, incl) : incl;
        }
        const int pj = REVERSE ? VEC - 1 - j : j;
        y[pj] = val;
        // row extras (scan' totals, scanr z's seed slot): the total at the
        // row's last traversal element, the seed at its first
        if (a.total_out || a.seed_out) {
          const int64_t p = e0 + pj;
          if (p < a.n) {
            uint32_t vc, vd, c;
            int64_t r;
            vec_pos(k, &vc, &vd);
            slot_pos(vc, brow + vd, pj, &c, &r);
            if (a.total_out && c == c_end) a.total_out[r * a.total_stride] = a.has_seed ? op(a.seed, incl) : incl;
            if (a.seed_out && c == c_start) a.seed_out[r * a.seed_stride] = a.seed;
          }
        }
      }
      if (ALIGNED && a.out_row == a.inner && e0 + VEC <= a.n) {
        Vec<T> r;
#pragma unroll
        for (int j = 0; j < VEC; ++j) r.e[j] = y[j];
        store16_stream(a.out + e0, r);  // written once: non-temporal
      } else {
#pragma unroll
        for (int j = 0; j < VEC; ++j) {
          const int64_t p = e0 + j;
          if (p < a.n) {
            uint32_t vc, vd, c;
            int64_t r;
            vec_pos(k, &vc, &vd);
            slot_pos(vc, brow + vd, j, &c, &r);
            a.out[r * a.out_row + c] = y[j];
          }
        }
      }
    }
    if (gridDim.x >= a.ntiles) break;  // one tile per workgroup
    __syncthreads();
    if (threadIdx.x == 0) s_tick = take_ticket(a.ctl);
    __syncthreads();
    t = (uint32_t)s_tick;
  }
  block_retire(a.ctl, gridDim.x, epoch);
}

template <typename T, typename Op, bool REVERSE, bool ALIGNED>
int scan_rows_launch(acc_stream_t s, RowScanArgs<T> a) {
#ifndef ACC_ROWS_K
#define ACC_ROWS_K 8
#endif
  constexpr int BLOCK = 256, K = ACC_ROWS_K;
  constexpr int64_t TILE = (int64_t)BLOCK * K * Vec<T>::N;
  if (a.inner >= ((int64_t)1 << 20)) return fail(ACC_ERROR_UNSUPPORTED, "scan rows: row of %lld", (long long)a.inner);
  constexpr int NW = Tagged<SegV<T>>::NW;
  const int64_t ntiles = ceil_div(a.n, TILE);
  if (ntiles > 0x7fffffffLL) return fail(ACC_ERROR_UNSUPPORTED, "scan: %lld elements", (long long)a.n);
  SyncCtl* ctl;
  uint64_t* status;
  ACC_CHECK(stream_workspace(s, (size_t)ntiles * NW * 8, &ctl, &status));
  a.ntiles = (uint32_t)ntiles;
  a.step_r = (uint32_t)(64 * Vec<T>::N % a.inner);
  a.ctl = ctl;
  a.status = status;
  auto kern = k_scan_rows<T, Op, BLOCK, K, REVERSE, ALIGNED>;
  const DeviceInfo* di;
  ACC_CHECK(device_info(s->device, &di));
  int64_t grid = (int64_t)di->compute_units * blocks_per_cu((const void*)kern, BLOCK);
  if (grid > ntiles) grid = ntiles;
  hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(BLOCK), 0, s->stream, a);
  ACC_HIP_CHECK(hipGetLastError());
  return ACC_SUCCESS;
}

}  // namespace acc

// All rows of a rank >= 2 scan in one launch (see the top of this file for the
// layout).  exclusive: emit exclusive values; sums: scan' totals;
// seed_at_end: scanr z (the seed closes each output row).
int acc_scan_rows(acc_stream_t s, acc_type t, acc_op op, bool reverse, bool exclusive, const void* seed,
                  const void* in, int64_t outer, int64_t inner, void* out, int64_t out_row, void* sums,
                  bool seed_at_end) {
  using namespace acc;
  return with_type(t, [&](auto z) {
    using T = decltype(z);
    return with_op<T>(op, [&](auto f) {
      using Op = decltype(f);
      if ((uintptr_t)in % sizeof(T) || (uintptr_t)out % sizeof(T))
        return fail(ACC_ERROR_INVALID_VALUE, "scan: arrays not element-aligned");
      RowScanArgs<T> a{};
      a.in = (const T*)in;
      a.out = (T*)out;
      a.n = outer * inner;
      a.inner = inner;
      a.out_row = out_row;
      a.exclusive = exclusive ? 1u : 0u;
      a.has_seed = seed ? 1u : 0u;
      a.seed = seed ? *(const T*)seed : T();
      if (sums) {
        a.total_out = (T*)sums;
        a.total_stride = 1;
      } else if (exclusive && !reverse) {  // scanl z: the total closes each output row
        a.total_out = (T*)out + inner;
        a.total_stride = out_row;
      }
      if (seed_at_end) {
        a.seed_out = (T*)out + inner;
        a.seed_stride = out_row;
      }
      const bool aligned = (uintptr_t)in % 16 == 0 && (uintptr_t)out % 16 == 0;
      if (reverse)
        return aligned ? scan_rows_launch<T, Op, true, true>(s, a) : scan_rows_launch<T, Op, true, false>(s, a);
      return aligned ? scan_rows_launch<T, Op, false, true>(s, a) : scan_rows_launch<T, Op, false, false>(s, a);
    });
  });
}
