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
#pragma once
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
  uint32_t indep;   // rows divide the tile: every tile starts a row (see below)
  uint32_t step_r;  // (64 * VEC) % inner: column step to one lane's next vector
  SyncCtl* ctl;
  uint64_t* status;
};

// Persistent workgroups taking tiles by ticket when ready (as the plain scan).
// Tile t covers traversal positions [t*TILE, (t+1)*TILE); right to left, it
// is physical tile ntiles-1-t read backwards.  Positions past the array are
// padding: identity values that write nothing.  They need no segment heads:
// left to right they follow the last row's end, right to left they precede
// the last row's start (its traversal start), which cuts them off.
// PLAIN: inclusive values only (scanl1 / scanr1), no seed and no row extras.
template <typename T, typename Op, int BLOCK, int K, bool REVERSE, bool ALIGNED, bool PLAIN>
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

  // Rows that divide the tile length make every tile start at a row start:
  // no tile needs a carry, so each workgroup takes the physical tile
  // blockIdx.x (ascending addresses in both directions) and no ticket,
  // status word or look-back is involved (one workgroup per tile).
  uint32_t epoch = 0, t = REVERSE ? a.ntiles - 1u - blockIdx.x : blockIdx.x;
  if (!a.indep) {
    if (threadIdx.x == 0) s_tick = take_ticket(a.ctl);
    __syncthreads();
    epoch = (uint32_t)(s_tick >> 32);
    t = (uint32_t)s_tick;
  }
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
    // the ragged tile (the physically last) has padding: identity values
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
          if (e0 + pj >= a.n) raw[k][pj] = id;
        }
      }
#pragma unroll
      for (int j = 0; j < VEC; ++j) x[k][j] = raw[k][REVERSE ? VEC - 1 - j : j];
      hb[k] = h;
#pragma unroll
      for (int j = 1; j < VEC; ++j) x[k][j] = (h >> j & 1u) ? x[k][j] : op(x[k][j - 1], x[k][j]);
    }

    // ---- segmented scan of the vector totals across the wave and the block:
    // pre[k] = everything before vector k in the wave's traversal order
    S pre[K];
    S run = sid;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const uint64_t heads = __ballot(hb[k] != 0u);
      bool below;
      const T inc = wave_seg_inclusive_scan(x[k][VEC - 1], heads, op, &below);
      const T up = wave_shr1(inc, id);  // lane - 1's inclusive value; id at lane 0
      pre[k].v = below ? up : op(run.v, up);
      pre[k].f = run.f | (below ? 1u : 0u);
      const T last = read_lane(inc, 63);
      run.v = heads ? last : op(run.v, last);
      run.f |= heads ? 1u : 0u;
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
    if (!a.indep && w == 0) {
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
    if (!a.indep) __syncthreads();
    const S XW = a.indep ? wp : sop(s_carry, wp);

    // ---- outputs
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const S P = sop(XW, pre[k]);  // everything before this vector, segmented
      const int64_t e0 = base + vidx(k) * VEC;
      // first traversal slot holding a row start (VEC when none): the slots
      // from it on already hold their row values
      const int fh = __builtin_ctz(hb[k] | (1u << VEC));
      T incl[VEC];  // row value through traversal slot j
#pragma unroll
      for (int j = 0; j < VEC; ++j) incl[j] = j >= fh ? x[k][j] : op(P.v, x[k][j]);
      T y[VEC];  // by physical slot
#pragma unroll
      for (int j = 0; j < VEC; ++j) {
        T val;
        if (PLAIN) {
          val = incl[j];
        } else if (a.exclusive) {
          // row value before j: the seed alone at a row start
          const T prev = j == 0 ? P.v : incl[j - 1];
          val = (hb[k] >> j & 1u) ? a.seed : (a.has_seed ? op(a.seed, prev) : prev);
        } else {
          val = a.has_seed ? op(a.seed, incl[j]) : incl[j];
        }
        const int pj = REVERSE ? VEC - 1 - j : j;
        y[pj] = val;
        // row extras (scan' totals, scanr z's seed slot): the total at the
        // row's last traversal element, the seed at its first
        if (!PLAIN && (a.total_out || a.seed_out)) {
          const int64_t p = e0 + pj;
          if (p < a.n) {
            uint32_t vc, vd, c;
            int64_t r;
            vec_pos(k, &vc, &vd);
            slot_pos(vc, brow + vd, pj, &c, &r);
            if (a.total_out && c == c_end) a.total_out[r * a.total_stride] = a.has_seed ? op(a.seed, incl[j]) : incl[j];
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
    if (a.indep || gridDim.x >= a.ntiles) break;  // one tile per workgroup
    __syncthreads();
    if (threadIdx.x == 0) s_tick = take_ticket(a.ctl);
    __syncthreads();
    t = (uint32_t)s_tick;
  }
  if (!a.indep) block_retire(a.ctl, gridDim.x, epoch);
}

template <typename T, typename Op, bool REVERSE, bool ALIGNED, bool PLAIN>
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
  auto kern = k_scan_rows<T, Op, BLOCK, K, REVERSE, ALIGNED, PLAIN>;
  const DeviceInfo* di;
  ACC_CHECK(device_info(s->device, &di));
  a.indep = TILE % a.inner == 0 ? 1u : 0u;
  int64_t grid = (int64_t)di->compute_units * blocks_per_cu((const void*)kern, BLOCK);
  if (a.indep || grid > ntiles) grid = ntiles;
  hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(BLOCK), 0, s->stream, a);
  ACC_HIP_CHECK(hipGetLastError());
  return ACC_SUCCESS;
}

// scanl1 / scanr1 along rows of 16-byte-aligned arrays: the PLAIN kernels,
// built in their own translation unit.
int scan_rows_plain(acc_stream_t s, acc_type t, acc_op op, bool reverse, const void* in, int64_t outer,
                    int64_t inner, void* out);

}  // namespace acc
