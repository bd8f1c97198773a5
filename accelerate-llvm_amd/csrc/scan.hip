// scanl/scanl1/scanl'/scanr/scanr1/scanr' over a vector in ONE pass.
//
// The reference runs three kernels (scanP1 block scans, scanP2 scan of block
// totals, scanP3 carry add; PTX/Execute.hs:422-468, PTX/CodeGen/Scan.hs:141-439)
// and moves ~4n elements.  This kernel is a single-pass decoupled look-back
// scan: each workgroup takes a tile by ticket, scans it in registers (thread
// scan of dwordx4 vectors, DPP wave scan, LDS wave-total scan), publishes the
// tile aggregate, looks back over predecessor tiles with one wave, and writes
// its outputs: n reads + n writes.
//
// Element semantics (NAT/CodeGen/Scan.hs:136-270):
//   left-to-right combines `acc (+) x`, right-to-left `x (+) acc`; a
//   right-to-left scan is the left-to-right scan of the reversed sequence with
//   the arguments flipped, so both directions share this kernel.  In traversal
//   order with prefix seed z (when given):
//     scanl1/scanr1 : out[i] = inclusive value at i
//     scanl z       : out[i] = exclusive value at i, out[n] = total
//     scanr z       : out[i] = inclusive value at i, out[n] = z
//     scanl'/scanr' : out[i] = exclusive value at i, sums = total
//   where "at i" is the element's own position in the array.
#include "scan_tile.hpp"

namespace acc {

template <typename T> struct ScanArgs {
  TileIO<T> io;
  uint32_t ntiles;  // all tiles of the scan, the ragged one included
  uint32_t t0;      // status index of the first tile this launch processes
  uint32_t nfull;   // tiles this launch processes (all full)
  SyncCtl* ctl;
};

// Persistent workgroups over the full tiles: each takes a tile by ticket when
// it is ready to start it, so tickets are started in order and a tile's
// predecessors have started no later than it did.  (Taking the next ticket
// ahead of time lets consecutive tiles start up to a whole tile apart, and
// every look-back then waits for that skew.)
// Minimum waves per SIMD asked of the persistent kernel.  Left-to-right
// Float32 (+) and (*) otherwise take 169 VGPRs (2 waves/SIMD); at 3 they
// spill only the seed, reloaded on the first tile (2^28 Float32 scanl1:
// 400 -> 394 us).  Elsewhere the default: Int32/Word32 already fit 3 waves,
// and Float32 min/max, right-to-left Float32 and the 8-byte types would
// spill inside the tile loop (2^27 Int64 424 -> 453 us, Float64 417 ->
// 599 us at 3 waves).  The look-back stays a called function: inlined it
// saves 10-12 VGPRs but measured slower (2^28 Int32 394 -> 403 us, Float32
// at 3 waves 405 us).
template <typename T, typename Op, bool REVERSE> struct ScanOccupancy {
  static constexpr int MIN_WAVES =
      std::is_same<T, float>::value && !REVERSE &&
              (std::is_same<Op, OpAdd>::value || std::is_same<Op, OpMul>::value)
          ? 3
          : 1;
};

template <typename T, typename Op, int BLOCK, int K, bool REVERSE, bool ALIGNED>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(ScanOccupancy<T, Op, REVERSE>::MIN_WAVES)))
void k_scan(ScanArgs<T> a) {
  constexpr int VEC = Vec<T>::N;
  constexpr int TILE = BLOCK * K * VEC;
  __shared__ uint64_t s_tick;
  __shared__ T s_wave[BLOCK / 64];
  __shared__ T s_prefix;
  using OpT = typename std::conditional<REVERSE, Flip<Op>, Op>::type;
  OpT op{};
  if (threadIdx.x == 0) s_tick = take_ticket(a.ctl);
  __syncthreads();
  const uint64_t tk0 = s_tick;
  const uint32_t epoch = (uint32_t)(tk0 >> 32);
  uint32_t i = (uint32_t)tk0;
  while (i < a.nfull) {
    const uint32_t t = a.t0 + i;
    const int64_t p = REVERSE ? (int64_t)(a.ntiles - 1 - t) : (int64_t)t;
    scan_tile<T, OpT, BLOCK, K, REVERSE, ALIGNED, false>(a.io, op, t, p * TILE, epoch, s_wave, &s_prefix);
    __syncthreads();
    if (threadIdx.x == 0) s_tick = take_ticket(a.ctl);
    __syncthreads();
    i = (uint32_t)s_tick;
  }
  block_retire(a.ctl, gridDim.x, epoch);
}

// The ragged tile (the last of a left-to-right scan, the first of a
// right-to-left one) in a one-workgroup launch of its own, so that its masked
// accesses do not raise the register budget of the persistent kernel.  It
// runs before k_scan when it is the first tile (publishing its prefix under
// the epoch k_scan will use) and after k_scan when it is the last (whose
// epoch k_scan has just advanced: epoch_back = 1).
template <typename T, typename Op, int BLOCK, int K, bool REVERSE, bool ALIGNED>
__global__ __launch_bounds__(BLOCK) void k_scan_ragged(ScanArgs<T> a, uint32_t t, uint32_t epoch_back) {
  __shared__ uint32_t s_epoch;
  __shared__ T s_wave[BLOCK / 64];
  __shared__ T s_prefix;
  using OpT = typename std::conditional<REVERSE, Flip<Op>, Op>::type;
  const uint32_t epoch = block_epoch(a.ctl, &s_epoch) - epoch_back;
  const int64_t p = REVERSE ? (int64_t)(a.ntiles - 1 - t) : (int64_t)t;
  scan_tile<T, OpT, BLOCK, K, REVERSE, ALIGNED, true>(a.io, OpT{}, t, p * (int64_t)(BLOCK * K * Vec<T>::N), epoch,
                                                      s_wave, &s_prefix);
}

#ifndef ACC_SCAN_K
#define ACC_SCAN_K 16
#endif
#ifndef ACC_SCAN_BLOCK
#define ACC_SCAN_BLOCK 256
#endif
template <typename T> struct ScanTune {
  static constexpr int BLOCK = ACC_SCAN_BLOCK;
  static constexpr int K = ACC_SCAN_K;   // dwordx4 vectors per thread
  static constexpr int K_UNALIGNED = 4;  // element accesses: a smaller tile
};

// Scans of fewer than kSmallTiles full-size tiles (below 2^20 4-byte or
// 2^19 8-byte elements) use 2 vectors per thread: eight times the
// workgroups for the same data, so a short scan is spread over many CUs
// instead of waiting on the few loads a handful of large tiles put in flight
// (2^16 Int64, the offsets of a 2^16-segment foldSeg: 8 tiles otherwise).
constexpr int64_t kSmallTiles = 64;
constexpr int kSmallK = 2;

template <typename T, typename Op, bool REVERSE, bool ALIGNED, int K>
int scan_launch(acc_stream_t s, ScanArgs<T> a) {
  constexpr int BLOCK = ScanTune<T>::BLOCK;
  constexpr int64_t TILE = (int64_t)BLOCK * K * Vec<T>::N;
  const bool ragged = a.io.n % TILE != 0;
  const uint32_t nfull = a.ntiles - (ragged ? 1u : 0u);
  // the ragged tile's traversal index: the last tile, or the first when reversed
  const uint32_t tr = REVERSE ? 0u : a.ntiles - 1;
  if (ragged && REVERSE) {
    hipLaunchKernelGGL((k_scan_ragged<T, Op, BLOCK, K, REVERSE, ALIGNED>), dim3(1), dim3(BLOCK), 0, s->stream, a,
                       tr, 0u);
    ACC_HIP_CHECK(hipGetLastError());
  }
  if (nfull > 0) {
    auto kern = k_scan<T, Op, BLOCK, K, REVERSE, ALIGNED>;
    const DeviceInfo* di;
    ACC_CHECK(device_info(s->device, &di));
    int64_t grid = (int64_t)di->compute_units * blocks_per_cu((const void*)kern, BLOCK);
    if (grid > (int64_t)nfull) grid = nfull;
    a.t0 = ragged && REVERSE ? 1u : 0u;
    a.nfull = nfull;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(BLOCK), 0, s->stream, a);
    ACC_HIP_CHECK(hipGetLastError());
  }
  if (ragged && !REVERSE) {
    hipLaunchKernelGGL((k_scan_ragged<T, Op, BLOCK, K, REVERSE, ALIGNED>), dim3(1), dim3(BLOCK), 0, s->stream, a,
                       tr, nfull > 0 ? 1u : 0u);
    ACC_HIP_CHECK(hipGetLastError());
  }
  return ACC_SUCCESS;
}

template <typename T, typename Op>
int scan_all(acc_stream_t s, bool reverse, bool exclusive, const T* seed, const T* in, int64_t n,
             T* out, T* total_out, T* seed_out) {
  constexpr int BLOCK = ScanTune<T>::BLOCK, VEC = Vec<T>::N;
  constexpr int NW = Tagged<T>::NW;
  if ((uintptr_t)in % sizeof(T) || (uintptr_t)out % sizeof(T))
    return fail(ACC_ERROR_INVALID_VALUE, "scan: arrays not element-aligned");
  const bool aligned = ((uintptr_t)in % 16 == 0) && ((uintptr_t)out % 16 == 0);
  const bool small = aligned && !reverse && ceil_div(n, (int64_t)BLOCK * ScanTune<T>::K * VEC) < kSmallTiles;
  const int64_t TILE =
      (int64_t)BLOCK * (small ? kSmallK : (aligned ? ScanTune<T>::K : ScanTune<T>::K_UNALIGNED)) * VEC;
  const int64_t ntiles = ceil_div(n, TILE);
  if (ntiles > 0x7fffffffLL) return fail(ACC_ERROR_UNSUPPORTED, "scan: %lld elements", (long long)n);
  SyncCtl* ctl;
  uint64_t* status;
  ACC_CHECK(stream_workspace(s, (size_t)ntiles * NW * 8, &ctl, &status));
  TileIO<T> io{in, out, n, 0u, (uint32_t)(ntiles - 1), exclusive ? 1u : 0u, seed ? 1u : 0u, seed ? *seed : T(),
               total_out, seed_out, status};
  ScanArgs<T> a{io, (uint32_t)ntiles, 0u, 0u, ctl};
  constexpr int KA = ScanTune<T>::K, KU = ScanTune<T>::K_UNALIGNED;
  if (reverse)
    return aligned ? scan_launch<T, Op, true, true, KA>(s, a) : scan_launch<T, Op, true, false, KU>(s, a);
  if (small) return scan_launch<T, Op, false, true, kSmallK>(s, a);
  return aligned ? scan_launch<T, Op, false, true, KA>(s, a) : scan_launch<T, Op, false, false, KU>(s, a);
}

// rank >= 2 scans: rows shorter than 2^20 elements in one fused launch
// (scan_rows.hip); longer rows one plain scan each, where the launch is
// amortised and the unsegmented tile is faster
template <typename T, typename Op>
int scan_rows(acc_stream_t s, acc_type ty, acc_op opc, bool reverse, bool exclusive, const T* seed, const T* in,
              int64_t outer, int64_t inner, T* out, int64_t out_row, T* sums, bool seed_at_end) {
  if (outer > 1 && inner < ((int64_t)1 << 20))
    return acc_scan_rows(s, ty, opc, reverse, exclusive, seed, in, outer, inner, out, out_row, sums, seed_at_end);
  for (int64_t r = 0; r < outer; ++r) {
    T* o = out + r * out_row;
    T* total = sums ? sums + r : (exclusive && !reverse && !sums ? o + inner : nullptr);
    T* seed_out = seed_at_end ? o + inner : nullptr;
    ACC_CHECK((scan_all<T, Op>(s, reverse, exclusive, seed, in + r * inner, inner, o, total, seed_out)));
  }
  return ACC_SUCCESS;
}

}  // namespace acc

using namespace acc;

extern "C" int acc_hip_scan(acc_stream_t s, acc_type t, acc_op op, acc_direction dir, const void* seed,
                            const void* in, int64_t outer, int64_t inner, void* out) {
  if (!s) return fail(ACC_ERROR_INVALID_VALUE, "scan: null stream");
  if (outer < 0 || inner < 0) return fail(ACC_ERROR_INVALID_VALUE, "scan: negative extent");
  const size_t B = type_size(t);
  if (!B) return fail(ACC_ERROR_UNSUPPORTED, "scan: element type %d", (int)t);
  if (outer == 0) return ACC_SUCCESS;
  if (inner == 0) {
    // scanl1 of an empty row is empty; scanl z gives [z] per row (PTX/Execute.hs:386-389)
    if (!seed) return ACC_SUCCESS;
    return acc_hip_fill_async(s, out, seed, B, (size_t)outer);
  }
  if (!in || !out) return fail(ACC_ERROR_INVALID_VALUE, "scan: null array");
  ACC_CHECK(set_device(s->device));
  const bool reverse = dir == ACC_RIGHT_TO_LEFT;
  return with_type(t, [&](auto z) {
    using T = decltype(z);
    return with_op<T>(op, [&](auto f) {
      using Op = decltype(f);
      const T* sd = (const T*)seed;
      if (!sd) return scan_rows<T, Op>(s, t, op, reverse, false, nullptr, (const T*)in, outer, inner, (T*)out, inner, nullptr, false);
      // scanl z: exclusive values + total at out[n]; scanr z: inclusive values + z at out[n]
      return scan_rows<T, Op>(s, t, op, reverse, !reverse, sd, (const T*)in, outer, inner, (T*)out, inner + 1,
                              nullptr, reverse);
    });
  });
}

extern "C" int acc_hip_scan_prime(acc_stream_t s, acc_type t, acc_op op, acc_direction dir, const void* seed,
                                  const void* in, int64_t outer, int64_t inner, void* out, void* sums) {
  if (!s) return fail(ACC_ERROR_INVALID_VALUE, "scan': null stream");
  if (!seed) return fail(ACC_ERROR_INVALID_VALUE, "scan': seed required");
  if (outer < 0 || inner < 0) return fail(ACC_ERROR_INVALID_VALUE, "scan': negative extent");
  const size_t B = type_size(t);
  if (!B) return fail(ACC_ERROR_UNSUPPORTED, "scan': element type %d", (int)t);
  if (outer == 0) return ACC_SUCCESS;
  // empty rows: no prefixes, the sums are the seed (PTX/Execute.hs:500-509)
  if (inner == 0) return acc_hip_fill_async(s, sums, seed, B, (size_t)outer);
  if (!in || !out || !sums) return fail(ACC_ERROR_INVALID_VALUE, "scan': null array");
  ACC_CHECK(set_device(s->device));
  const bool reverse = dir == ACC_RIGHT_TO_LEFT;
  return with_type(t, [&](auto z) {
    using T = decltype(z);
    return with_op<T>(op, [&](auto f) {
      using Op = decltype(f);
      return scan_rows<T, Op>(s, t, op, reverse, true, (const T*)seed, (const T*)in, outer, inner, (T*)out, inner,
                              (T*)sums, false);
    });
  });
}
