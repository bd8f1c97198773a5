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
  uint32_t ntiles;
  SyncCtl* ctl;
};

// Persistent workgroups: each takes tiles by ticket until none are left,
// prefetching the next ticket while it scans the current tile.
template <typename T, typename Op, int BLOCK, int K, bool REVERSE, bool ALIGNED>
__global__ __launch_bounds__(BLOCK) void k_scan(ScanArgs<T> a) {
  constexpr int VEC = Vec<T>::N;
  constexpr int TILE = BLOCK * K * VEC;
  __shared__ uint64_t s_tick[2];
  __shared__ Maybe<T> s_wave[BLOCK / 64];
  __shared__ Maybe<T> s_prefix;
  using OpT = typename std::conditional<REVERSE, Flip<Op>, Op>::type;
  OpT op{};
  if (threadIdx.x == 0) s_tick[0] = take_ticket(a.ctl);
  __syncthreads();
  const uint64_t tk0 = s_tick[0];
  const uint32_t epoch = (uint32_t)(tk0 >> 32);
  uint32_t t = (uint32_t)tk0;
  int cur = 0;
  while (t < a.ntiles) {
    if (threadIdx.x == 0) s_tick[cur ^ 1] = take_ticket(a.ctl);
    const int64_t p = REVERSE ? (int64_t)(a.ntiles - 1 - t) : (int64_t)t;
    const int64_t base = p * TILE;
    if (base + TILE <= a.io.n)
      scan_tile<T, OpT, BLOCK, K, REVERSE, ALIGNED, false>(a.io, op, t, base, epoch, s_wave, &s_prefix);
    else
      scan_tile<T, OpT, BLOCK, K, REVERSE, ALIGNED, true>(a.io, op, t, base, epoch, s_wave, &s_prefix);
    __syncthreads();
    cur ^= 1;
    t = (uint32_t)s_tick[cur];
  }
  block_retire(a.ctl, gridDim.x, epoch);
}

#ifndef ACC_SCAN_K
#define ACC_SCAN_K 8
#endif
template <typename T> struct ScanTune {
  static constexpr int BLOCK = 256;
  static constexpr int K = ACC_SCAN_K;  // dwordx4 vectors per thread
};

template <typename T, typename Op, bool REVERSE, bool ALIGNED>
int scan_launch(acc_stream_t s, const ScanArgs<T>& a) {
  constexpr int BLOCK = ScanTune<T>::BLOCK, K = ScanTune<T>::K;
  auto kern = k_scan<T, Op, BLOCK, K, REVERSE, ALIGNED>;
  const DeviceInfo* di;
  ACC_CHECK(device_info(s->device, &di));
  int64_t grid = (int64_t)di->compute_units * blocks_per_cu((const void*)kern, BLOCK);
  if (grid > (int64_t)a.ntiles) grid = a.ntiles;
  hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(BLOCK), 0, s->stream, a);
  ACC_HIP_CHECK(hipGetLastError());
  return ACC_SUCCESS;
}

template <typename T, typename Op>
int scan_all(acc_stream_t s, bool reverse, bool exclusive, const T* seed, const T* in, int64_t n,
             T* out, T* total_out, T* seed_out) {
  constexpr int BLOCK = ScanTune<T>::BLOCK, K = ScanTune<T>::K, VEC = Vec<T>::N;
  constexpr int64_t TILE = (int64_t)BLOCK * K * VEC;
  constexpr int NW = Tagged<T>::NW;
  const int64_t ntiles = ceil_div(n, TILE);
  if (ntiles > 0x7fffffffLL) return fail(ACC_ERROR_UNSUPPORTED, "scan: %lld elements", (long long)n);
  if ((uintptr_t)in % sizeof(T) || (uintptr_t)out % sizeof(T))
    return fail(ACC_ERROR_INVALID_VALUE, "scan: arrays not element-aligned");
  const bool aligned = ((uintptr_t)in % 16 == 0) && ((uintptr_t)out % 16 == 0);
  SyncCtl* ctl;
  uint64_t* status;
  ACC_CHECK(stream_workspace(s, (size_t)ntiles * NW * 8, &ctl, &status));
  TileIO<T> io{in, out, n, 0u, (uint32_t)(ntiles - 1), exclusive ? 1u : 0u, seed ? 1u : 0u, seed ? *seed : T(),
               total_out, seed_out, status};
  ScanArgs<T> a{io, (uint32_t)ntiles, ctl};
  if (reverse) return aligned ? scan_launch<T, Op, true, true>(s, a) : scan_launch<T, Op, true, false>(s, a);
  return aligned ? scan_launch<T, Op, false, true>(s, a) : scan_launch<T, Op, false, false>(s, a);
}

// rank >= 2 scans: one call per row for now (each row is an independent scan)
template <typename T, typename Op>
int scan_rows(acc_stream_t s, bool reverse, bool exclusive, const T* seed, const T* in,
              int64_t outer, int64_t inner, T* out, int64_t out_row, T* sums, bool seed_at_end) {
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
      if (!sd) return scan_rows<T, Op>(s, reverse, false, nullptr, (const T*)in, outer, inner, (T*)out, inner, nullptr, false);
      // scanl z: exclusive values + total at out[n]; scanr z: inclusive values + z at out[n]
      return scan_rows<T, Op>(s, reverse, !reverse, sd, (const T*)in, outer, inner, (T*)out, inner + 1,
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
      return scan_rows<T, Op>(s, reverse, true, (const T*)seed, (const T*)in, outer, inner, (T*)out, inner,
                              (T*)sums, false);
    });
  });
}
