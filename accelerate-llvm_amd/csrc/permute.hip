// permute: forward permutation with an atomic read-modify-write combine.
//
// The reference issues one global `atomicrmw` per source element
// (permute_rmw, PTX/CodeGen/Permute.hs:117-213), so every update of a hot
// output slot (a Zipf histogram's top bin takes ~16% of them) serialises at
// the L2 of one channel.  Here, when the output fits in LDS, each workgroup
// keeps private copies of the output in LDS (several, so that lanes hitting
// one hot slot spread over several addresses), folds its share of the source
// into them with LDS atomics, and flushes the merged non-neutral slots to the
// output with one global atomic each.  Larger outputs take one global atomic
// per element, as the reference does.
//
// Combine functions are those the reference turns into atomics
// (LLVM/CodeGen/Permute.hs:117-145): new (+) old for + min max and or xor,
// old - new for sub, and new for a combine that ignores the old value
// (Exchange, :103-106).  Floating min/max use a compare-and-swap loop with
// fminf/fmaxf semantics, as the reference's CAS emulation does (:187-342).
// Integer and Bool results are exact (the combines commute); floating (+)
// depends on the order of atomics, as in the reference.
//
// HBM roofline: n*sizeof(I) (+ n*sizeof(T) for a manifest source) + 2*m*sizeof(T).
#include "dispatch.hpp"
#include "vecmem.hpp"
#include "wave64.hpp"

namespace acc {

// ---------------------------------------------------------- atomic combines
template <typename T> struct AsU;
template <> struct AsU<float> { using type = uint32_t; };
template <> struct AsU<double> { using type = uint64_t; };

template <typename T, typename F>
__device__ __forceinline__ void cas_loop(T* p, T v, F f) {
  using U = typename AsU<T>::type;
  U* q = (U*)p;
  U old = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (true) {
    T o;
    __builtin_memcpy(&o, &old, sizeof(T));
    T n = f(v, o);
    U nu;
    __builtin_memcpy(&nu, &n, sizeof(T));
    if (nu == old) return;
    if (__hip_atomic_compare_exchange_strong(q, &old, nu, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT))
      return;
  }
}

// Each Rmw has: identity() (neutral for LDS privatisation, when one exists),
// apply(p, v) at agent scope (global) and lds(p, v) at workgroup scope,
// flush(p, v) applying a private partial to the output, and merge(x, y)
// combining two private partials of one slot.
template <typename T> struct RmwAdd {
  static constexpr bool privatise = true;
  static __device__ __forceinline__ T identity() { return T(0); }
  static __device__ __forceinline__ void apply(T* p, T v) {
    if constexpr (std::is_floating_point<T>::value) unsafeAtomicAdd(p, v);
    else __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  static __device__ __forceinline__ void lds(T* p, T v) {
    if constexpr (std::is_floating_point<T>::value) atomicAdd(p, v);
    else __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  static __device__ __forceinline__ void flush(T* p, T v) { apply(p, v); }
  static __device__ __forceinline__ T merge(T x, T y) { return x + y; }
};

// old - new: privatised as a sum of the new values, flushed by subtraction
template <typename T> struct RmwSub {
  static constexpr bool privatise = true;
  static __device__ __forceinline__ T identity() { return T(0); }
  static __device__ __forceinline__ void apply(T* p, T v) {
    if constexpr (std::is_floating_point<T>::value) unsafeAtomicAdd(p, -v);
    else __hip_atomic_fetch_sub(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  static __device__ __forceinline__ void lds(T* p, T v) { RmwAdd<T>::lds(p, v); }
  static __device__ __forceinline__ void flush(T* p, T v) { apply(p, v); }
  static __device__ __forceinline__ T merge(T x, T y) { return x + y; }
};

template <typename T, bool IS_MAX> struct RmwMinMax {
  static constexpr bool privatise = std::is_integral<T>::value;
  static __device__ __forceinline__ T identity() {
    if constexpr (std::is_integral<T>::value)
      return IS_MAX ? std::numeric_limits<T>::lowest() : std::numeric_limits<T>::max();
    else
      return T(0);
  }
  static __device__ __forceinline__ void apply(T* p, T v) {
    if constexpr (std::is_integral<T>::value) {
      if (IS_MAX) __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      else __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      cas_loop(p, v, [](T n, T o) { return IS_MAX ? OpMax{}(n, o) : OpMin{}(n, o); });
    }
  }
  static __device__ __forceinline__ void lds(T* p, T v) {
    if constexpr (std::is_integral<T>::value) {
      if (IS_MAX) __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      else __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  }
  static __device__ __forceinline__ void flush(T* p, T v) { apply(p, v); }
  static __device__ __forceinline__ T merge(T x, T y) { return IS_MAX ? (x < y ? y : x) : (y < x ? y : x); }
};

template <typename T, int K> struct RmwBits {  // K: 0 and, 1 or, 2 xor
  static constexpr bool privatise = true;
  static __device__ __forceinline__ T identity() { return K == 0 ? (T)~(T)0 : T(0); }
  template <int SCOPE> static __device__ __forceinline__ void op(T* p, T v) {
    if (K == 0) __hip_atomic_fetch_and(p, v, __ATOMIC_RELAXED, SCOPE);
    else if (K == 1) __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, SCOPE);
    else __hip_atomic_fetch_xor(p, v, __ATOMIC_RELAXED, SCOPE);
  }
  static __device__ __forceinline__ void apply(T* p, T v) { op<__HIP_MEMORY_SCOPE_AGENT>(p, v); }
  static __device__ __forceinline__ void lds(T* p, T v) { op<__HIP_MEMORY_SCOPE_WORKGROUP>(p, v); }
  static __device__ __forceinline__ void flush(T* p, T v) { apply(p, v); }
  static __device__ __forceinline__ T merge(T x, T y) { return K == 0 ? (x & y) : K == 1 ? (x | y) : (x ^ y); }
};

template <typename T> struct RmwExchange {
  static constexpr bool privatise = false;
  static __device__ __forceinline__ T identity() { return T(0); }
  static __device__ __forceinline__ void apply(T* p, T v) { *p = v; }
  static __device__ __forceinline__ void lds(T*, T) {}
  static __device__ __forceinline__ void flush(T* p, T v) { *p = v; }
  static __device__ __forceinline__ T merge(T x, T) { return x; }
};

// The type the private LDS copies accumulate in.  LDS float32 atomic adds
// run ~20x slower than integer or float64 ones on gfx950 (measured: 0.2 vs
// 4.3 and 1.7 Tadd/s over 4096 bins), so float32 sums accumulate in float64
// copies and are rounded to float32 once, at the flush.
template <typename T, typename R> struct LdsAcc { using type = T; };
template <> struct LdsAcc<float, RmwAdd<float>> { using type = double; };
template <> struct LdsAcc<float, RmwSub<float>> { using type = double; };

template <typename T, typename R, typename L>
__device__ __forceinline__ void lds_update(L* p, T v) {
  if constexpr (std::is_same<L, T>::value) R::lds(p, v);
  else atomicAdd(p, (L)v);  // a float64 copy of a float32 sum
}

template <typename T, typename I> struct PermArgs {
  const I* idx;
  const T* src;  // null: constant source
  T fill;
  int64_t n;
  T* out;
  int64_t m;
  int64_t head;   // index elements before the first 16-byte aligned one
  int64_t nvec;   // whole 16-byte index vectors after the head
  int64_t chunk;  // vectors per workgroup
  int32_t copies; // private copies of the output per workgroup (k_permute_lds)
};

template <typename I> __device__ __forceinline__ bool in_range(I j, int64_t m) {
  return (int64_t)j >= 0 && (int64_t)j < m;
}

// Private LDS copies of the whole output per workgroup: lane l updates copy
// l % copies, so lanes of one wave that hit the same hot slot (Zipf bins)
// spread over several addresses, and the copy stride m + 1 puts one slot's
// copies in different banks.  An index outside the output updates a dummy
// word of the lane's own (after the copies) instead of branching, so an
// update is an address select and one LDS atomic.  At the end the copies of
// each slot are merged and flushed.
template <typename T, typename I, typename R, bool CONST_SRC, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_permute_lds(PermArgs<T, I> a) {
  extern __shared__ __align__(16) unsigned char s_raw[];
  using L = typename LdsAcc<T, R>::type;
  L* bins = (L*)s_raw;
  constexpr int VI = 16 / sizeof(I);
  const L ident = (L)R::identity();
  const int64_t stride = a.m + 1;
  const int64_t words = stride * a.copies;  // the dummy words follow
  for (int64_t b = threadIdx.x; b < words + 64; b += BLOCK) bins[b] = ident;
  __syncthreads();

  const uint32_t m32 = (uint32_t)a.m;  // m < 2^31: the output fits in LDS
  L* const mine = bins + (int64_t)(lane_id() % a.copies) * stride;
  L* const dummy = bins + words + lane_id();
  auto update = [&](I j, T v) {
    L* p = (uint64_t)(int64_t)j < (uint64_t)m32 ? mine + (int64_t)j : dummy;
    lds_update<T, R>(p, v);
  };

  // U index vectors (and their source values) in flight per thread before
  // the LDS updates: one load at a time leaves the loop latency-bound.  The
  // loads are unconditional (a vector past the chunk re-reads its last one)
  // so that each wait names only the loads it needs; the updates skip them.
  constexpr int U = CONST_SRC ? 8 : 4;
  const I* body = a.idx + a.head;
  const int64_t vb = (int64_t)blockIdx.x * a.chunk;
  const int64_t ve = min(vb + a.chunk, a.nvec);
  for (int64_t i0 = vb + threadIdx.x; i0 < ve; i0 += (int64_t)BLOCK * U) {
    Vec<I> iv[U];
    T sv[U][VI];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = min(i0 + (int64_t)u * BLOCK, ve - 1);
      iv[u] = load16_stream(body + i * VI);
      if (!CONST_SRC) {
#pragma unroll
        for (int k = 0; k < VI; ++k) sv[u][k] = a.src[a.head + i * VI + k];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = i0 + (int64_t)u * BLOCK;
      if (i < ve) {
#pragma unroll
        for (int k = 0; k < VI; ++k) update(iv[u].e[k], CONST_SRC ? a.fill : sv[u][k]);
      }
    }
  }
  // head and tail elements (outside whole vectors): the last workgroup
  if (blockIdx.x == gridDim.x - 1) {
    const int64_t tail0 = a.head + a.nvec * VI;
    for (int64_t e = threadIdx.x; e < a.head; e += BLOCK) {
      const I j = a.idx[e];
      if (in_range(j, a.m)) lds_update<T, R>(&bins[j], CONST_SRC ? a.fill : a.src[e]);
    }
    for (int64_t e = tail0 + threadIdx.x; e < a.n; e += BLOCK) {
      const I j = a.idx[e];
      if (in_range(j, a.m)) lds_update<T, R>(&bins[j], CONST_SRC ? a.fill : a.src[e]);
    }
  }
  __syncthreads();
  for (int64_t b = threadIdx.x; b < a.m; b += BLOCK) {
    L v = bins[b];
    for (int c = 1; c < a.copies; ++c) {
      if constexpr (std::is_same<L, T>::value) v = R::merge(v, bins[c * stride + b]);
      else v += bins[c * stride + b];
    }
    if (v != ident) R::flush(&a.out[b], (T)v);
  }
}

// One global atomic per element (large outputs, Exchange, floating min/max).
template <typename T, typename I, typename R, bool CONST_SRC, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_permute_direct(PermArgs<T, I> a) {
  const int64_t stride = (int64_t)gridDim.x * BLOCK;
  for (int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x; e < a.n; e += stride) {
    const I j = a.idx[e];
    if (in_range(j, a.m)) R::apply(&a.out[j], CONST_SRC ? a.fill : a.src[e]);
  }
}

template <typename T, typename I, typename R>
int permute_launch(acc_stream_t s, T* out, int64_t m, const T* src, T fill, const I* idx, int64_t n) {
  constexpr int BLOCK = 512;
  constexpr int VI = 16 / sizeof(I);
  const DeviceInfo* di;
  ACC_CHECK(device_info(s->device, &di));
  if ((uintptr_t)idx % sizeof(I)) return fail(ACC_ERROR_INVALID_VALUE, "permute: index not element-aligned");
  int64_t head = (int64_t)(((16 - (uintptr_t)idx % 16) % 16) / sizeof(I));
  if (head > n) head = n;
  const int64_t nvec = (n - head) / VI;
  PermArgs<T, I> a{idx, src, fill, n, out, m, head, nvec, 0, 1};
  // private copies: as many as fit in 64 KiB of LDS (with the 64 dummy
  // words), at most ACC_PERM_COPIES
#ifndef ACC_PERM_COPIES
#define ACC_PERM_COPIES 2
#endif
  using L = typename LdsAcc<T, R>::type;
  constexpr int64_t LDS_WORDS = 64 * 1024 / sizeof(L);
  const int64_t copies = min((int64_t)ACC_PERM_COPIES, (LDS_WORDS - 64) / (m + 1));
  const bool privatise = R::privatise && copies >= 1;
  if (privatise) {
    a.copies = (int32_t)copies;
    const size_t lds = (size_t)(copies * (m + 1) + 64) * sizeof(L);
    int64_t nb = ceil_div(nvec, (int64_t)BLOCK * 8);
#ifndef ACC_PERM_WGS
#define ACC_PERM_WGS 2
#endif
    const int64_t maxb = (int64_t)di->compute_units * ACC_PERM_WGS;
    if (nb > maxb) nb = maxb;
    if (nb < 1) nb = 1;
    a.chunk = nvec ? ceil_div(nvec, nb) : 0;
    if (a.chunk) nb = max((int64_t)1, ceil_div(nvec, a.chunk));
    if (src) hipLaunchKernelGGL((k_permute_lds<T, I, R, false, BLOCK>), dim3((unsigned)nb), dim3(BLOCK), lds, s->stream, a);
    else hipLaunchKernelGGL((k_permute_lds<T, I, R, true, BLOCK>), dim3((unsigned)nb), dim3(BLOCK), lds, s->stream, a);
  } else {
    int64_t nb = ceil_div(n, (int64_t)BLOCK);
    const int64_t maxb = (int64_t)di->compute_units * 8;
    if (nb > maxb) nb = maxb;
    if (src) hipLaunchKernelGGL((k_permute_direct<T, I, R, false, BLOCK>), dim3((unsigned)nb), dim3(BLOCK), 0, s->stream, a);
    else hipLaunchKernelGGL((k_permute_direct<T, I, R, true, BLOCK>), dim3((unsigned)nb), dim3(BLOCK), 0, s->stream, a);
  }
  ACC_HIP_CHECK(hipGetLastError());
  return ACC_SUCCESS;
}

template <typename T, typename F>
int with_rmw(acc_op op, F&& f) {
  switch (op) {
    case ACC_OP_ADD: return f(RmwAdd<T>{});
    case ACC_OP_SUB: return f(RmwSub<T>{});
    case ACC_OP_MIN: return f(RmwMinMax<T, false>{});
    case ACC_OP_MAX: return f(RmwMinMax<T, true>{});
    case ACC_OP_EXCHANGE: return f(RmwExchange<T>{});
    default: break;
  }
  if constexpr (std::is_integral<T>::value) {
    switch (op) {
      case ACC_OP_AND: return f(RmwBits<T, 0>{});
      case ACC_OP_OR: return f(RmwBits<T, 1>{});
      case ACC_OP_XOR: return f(RmwBits<T, 2>{});
      default: break;
    }
  }
  return fail(ACC_ERROR_UNSUPPORTED, "permute: combine function %d for this element type", (int)op);
}

}  // namespace acc

using namespace acc;

extern "C" int acc_hip_permute(acc_stream_t s, acc_type t, acc_op combine, const void* defaults, void* out,
                               int64_t m, const void* src, const void* src_fill, acc_type index_type,
                               const void* index, int64_t n) {
  if (!s) return fail(ACC_ERROR_INVALID_VALUE, "permute: null stream");
  if (m < 0 || n < 0) return fail(ACC_ERROR_INVALID_VALUE, "permute: negative extent");
  if (index_type != ACC_INT32 && index_type != ACC_INT64)
    return fail(ACC_ERROR_UNSUPPORTED, "permute: index type must be Int32 or Int64");
  const size_t B = type_size(t);
  if (!B) return fail(ACC_ERROR_UNSUPPORTED, "permute: element type %d", (int)t);
  if (m > 0 && (!out || !defaults)) return fail(ACC_ERROR_INVALID_VALUE, "permute: null output");
  if (n > 0 && !index) return fail(ACC_ERROR_INVALID_VALUE, "permute: null index");
  if (n > 0 && !src && !src_fill) return fail(ACC_ERROR_INVALID_VALUE, "permute: no source");
  ACC_CHECK(set_device(s->device));
  // clone the defaults unless updating in place (PTX/Execute.hs:623-625)
  if (out != defaults && m > 0)
    ACC_HIP_CHECK(hipMemcpyAsync(out, defaults, (size_t)m * B, hipMemcpyDeviceToDevice, s->stream));
  if (n == 0 || m == 0) return ACC_SUCCESS;
  return with_type(t, [&](auto z) {
    using T = decltype(z);
    const T fill = src_fill ? *(const T*)src_fill : T();
    return with_rmw<T>(combine, [&](auto r) {
      using R = decltype(r);
      if (index_type == ACC_INT32)
        return permute_launch<T, int32_t, R>(s, (T*)out, m, (const T*)src, fill, (const int32_t*)index, n);
      return permute_launch<T, int64_t, R>(s, (T*)out, m, (const T*)src, fill, (const int64_t*)index, n);
    });
  });
}
