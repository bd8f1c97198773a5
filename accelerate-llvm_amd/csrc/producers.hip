// Producers: map, zipWith, generate, transform and backpermute.
//
// The reference compiles each producer from the program's own functions into
// a grid-stride loop over the output's linear index (map: PTX/CodeGen/Map.hs
// :37-63, generate: PTX/CodeGen/Generate.hs:37-62, transform:
// PTX/CodeGen/Transform.hs:37-67) and launches it from mapOp / generateOp /
// transformOp / backpermuteOp (PTX/Execute.hs:117-206); backpermute is
// transform with the identity (:206).  Without run-time compilation the
// element functions come from a fixed set compiled ahead of time:
//   map f       f in {id, negate, abs, a*x + b, conversion}
//   zipWith f   f a binary combine (+ * min max and or xor, and a - b)
//   generate g  g in {fill a, iota a + i*b}
//   transform   index maps {gather by an index array, reverse, transpose of
//               a rank-2 array}, then f in {id, negate, abs, a*x + b}
// Integer arithmetic wraps; floating a*x + b and a + i*b round after the
// product and after the sum (no fused multiply-add), so results are the
// IEEE single-rounding results of each step.
//
// All of these are HBM streams: every thread moves whole 16-byte vectors on
// both sides (several in flight), with non-temporal loads and stores; the
// gather reads its indices as vectors and the gathered elements one by one,
// and the transpose goes through 64 x 64 LDS tiles so that both its reads
// and its writes are row-contiguous.  Unaligned views take element-wise
// kernels.
//
// HBM roofline: map n*(Bin + Bout), zipWith 3*n*B, generate n*B, gather
// n*(I + Bin + Bout), reverse and transpose n*(Bin + Bout).
#include <type_traits>
#include "dispatch.hpp"
#include "host.hpp"
#include "vecmem.hpp"

namespace acc {

// ------------------------------------------------------------ element functions
// integer arithmetic goes through the unsigned type of the same width (wraps)
template <typename T, bool = std::is_integral<T>::value> struct UnsignedOfT {
  using type = T;
};
template <typename T> struct UnsignedOfT<T, true> {
  using type = typename std::make_unsigned<T>::type;
};
template <typename T> using UnsignedOf = typename UnsignedOfT<T>::type;

// negate and abs as the reference generates them
// (LLVM/CodeGen/Arithmetic.hs:69-86): negate x = x * (-1) (wrapping for
// integers), abs = fabs for floating types, llvm.abs (abs minBound ==
// minBound) for signed and the identity for unsigned integers.

template <typename T> struct FId {
  __device__ __forceinline__ T operator()(T x) const { return x; }
};
template <typename T> struct FNeg {
  __device__ __forceinline__ T operator()(T x) const {
    if constexpr (std::is_integral<T>::value) return (T)((UnsignedOf<T>)0 - (UnsignedOf<T>)x);
    else return -x;
  }
};
template <typename T> struct FAbs {
  __device__ __forceinline__ T operator()(T x) const {
    if constexpr (std::is_unsigned<T>::value) return x;
    else if constexpr (std::is_integral<T>::value) return x < 0 ? (T)((UnsignedOf<T>)0 - (UnsignedOf<T>)x) : x;
    else if constexpr (sizeof(T) == 4) return __builtin_fabsf(x);
    else return __builtin_fabs(x);
  }
};
template <typename T> struct FAffine {
  T a, b;
  __device__ __forceinline__ T operator()(T x) const {
    if constexpr (std::is_integral<T>::value) {
      using U = UnsignedOf<T>;
      return (T)((U)a * (U)x + (U)b);
    } else {
#pragma clang fp contract(off)
      return a * x + b;
    }
  }
};
template <typename Tin, typename Tout> struct FCast {
  __device__ __forceinline__ Tout operator()(Tin x) const { return (Tout)x; }
};

// a - b (zipWith's subtraction; integer forms wrap)
struct OpSubZ {
  template <typename T> __device__ __forceinline__ T operator()(T a, T b) const {
    if constexpr (std::is_integral<T>::value) return (T)((UnsignedOf<T>)a - (UnsignedOf<T>)b);
    else return a - b;
  }
};

template <typename T> struct GFill {
  T a;
  __device__ __forceinline__ T operator()(int64_t) const { return a; }
};
template <typename T> struct GIota {
  T a, b;
  __device__ __forceinline__ T operator()(int64_t i) const {
    if constexpr (std::is_integral<T>::value) {
      using U = UnsignedOf<T>;
      return (T)((U)a + (U)i * (U)b);
    } else {
#pragma clang fp contract(off)
      return a + (T)i * b;
    }
  }
};

// ------------------------------------------------------------------ kernels
constexpr int PB = 256;  // threads per workgroup
constexpr int PU = 4;    // vector chunks in flight per thread

template <int BYTES> struct Raw {
  static_assert(BYTES % 4 == 0, "whole dwords");
  uint32_t w[BYTES / 4];
};
template <int BYTES> __device__ __forceinline__ Raw<BYTES> ld_stream(const void* p) {
  Raw<BYTES> r;
  if constexpr (BYTES % 16 == 0) {
#pragma unroll
    for (int v = 0; v < BYTES / 16; ++v) {
      const u32x4 x = __builtin_nontemporal_load((const u32x4*)p + v);
      r.w[4 * v] = x[0], r.w[4 * v + 1] = x[1], r.w[4 * v + 2] = x[2], r.w[4 * v + 3] = x[3];
    }
  } else {
    r = *(const Raw<BYTES>*)p;
  }
  return r;
}
template <int BYTES> __device__ __forceinline__ void st_stream(void* p, const Raw<BYTES>& r) {
  if constexpr (BYTES % 16 == 0) {
#pragma unroll
    for (int v = 0; v < BYTES / 16; ++v) {
      u32x4 x;
      x[0] = r.w[4 * v], x[1] = r.w[4 * v + 1], x[2] = r.w[4 * v + 2], x[3] = r.w[4 * v + 3];
      __builtin_nontemporal_store(x, (u32x4*)p + v);
    }
  } else {
    *(Raw<BYTES>*)p = r;  // 4-byte aligned (an 8-byte output chunk)
  }
}

// elements per chunk: 16 bytes of the narrower of two types
template <typename A, typename B> constexpr int chunk_elems() {
  return 16 / (int)(sizeof(A) < sizeof(B) ? sizeof(A) : sizeof(B));
}

// map: out[i] = f(in[i]); chunks of E elements, the last (n mod E) elements
// by workgroup 0.  Loads are unconditional (a chunk past the end re-reads the
// last one), so every load of an iteration is in flight before the first use.
template <typename Tin, typename Tout, typename F>
__global__ __launch_bounds__(PB) void k_map(const Tin* in, Tout* out, int64_t n, F f) {
  constexpr int E = chunk_elems<Tin, Tout>();
  constexpr int BI = E * (int)sizeof(Tin), BO = E * (int)sizeof(Tout);
  const int64_t nchunk = n / E;
  const int64_t stride = (int64_t)gridDim.x * PB;
  for (int64_t c0 = (int64_t)blockIdx.x * PB * PU + threadIdx.x; c0 < nchunk; c0 += stride * PU) {
    Raw<BI> r[PU];
#pragma unroll
    for (int u = 0; u < PU; ++u) r[u] = ld_stream<BI>(in + min(c0 + u * PB, nchunk - 1) * E);
#pragma unroll
    for (int u = 0; u < PU; ++u) {
      const int64_t c = c0 + u * PB;
      if (c < nchunk) {
        Tin x[E];
        __builtin_memcpy(x, &r[u], BI);
        Tout y[E];
#pragma unroll
        for (int j = 0; j < E; ++j) y[j] = f(x[j]);
        Raw<BO> w;
        __builtin_memcpy(&w, y, BO);
        st_stream<BO>(out + c * E, w);
      }
    }
  }
  if (blockIdx.x == 0) {
    const int64_t i = nchunk * E + threadIdx.x;
    if (i < n) out[i] = f(in[i]);
  }
}

// element-wise forms for unaligned views
template <typename Tin, typename Tout, typename F>
__global__ __launch_bounds__(PB) void k_map1(const Tin* in, Tout* out, int64_t n, F f) {
  const int64_t stride = (int64_t)gridDim.x * PB;
  for (int64_t i = (int64_t)blockIdx.x * PB + threadIdx.x; i < n; i += stride) out[i] = f(in[i]);
}

template <typename T, typename Op>
__global__ __launch_bounds__(PB) void k_zip(const T* a, const T* b, T* out, int64_t n, Op op) {
  constexpr int E = 16 / (int)sizeof(T);
  const int64_t nchunk = n / E;
  const int64_t stride = (int64_t)gridDim.x * PB;
  for (int64_t c0 = (int64_t)blockIdx.x * PB * PU + threadIdx.x; c0 < nchunk; c0 += stride * PU) {
    Vec<T> va[PU], vb[PU];
#pragma unroll
    for (int u = 0; u < PU; ++u) {
      const int64_t c = min(c0 + u * PB, nchunk - 1);
      va[u] = load16_stream(a + c * E);
      vb[u] = load16_stream(b + c * E);
    }
#pragma unroll
    for (int u = 0; u < PU; ++u) {
      const int64_t c = c0 + u * PB;
      if (c < nchunk) {
        Vec<T> y;
#pragma unroll
        for (int j = 0; j < E; ++j) y.e[j] = op(va[u].e[j], vb[u].e[j]);
        store16_stream(out + c * E, y);
      }
    }
  }
  if (blockIdx.x == 0) {
    const int64_t i = nchunk * E + threadIdx.x;
    if (i < n) out[i] = op(a[i], b[i]);
  }
}

template <typename T, typename Op>
__global__ __launch_bounds__(PB) void k_zip1(const T* a, const T* b, T* out, int64_t n, Op op) {
  const int64_t stride = (int64_t)gridDim.x * PB;
  for (int64_t i = (int64_t)blockIdx.x * PB + threadIdx.x; i < n; i += stride) out[i] = op(a[i], b[i]);
}

template <typename T, typename G>
__global__ __launch_bounds__(PB) void k_gen(T* out, int64_t n, G g, int aligned) {
  constexpr int E = 16 / (int)sizeof(T);
  const int64_t stride = (int64_t)gridDim.x * PB;
  if (!aligned) {
    for (int64_t i = (int64_t)blockIdx.x * PB + threadIdx.x; i < n; i += stride) out[i] = g(i);
    return;
  }
  const int64_t nchunk = n / E;
  for (int64_t c0 = (int64_t)blockIdx.x * PB * PU + threadIdx.x; c0 < nchunk; c0 += stride * PU) {
#pragma unroll
    for (int u = 0; u < PU; ++u) {
      const int64_t c = c0 + u * PB;
      if (c < nchunk) {
        Vec<T> y;
#pragma unroll
        for (int j = 0; j < E; ++j) y.e[j] = g(c * E + j);
        // ordinary stores: for this write-only stream they measured faster
        // than non-temporal ones (2^28 Int32 203 -> 193 us), which win for
        // the kernels that also read
        store16(out + c * E, y);
      }
    }
  }
  if (blockIdx.x == 0) {
    const int64_t i = nchunk * E + threadIdx.x;
    if (i < n) out[i] = g(i);
  }
}

// gather: out[i] = f(in[index[i]]) for 0 <= index[i] < n_in, else 0 (no
// read).  Index chunks of 16 bytes; the gathered element loads are
// unconditional (an out-of-range index reads element 0), all of an
// iteration's loads in flight together.
template <typename T, typename I, typename F>
__global__ __launch_bounds__(PB) void k_gather(const T* in, int64_t n_in, const I* idx, T* out, int64_t n, F f) {
  constexpr int E = 16 / (int)sizeof(I);
  constexpr int BO = E * (int)sizeof(T);
  const int64_t nchunk = n / E;
  const int64_t stride = (int64_t)gridDim.x * PB;
  for (int64_t c0 = (int64_t)blockIdx.x * PB * PU + threadIdx.x; c0 < nchunk; c0 += stride * PU) {
    Vec<I> iv[PU];
#pragma unroll
    for (int u = 0; u < PU; ++u) iv[u] = load16_stream(idx + min(c0 + u * PB, nchunk - 1) * E);
    T x[PU][E];
#pragma unroll
    for (int u = 0; u < PU; ++u)
#pragma unroll
      for (int j = 0; j < E; ++j) {
        const int64_t k = (int64_t)iv[u].e[j];
        x[u][j] = in[(uint64_t)k < (uint64_t)n_in ? k : 0];
      }
#pragma unroll
    for (int u = 0; u < PU; ++u) {
      const int64_t c = c0 + u * PB;
      if (c < nchunk) {
        T y[E];
#pragma unroll
        for (int j = 0; j < E; ++j) y[j] = (uint64_t)(int64_t)iv[u].e[j] < (uint64_t)n_in ? f(x[u][j]) : T();
        Raw<BO> w;
        __builtin_memcpy(&w, y, BO);
        st_stream<BO>(out + c * E, w);
      }
    }
  }
  if (blockIdx.x == 0) {
    const int64_t i = nchunk * E + threadIdx.x;
    if (i < n) {
      const int64_t k = (int64_t)idx[i];
      out[i] = (uint64_t)k < (uint64_t)n_in ? f(in[k]) : T();
    }
  }
}

template <typename T, typename I, typename F>
__global__ __launch_bounds__(PB) void k_gather1(const T* in, int64_t n_in, const I* idx, T* out, int64_t n, F f) {
  const int64_t stride = (int64_t)gridDim.x * PB;
  for (int64_t i = (int64_t)blockIdx.x * PB + threadIdx.x; i < n; i += stride) {
    const int64_t k = (int64_t)idx[i];
    out[i] = (uint64_t)k < (uint64_t)n_in ? f(in[k]) : T();
  }
}

// reverse: out[i] = f(in[n - 1 - i]); output chunks of 16 bytes.  VIN (the
// input 16-byte aligned and n a multiple of the chunk): each chunk's source
// is one aligned 16-byte block, read whole and reversed in registers;
// otherwise the reads go element by element (a wave's reads are still one
// contiguous, descending range).
template <typename T, typename F, bool VIN>
__global__ __launch_bounds__(PB) void k_reverse(const T* in, T* out, int64_t n, F f, int aligned) {
  constexpr int E = 16 / (int)sizeof(T);
  const int64_t stride = (int64_t)gridDim.x * PB;
  if (!aligned) {
    for (int64_t i = (int64_t)blockIdx.x * PB + threadIdx.x; i < n; i += stride) out[i] = f(in[n - 1 - i]);
    return;
  }
  const int64_t nchunk = n / E;
  for (int64_t c0 = (int64_t)blockIdx.x * PB * PU + threadIdx.x; c0 < nchunk; c0 += stride * PU) {
    T x[PU][E];
#pragma unroll
    for (int u = 0; u < PU; ++u) {
      const int64_t c = min(c0 + u * PB, nchunk - 1);
      if (VIN) {
        const Vec<T> v = load16_stream(in + (n - (c + 1) * E));  // elements n-(c+1)E .. n-1-cE
#pragma unroll
        for (int j = 0; j < E; ++j) x[u][j] = v.e[E - 1 - j];
      } else {
#pragma unroll
        for (int j = 0; j < E; ++j) x[u][j] = in[n - 1 - (c * E + j)];
      }
    }
#pragma unroll
    for (int u = 0; u < PU; ++u) {
      const int64_t c = c0 + u * PB;
      if (c < nchunk) {
        Vec<T> y;
#pragma unroll
        for (int j = 0; j < E; ++j) y.e[j] = f(x[u][j]);
        store16_stream(out + c * E, y);
      }
    }
  }
  if (blockIdx.x == 0) {
    const int64_t i = nchunk * E + threadIdx.x;
    if (i < n) out[i] = f(in[n - 1 - i]);
  }
}

// transpose of a rows x cols array: out[c][r] = f(in[r][c]).  One 64 x 64
// tile per workgroup through LDS (row stride 65: the column reads of the
// transposed write are free of bank conflicts), read and written along rows.
// Tiles inside the array move 16-byte vectors when every row starts 16-byte
// aligned and both row pitches are powers of two (vec_ok); edge tiles and
// other shapes move one element per lane.
constexpr int TT = 64;
template <typename T, typename F>
__global__ __launch_bounds__(PB) void k_transpose(const T* in, T* out, int64_t rows, int64_t cols, int64_t tiles_c,
                                                  uint32_t vec_ok, uint32_t diag, F f) {
  __shared__ T tile[TT][TT + 1];
  // diag: tile rows in diagonal order (row band shifted by the column tile),
  // so the workgroups in flight write different column ranges of their
  // output rows; with a power-of-two output row pitch the plain order puts
  // all of them on the same HBM channels
  const int64_t tiles_r = gridDim.x / tiles_c;
  const int64_t tc = (int64_t)blockIdx.x % tiles_c;
  const int64_t tr = diag ? ((int64_t)blockIdx.x / tiles_c + tc) % tiles_r : (int64_t)blockIdx.x / tiles_c;
  const int64_t r0 = tr * TT, c0 = tc * TT;
  if (vec_ok && r0 + TT <= rows && c0 + TT <= cols) {
    // a tile inside the array with 16-byte-aligned rows on both sides: one
    // dwordx4 per thread and pass each way (all loads issued first); the
    // 65-element LDS row stride keeps both the row writes and the column
    // reads at two lanes per bank
    constexpr int VEC = Vec<T>::N;
    constexpr int TPR = TT / VEC;  // threads per tile row
    constexpr int RPP = PB / TPR;  // tile rows per pass
    constexpr int NP = TT / RPP;   // passes
    const int cv = threadIdx.x % TPR, rr = threadIdx.x / TPR;
    Vec<T> v[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) v[p] = load16_stream(in + (r0 + rr + p * RPP) * cols + c0 + cv * VEC);
#pragma unroll
    for (int p = 0; p < NP; ++p)
#pragma unroll
      for (int j = 0; j < VEC; ++j) tile[rr + p * RPP][cv * VEC + j] = v[p].e[j];
    __syncthreads();
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const int c = rr + p * RPP;  // output row c0 + c (input column)
      Vec<T> o;
#pragma unroll
      for (int j = 0; j < VEC; ++j) o.e[j] = f(tile[cv * VEC + j][c]);
      store16_stream(out + (c0 + c) * rows + r0 + cv * VEC, o);
    }
    return;
  }
  const int tx = threadIdx.x % TT, ty = threadIdx.x / TT;  // 64 x 4
  T v[TT / 4];
#pragma unroll
  for (int k = 0; k < TT / 4; ++k) {
    const int64_t r = r0 + ty + 4 * k, c = c0 + tx;
    v[k] = r < rows && c < cols ? in[r * cols + c] : T();
  }
#pragma unroll
  for (int k = 0; k < TT / 4; ++k) tile[ty + 4 * k][tx] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < TT / 4; ++k) {
    const int64_t c = c0 + ty + 4 * k, r = r0 + tx;  // output row c, column r
    if (c < cols && r < rows) out[c * rows + r] = f(tile[tx][ty + 4 * k]);
  }
}

// ------------------------------------------------------------------- launch
inline unsigned stream_grid(int64_t work, int cus) {
  const int64_t want = ceil_div(work, (int64_t)PB * PU);
  const int64_t cap = (int64_t)cus * 32;
  return (unsigned)max((int64_t)1, min(want, cap));
}
inline bool al16(const void* p) { return (uintptr_t)p % 16 == 0; }

template <typename T, typename F> int with_unop(acc_unop f, const void* params, F&& k) {
  switch (f) {
    case ACC_UNOP_ID: return k(FId<T>{});
    case ACC_UNOP_NEG: return k(FNeg<T>{});
    case ACC_UNOP_ABS: return k(FAbs<T>{});
    case ACC_UNOP_AFFINE: {
      if (!params) return fail(ACC_ERROR_INVALID_VALUE, "affine map: null params");
      FAffine<T> a;
      __builtin_memcpy(&a.a, params, sizeof(T));
      __builtin_memcpy(&a.b, (const char*)params + sizeof(T), sizeof(T));
      return k(a);
    }
    default: break;
  }
  return fail(ACC_ERROR_UNSUPPORTED, "element function %d", (int)f);
}

template <typename Tin, typename Tout, typename F>
int launch_map(acc_stream_t s, const Tin* in, Tout* out, int64_t n, F f, int cus) {
  constexpr int E = chunk_elems<Tin, Tout>();
  if (al16(in) && al16(out))
    hipLaunchKernelGGL((k_map<Tin, Tout, F>), dim3(stream_grid(n / E, cus)), dim3(PB), 0, s->stream, in, out, n, f);
  else
    hipLaunchKernelGGL((k_map1<Tin, Tout, F>), dim3(stream_grid(n, cus)), dim3(PB), 0, s->stream, in, out, n, f);
  ACC_HIP_CHECK(hipGetLastError());
  return ACC_SUCCESS;
}

static int begin(acc_stream_t s, const DeviceInfo** di) {
  if (!s) return fail(ACC_ERROR_INVALID_VALUE, "null stream");
  ACC_CHECK(set_device(s->device));
  return device_info(s->device, di);
}

}  // namespace acc

using namespace acc;

extern "C" int acc_hip_map(acc_stream_t s, acc_type t_in, acc_type t_out, acc_unop f, const void* params,
                           const void* in, int64_t n, void* out) {
  const DeviceInfo* di;
  ACC_CHECK(begin(s, &di));
  if (n < 0) return fail(ACC_ERROR_INVALID_VALUE, "map: negative size");
  if (n == 0) return ACC_SUCCESS;
  if (!in || !out) return fail(ACC_ERROR_INVALID_VALUE, "map: null array");
  if (f == ACC_UNOP_CAST) {
    return with_type(t_in, [&](auto zi) -> int {
      using Tin = decltype(zi);
      return with_type(t_out, [&](auto zo) -> int {
        using Tout = decltype(zo);
        if ((uintptr_t)in % sizeof(Tin) || (uintptr_t)out % sizeof(Tout))
          return fail(ACC_ERROR_INVALID_VALUE, "map: arrays not element-aligned");
        return launch_map<Tin, Tout>(s, (const Tin*)in, (Tout*)out, n, FCast<Tin, Tout>{}, di->compute_units);
      });
    });
  }
  if (t_in != t_out) return fail(ACC_ERROR_UNSUPPORTED, "map: only the conversion changes the element type");
  return with_type(t_in, [&](auto z) -> int {
    using T = decltype(z);
    if ((uintptr_t)in % sizeof(T) || (uintptr_t)out % sizeof(T))
      return fail(ACC_ERROR_INVALID_VALUE, "map: arrays not element-aligned");
    return with_unop<T>(f, params, [&](auto fn) -> int {
      return launch_map<T, T>(s, (const T*)in, (T*)out, n, fn, di->compute_units);
    });
  });
}

extern "C" int acc_hip_zipwith(acc_stream_t s, acc_type t, acc_op f, const void* a, const void* b, int64_t n,
                               void* out) {
  const DeviceInfo* di;
  ACC_CHECK(begin(s, &di));
  if (n < 0) return fail(ACC_ERROR_INVALID_VALUE, "zipWith: negative size");
  if (n == 0) return ACC_SUCCESS;
  if (!a || !b || !out) return fail(ACC_ERROR_INVALID_VALUE, "zipWith: null array");
  return with_type(t, [&](auto z) -> int {
    using T = decltype(z);
    if ((uintptr_t)a % sizeof(T) || (uintptr_t)b % sizeof(T) || (uintptr_t)out % sizeof(T))
      return fail(ACC_ERROR_INVALID_VALUE, "zipWith: arrays not element-aligned");
    auto go = [&](auto op) -> int {
      using Op = decltype(op);
      constexpr int E = 16 / (int)sizeof(T);
      if (al16(a) && al16(b) && al16(out))
        hipLaunchKernelGGL((k_zip<T, Op>), dim3(stream_grid(n / E, di->compute_units)), dim3(PB), 0, s->stream,
                           (const T*)a, (const T*)b, (T*)out, n, op);
      else
        hipLaunchKernelGGL((k_zip1<T, Op>), dim3(stream_grid(n, di->compute_units)), dim3(PB), 0, s->stream,
                           (const T*)a, (const T*)b, (T*)out, n, op);
      ACC_HIP_CHECK(hipGetLastError());
      return ACC_SUCCESS;
    };
    if (f == ACC_OP_SUB) return go(OpSubZ{});
    if (f == ACC_OP_EXCHANGE) return fail(ACC_ERROR_UNSUPPORTED, "zipWith: combine %d", (int)f);
    return with_op<T>(f, go);
  });
}

extern "C" int acc_hip_generate(acc_stream_t s, acc_type t, acc_genop g, const void* params, int64_t n,
                                void* out) {
  const DeviceInfo* di;
  ACC_CHECK(begin(s, &di));
  if (n < 0) return fail(ACC_ERROR_INVALID_VALUE, "generate: negative size");
  if (n == 0) return ACC_SUCCESS;
  if (!out || !params) return fail(ACC_ERROR_INVALID_VALUE, "generate: null argument");
  return with_type(t, [&](auto z) -> int {
    using T = decltype(z);
    if ((uintptr_t)out % sizeof(T)) return fail(ACC_ERROR_INVALID_VALUE, "generate: array not element-aligned");
    T a, b = T();
    __builtin_memcpy(&a, params, sizeof(T));
    if (g == ACC_GEN_IOTA) __builtin_memcpy(&b, (const char*)params + sizeof(T), sizeof(T));
    const int aligned = al16(out) ? 1 : 0;
    const unsigned grid = stream_grid(n / (16 / (int64_t)sizeof(T)) + 1, di->compute_units);  // PU chunks per thread
    if (g == ACC_GEN_FILL)
      hipLaunchKernelGGL((k_gen<T, GFill<T>>), dim3(grid), dim3(PB), 0, s->stream, (T*)out, n, GFill<T>{a}, aligned);
    else if (g == ACC_GEN_IOTA)
      hipLaunchKernelGGL((k_gen<T, GIota<T>>), dim3(grid), dim3(PB), 0, s->stream, (T*)out, n, GIota<T>{a, b},
                         aligned);
    else
      return fail(ACC_ERROR_UNSUPPORTED, "generate: generator %d", (int)g);
    ACC_HIP_CHECK(hipGetLastError());
    return ACC_SUCCESS;
  });
}

extern "C" int acc_hip_transform(acc_stream_t s, acc_type t, acc_ixmap p, acc_unop f, const void* params,
                                 const void* in, int64_t in_outer, int64_t in_inner, acc_type index_type,
                                 const void* index, int64_t n, void* out) {
  const DeviceInfo* di;
  ACC_CHECK(begin(s, &di));
  if (n < 0 || in_outer < 0 || in_inner < 0) return fail(ACC_ERROR_INVALID_VALUE, "transform: negative extent");
  const int64_t n_in = in_outer * in_inner;
  if (p != ACC_IX_GATHER && n != n_in)
    return fail(ACC_ERROR_INVALID_VALUE, "transform: reverse/transpose keep the element count");
  if (n == 0) return ACC_SUCCESS;
  if (!out || (n_in > 0 && !in)) return fail(ACC_ERROR_INVALID_VALUE, "transform: null array");
  if (f == ACC_UNOP_CAST) return fail(ACC_ERROR_UNSUPPORTED, "transform: conversions go through map");
  return with_type(t, [&](auto z) -> int {
    using T = decltype(z);
    if ((uintptr_t)in % sizeof(T) || (uintptr_t)out % sizeof(T))
      return fail(ACC_ERROR_INVALID_VALUE, "transform: arrays not element-aligned");
    return with_unop<T>(f, params, [&](auto fn) -> int {
      using F = decltype(fn);
      const int cus = di->compute_units;
      if (p == ACC_IX_GATHER) {
        if (!index) return fail(ACC_ERROR_INVALID_VALUE, "backpermute: null index array");
        if (n_in == 0) {  // every index is out of range: zeros, no read
          hipLaunchKernelGGL((k_gen<T, GFill<T>>), dim3(stream_grid(n, cus)), dim3(PB), 0, s->stream, (T*)out, n,
                             GFill<T>{T()}, 0);
          ACC_HIP_CHECK(hipGetLastError());
          return ACC_SUCCESS;
        }
        auto gather = [&](auto zi) -> int {
          using I = decltype(zi);
          const I* idx = (const I*)index;
          if ((uintptr_t)idx % sizeof(I)) return fail(ACC_ERROR_INVALID_VALUE, "backpermute: index not element-aligned");
          constexpr int E = 16 / (int)sizeof(I);
          // vector form: 16-byte index chunks and E-element output chunks
          if (al16(idx) && (uintptr_t)out % (E * sizeof(T)) == 0)
            hipLaunchKernelGGL((k_gather<T, I, F>), dim3(stream_grid(n / E, cus)), dim3(PB), 0, s->stream,
                               (const T*)in, n_in, idx, (T*)out, n, fn);
          else
            hipLaunchKernelGGL((k_gather1<T, I, F>), dim3(stream_grid(n, cus)), dim3(PB), 0, s->stream,
                               (const T*)in, n_in, idx, (T*)out, n, fn);
          ACC_HIP_CHECK(hipGetLastError());
          return ACC_SUCCESS;
        };
        if (index_type == ACC_INT32) return gather((int32_t)0);
        if (index_type == ACC_INT64) return gather((int64_t)0);
        return fail(ACC_ERROR_UNSUPPORTED, "backpermute: index type %d", (int)index_type);
      }
      if (p == ACC_IX_REVERSE) {
        constexpr int E = 16 / (int)sizeof(T);
        if (al16(out) && al16(in) && n % E == 0)
          hipLaunchKernelGGL((k_reverse<T, F, true>), dim3(stream_grid(n, cus)), dim3(PB), 0, s->stream,
                             (const T*)in, (T*)out, n, fn, 1);
        else
          hipLaunchKernelGGL((k_reverse<T, F, false>), dim3(stream_grid(n, cus)), dim3(PB), 0, s->stream,
                             (const T*)in, (T*)out, n, fn, al16(out) ? 1 : 0);
        ACC_HIP_CHECK(hipGetLastError());
        return ACC_SUCCESS;
      }
      if (p == ACC_IX_TRANSPOSE) {
        const int64_t tiles_r = ceil_div(in_outer, (int64_t)TT), tiles_c = ceil_div(in_inner, (int64_t)TT);
        if (tiles_r * tiles_c > 0x7fffffffLL) return fail(ACC_ERROR_UNSUPPORTED, "transpose: array too large");
        constexpr int VEC = Vec<T>::N;
        // vector tiles in diagonal order where both row pitches are powers of
        // two of at least 1 KiB (measured faster there: 16384^2 Int32 461 ->
        // 410 us); other shapes keep element tiles in row order, which
        // measured as fast or faster
        auto pow2k = [](uint64_t b) { return b >= 1024 && (b & (b - 1)) == 0; };
        const uint32_t vec_ok = al16(in) && al16(out) && in_outer % VEC == 0 && in_inner % VEC == 0 &&
                                pow2k((uint64_t)in_outer * sizeof(T)) && pow2k((uint64_t)in_inner * sizeof(T));
        const uint32_t diag = vec_ok;
        hipLaunchKernelGGL((k_transpose<T, F>), dim3((unsigned)(tiles_r * tiles_c)), dim3(PB), 0, s->stream,
                           (const T*)in, (T*)out, in_outer, in_inner, tiles_c, vec_ok, diag, fn);
        ACC_HIP_CHECK(hipGetLastError());
        return ACC_SUCCESS;
      }
      return fail(ACC_ERROR_UNSUPPORTED, "transform: index map %d", (int)p);
    });
  });
}
