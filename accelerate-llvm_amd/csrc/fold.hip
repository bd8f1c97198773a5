// fold / fold1 over a whole vector in ONE launch.
//
// Replaces the reference's multi-launch tree (foldAllM1 then foldAllM2
// relaunched until one element remains, PTX/Execute.hs:267-318;
// PTX/CodeGen/Fold.hs:175-285).  Here each workgroup reduces a contiguous
// chunk with dwordx4 loads and a DPP wave reduction, publishes its partial as
// a tagged word, and the last workgroup to finish (a done-counter ticket)
// combines the partials and applies the seed on the left
// (PTX/CodeGen/Fold.hs:282; NAT/CodeGen/Fold.hs:160-163).
//
// HBM roofline: n*sizeof(T) bytes read, sizeof(T) written.
#include "dispatch.hpp"
#include "lookback.hpp"
#include "vecmem.hpp"

namespace acc {

template <typename T> struct FoldAllArgs {
  const T* in;
  int64_t n;
  int64_t head;   // elements before the first 16-byte aligned element
  int64_t nvec;   // whole 16-byte vectors after the head
  int64_t chunk;  // vectors per workgroup
  T* out;
  T seed;
  uint32_t has_seed;
  SyncCtl* ctl;
  uint64_t* status;
};

// Requires a commutative combine function: each lane accumulates vectors
// strided by the workgroup size.  All functions the C ABI offers are
// commutative; their results are exact for integers and order-independent
// up to rounding for floating point.
template <typename T, typename Op, int BLOCK, int U>
__global__ __launch_bounds__(BLOCK) void k_fold_all(FoldAllArgs<T> a) {
  constexpr int VEC = Vec<T>::N;
  constexpr int WAVES = BLOCK / 64;
  constexpr int NW = Tagged<Maybe<T>>::NW;
  Op op{};
  __shared__ uint32_t s_epoch, s_last;
  __shared__ Maybe<T> s_w[WAVES];
  const uint32_t epoch = block_epoch(a.ctl, &s_epoch);
  const uint32_t tag = make_tag(epoch, ST_PREFIX);
  const int lane = lane_id(), wid = threadIdx.x / 64;

  const T* body = a.in + a.head;
  const int64_t vb = (int64_t)blockIdx.x * a.chunk;
  const int64_t ve = min(vb + a.chunk, a.nvec);
  int64_t i = vb + threadIdx.x;
  Maybe<T> acc;
  acc.ok = 0;
  acc.v = T();
  if (i < ve) {
    Vec<T> v0 = load16_stream(body + i * VEC);
    T r = v0.e[0];
#pragma unroll
    for (int j = 1; j < VEC; ++j) r = op(r, v0.e[j]);
    i += BLOCK;
    for (; i + (U - 1) * BLOCK < ve; i += U * BLOCK) {
      Vec<T> v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = load16_stream(body + (i + u * BLOCK) * VEC);
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int j = 0; j < VEC; ++j) r = op(r, v[u].e[j]);
    }
    for (; i < ve; i += BLOCK) {
      Vec<T> v = load16_stream(body + i * VEC);
#pragma unroll
      for (int j = 0; j < VEC; ++j) r = op(r, v.e[j]);
    }
    acc.v = r;
    acc.ok = 1;
  }

  Maybe<T> w = wave_reduce_m(acc, op);
  if (lane == 0) s_w[wid] = w;
  __syncthreads();
  if (threadIdx.x == 0) {
    Maybe<T> b = s_w[0];
#pragma unroll
    for (int k = 1; k < WAVES; ++k) b = mcombine(op, b, s_w[k]);
    Tagged<Maybe<T>>::publish(a.status + (size_t)blockIdx.x * NW, b, tag);
    s_last = (add_relaxed(&a.ctl->done, 1u) == gridDim.x - 1) ? 1u : 0u;
  }
  __syncthreads();
  if (!s_last) return;

  // The last workgroup combines every partial.
  Maybe<T> g;
  g.ok = 0;
  g.v = T();
  for (uint32_t b = threadIdx.x; b < gridDim.x; b += BLOCK) {
    Maybe<T> p;
    while (Tagged<Maybe<T>>::peek(a.status + (size_t)b * NW, p) != tag) __builtin_amdgcn_s_sleep(1);
    g = mcombine(op, g, p);
  }
  g = wave_reduce_m(g, op);
  __syncthreads();
  if (lane == 0) s_w[wid] = g;
  __syncthreads();
  if (threadIdx.x == 0) {
    Maybe<T> r;
    r.ok = 0;
    r.v = T();
    for (int64_t k = 0; k < a.head; ++k) {
      Maybe<T> x{a.in[k], 1u};
      r = mcombine(op, r, x);
    }
#pragma unroll
    for (int k = 0; k < WAVES; ++k) r = mcombine(op, r, s_w[k]);
    for (int64_t k = a.head + a.nvec * VEC; k < a.n; ++k) {
      Maybe<T> x{a.in[k], 1u};
      r = mcombine(op, r, x);
    }
    if (a.has_seed) {
      Maybe<T> z{a.seed, 1u};
      r = mcombine(op, z, r);  // seed on the left
    }
    *a.out = r.v;
    st_relaxed(&a.ctl->done, 0u);
    advance_epoch(a.ctl, epoch);
  }
}

template <typename T, typename Op>
int fold_all(acc_stream_t s, const T* in, int64_t n, const T* seed, T* out) {
  // 2 workgroups per CU with 4 vectors in flight per lane read fastest
  // (1 GiB Int32: 152 us = 7.05 TB/s; 4 per CU 158 us, 8 per CU 174 us; a
  // read-only sweep agrees: more concurrent streams per CU read slower)
  constexpr int BLOCK = 256, U = 4, VEC = Vec<T>::N, BPC = 2;
  constexpr int NW = Tagged<Maybe<T>>::NW;
  const DeviceInfo* di;
  ACC_CHECK(device_info(s->device, &di));
  const uintptr_t addr = (uintptr_t)in;
  if (addr % sizeof(T)) return fail(ACC_ERROR_INVALID_VALUE, "fold: input not element-aligned");
  int64_t head = (int64_t)(((16 - addr % 16) % 16) / sizeof(T));
  if (head > n) head = n;
  const int64_t nvec = (n - head) / VEC;
  int64_t nb = ceil_div(nvec, (int64_t)BLOCK * U);
  const int64_t maxb = (int64_t)di->compute_units * BPC;
  if (nb > maxb) nb = maxb;
  if (nb < 1) nb = 1;
  int64_t chunk = nvec ? ceil_div(nvec, nb) : 0;
  if (chunk) nb = ceil_div(nvec, chunk);
  SyncCtl* ctl;
  uint64_t* status;
  ACC_CHECK(stream_workspace(s, (size_t)nb * NW * 8, &ctl, &status));
  FoldAllArgs<T> a{in, n, head, nvec, chunk, out, seed ? *seed : T(), seed ? 1u : 0u, ctl, status};
  hipLaunchKernelGGL((k_fold_all<T, Op, BLOCK, U>), dim3((unsigned)nb), dim3(BLOCK), 0, s->stream, a);
  ACC_HIP_CHECK(hipGetLastError());
  return ACC_SUCCESS;
}

int fold_dim(acc_stream_t s, acc_type t, acc_op op, const void* seed, const void* in, int64_t outer,
             int64_t inner, void* out);

}  // namespace acc

using namespace acc;

extern "C" int acc_hip_fold(acc_stream_t s, acc_type t, acc_op op, const void* seed, const void* in,
                            int64_t outer, int64_t inner, void* out) {
  if (!s) return fail(ACC_ERROR_INVALID_VALUE, "fold: null stream");
  if (outer < 0 || inner < 0) return fail(ACC_ERROR_INVALID_VALUE, "fold: negative extent");
  if (!seed && inner == 0) return fail(ACC_ERROR_EMPTY_ARRAY, "fold1: empty array");
  if (outer == 0) return ACC_SUCCESS;
  const size_t B = type_size(t);
  if (!B) return fail(ACC_ERROR_UNSUPPORTED, "fold: element type %d", (int)t);
  if (inner == 0) return acc_hip_fill_async(s, out, seed, B, (size_t)outer);
  if (!in || !out) return fail(ACC_ERROR_INVALID_VALUE, "fold: null array");
  ACC_CHECK(set_device(s->device));
  if (outer > 1) return fold_dim(s, t, op, seed, in, outer, inner, out);
  return with_type(t, [&](auto z) {
    using T = decltype(z);
    return with_op<T>(op, [&](auto f) {
      using Op = decltype(f);
      return fold_all<T, Op>(s, (const T*)in, inner, (const T*)seed, (T*)out);
    });
  });
}
