// stencil1 on DIM3 arrays with the weighted-sum stencil function.
//
// The reference generates the same kernels for every rank: one interior
// kernel and 2*rank border kernels, every tap a separate global load
// (PTX/Execute.hs:686-776, PTX/CodeGen/Stencil.hs:113-177), the boundary map
// applied per coordinate (LLVM/CodeGen/Stencil.hs:224-253).  Here a 2.5-D
// march: a workgroup owns a 64 x 32 (x, y) block of outputs (64 x 16 when the
// LDS ring of 8-byte planes would not fit) over a run of ZC
// planes; it keeps the 2*RZ+1 halo planes the current output plane needs in
// an LDS ring and brings in one new plane per step, so each input plane is
// read once per run (plus 2*RZ planes at the start of each run).  The next
// plane's loads are issued into registers before the current plane is
// computed, so their latency overlaps the arithmetic.  Each thread computes
// 8 (or 4) consecutive y outputs at one x: per plane and per dx it reads one
// column of 8 + 2*RY values from LDS and applies every dy tap to it.  Staging maps
// coordinates outside the array with the boundary map (the constant boundary
// substitutes its value), so interior and border blocks share one kernel.
//
// The stencil function is f = sum over (dz, dy, dx) in tuple order (planes,
// rows, columns) of w[dz][dy][dx] * x[z+dz][y+dy][x+dx]; floating results may
// differ from a left-to-right sum by rounding (fused multiply-add), integer
// results wrap.
//
// HBM roofline: 2*D*H*W*sizeof(T) bytes.
#include "boundary.hpp"
#include "dispatch.hpp"
#include "vecmem.hpp"

namespace acc {

constexpr int S3_TX = 64;     // block width (x)
constexpr int S3_BLOCK = 256;  // 64 x 4 threads

// Block height 32 (8 y outputs per thread), or 16 (4 per thread) when the
// ring of 32-row planes would pass 64 KiB of LDS (8-byte types with RZ = 2).
template <typename T, int RZ, int RY, int RX> struct S3Shape {
#ifndef ACC_S3_TY
#define ACC_S3_TY 16
#endif
  static constexpr int TY =
      (int)sizeof(T) * (2 * RZ + 1) * (ACC_S3_TY + 2 * RY) * (S3_TX + 2 * RX) > 64 * 1024 ? 16 : ACC_S3_TY;
  static constexpr int PY = TY / 4;
};

template <typename T> struct St3Args {
  const T* in;
  T* out;
  int64_t d, h, w;
  uint32_t tiles_x, tiles_y, zc;  // blocks in x and y, planes per run
  int boundary;
  T cval;
  T wt[125];  // (2RZ+1) x (2RY+1) x (2RX+1), tuple order
};

template <typename T, int RZ, int RY, int RX>
__global__ __launch_bounds__(S3_BLOCK) void k_stencil3d(St3Args<T> a) {
  constexpr int S3_TY = S3Shape<T, RZ, RY, RX>::TY, S3_PY = S3Shape<T, RZ, RY, RX>::PY;
  constexpr int NP = 2 * RZ + 1;           // planes in the ring
  constexpr int LH = S3_TY + 2 * RY;       // halo rows per plane
  constexpr int LW = S3_TX + 2 * RX;       // halo columns per plane
  constexpr int PLANE = LH * LW;
  constexpr int NI = (PLANE + S3_BLOCK - 1) / S3_BLOCK;  // loads per thread per plane
  constexpr int KW = 2 * RX + 1, KH = 2 * RY + 1;
  __shared__ T ring[NP][PLANE];

  // XCD-aware block order: workgroup g runs on XCD g % 8, and XCD k takes the
  // contiguous band of blocks [start(k), start(k+1)) (x fastest, then y, then
  // the z run), so blocks that share halo rows share one L2
  const uint32_t nb = gridDim.x, q = nb / 8u, rr = nb % 8u, xcd = blockIdx.x % 8u;
  const uint32_t b = xcd * q + min(xcd, rr) + blockIdx.x / 8u;
  const uint32_t txb = b % a.tiles_x, tyb = (b / a.tiles_x) % a.tiles_y, zr = b / (a.tiles_x * a.tiles_y);
  // plain locals (no references to the kernel argument) and 32-bit row and
  // column coordinates (the host checks h, w < 2^31)
  const T* const in = a.in;
  T* const out = a.out;
  const int64_t D = a.d;
  const int H = (int)a.h, W = (int)a.w;
  const int bnd = a.boundary;
  const T cval = a.cval;
  const int x0 = (int)txb * S3_TX, y0 = (int)tyb * S3_TY;
  const int64_t z0 = (int64_t)zr * a.zc, z1 = min(z0 + (int64_t)a.zc, D);
  const bool constant = bnd == ACC_BOUNDARY_CONSTANT;

  // the boundary map (as bound_index) as selects: only the boundary kind,
  // uniform, branches
  const bool mirror = bnd == ACC_BOUNDARY_MIRROR, wrap = bnd == ACC_BOUNDARY_WRAP;
  auto map = [mirror, wrap](int i, int n) -> int {
    const int mi = i < 0 ? -i : (i >= n ? 2 * n - 2 - i : i);
    const int wi = i < 0 ? i + n : (i >= n ? i - n : i);
    const int m = mirror ? mi : (wrap ? wi : i);
    return min(max(m, 0), n - 1);
  };

  // staging: this thread's elements of plane gz into registers; every load
  // unconditional (indices past the plane re-read its last element)
  auto fetch = [&](int64_t gz, T (&v)[NI]) {
    const bool zout = gz < 0 || gz >= D;
    const int64_t mz = zout ? bound_index(bnd, gz, D) : gz;
    const T* pl = in + mz * (int64_t)H * W;
#pragma unroll
    for (int u = 0; u < NI; ++u) {
      const int e = min((int)threadIdx.x + u * S3_BLOCK, PLANE - 1);
      const int r = e / LW, c = e - r * LW;
      const int gy = y0 - RY + r, gx = x0 - RX + c;
      const bool outside = zout || (unsigned)gy >= (unsigned)H || (unsigned)gx >= (unsigned)W;
      const T x = pl[(int64_t)map(gy, H) * W + map(gx, W)];
      v[u] = constant && outside ? cval : x;
    }
  };
  auto put = [&](int slot, const T (&v)[NI]) {
#pragma unroll
    for (int u = 0; u < NI; ++u) {
      const int e = threadIdx.x + u * S3_BLOCK;
      if (e < PLANE) ring[slot][e] = v[u];
    }
  };

  // prologue: planes z0 - RZ .. z0 + RZ - 1 into slots 0 .. 2*RZ - 1; plane
  // z - RZ + k then sits in slot (s0 + k) mod NP, s0 advancing by one per step
  for (int k = 0; k < 2 * RZ; ++k) {
    T v[NI];
    fetch(z0 - RZ + k, v);
    put(k, v);
  }
  T nxt[NI];
  fetch(z0 + RZ, nxt);

  const int tx = threadIdx.x % S3_TX, ty = threadIdx.x / S3_TX;  // 64 x 4
  const int gxo = x0 + tx;
  int s0 = 0;  // slot of plane z - RZ
  for (int64_t z = z0; z < z1; ++z) {
    put(s0 + 2 * RZ >= NP ? s0 + 2 * RZ - NP : s0 + 2 * RZ, nxt);
    __syncthreads();
    if (z + 1 < z1) fetch(z + 1 + RZ, nxt);  // in flight during the arithmetic
    T acc[S3_PY];
#pragma unroll
    for (int dz = 0; dz < NP; ++dz) {
      const int sl = s0 + dz >= NP ? s0 + dz - NP : s0 + dz;
      const T* pl = ring[sl];
#pragma unroll
      for (int dx = 0; dx < KW; ++dx) {
        T col[S3_PY + 2 * RY];
#pragma unroll
        for (int j = 0; j < S3_PY + 2 * RY; ++j) col[j] = pl[(ty * S3_PY + j) * LW + tx + dx];
#pragma unroll
        for (int k = 0; k < S3_PY; ++k)
#pragma unroll
          for (int dy = 0; dy < KH; ++dy) {
            const T p = OpMul{}(a.wt[(dz * KH + dy) * KW + dx], col[k + dy]);
            // taps in tuple order: (dz, dy, dx) row-major; the first starts the sum
            acc[k] = (dz == 0 && dy == 0 && dx == 0) ? p : OpAdd{}(acc[k], p);
          }
      }
      // one plane at a time: otherwise the scheduler hoists every plane's
      // column reads and keeps them all live
      __builtin_amdgcn_sched_barrier(0);
    }
    if (gxo < W) {
#pragma unroll
      for (int k = 0; k < S3_PY; ++k) {
        const int gy = y0 + ty * S3_PY + k;
        if (gy < H) out[(z * H + gy) * (int64_t)W + gxo] = acc[k];
      }
    }
    s0 = s0 + 1 == NP ? 0 : s0 + 1;
    __syncthreads();  // the oldest plane's slot is refilled next step
  }
}

template <typename T, int RZ, int RY, int RX>
int stencil3d_launch(acc_stream_t s, const T* w, int boundary, const T* cval, const T* in, int64_t d, int64_t h,
                     int64_t wd, T* out, int cus) {
  St3Args<T> a;
  a.in = in;
  a.out = out;
  a.d = d;
  a.h = h;
  a.w = wd;
  const int64_t tx = ceil_div(wd, (int64_t)S3_TX), ty = ceil_div(h, (int64_t)S3Shape<T, RZ, RY, RX>::TY);
  // planes per run: long runs amortise the 2*RZ extra planes, enough runs
  // keep every CU busy (about 8 workgroups per CU)
  int64_t zc = d;
  const int64_t want = (int64_t)cus * 8;
  while (zc > 16 && tx * ty * ceil_div(d, zc) < want) zc = (zc + 1) / 2;
  if (tx * ty * ceil_div(d, zc) > 0x7fffffffLL) return fail(ACC_ERROR_UNSUPPORTED, "stencil3d: array too large");
  a.tiles_x = (uint32_t)tx;
  a.tiles_y = (uint32_t)ty;
  a.zc = (uint32_t)zc;
  a.boundary = boundary;
  a.cval = cval ? *cval : T();
  const int taps = (2 * RZ + 1) * (2 * RY + 1) * (2 * RX + 1);
  for (int i = 0; i < 125; ++i) a.wt[i] = i < taps ? w[i] : T();
  hipLaunchKernelGGL((k_stencil3d<T, RZ, RY, RX>), dim3((unsigned)(tx * ty * ceil_div(d, zc))), dim3(S3_BLOCK), 0,
                     s->stream, a);
  ACC_HIP_CHECK(hipGetLastError());
  return ACC_SUCCESS;
}

template <typename T>
int stencil3d_any(int rz, int ry, int rx, acc_stream_t s, const T* w, int b, const T* c, const T* in, int64_t d,
                  int64_t h, int64_t wd, T* out, int cus) {
  // radii 1 and 2 on each axis
  auto go = [&](auto RZc, auto RYc, auto RXc) {
    return stencil3d_launch<T, decltype(RZc)::value, decltype(RYc)::value, decltype(RXc)::value>(
        s, w, b, c, in, d, h, wd, out, cus);
  };
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  auto gx = [&](auto RZc, auto RYc) { return rx == 1 ? go(RZc, RYc, I1{}) : go(RZc, RYc, I2{}); };
  auto gy = [&](auto RZc) { return ry == 1 ? gx(RZc, I1{}) : gx(RZc, I2{}); };
  return rz == 1 ? gy(I1{}) : gy(I2{});
}

}  // namespace acc

using namespace acc;

extern "C" int acc_hip_stencil3d(acc_stream_t s, acc_type t, int rz, int ry, int rx, const void* weights,
                                 acc_boundary boundary, const void* constant, const void* in, int64_t depth,
                                 int64_t height, int64_t width, void* out) {
  if (!s) return fail(ACC_ERROR_INVALID_VALUE, "stencil3d: null stream");
  if (depth < 0 || height < 0 || width < 0) return fail(ACC_ERROR_INVALID_VALUE, "stencil3d: negative extent");
  if (rz < 1 || rz > 2 || ry < 1 || ry > 2 || rx < 1 || rx > 2)
    return fail(ACC_ERROR_UNSUPPORTED, "stencil3d: radius must be 1 or 2");
  if (boundary == ACC_BOUNDARY_FUNCTION)
    return fail(ACC_ERROR_UNSUPPORTED, "stencil3d: the function boundary is for the 2-D stencils");
  if (boundary < ACC_BOUNDARY_CLAMP || boundary > ACC_BOUNDARY_CONSTANT)
    return fail(ACC_ERROR_INVALID_VALUE, "stencil3d: boundary %d", (int)boundary);
  if (boundary == ACC_BOUNDARY_CONSTANT && !constant)
    return fail(ACC_ERROR_INVALID_VALUE, "stencil3d: constant boundary needs a value");
  if (!weights) return fail(ACC_ERROR_INVALID_VALUE, "stencil3d: null weights");
  if (depth == 0 || height == 0 || width == 0) return ACC_SUCCESS;
  if (!in || !out) return fail(ACC_ERROR_INVALID_VALUE, "stencil3d: null array");
  if (in == out) return fail(ACC_ERROR_INVALID_VALUE, "stencil3d: output must not alias the input");
  if (height >= ((int64_t)1 << 31) || width >= ((int64_t)1 << 31))
    return fail(ACC_ERROR_UNSUPPORTED, "stencil3d: rows and columns below 2^31");
  ACC_CHECK(set_device(s->device));
  const DeviceInfo* di;
  ACC_CHECK(device_info(s->device, &di));
  const int cus = di->compute_units;
  // (+) and (*) give the same bits for signed and unsigned words
  switch (t) {
    case ACC_INT32: case ACC_WORD32:
      return stencil3d_any<int32_t>(rz, ry, rx, s, (const int32_t*)weights, boundary, (const int32_t*)constant,
                                    (const int32_t*)in, depth, height, width, (int32_t*)out, cus);
    case ACC_INT64: case ACC_WORD64:
      return stencil3d_any<int64_t>(rz, ry, rx, s, (const int64_t*)weights, boundary, (const int64_t*)constant,
                                    (const int64_t*)in, depth, height, width, (int64_t*)out, cus);
    case ACC_FLOAT32:
      return stencil3d_any<float>(rz, ry, rx, s, (const float*)weights, boundary, (const float*)constant,
                                  (const float*)in, depth, height, width, (float*)out, cus);
    case ACC_FLOAT64:
      return stencil3d_any<double>(rz, ry, rx, s, (const double*)weights, boundary, (const double*)constant,
                                   (const double*)in, depth, height, width, (double*)out, cus);
  }
  return fail(ACC_ERROR_UNSUPPORTED, "stencil3d: element type %d", (int)t);
}
