// stencil1 on DIM2 arrays with the weighted-sum stencil function.
//
// The reference launches one interior kernel with no bounds checks plus one
// border kernel per face (2*rank launches on forked streams joined by
// events, PTX/Execute.hs:686-776), and every tap is a separate global load
// (PTX/CodeGen/Stencil.hs:113-177).  Here one launch covers interior and
// border: each workgroup stages a (TH+2RY) x (TW+2RX) halo tile in LDS,
// applying the boundary map (LLVM/CodeGen/Stencil.hs:224-253) to the
// coordinates outside the array as it loads, then each thread computes a
// 4x4 block of outputs from a register window over the LDS tile and writes
// it with dwordx4 stores.
//
// The stencil function is f(tuple) = sum_{dy,dx} w[dy][dx] * x[y+dy][x+dx]
// with taps in tuple order (rows outer, columns inner,
// LLVM/CodeGen/Stencil.hs:126-180).  Floating results may differ from a
// left-to-right sum by rounding (fused multiply-add); integer results wrap.
//
// HBM roofline: 2*H*W*sizeof(T) bytes (each input read once, each output
// written once; halo re-reads are served by L2).
#include "dispatch.hpp"
#include "vecmem.hpp"

namespace acc {

constexpr int ST_TW = 64;  // tile width (columns)
constexpr int ST_TH = 64;  // tile height (rows)
constexpr int ST_PX = 4;   // outputs per thread in x
constexpr int ST_PY = 4;   // outputs per thread in y
constexpr int ST_BLOCK = (ST_TW / ST_PX) * (ST_TH / ST_PY);  // 256

template <typename T> struct StArgs {
  const T* in;
  T* out;
  int64_t h, w;
  uint32_t tiles_x, tiles_y;
  int boundary;
  T cval;
  T wt[81];  // (2RY+1) x (2RX+1) weights, row-major
};

__device__ __forceinline__ int64_t bound_index(int b, int64_t i, int64_t n) {
  if (i < 0) {
    if (b == ACC_BOUNDARY_CLAMP) i = 0;
    else if (b == ACC_BOUNDARY_MIRROR) i = -i;
    else if (b == ACC_BOUNDARY_WRAP) i = n + i;
  } else if (i >= n) {
    if (b == ACC_BOUNDARY_CLAMP) i = n - 1;
    else if (b == ACC_BOUNDARY_MIRROR) i = n - (i - n + 2);
    else if (b == ACC_BOUNDARY_WRAP) i = i - n;
  }
  // arrays smaller than the radius: keep the single-step map in bounds
  return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

template <typename T, int RY, int RX>
__global__ __launch_bounds__(ST_BLOCK) void k_stencil2d(StArgs<T> a) {
  constexpr int LW = ST_TW + 2 * RX;  // LDS row length
  constexpr int LH = ST_TH + 2 * RY;
  constexpr int KW = 2 * RX + 1;
  __shared__ __align__(16) T tile[LH * LW];

  // XCD-aware tile order: consecutive workgroups land on different XCDs, so
  // when the tile count allows, each XCD gets a contiguous band of tiles,
  // walked row by row.  Horizontal neighbours then run together on one XCD
  // and share, in its L2, the 128-byte lines that hold each row's horizontal
  // halo (a column-major walk fetched every such line twice from HBM); the
  // vertical halo is 2*RY of 64 rows.
  const uint32_t nt = a.tiles_x * a.tiles_y;
  uint32_t b = blockIdx.x;
  if (nt % 8u == 0u) b = (b % 8u) * (nt / 8u) + b / 8u;
  const uint32_t ty = b / a.tiles_x;  // row-major
  const uint32_t tx = b - ty * a.tiles_x;
  const int64_t x0 = (int64_t)tx * ST_TW, y0 = (int64_t)ty * ST_TH;

  // ---- stage the halo tile; coordinates outside the array go through the
  // boundary map (constant: the value itself)
  const bool constant = a.boundary == ACC_BOUNDARY_CONSTANT;
  for (int i = threadIdx.x; i < LH * LW; i += ST_BLOCK) {
    const int r = i / LW, c = i - r * LW;
    int64_t gy = y0 - RY + r, gx = x0 - RX + c;
    const bool inside = (uint64_t)gy < (uint64_t)a.h && (uint64_t)gx < (uint64_t)a.w;
    if (!inside) {
      gy = bound_index(a.boundary, gy, a.h);
      gx = bound_index(a.boundary, gx, a.w);
    }
    T v = a.in[gy * a.w + gx];
    tile[i] = (inside || !constant) ? v : a.cval;
  }
  __syncthreads();

  // ---- 4x4 outputs per thread from a register window
  const int px = threadIdx.x % (ST_TW / ST_PX), py = threadIdx.x / (ST_TW / ST_PX);
  const int cx = px * ST_PX, cy = py * ST_PY;  // tile-local output origin
  T acc[ST_PY][ST_PX];
#pragma unroll
  for (int lr = 0; lr < ST_PY + 2 * RY; ++lr) {
    T v[ST_PX + 2 * RX];
    const T* row = tile + (cy + lr) * LW + cx;
#pragma unroll
    for (int c = 0; c < ST_PX + 2 * RX; ++c) v[c] = row[c];
#pragma unroll
    for (int oy = 0; oy < ST_PY; ++oy) {
      const int dy = lr - oy;  // tap row 0..2RY
      if (dy >= 0 && dy <= 2 * RY) {
#pragma unroll
        for (int ox = 0; ox < ST_PX; ++ox) {
#pragma unroll
          for (int dx = 0; dx < KW; ++dx) {
            const T p = OpMul{}(a.wt[dy * KW + dx], v[ox + dx]);
            // the first tap (dy == 0, dx == 0) starts the sum
            acc[oy][ox] = (dx == 0 && dy == 0) ? p : OpAdd{}(acc[oy][ox], p);
          }
        }
      }
    }
  }

  // ---- store
#pragma unroll
  for (int oy = 0; oy < ST_PY; ++oy) {
    const int64_t gy = y0 + cy + oy;
    const int64_t gx = x0 + cx;
    T* o = a.out + gy * a.w + gx;
    if (gy < a.h && gx + ST_PX <= a.w && ((uintptr_t)o % 16) == 0 && ST_PX * sizeof(T) % 16 == 0) {
#pragma unroll
      for (int q = 0; q < (int)(ST_PX * sizeof(T) / 16); ++q) {
        Vec<T> r;
#pragma unroll
        for (int j = 0; j < Vec<T>::N; ++j) r.e[j] = acc[oy][q * Vec<T>::N + j];
        store16(o + q * Vec<T>::N, r);
      }
    } else if (gy < a.h) {
#pragma unroll
      for (int ox = 0; ox < ST_PX; ++ox)
        if (gx + ox < a.w) o[ox] = acc[oy][ox];
    }
  }
}

template <typename T, int RY, int RX>
int stencil_launch(acc_stream_t s, const T* w, int boundary, const T* cval, const T* in, int64_t h,
                   int64_t wd, T* out) {
  StArgs<T> a;
  a.in = in;
  a.out = out;
  a.h = h;
  a.w = wd;
  const int64_t tiles_x = ceil_div(wd, ST_TW), tiles_y = ceil_div(h, ST_TH);
  if (tiles_x * tiles_y > 0x7fffffffLL) return fail(ACC_ERROR_UNSUPPORTED, "stencil: array too large");
  a.tiles_x = (uint32_t)tiles_x;
  a.tiles_y = (uint32_t)tiles_y;
  a.boundary = boundary;
  a.cval = cval ? *cval : T();
  const int taps = (2 * RY + 1) * (2 * RX + 1);
  for (int i = 0; i < 81; ++i) a.wt[i] = i < taps ? w[i] : T();
  const int64_t nt = tiles_x * tiles_y;
  hipLaunchKernelGGL((k_stencil2d<T, RY, RX>), dim3((unsigned)nt), dim3(ST_BLOCK), 0, s->stream, a);
  ACC_HIP_CHECK(hipGetLastError());
  return ACC_SUCCESS;
}

template <typename T, int RY>
int stencil_rx(int rx, acc_stream_t s, const T* w, int b, const T* c, const T* in, int64_t h, int64_t wd, T* out) {
  switch (rx) {
    case 1: return stencil_launch<T, RY, 1>(s, w, b, c, in, h, wd, out);
    case 2: return stencil_launch<T, RY, 2>(s, w, b, c, in, h, wd, out);
    case 3: return stencil_launch<T, RY, 3>(s, w, b, c, in, h, wd, out);
    case 4: return stencil_launch<T, RY, 4>(s, w, b, c, in, h, wd, out);
  }
  return fail(ACC_ERROR_UNSUPPORTED, "stencil: x radius %d", rx);
}

template <typename T>
int stencil_any(int ry, int rx, acc_stream_t s, const T* w, int b, const T* c, const T* in, int64_t h, int64_t wd,
                T* out) {
  switch (ry) {
    case 1: return stencil_rx<T, 1>(rx, s, w, b, c, in, h, wd, out);
    case 2: return stencil_rx<T, 2>(rx, s, w, b, c, in, h, wd, out);
    case 3: return stencil_rx<T, 3>(rx, s, w, b, c, in, h, wd, out);
    case 4: return stencil_rx<T, 4>(rx, s, w, b, c, in, h, wd, out);
  }
  return fail(ACC_ERROR_UNSUPPORTED, "stencil: y radius %d", ry);
}

}  // namespace acc

using namespace acc;

extern "C" int acc_hip_stencil2d(acc_stream_t s, acc_type t, int ry, int rx, const void* weights,
                                 acc_boundary boundary, const void* constant, const void* in, int64_t height,
                                 int64_t width, void* out) {
  if (!s) return fail(ACC_ERROR_INVALID_VALUE, "stencil: null stream");
  if (height < 0 || width < 0) return fail(ACC_ERROR_INVALID_VALUE, "stencil: negative extent");
  if (ry < 1 || ry > 4 || rx < 1 || rx > 4) return fail(ACC_ERROR_UNSUPPORTED, "stencil: radius must be 1..4");
  if (boundary < ACC_BOUNDARY_CLAMP || boundary > ACC_BOUNDARY_CONSTANT)
    return fail(ACC_ERROR_INVALID_VALUE, "stencil: boundary %d", (int)boundary);
  if (boundary == ACC_BOUNDARY_CONSTANT && !constant)
    return fail(ACC_ERROR_INVALID_VALUE, "stencil: constant boundary needs a value");
  if (!weights) return fail(ACC_ERROR_INVALID_VALUE, "stencil: null weights");
  if (height == 0 || width == 0) return ACC_SUCCESS;
  if (!in || !out) return fail(ACC_ERROR_INVALID_VALUE, "stencil: null array");
  if (in == out) return fail(ACC_ERROR_INVALID_VALUE, "stencil: output must not alias the input");
  ACC_CHECK(set_device(s->device));
  // (+) and (*) give the same bits for signed and unsigned words
  switch (t) {
    case ACC_INT32: case ACC_WORD32:
      return stencil_any<int32_t>(ry, rx, s, (const int32_t*)weights, boundary, (const int32_t*)constant,
                                  (const int32_t*)in, height, width, (int32_t*)out);
    case ACC_INT64: case ACC_WORD64:
      return stencil_any<int64_t>(ry, rx, s, (const int64_t*)weights, boundary, (const int64_t*)constant,
                                  (const int64_t*)in, height, width, (int64_t*)out);
    case ACC_FLOAT32:
      return stencil_any<float>(ry, rx, s, (const float*)weights, boundary, (const float*)constant,
                                (const float*)in, height, width, (float*)out);
    case ACC_FLOAT64:
      return stencil_any<double>(ry, rx, s, (const double*)weights, boundary, (const double*)constant,
                                 (const double*)in, height, width, (double*)out);
  }
  return fail(ACC_ERROR_UNSUPPORTED, "stencil: element type %d", (int)t);
}
