// stencil1 on DIM2 arrays with the weighted-sum stencil function.
//
// The reference launches one interior kernel with no bounds checks plus one
// border kernel per face (2*rank launches on forked streams joined by
// events, PTX/Execute.hs:686-776), and every tap is a separate global load
// (PTX/CodeGen/Stencil.hs:113-177).  Here one launch covers the interior
// tiles and the frame of border tiles: each workgroup stages a
// (TH+2RY) x (TW+2RX) halo tile in LDS (interior tiles with dwordx4 copies,
// border tiles applying the boundary map, LLVM/CodeGen/Stencil.hs:224-253, to
// the coordinates outside the array as they load), then each thread computes
// an 8x2 block of outputs from a register window read from LDS with 16-byte
// reads and writes it with dwordx4 stores.
//
// The stencil function is f(tuple) = sum_{dy,dx} w[dy][dx] * x[y+dy][x+dx]
// with taps in tuple order (rows outer, columns inner,
// LLVM/CodeGen/Stencil.hs:126-180).  Floating results may differ from a
// left-to-right sum by rounding (fused multiply-add); integer results wrap.
//
// HBM roofline: 2*H*W*sizeof(T) bytes (each input read once, each output
// written once; halo re-reads are served by L2).
#include "boundary.hpp"
#include "dispatch.hpp"
#include "vecmem.hpp"

namespace acc {

// Wide, flat tiles with two 16-byte output vectors per thread and row:
// 8192^2 Float32 3x3 106 -> 91 us and 5x5 109 -> 95 us against 64 x 64 tiles
// of 4 x 4 outputs (128 x 32 and 256 x 32 tiles of 8 x 2 or 8 x 4 outputs were
// within 2 us of this; 128 x 64 of 4 x 8 outputs no better than 64 x 64).
constexpr int ST_TW = 256;  // tile width (columns)
constexpr int ST_TH = 16;   // tile height (rows)
constexpr int ST_PX = 8;    // outputs per thread in x
constexpr int ST_PY = 2;    // outputs per thread in y
constexpr int ST_BLOCK = (ST_TW / ST_PX) * (ST_TH / ST_PY);  // 256

template <typename T> struct StArgs {
  const T* in;
  T* out;
  int64_t h, w;
  uint32_t tiles_x, tiles_y;
  // interior tiles: tx in [ix0, ix1), ty in [iy0, iy1) (empty when the
  // rows are not 16-byte aligned); the border launch covers the others
  uint32_t ix0, ix1, iy0, iy1;
  int boundary;
  T cval;
  T fc[3];   // Function boundary: (c0 + cy*y) + cx*x
  T wt[81];  // (2RY+1) x (2RX+1) weights, row-major
};

// Interior tiles: the halo, rounded out to 16-byte boundaries, lies inside
// the array and is copied with dwordx4 vectors.  Border tiles: coordinates
// outside the array go through the boundary map (constant: the value
// itself), element by element.  MODE 2 (the one launched) holds both, blocks
// [0, ni) taking the interior tiles; each staging path is straight-line code
// (an earlier form that looped over both in one kernel was compiled with the
// register holding the thread index undefined on one path).
template <typename T, int RY, int RX, int MODE>  // 0 frame, 1 interior, 2 both
__global__ __launch_bounds__(ST_BLOCK) void k_stencil2d(StArgs<T> a) {
  constexpr int VEC = Vec<T>::N;
  constexpr int PADL = (RX + VEC - 1) / VEC * VEC;  // halo rounded out to 16-byte boundaries
  constexpr int CO = PADL - RX;                     // halo column c sits at LDS column c + CO
  constexpr int LS = ST_TW + 2 * PADL;              // LDS row stride: a 16-byte multiple
  constexpr int LW = ST_TW + 2 * RX;                // halo row length
  constexpr int LH = ST_TH + 2 * RY;
  constexpr int KW = 2 * RX + 1;
  constexpr int WIN = ST_PX + 2 * RX;               // register window row
  constexpr int NVW = (CO + WIN + VEC - 1) / VEC;   // LDS vectors covering it
  __shared__ __align__(16) T tile[LH * LS];

  // XCD-aware tile order: consecutive workgroups land on different XCDs, so
  // when the tile count allows, each XCD gets a contiguous band of tiles,
  // walked row by row.  Horizontal neighbours then run together on one XCD
  // and share, in its L2, the 128-byte lines that hold each row's horizontal
  // halo (a column-major walk fetched every such line twice from HBM); the
  // vertical halo is 2*RY of 64 rows.
  uint32_t tx, ty;
  const uint32_t nint = (a.ix1 - a.ix0) * (a.iy1 - a.iy0);  // interior tiles
  const bool interior = MODE == 1 || (MODE == 2 && blockIdx.x < nint);
  if (interior) {
    // workgroup b runs on XCD b % 8; XCD k takes tiles [start(k), start(k+1))
    // with nt / 8 tiles each and one more for the first nt % 8 XCDs
    const uint32_t nx = a.ix1 - a.ix0, nt = nx * (a.iy1 - a.iy0);
    const uint32_t q = nt / 8u, r = nt % 8u, xcd = blockIdx.x % 8u;
    const uint32_t b = xcd * q + min(xcd, r) + blockIdx.x / 8u;
    ty = a.iy0 + b / nx;
    tx = a.ix0 + b % nx;
  } else {
    // the frame around the interior rectangle: rows above it, rows below
    // it, then the left and right parts of the rows beside it
    uint32_t b = MODE == 2 ? blockIdx.x - nint : blockIdx.x;
    const uint32_t top = a.iy0 * a.tiles_x, bottom = (a.tiles_y - a.iy1) * a.tiles_x;
    if (b < top) {
      ty = b / a.tiles_x;
      tx = b % a.tiles_x;
    } else if (b - top < bottom) {
      b -= top;
      ty = a.iy1 + b / a.tiles_x;
      tx = b % a.tiles_x;
    } else {
      b -= top + bottom;
      const uint32_t left = a.ix0, side = a.ix0 + (a.tiles_x - a.ix1);
      ty = a.iy0 + b / side;
      const uint32_t k = b % side;
      tx = k < left ? k : a.ix1 + (k - left);
    }
  }
  const int64_t x0 = (int64_t)tx * ST_TW, y0 = (int64_t)ty * ST_TH;

  // ---- stage the halo tile.  Every load of the tile is issued before the
  // first LDS write (indexes past the tile re-read its last element, so the
  // loads need no branch and each wait names only the loads it needs).
  if (interior) {
    constexpr int NV = LS / VEC;  // vectors per LDS row
    constexpr int TOT = LH * NV;
    constexpr int NI = (TOT + ST_BLOCK - 1) / ST_BLOCK;
    const T* src = a.in + (y0 - RY) * a.w + (x0 - PADL);
    Vec<T> vv[NI];
#pragma unroll
    for (int u = 0; u < NI; ++u) {
      const int i = min((int)threadIdx.x + u * ST_BLOCK, TOT - 1);
      const int r = i / NV, q = i - r * NV;
      vv[u] = load16(src + (int64_t)r * a.w + q * VEC);
    }
#pragma unroll
    for (int u = 0; u < NI; ++u) {
      const int i = threadIdx.x + u * ST_BLOCK;
      const int r = i / NV, q = i - r * NV;
      if (i < TOT) *(Vec<T>*)&tile[r * LS + q * VEC] = vv[u];
    }
  } else {
    constexpr int TOT = LH * LW;
    constexpr int NI = (TOT + ST_BLOCK - 1) / ST_BLOCK;
    const bool constant = a.boundary == ACC_BOUNDARY_CONSTANT;
    T vb[NI];
    uint32_t outside = 0;  // bit u: element u of this thread lies outside the array
    static_assert(NI <= 32, "halo tile too large for the outside mask");
#pragma unroll
    for (int u = 0; u < NI; ++u) {
      const int i = min((int)threadIdx.x + u * ST_BLOCK, TOT - 1);
      const int r = i / LW, c = i - r * LW;
      int64_t gy = y0 - RY + r, gx = x0 - RX + c;
      const bool inside = (uint64_t)gy < (uint64_t)a.h && (uint64_t)gx < (uint64_t)a.w;
      if (!inside) {
        gy = bound_index(a.boundary, gy, a.h);
        gx = bound_index(a.boundary, gx, a.w);
      }
      outside |= (inside ? 0u : 1u) << u;
      vb[u] = a.in[gy * a.w + gx];
    }
#pragma unroll
    for (int u = 0; u < NI; ++u) {
      const int i = threadIdx.x + u * ST_BLOCK;
      const int r = i / LW, c = i - r * LW;
      const bool use_cval = constant && (outside >> u & 1u);
      if (i < TOT) tile[r * LS + c + CO] = use_cval ? a.cval : vb[u];
    }
  }
  __syncthreads();
  if (!interior && a.boundary == ACC_BOUNDARY_FUNCTION) {
    // the function boundary: a second pass gives each cell outside the array
    // f(index) (staged above from a clamped index), so the staging keeps the
    // registers of the other boundaries
    constexpr int TOT = LH * LW;
    for (int i = threadIdx.x; i < TOT; i += ST_BLOCK) {
      const int r = i / LW, c = i - r * LW;
      const int64_t gy = y0 - RY + r, gx = x0 - RX + c;
      if ((uint64_t)gy >= (uint64_t)a.h || (uint64_t)gx >= (uint64_t)a.w)
        tile[r * LS + c + CO] = bound_function(a.fc, (int32_t)gy, (int32_t)gx);
    }
    __syncthreads();
  }

  // ---- ST_PX x ST_PY outputs per thread from a register window
  const int px = threadIdx.x % (ST_TW / ST_PX), py = threadIdx.x / (ST_TW / ST_PX);
  const int cx = px * ST_PX, cy = py * ST_PY;  // tile-local output origin
  T acc[ST_PY][ST_PX];
#pragma unroll
  for (int lr = 0; lr < ST_PY + 2 * RY; ++lr) {
    // the window row from 16-byte LDS reads (cx is a multiple of VEC)
    T wv[NVW * VEC];
#pragma unroll
    for (int q = 0; q < NVW; ++q) {
      const Vec<T> r = *(const Vec<T>*)&tile[(cy + lr) * LS + cx + q * VEC];
#pragma unroll
      for (int j = 0; j < VEC; ++j) wv[q * VEC + j] = r.e[j];
    }
    T v[WIN];
#pragma unroll
    for (int c = 0; c < WIN; ++c) v[c] = wv[CO + c];
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
        store16_stream(o + q * Vec<T>::N, r);  // written once: non-temporal
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
  constexpr int VEC = Vec<T>::N;
  constexpr int64_t PADL = (RX + VEC - 1) / VEC * VEC;
  StArgs<T> a;
  a.in = in;
  a.out = out;
  a.h = h;
  a.w = wd;
  const int64_t tiles_x = ceil_div(wd, ST_TW), tiles_y = ceil_div(h, ST_TH);
  if (tiles_x * tiles_y > 0x7fffffffLL) return fail(ACC_ERROR_UNSUPPORTED, "stencil: array too large");
  a.tiles_x = (uint32_t)tiles_x;
  a.tiles_y = (uint32_t)tiles_y;
  // interior tiles: x0 >= PADL, x0 + TW + PADL <= w, y0 >= RY, y0 + TH + RY <= h,
  // with 16-byte aligned rows
  int64_t ix0 = ceil_div(PADL, (int64_t)ST_TW), ix1 = wd >= ST_TW + PADL ? (wd - ST_TW - PADL) / ST_TW + 1 : 0;
  int64_t iy0 = ceil_div((int64_t)RY, (int64_t)ST_TH), iy1 = h >= ST_TH + RY ? (h - ST_TH - RY) / ST_TH + 1 : 0;
  const bool aligned = (uintptr_t)in % 16 == 0 && (wd * (int64_t)sizeof(T)) % 16 == 0;
  if (!aligned || ix1 <= ix0 || iy1 <= iy0) ix0 = ix1 = iy0 = iy1 = 0;
  a.ix0 = (uint32_t)ix0;
  a.ix1 = (uint32_t)ix1;
  a.iy0 = (uint32_t)iy0;
  a.iy1 = (uint32_t)iy1;
  a.boundary = boundary;
  a.cval = cval && boundary == ACC_BOUNDARY_CONSTANT ? *cval : T();
  for (int k = 0; k < 3; ++k) a.fc[k] = cval && boundary == ACC_BOUNDARY_FUNCTION ? cval[k] : T();
  const int taps = (2 * RY + 1) * (2 * RX + 1);
  for (int i = 0; i < 81; ++i) a.wt[i] = i < taps ? w[i] : T();
  const int64_t nt = tiles_x * tiles_y;
  // one launch: blocks [0, (ix1-ix0)*(iy1-iy0)) take the interior tiles, the
  // rest the frame
  hipLaunchKernelGGL((k_stencil2d<T, RY, RX, 2>), dim3((unsigned)nt), dim3(ST_BLOCK), 0, s->stream, a);
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
  if (boundary < ACC_BOUNDARY_CLAMP || boundary > ACC_BOUNDARY_FUNCTION)
    return fail(ACC_ERROR_INVALID_VALUE, "stencil: boundary %d", (int)boundary);
  if ((boundary == ACC_BOUNDARY_CONSTANT || boundary == ACC_BOUNDARY_FUNCTION) && !constant)
    return fail(ACC_ERROR_INVALID_VALUE, "stencil: constant/function boundary needs its values");
  if (boundary == ACC_BOUNDARY_FUNCTION && (height >= 0x7ffffff8LL || width >= 0x7ffffff8LL))
    return fail(ACC_ERROR_UNSUPPORTED, "stencil: function boundary on extents of 2^31 or more");
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
