// stencil2 on DIM2 arrays with the weighted-sum stencil function of two
// inputs (PTX/Execute.hs:669-684 runs it through the same interior + border
// kernels as stencil1; NAT/CodeGen/Stencil.hs:75-107 reads both stencils,
// each through its own boundary, and applies the function to the pair):
//
//   out[y][x] = sum_{dy,dx} w1[dy][dx] * a[y+dy][x+dx] + sum_{dy,dx} w2[dy][dx] * b[y+dy][x+dx]
//
// over the intersection of the two shapes (PTX/Execute.hs:682), each input
// indexed through its own boundary against its own extents.  The two stencils
// share the radii (a smaller one is written with zero weights).  Taps are
// summed in tuple order, all of a's taps before b's; floating results may
// differ from a left-to-right sum by rounding (fused multiply-add); integers
// wrap.
//
// One launch: each workgroup stages the (TH+2RY) x (TW+2RX) halo tiles of
// both inputs in LDS (interior tiles with dwordx4 copies, border tiles
// through the boundary maps), then each thread computes a 4x2 block of
// outputs from register windows read from LDS.  64 x 32 tiles keep both halo
// tiles of Float64 with radius 2 within 40 KiB of LDS.
//
// HBM roofline: 3*H*W*sizeof(T) bytes for equal shapes (each input read once,
// the output written once; halo re-reads are served by L2).
#include <algorithm>
#include "boundary.hpp"
#include "dispatch.hpp"
#include "vecmem.hpp"

namespace acc {

constexpr int S2_TW = 64;  // tile width (columns)
constexpr int S2_TH = 32;  // tile height (rows)
constexpr int S2_PX = 4;   // outputs per thread in x
constexpr int S2_PY = 2;   // outputs per thread in y
constexpr int S2_BLOCK = (S2_TW / S2_PX) * (S2_TH / S2_PY);  // 256

template <typename T> struct St2Src {
  const T* p;
  int64_t h, w;  // extents; w is also the row pitch
  int boundary;
  T cval;
  T fc[3];   // Function boundary: (c0 + cy*y) + cx*x
  T wt[25];  // (2RY+1) x (2RX+1) weights, row-major (radii 1..2)
};

template <typename T> struct St2Args {
  St2Src<T> src[2];
  T* out;
  int64_t h, w;  // output extents (the intersection); w is the output row pitch
  uint32_t tiles_x, tiles_y;
  // interior tiles: tx in [ix0, ix1), ty in [iy0, iy1), where both inputs'
  // halos lie inside and their rows are 16-byte aligned (empty otherwise)
  uint32_t ix0, ix1, iy0, iy1;
};

template <typename T, int RY, int RX>
__global__ __launch_bounds__(S2_BLOCK) void k_stencil2(St2Args<T> a) {
  constexpr int VEC = Vec<T>::N;
  constexpr int PADL = (RX + VEC - 1) / VEC * VEC;  // halo rounded out to 16-byte boundaries
  constexpr int CO = PADL - RX;                     // halo column c sits at LDS column c + CO
  constexpr int LS = S2_TW + 2 * PADL;              // LDS row stride: a 16-byte multiple
  constexpr int LW = S2_TW + 2 * RX;                // halo row length
  constexpr int LH = S2_TH + 2 * RY;
  constexpr int KW = 2 * RX + 1;
  constexpr int WIN = S2_PX + 2 * RX;               // register window row
  constexpr int NVW = (CO + WIN + VEC - 1) / VEC;   // LDS vectors covering it
  __shared__ __align__(16) T tile[2][LH * LS];

  // Tile order: blocks [0, nint) take the interior tiles, each XCD (block b
  // runs on XCD b % 8) a contiguous band walked row by row, so horizontal
  // neighbours share their halo lines in one L2; the other blocks take the
  // frame (rows above, rows below, then the left and right parts of the rows
  // beside the interior).
  uint32_t tx, ty;
  const uint32_t nx = a.ix1 - a.ix0, nint = nx * (a.iy1 - a.iy0);
  const bool interior = blockIdx.x < nint;
  if (interior) {
    const uint32_t q = nint / 8u, r = nint % 8u, xcd = blockIdx.x % 8u;
    const uint32_t b = xcd * q + min(xcd, r) + blockIdx.x / 8u;
    ty = a.iy0 + b / nx;
    tx = a.ix0 + b % nx;
  } else {
    uint32_t b = blockIdx.x - nint;
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
  const int64_t x0 = (int64_t)tx * S2_TW, y0 = (int64_t)ty * S2_TH;

  // ---- stage both halo tiles; every load is issued before the first LDS
  // write (indexes past the tile re-read its last element, so the loads
  // need no branch)
  if (interior) {
    constexpr int NV = LS / VEC;  // vectors per LDS row
    constexpr int TOT = LH * NV;
    constexpr int NI = (TOT + S2_BLOCK - 1) / S2_BLOCK;
    Vec<T> vv[2][NI];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int64_t pitch = a.src[s].w;
      const T* src = a.src[s].p + (y0 - RY) * pitch + (x0 - PADL);
#pragma unroll
      for (int u = 0; u < NI; ++u) {
        const int i = min((int)threadIdx.x + u * S2_BLOCK, TOT - 1);
        const int r = i / NV, q = i - r * NV;
        vv[s][u] = load16(src + (int64_t)r * pitch + q * VEC);
      }
    }
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int u = 0; u < NI; ++u) {
        const int i = threadIdx.x + u * S2_BLOCK;
        const int r = i / NV, q = i - r * NV;
        if (i < TOT) *(Vec<T>*)&tile[s][r * LS + q * VEC] = vv[s][u];
      }
  } else {
    constexpr int TOT = LH * LW;
    constexpr int NI = (TOT + S2_BLOCK - 1) / S2_BLOCK;
    static_assert(NI <= 32, "halo tile too large for the outside mask");
    T vb[2][NI];
    uint32_t outside[2] = {0u, 0u};  // bit u: element u lies outside that input
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int64_t sh = a.src[s].h, sw = a.src[s].w;
      const int bnd = a.src[s].boundary;
      const T* sp = a.src[s].p;
#pragma unroll
      for (int u = 0; u < NI; ++u) {
        const int i = min((int)threadIdx.x + u * S2_BLOCK, TOT - 1);
        const int r = i / LW, c = i - r * LW;
        int64_t gy = y0 - RY + r, gx = x0 - RX + c;
        const bool inside = (uint64_t)gy < (uint64_t)sh && (uint64_t)gx < (uint64_t)sw;
        if (!inside) {
          gy = bound_index(bnd, gy, sh);
          gx = bound_index(bnd, gx, sw);
        }
        outside[s] |= (inside ? 0u : 1u) << u;
        vb[s][u] = sp[gy * sw + gx];
      }
    }
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const bool constant = a.src[s].boundary == ACC_BOUNDARY_CONSTANT;
#pragma unroll
      for (int u = 0; u < NI; ++u) {
        const int i = threadIdx.x + u * S2_BLOCK;
        const int r = i / LW, c = i - r * LW;
        const bool use_cval = constant && (outside[s] >> u & 1u);
        if (i < TOT) tile[s][r * LS + c + CO] = use_cval ? a.src[s].cval : vb[s][u];
      }
    }
  }
  __syncthreads();
  if (!interior && (a.src[0].boundary == ACC_BOUNDARY_FUNCTION || a.src[1].boundary == ACC_BOUNDARY_FUNCTION)) {
    // the function boundary: a second pass gives each cell outside that
    // input f(index) (staged above from a clamped index), so the staging
    // keeps the registers of the other boundaries
    constexpr int TOT = LH * LW;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      if (a.src[s].boundary != ACC_BOUNDARY_FUNCTION) continue;
      for (int i = threadIdx.x; i < TOT; i += S2_BLOCK) {
        const int r = i / LW, c = i - r * LW;
        const int64_t gy = y0 - RY + r, gx = x0 - RX + c;
        if ((uint64_t)gy >= (uint64_t)a.src[s].h || (uint64_t)gx >= (uint64_t)a.src[s].w)
          tile[s][r * LS + c + CO] = bound_function(a.src[s].fc, (int32_t)gy, (int32_t)gx);
      }
    }
    __syncthreads();
  }

  // ---- 4x2 outputs per thread: a's taps, then b's, in tuple order
  const int px = threadIdx.x % (S2_TW / S2_PX), py = threadIdx.x / (S2_TW / S2_PX);
  const int cx = px * S2_PX, cy = py * S2_PY;  // tile-local output origin
  T acc[S2_PY][S2_PX];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
#pragma unroll
    for (int lr = 0; lr < S2_PY + 2 * RY; ++lr) {
      T wv[NVW * VEC];
#pragma unroll
      for (int q = 0; q < NVW; ++q) {
        const Vec<T> r = *(const Vec<T>*)&tile[s][(cy + lr) * LS + cx + q * VEC];
#pragma unroll
        for (int j = 0; j < VEC; ++j) wv[q * VEC + j] = r.e[j];
      }
#pragma unroll
      for (int oy = 0; oy < S2_PY; ++oy) {
        const int dy = lr - oy;  // tap row 0..2RY
        if (dy >= 0 && dy <= 2 * RY) {
#pragma unroll
          for (int ox = 0; ox < S2_PX; ++ox) {
#pragma unroll
            for (int dx = 0; dx < KW; ++dx) {
              const T p = OpMul{}(a.src[s].wt[dy * KW + dx], wv[CO + ox + dx]);
              // a's first tap (dy == 0, dx == 0) starts the sum
              acc[oy][ox] = (s == 0 && dx == 0 && dy == 0) ? p : OpAdd{}(acc[oy][ox], p);
            }
          }
        }
      }
    }
  }

  // ---- store
#pragma unroll
  for (int oy = 0; oy < S2_PY; ++oy) {
    const int64_t gy = y0 + cy + oy;
    const int64_t gx = x0 + cx;
    T* o = a.out + gy * a.w + gx;
    if (gy < a.h && gx + S2_PX <= a.w && ((uintptr_t)o % 16) == 0 && S2_PX * sizeof(T) % 16 == 0) {
#pragma unroll
      for (int q = 0; q < (int)(S2_PX * sizeof(T) / 16); ++q) {
        Vec<T> r;
#pragma unroll
        for (int j = 0; j < Vec<T>::N; ++j) r.e[j] = acc[oy][q * Vec<T>::N + j];
        store16_stream(o + q * Vec<T>::N, r);  // written once: non-temporal
      }
    } else if (gy < a.h) {
#pragma unroll
      for (int ox = 0; ox < S2_PX; ++ox)
        if (gx + ox < a.w) o[ox] = acc[oy][ox];
    }
  }
}

template <typename T, int RY, int RX>
int stencil2_launch(acc_stream_t s, St2Args<T> a) {
  constexpr int VEC = Vec<T>::N;
  constexpr int64_t PADL = (RX + VEC - 1) / VEC * VEC;
  const int64_t h = a.h, wd = a.w;
  const int64_t tiles_x = ceil_div(wd, S2_TW), tiles_y = ceil_div(h, S2_TH);
  if (tiles_x * tiles_y > 0x7fffffffLL) return fail(ACC_ERROR_UNSUPPORTED, "stencil2: array too large");
  a.tiles_x = (uint32_t)tiles_x;
  a.tiles_y = (uint32_t)tiles_y;
  // interior tiles against the output extents, which neither input's
  // extents undercut: x0 >= PADL, x0 + TW + PADL <= w, y0 >= RY, y0 + TH + RY <= h
  int64_t ix0 = ceil_div(PADL, (int64_t)S2_TW), ix1 = wd >= S2_TW + PADL ? (wd - S2_TW - PADL) / S2_TW + 1 : 0;
  int64_t iy0 = ceil_div((int64_t)RY, (int64_t)S2_TH), iy1 = h >= S2_TH + RY ? (h - S2_TH - RY) / S2_TH + 1 : 0;
  bool aligned = true;
  for (int k = 0; k < 2; ++k)
    aligned = aligned && (uintptr_t)a.src[k].p % 16 == 0 && (a.src[k].w * (int64_t)sizeof(T)) % 16 == 0;
  if (!aligned || ix1 <= ix0 || iy1 <= iy0) ix0 = ix1 = iy0 = iy1 = 0;
  a.ix0 = (uint32_t)ix0;
  a.ix1 = (uint32_t)ix1;
  a.iy0 = (uint32_t)iy0;
  a.iy1 = (uint32_t)iy1;
  hipLaunchKernelGGL((k_stencil2<T, RY, RX>), dim3((unsigned)(tiles_x * tiles_y)), dim3(S2_BLOCK), 0, s->stream, a);
  ACC_HIP_CHECK(hipGetLastError());
  return ACC_SUCCESS;
}

template <typename T>
int stencil2_any(acc_stream_t s, int ry, int rx, const void* const* weights, const acc_boundary* boundary,
                 const void* const* constant, const void* const* in, const int64_t* height, const int64_t* width,
                 void* out) {
  St2Args<T> a{};
  const int taps = (2 * ry + 1) * (2 * rx + 1);
  for (int k = 0; k < 2; ++k) {
    St2Src<T>& src = a.src[k];
    src.p = (const T*)in[k];
    src.h = height[k];
    src.w = width[k];
    src.boundary = (int)boundary[k];
    const T* cv = (const T*)constant[k];
    src.cval = cv && boundary[k] == ACC_BOUNDARY_CONSTANT ? cv[0] : T();
    for (int j = 0; j < 3; ++j) src.fc[j] = cv && boundary[k] == ACC_BOUNDARY_FUNCTION ? cv[j] : T();
    for (int i = 0; i < 25; ++i) src.wt[i] = i < taps ? ((const T*)weights[k])[i] : T();
  }
  a.out = (T*)out;
  a.h = std::min(height[0], height[1]);
  a.w = std::min(width[0], width[1]);
  if (ry == 1) return rx == 1 ? stencil2_launch<T, 1, 1>(s, a) : stencil2_launch<T, 1, 2>(s, a);
  return rx == 1 ? stencil2_launch<T, 2, 1>(s, a) : stencil2_launch<T, 2, 2>(s, a);
}

}  // namespace acc

using namespace acc;

extern "C" int acc_hip_stencil2_2d(acc_stream_t s, acc_type t, int ry, int rx, const void* weights1,
                                   acc_boundary boundary1, const void* constant1, const void* in1,
                                   int64_t height1, int64_t width1, const void* weights2,
                                   acc_boundary boundary2, const void* constant2, const void* in2,
                                   int64_t height2, int64_t width2, void* out) {
  if (!s) return fail(ACC_ERROR_INVALID_VALUE, "stencil2: null stream");
  if (height1 < 0 || width1 < 0 || height2 < 0 || width2 < 0)
    return fail(ACC_ERROR_INVALID_VALUE, "stencil2: negative extent");
  if (ry < 1 || ry > 2 || rx < 1 || rx > 2) return fail(ACC_ERROR_UNSUPPORTED, "stencil2: radius must be 1..2");
  const acc_boundary bnd[2] = {boundary1, boundary2};
  const void* cst[2] = {constant1, constant2};
  for (int k = 0; k < 2; ++k) {
    if (bnd[k] < ACC_BOUNDARY_CLAMP || bnd[k] > ACC_BOUNDARY_FUNCTION)
      return fail(ACC_ERROR_INVALID_VALUE, "stencil2: boundary %d", (int)bnd[k]);
    if ((bnd[k] == ACC_BOUNDARY_CONSTANT || bnd[k] == ACC_BOUNDARY_FUNCTION) && !cst[k])
      return fail(ACC_ERROR_INVALID_VALUE, "stencil2: constant/function boundary needs its values");
    if (bnd[k] == ACC_BOUNDARY_FUNCTION && (std::min(height1, height2) >= 0x7ffffff8LL ||
                                            std::min(width1, width2) >= 0x7ffffff8LL))
      return fail(ACC_ERROR_UNSUPPORTED, "stencil2: function boundary on extents of 2^31 or more");
  }
  if (!weights1 || !weights2) return fail(ACC_ERROR_INVALID_VALUE, "stencil2: null weights");
  if (std::min(height1, height2) == 0 || std::min(width1, width2) == 0) return ACC_SUCCESS;
  if (!in1 || !in2 || !out) return fail(ACC_ERROR_INVALID_VALUE, "stencil2: null array");
  if (out == in1 || out == in2) return fail(ACC_ERROR_INVALID_VALUE, "stencil2: output must not alias an input");
  ACC_CHECK(set_device(s->device));
  const void* w[2] = {weights1, weights2};
  const void* in[2] = {in1, in2};
  const int64_t hh[2] = {height1, height2}, ww[2] = {width1, width2};
  // (+) and (*) give the same bits for signed and unsigned words
  switch (t) {
    case ACC_INT32: case ACC_WORD32: return stencil2_any<int32_t>(s, ry, rx, w, bnd, cst, in, hh, ww, out);
    case ACC_INT64: case ACC_WORD64: return stencil2_any<int64_t>(s, ry, rx, w, bnd, cst, in, hh, ww, out);
    case ACC_FLOAT32: return stencil2_any<float>(s, ry, rx, w, bnd, cst, in, hh, ww, out);
    case ACC_FLOAT64: return stencil2_any<double>(s, ry, rx, w, bnd, cst, in, hh, ww, out);
  }
  return fail(ACC_ERROR_UNSUPPORTED, "stencil2: element type %d", (int)t);
}
