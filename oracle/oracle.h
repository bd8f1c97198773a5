/*
 * oracle.h - CPU restatement of accelerate-llvm-native's hot-path skeletons.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; it is used by the tests (as the checker), by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg.
 *
 * Three families of functions:
 *   or_*        sequential semantics of the native backend's kernels, element
 *               for element (exact for integers);
 *   or_*_f64    FP64 "truth" for Float32 inputs (compensated summation) plus
 *               the sum of magnitudes needed by the normwise error bound;
 *   or_native_* the native backend's parallel schedule (divideWork ranges,
 *               3-phase scan, atomic permute, per-range stencil) on pthreads,
 *               timed as the CPU baseline.
 *
 * Op codes and boundary codes are those of include/accelerate_hip.h.
 * Reference files are cited relative to the accelerate-llvm repository:
 *   NAT/ = accelerate-llvm-native/src/Data/Array/Accelerate/LLVM/Native/
 *   LLVM/ = accelerate-llvm/src/Data/Array/Accelerate/LLVM/
 */
#ifndef ACC_ORACLE_H
#define ACC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_DECL(T, N)                                                                              \
  int or_fold_##N(int op, const T* seed, const T* in, int64_t outer, int64_t inner, T* out);      \
  int or_fold_seg_##N(int op, const T* seed, const T* in, int64_t outer, int64_t inner,           \
                      const int64_t* offsets, int64_t nseg, T* out);                               \
  int or_scan_##N(int op, int dir, const T* seed, const T* in, int64_t outer, int64_t inner,       \
                  T* out);                                                                         \
  int or_scan_prime_##N(int op, int dir, const T* seed, const T* in, int64_t outer,               \
                        int64_t inner, T* out, T* sums);                                           \
  int or_permute_##N(int op, T* out, int64_t m, const T* src, const T* fill, const int64_t* idx,  \
                     int64_t n);                                                                   \
  int or_stencil2d_##N(int ry, int rx, const T* w, int boundary, const T* cval, const T* in,      \
                       int64_t h, int64_t wd, T* out);

OR_DECL(int32_t, i32)
OR_DECL(int64_t, i64)
OR_DECL(uint32_t, u32)
OR_DECL(uint64_t, u64)
OR_DECL(float, f32)
OR_DECL(double, f64)

/* FP64 truth for Float32 data: value and sum of |terms| per result. */
void or_fold_f32_f64(const float* in, int64_t n, const float* seed, double* out, double* abs_sum);
void or_fold_seg_f32_f64(const float* in, const int64_t* offsets, int64_t nseg, const float* seed,
                         double* out, double* abs_sum);
void or_scan_f32_f64(const float* in, int64_t n, double* out, double* abs_sum);
void or_stencil2d_f32_f64(int ry, int rx, const float* w, int boundary, const float* cval,
                          const float* in, int64_t h, int64_t wd, double* out, double* abs_sum);

/* divideWork1 (NAT/Execute/Divide.hs:67-104): pieces in piece-index order.
 * ranges[2i], ranges[2i+1] = [from, to) of piece i.  Returns the count. */
int64_t or_divide_work1(int64_t splits, int64_t minsize, int64_t from, int64_t to, int64_t* ranges,
                        int64_t max_pieces);

/* Native-schedule parallel versions (CPU baseline). threads >= 1. */
int32_t or_native_fold_add_i32(int threads, const int32_t* in, int64_t n, int32_t seed);
void or_native_scanl1_add_i32(int threads, const int32_t* in, int64_t n, int32_t* out);
void or_native_fold1_seg_f32(int threads, int op, const float* in, const int64_t* offsets,
                             int64_t nseg, float* out);
void or_native_histogram_i32(int threads, const int32_t* idx, int64_t n, int32_t* bins, int64_t m);
void or_native_stencil2d_f32(int threads, int ry, int rx, const float* w, int boundary,
                             const float* in, int64_t h, int64_t wd, float* out);

/* Seeded generators (splitmix64 -> xoshiro256**), shared with bench/tests. */
const char* or_build_info(void);
void or_gen_uniform_i32(uint64_t seed, int32_t lo, int32_t hi, int32_t* out, int64_t n);
void or_gen_uniform_f32(uint64_t seed, float lo, float hi, float* out, int64_t n);
void or_gen_normal_f32(uint64_t seed, float* out, int64_t n);
void or_gen_zipf_i32(uint64_t seed, double s, int32_t nbins, int32_t* out, int64_t n);
/* log-uniform segment lengths in [1, maxlen] normalised to sum to total */
void or_gen_segment_lengths(uint64_t seed, int64_t nseg, int64_t maxlen, int64_t total, int64_t* out);

#ifdef __cplusplus
}
#endif
#endif
