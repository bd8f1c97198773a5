/*
 * oracle.c - CPU restatement of accelerate-llvm-native's hot-path skeletons.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  The product library never links,
 * loads or calls this file.
 *
 * Scalar semantics (shared with the GPU functors):
 *   integer + - * wrap (two's complement; no nsw/nuw on the reference's
 *   add/sub/mul, accelerate-llvm/src/LLVM/AST/Type/Instruction.hs:425-434);
 *   floating min/max are fminf/fmaxf (LLVM/CodeGen/Arithmetic.hs:456-466).
 */
#include "oracle.h"
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

enum { OP_ADD = 0, OP_MUL = 1, OP_MIN = 2, OP_MAX = 3, OP_AND = 4, OP_OR = 5, OP_XOR = 6,
       OP_SUB = 7, OP_EXCHANGE = 8 };
enum { B_CLAMP = 0, B_MIRROR = 1, B_WRAP = 2, B_CONSTANT = 3, B_FUNCTION = 4 };

/* ------------------------------------------------------------ combine */
#define INT_COMBINE(T, U, N)                                                            \
  static inline int comb_##N(int op, T a, T b, T* r) {                                  \
    switch (op) {                                                                       \
      case OP_ADD: *r = (T)((U)a + (U)b); return 0;                                     \
      case OP_MUL: *r = (T)((U)a * (U)b); return 0;                                     \
      case OP_MIN: *r = a < b ? a : b; return 0;                                        \
      case OP_MAX: *r = a < b ? b : a; return 0;                                        \
      case OP_AND: *r = a & b; return 0;                                                \
      case OP_OR: *r = a | b; return 0;                                                 \
      case OP_XOR: *r = a ^ b; return 0;                                                \
    }                                                                                   \
    return -1;                                                                          \
  }                                                                                     \
  /* permute: combine new old (NAT/CodeGen/Permute.hs:110-117; RMW forms  */           \
  /* LLVM/CodeGen/Permute.hs:117-145: sub is old - new, exchange is new)  */           \
  static inline int pcomb_##N(int op, T nw, T old, T* r) {                              \
    if (op == OP_SUB) { *r = (T)((U)old - (U)nw); return 0; }                           \
    if (op == OP_EXCHANGE) { *r = nw; return 0; }                                       \
    return comb_##N(op, nw, old, r);                                                    \
  }

INT_COMBINE(int32_t, uint32_t, i32)
INT_COMBINE(int64_t, uint64_t, i64)
INT_COMBINE(uint32_t, uint32_t, u32)
INT_COMBINE(uint64_t, uint64_t, u64)

#define FLT_COMBINE(T, N, FMIN, FMAX)                                   \
  static inline int comb_##N(int op, T a, T b, T* r) {                  \
    switch (op) {                                                       \
      case OP_ADD: *r = a + b; return 0;                                \
      case OP_MUL: *r = a * b; return 0;                                \
      case OP_MIN: *r = FMIN(a, b); return 0;                           \
      case OP_MAX: *r = FMAX(a, b); return 0;                           \
    }                                                                   \
    return -1;                                                          \
  }                                                                     \
  static inline int pcomb_##N(int op, T nw, T old, T* r) {              \
    if (op == OP_SUB) { *r = old - nw; return 0; }                      \
    if (op == OP_EXCHANGE) { *r = nw; return 0; }                       \
    return comb_##N(op, nw, old, r);                                    \
  }

FLT_COMBINE(float, f32, fminf, fmaxf)
FLT_COMBINE(double, f64, fmin, fmax)

/* Stencil boundary index map, one dimension (LLVM/CodeGen/Stencil.hs:224-253).
 * The reference maps once; for arrays smaller than the radius the mapped index
 * can still be outside, which the reference does not guard.  Here it is then
 * clamped so the restatement stays in bounds. */
static inline int64_t bmap(int b, int64_t i, int64_t n) {
  if (i < 0) {
    if (b == B_CLAMP) i = 0;
    else if (b == B_MIRROR) i = -i;
    else if (b == B_WRAP) i = n + i;
  } else if (i >= n) {
    if (b == B_CLAMP) i = n - 1;
    else if (b == B_MIRROR) i = n - (i - n + 2);
    else if (b == B_WRAP) i = i - n;
  }
  if (i < 0) i = 0;
  if (i >= n) i = n - 1;
  return i;
}

#define OR_DEFINE(T, N)                                                                         \
  /* fold: foldAllS / foldAllP1+P2 (NAT/CodeGen/Fold.hs:141-228) and the per-row fold      */  \
  /* (Fold.hs:230-275): r = z; r = r (+) x[i] left to right; fold1 starts from x[0].        */  \
  int or_fold_##N(int op, const T* seed, const T* in, int64_t outer, int64_t inner, T* out) {   \
    if (!seed && inner == 0) return -2; /* fold1 of an empty array: boundsCheck */              \
    for (int64_t r = 0; r < outer; ++r) {                                                       \
      const T* x = in + r * inner;                                                              \
      T acc;                                                                                    \
      int64_t i = 0;                                                                            \
      if (seed) acc = *seed; else acc = x[i++];                                                 \
      for (; i < inner; ++i)                                                                    \
        if (comb_##N(op, acc, x[i], &acc)) return -1;                                           \
      out[r] = acc;                                                                             \
    }                                                                                           \
    return 0;                                                                                   \
  }                                                                                             \
  /* foldSegP (NAT/CodeGen/FoldSeg.hs:103-141): out[r][s] = fold over [off[s], off[s+1]).   */  \
  int or_fold_seg_##N(int op, const T* seed, const T* in, int64_t outer, int64_t inner,         \
                      const int64_t* off, int64_t nseg, T* out) {                               \
    for (int64_t r = 0; r < outer; ++r)                                                         \
      for (int64_t s = 0; s < nseg; ++s) {                                                      \
        const int64_t u = off[s], v = off[s + 1];                                               \
        if (!seed && v <= u) continue; /* fold1Seg: non-empty segments only */                  \
        T acc;                                                                                  \
        int64_t i = u;                                                                          \
        if (seed) acc = *seed; else acc = in[r * inner + i++];                                  \
        for (; i < v; ++i)                                                                      \
          if (comb_##N(op, acc, in[r * inner + i], &acc)) return -1;                            \
        out[r * nseg + s] = acc;                                                                \
      }                                                                                         \
    return 0;                                                                                   \
  }                                                                                             \
  /* scanS (NAT/CodeGen/Scan.hs:136-210): L->R combines acc (+) x, R->L x (+) acc;          */  \
  /* with a seed the output row has inner+1 elements, seed first (L->R) or last (R->L).    */  \
  int or_scan_##N(int op, int dir, const T* seed, const T* in, int64_t outer, int64_t inner,   \
                  T* out) {                                                                     \
    const int64_t on = seed ? inner + 1 : inner;                                                \
    for (int64_t r = 0; r < outer; ++r) {                                                       \
      const T* x = in + r * inner;                                                              \
      T* y = out + r * on;                                                                      \
      if (inner == 0) { if (seed) y[0] = *seed; continue; }                                     \
      if (dir == 0) {                                                                           \
        int64_t i = 0, j = 0;                                                                   \
        T acc;                                                                                  \
        if (seed) acc = *seed; else acc = x[i++];                                               \
        y[j++] = acc;                                                                           \
        for (; i < inner; ++i) { if (comb_##N(op, acc, x[i], &acc)) return -1; y[j++] = acc; } \
      } else {                                                                                  \
        int64_t i = inner - 1, j = seed ? inner : inner - 1;                                    \
        T acc;                                                                                  \
        if (seed) acc = *seed; else acc = x[i--];                                               \
        y[j--] = acc;                                                                           \
        for (; i >= 0; --i) { if (comb_##N(op, x[i], acc, &acc)) return -1; y[j--] = acc; }   \
      }                                                                                         \
    }                                                                                           \
    return 0;                                                                                   \
  }                                                                                             \
  /* scan'S (NAT/CodeGen/Scan.hs:213-270): out[i] is the carry before x[i] is combined,     */  \
  /* sums[r] the final carry.  The code (not the doc comment at :81-85) fixes the R->L     */  \
  /* result: scanr' (+) 10 [0..9] = ([55,54,52,49,45,40,34,27,19,10], 55).                 */  \
  int or_scan_prime_##N(int op, int dir, const T* seed, const T* in, int64_t outer,            \
                        int64_t inner, T* out, T* sums) {                                       \
    for (int64_t r = 0; r < outer; ++r) {                                                       \
      const T* x = in + r * inner;                                                              \
      T* y = out + r * inner;                                                                   \
      T acc = *seed;                                                                            \
      if (dir == 0) {                                                                           \
        for (int64_t i = 0; i < inner; ++i) { y[i] = acc; if (comb_##N(op, acc, x[i], &acc)) return -1; } \
      } else {                                                                                  \
        for (int64_t i = inner - 1; i >= 0; --i) { y[i] = acc; if (comb_##N(op, x[i], acc, &acc)) return -1; } \
      }                                                                                         \
      sums[r] = acc;                                                                            \
    }                                                                                           \
    return 0;                                                                                   \
  }                                                                                             \
  /* permuteS (NAT/CodeGen/Permute.hs:85-117): for each source index in order,              */  \
  /* out[j] = combine new old when the projection is Just j.  Indices outside [0, m)       */  \
  /* stand for Nothing.  src == NULL is the delayed constant source `fill sh v`.           */  \
  int or_permute_##N(int op, T* out, int64_t m, const T* src, const T* fill, const int64_t* idx, \
                     int64_t n) {                                                               \
    for (int64_t i = 0; i < n; ++i) {                                                           \
      const int64_t j = idx[i];                                                                 \
      if (j < 0 || j >= m) continue;                                                            \
      const T nw = src ? src[i] : *fill;                                                        \
      if (pcomb_##N(op, nw, out[j], &out[j])) return -1;                                        \
    }                                                                                           \
    return 0;                                                                                   \
  }                                                                                             \
  /* stencil1 with f = sum of w * x over the (2ry+1)x(2rx+1) taps in tuple order: rows      */  \
  /* (leftmost index) -ry..ry outer, columns -rx..rx inner (LLVM/CodeGen/Stencil.hs:54-180) */  \
  /* and boundary maps at :207-253 (constant: any coordinate outside gives c; function:   */  \
  /* an index outside gives f(index), :215-218, here f = (c0 + cy*y) + cx*x in T).        */  \
  int or_stencil2d_##N(int ry, int rx, const T* w, int boundary, const T* cval, const T* in,   \
                       int64_t h, int64_t wd, T* out) {                                         \
    for (int64_t y = 0; y < h; ++y)                                                             \
      for (int64_t x = 0; x < wd; ++x) {                                                        \
        T acc = 0;                                                                              \
        int first = 1;                                                                          \
        for (int dy = -ry; dy <= ry; ++dy)                                                      \
          for (int dx = -rx; dx <= rx; ++dx) {                                                  \
            const int64_t yy = y + dy, xx = x + dx;                                             \
            T v;                                                                                \
            const int outside = yy < 0 || yy >= h || xx < 0 || xx >= wd;                       \
            if (boundary == B_CONSTANT && outside) {                                            \
              v = *cval;                                                                        \
            } else if (boundary == B_FUNCTION && outside) {                                     \
              T fy, fx;                                                                         \
              comb_##N(OP_MUL, cval[1], (T)yy, &fy);                                            \
              comb_##N(OP_ADD, cval[0], fy, &fy);                                               \
              comb_##N(OP_MUL, cval[2], (T)xx, &fx);                                            \
              comb_##N(OP_ADD, fy, fx, &v);                                                     \
            } else {                                                                            \
              v = in[bmap(boundary, yy, h) * wd + bmap(boundary, xx, wd)];                      \
            }                                                                                   \
            T p;                                                                                \
            comb_##N(OP_MUL, w[(dy + ry) * (2 * rx + 1) + (dx + rx)], v, &p);                   \
            if (first) { acc = p; first = 0; } else comb_##N(OP_ADD, acc, p, &acc);             \
          }                                                                                     \
        out[y * wd + x] = acc;                                                                  \
      }                                                                                         \
    return 0;                                                                                   \
  }

OR_DEFINE(int32_t, i32)
OR_DEFINE(int64_t, i64)
OR_DEFINE(uint32_t, u32)
OR_DEFINE(uint64_t, u64)
OR_DEFINE(float, f32)
OR_DEFINE(double, f64)

/* --------------------------------------------------------- FP64 truth */
/* Neumaier compensated summation in double. */
typedef struct { double s, c; } ksum;
static inline void kadd(ksum* k, double x) {
  double t = k->s + x;
  if (fabs(k->s) >= fabs(x)) k->c += (k->s - t) + x; else k->c += (x - t) + k->s;
  k->s = t;
}
static inline double kval(const ksum* k) { return k->s + k->c; }

void or_fold_f32_f64(const float* in, int64_t n, const float* seed, double* out, double* abs_sum) {
  ksum k = {0, 0}, a = {0, 0};
  if (seed) { kadd(&k, *seed); kadd(&a, fabs(*seed)); }
  for (int64_t i = 0; i < n; ++i) { kadd(&k, in[i]); kadd(&a, fabs(in[i])); }
  *out = kval(&k);
  *abs_sum = kval(&a);
}

void or_fold_seg_f32_f64(const float* in, const int64_t* off, int64_t nseg, const float* seed,
                         double* out, double* abs_sum) {
  for (int64_t s = 0; s < nseg; ++s) {
    ksum k = {0, 0}, a = {0, 0};
    if (seed) { kadd(&k, *seed); kadd(&a, fabs(*seed)); }
    for (int64_t i = off[s]; i < off[s + 1]; ++i) { kadd(&k, in[i]); kadd(&a, fabs(in[i])); }
    out[s] = kval(&k);
    abs_sum[s] = kval(&a);
  }
}

void or_scan_f32_f64(const float* in, int64_t n, double* out, double* abs_sum) {
  ksum k = {0, 0}, a = {0, 0};
  for (int64_t i = 0; i < n; ++i) {
    kadd(&k, in[i]);
    kadd(&a, fabs(in[i]));
    out[i] = kval(&k);
    abs_sum[i] = kval(&a);
  }
}

void or_stencil2d_f32_f64(int ry, int rx, const float* w, int boundary, const float* cval,
                          const float* in, int64_t h, int64_t wd, double* out, double* abs_sum) {
  for (int64_t y = 0; y < h; ++y)
    for (int64_t x = 0; x < wd; ++x) {
      double acc = 0, ab = 0;
      for (int dy = -ry; dy <= ry; ++dy)
        for (int dx = -rx; dx <= rx; ++dx) {
          const int64_t yy = y + dy, xx = x + dx;
          double v;
          const int outside = yy < 0 || yy >= h || xx < 0 || xx >= wd;
          if (boundary == B_CONSTANT && outside) {
            v = *cval;
          } else if (boundary == B_FUNCTION && outside) {
            const float fy = cval[0] + cval[1] * (float)yy;  /* in Float32, rounded per step */
            v = (double)(fy + cval[2] * (float)xx);
          } else {
            v = in[bmap(boundary, yy, h) * wd + bmap(boundary, xx, wd)];
          }
          const double p = (double)w[(dy + ry) * (2 * rx + 1) + (dx + rx)] * v;
          acc += p;
          ab += fabs(p);
        }
      out[y * wd + x] = acc;
      abs_sum[y * wd + x] = ab;
    }
}

/* ------------------------------------------------------- divideWork1 */
/* NAT/Execute/Divide.hs:67-104: recursive bisection with split points rounded
 * up to a multiple of minsize (a power of two); pieces are numbered left to
 * right. */
typedef struct { int64_t* r; int64_t max, count; } pieces;

static void dw_split(int64_t s, int64_t u, int64_t v, int64_t minsize, pieces* p) {
  if (s == 0 || v - u <= minsize) {
    if (p->count < p->max) { p->r[2 * p->count] = u; p->r[2 * p->count + 1] = v; }
    p->count++;
    return;
  }
  /* findSplitPoint1 (Divide.hs:90-104) */
  const int64_t a = v - u;
  const int64_t b = (a + 1) >> 1;
  const int64_t c = minsize - 1;
  const int64_t d = (b + c) & ~c;
  dw_split(s >> 1, u, d + u, minsize, p);
  dw_split(s >> 1, v - a + d, v, minsize, p);
}

int64_t or_divide_work1(int64_t splits, int64_t minsize, int64_t from, int64_t to, int64_t* ranges,
                        int64_t max_pieces) {
  pieces p = {ranges, max_pieces, 0};
  dw_split(splits, from, to, minsize, &p);
  return p.count;
}

/* ---------------------------------------------------- thread helper */
typedef void (*task_fn)(void* ctx, int64_t i);
typedef struct { task_fn fn; void* ctx; int64_t count; int64_t next; pthread_mutex_t mu; } pool;

static void* pool_worker(void* arg) {
  pool* p = (pool*)arg;
  for (;;) {
    pthread_mutex_lock(&p->mu);
    int64_t i = p->next++;
    pthread_mutex_unlock(&p->mu);
    if (i >= p->count) return 0;
    p->fn(p->ctx, i);
  }
}

static void run_tasks(int threads, int64_t count, task_fn fn, void* ctx) {
  pool p = {fn, ctx, count, 0, PTHREAD_MUTEX_INITIALIZER};
  if (threads <= 1 || count <= 1) { for (int64_t i = 0; i < count; ++i) fn(ctx, i); return; }
  pthread_t* t = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
  for (int i = 0; i < threads; ++i) pthread_create(&t[i], 0, pool_worker, &p);
  for (int i = 0; i < threads; ++i) pthread_join(t[i], 0);
  free(t);
}

/* -------------------------------------------- native-schedule baseline */
/* foldAllOp (NAT/Execute.hs:293-333): divideWork1 (workers-1) splits, minsize
 * 4096; foldAllP1 per range into tmp[piece]; foldAllP2 folds tmp with the seed. */
typedef struct { const int32_t* in; const int64_t* r; int32_t* tmp; } fold_ctx;
static void fold_task(void* c, int64_t i) {
  fold_ctx* x = (fold_ctx*)c;
  uint32_t acc = 0;
  for (int64_t k = x->r[2 * i]; k < x->r[2 * i + 1]; ++k) acc += (uint32_t)x->in[k];
  x->tmp[i] = (int32_t)acc;
}

int32_t or_native_fold_add_i32(int threads, const int32_t* in, int64_t n, int32_t seed) {
  int64_t maxp = 4 * (int64_t)threads + 8;
  int64_t* r = (int64_t*)malloc(sizeof(int64_t) * 2 * (size_t)maxp);
  int64_t np = or_divide_work1(threads - 1, 4096, 0, n, r, maxp);
  if (np > maxp) np = maxp;
  int32_t* tmp = (int32_t*)malloc(sizeof(int32_t) * (size_t)np);
  fold_ctx c = {in, r, tmp};
  run_tasks(threads, np, fold_task, &c);
  uint32_t acc = (uint32_t)seed;
  for (int64_t i = 0; i < np; ++i) acc += (uint32_t)tmp[i];
  free(tmp);
  free(r);
  return (int32_t)acc;
}

/* scanOp, n >= 8192 (NAT/Execute.hs:462-487): scanP1 per range (local
 * inclusive scan, range total to tmp), scanP2 (sequential scan of tmp),
 * scanP3 (add carry to every range but the first). */
typedef struct { const int32_t* in; int32_t* out; const int64_t* r; int32_t* tmp; int phase; } scan_ctx;
static void scan_task(void* c, int64_t i) {
  scan_ctx* x = (scan_ctx*)c;
  const int64_t u = x->r[2 * i], v = x->r[2 * i + 1];
  if (x->phase == 1) {
    uint32_t acc = 0;
    for (int64_t k = u; k < v; ++k) { acc += (uint32_t)x->in[k]; x->out[k] = (int32_t)acc; }
    x->tmp[i] = (int32_t)acc;
  } else if (i > 0) {
    const uint32_t carry = (uint32_t)x->tmp[i - 1];
    for (int64_t k = u; k < v; ++k) x->out[k] = (int32_t)((uint32_t)x->out[k] + carry);
  }
}

void or_native_scanl1_add_i32(int threads, const int32_t* in, int64_t n, int32_t* out) {
  if (n < 8192 || threads <= 1) {
    uint32_t acc = 0;
    for (int64_t k = 0; k < n; ++k) { acc += (uint32_t)in[k]; out[k] = (int32_t)acc; }
    return;
  }
  int64_t maxp = 4 * (int64_t)threads + 8;
  int64_t* r = (int64_t*)malloc(sizeof(int64_t) * 2 * (size_t)maxp);
  int64_t np = or_divide_work1(threads - 1, 8192, 0, n, r, maxp);
  if (np > maxp) np = maxp;
  int32_t* tmp = (int32_t*)malloc(sizeof(int32_t) * (size_t)np);
  scan_ctx c = {in, out, r, tmp, 1};
  run_tasks(threads, np, scan_task, &c);
  for (int64_t i = 1; i < np; ++i) tmp[i] = (int32_t)((uint32_t)tmp[i] + (uint32_t)tmp[i - 1]);
  c.phase = 3;
  run_tasks(threads, np, scan_task, &c);
  free(tmp);
  free(r);
}

/* foldSegOp: segments split over workers (scheduleOp, DIM1 minsize 4096,
 * NAT/Execute.hs:818-824), each segment folded sequentially. */
typedef struct { int op; const float* in; const int64_t* off; float* out; const int64_t* r; } seg_ctx;
static void seg_task(void* c, int64_t i) {
  seg_ctx* x = (seg_ctx*)c;
  for (int64_t s = x->r[2 * i]; s < x->r[2 * i + 1]; ++s) {
    const int64_t u = x->off[s], v = x->off[s + 1];
    if (v <= u) continue;
    float acc = x->in[u];
    if (x->op == OP_MAX) for (int64_t k = u + 1; k < v; ++k) acc = fmaxf(acc, x->in[k]);
    else for (int64_t k = u + 1; k < v; ++k) acc += x->in[k];
    x->out[s] = acc;
  }
}

void or_native_fold1_seg_f32(int threads, int op, const float* in, const int64_t* off, int64_t nseg,
                             float* out) {
  int64_t maxp = 4 * (int64_t)threads + 8;
  int64_t* r = (int64_t*)malloc(sizeof(int64_t) * 2 * (size_t)maxp);
  int64_t np = or_divide_work1(threads - 1, 4096, 0, nseg, r, maxp);
  if (np > maxp) np = maxp;
  seg_ctx c = {op, in, off, out, r};
  run_tasks(threads, np, seg_task, &c);
  free(r);
}

/* permuteOp with an RMW combine (NAT/Execute.hs:595-654, permuteP_rmw
 * NAT/CodeGen/Permute.hs:140-213): one atomic fetch-add per source element. */
typedef struct { const int32_t* idx; int32_t* bins; int64_t m; const int64_t* r; } hist_ctx;
static void hist_task(void* c, int64_t i) {
  hist_ctx* x = (hist_ctx*)c;
  for (int64_t k = x->r[2 * i]; k < x->r[2 * i + 1]; ++k) {
    const int64_t j = x->idx[k];
    if (j >= 0 && j < x->m) __atomic_fetch_add(&x->bins[j], 1, __ATOMIC_RELAXED);
  }
}

void or_native_histogram_i32(int threads, const int32_t* idx, int64_t n, int32_t* bins, int64_t m) {
  int64_t maxp = 4 * (int64_t)threads + 8;
  int64_t* r = (int64_t*)malloc(sizeof(int64_t) * 2 * (size_t)maxp);
  int64_t np = or_divide_work1(threads - 1, 4096, 0, n, r, maxp);
  if (np > maxp) np = maxp;
  hist_ctx c = {idx, bins, m, r};
  run_tasks(threads, np, hist_task, &c);
  free(r);
}

/* stencilOp (NAT/Execute.hs:657-765; NAT/CodeGen/Stencil.hs:60-197): rows
 * split over workers; interior points without boundary checks. */
typedef struct { int ry, rx, b; const float* w; const float* in; float* out; int64_t h, wd; const int64_t* r; } st_ctx;
static void st_task(void* c, int64_t i) {
  st_ctx* x = (st_ctx*)c;
  const int ry = x->ry, rx = x->rx, kw = 2 * rx + 1;
  for (int64_t y = x->r[2 * i]; y < x->r[2 * i + 1]; ++y) {
    const int yin = y >= ry && y < x->h - ry;
    for (int64_t xx = 0; xx < x->wd; ++xx) {
      float acc = 0;
      if (yin && xx >= rx && xx < x->wd - rx) {
        for (int dy = -ry; dy <= ry; ++dy) {
          const float* row = x->in + (y + dy) * x->wd + xx;
          for (int dx = -rx; dx <= rx; ++dx) acc += x->w[(dy + ry) * kw + dx + rx] * row[dx];
        }
      } else {
        for (int dy = -ry; dy <= ry; ++dy)
          for (int dx = -rx; dx <= rx; ++dx)
            acc += x->w[(dy + ry) * kw + dx + rx] *
                   x->in[bmap(x->b, y + dy, x->h) * x->wd + bmap(x->b, xx + dx, x->wd)];
      }
      x->out[y * x->wd + xx] = acc;
    }
  }
}

void or_native_stencil2d_f32(int threads, int ry, int rx, const float* w, int boundary,
                             const float* in, int64_t h, int64_t wd, float* out) {
  int64_t maxp = 4 * (int64_t)threads + 8;
  int64_t* r = (int64_t*)malloc(sizeof(int64_t) * 2 * (size_t)maxp);
  int64_t np = or_divide_work1(threads - 1, 64, 0, h, r, maxp);
  if (np > maxp) np = maxp;
  st_ctx c = {ry, rx, boundary, w, in, out, h, wd, r};
  run_tasks(threads, np, st_task, &c);
  free(r);
}

/* ------------------------------------------------------- build */
#ifndef OR_MARCH
#define OR_MARCH "default"
#endif
/* compiler and target of this build (reported beside the CPU baseline) */
const char* or_build_info(void) { return "compiler " __VERSION__ ", -O3 -march=" OR_MARCH; }

/* ------------------------------------------------------- generators */
typedef struct { uint64_t s[4]; } xo;
static uint64_t splitmix64(uint64_t* x) {
  uint64_t z = (*x += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
static void xo_seed(xo* r, uint64_t seed) {
  uint64_t x = seed;
  for (int i = 0; i < 4; ++i) r->s[i] = splitmix64(&x);
}
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static inline uint64_t xo_next(xo* r) {
  uint64_t* s = r->s;
  const uint64_t res = rotl(s[1] * 5, 7) * 9;
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
  s[2] ^= t; s[3] = rotl(s[3], 45);
  return res;
}
static inline double xo_unit(xo* r) { return (double)(xo_next(r) >> 11) * 0x1.0p-53; } /* [0,1) */

void or_gen_uniform_i32(uint64_t seed, int32_t lo, int32_t hi, int32_t* out, int64_t n) {
  xo r; xo_seed(&r, seed);
  const uint64_t span = (uint64_t)((int64_t)hi - (int64_t)lo + 1);
  for (int64_t i = 0; i < n; ++i) out[i] = (int32_t)((int64_t)lo + (int64_t)(xo_next(&r) % span));
}

void or_gen_uniform_f32(uint64_t seed, float lo, float hi, float* out, int64_t n) {
  xo r; xo_seed(&r, seed);
  for (int64_t i = 0; i < n; ++i) out[i] = lo + (hi - lo) * (float)((xo_next(&r) >> 40) * 0x1.0p-24);
}

void or_gen_normal_f32(uint64_t seed, float* out, int64_t n) {
  xo r; xo_seed(&r, seed);
  for (int64_t i = 0; i < n; i += 2) {
    const double u1 = 1.0 - xo_unit(&r), u2 = xo_unit(&r);
    const double m = sqrt(-2.0 * log(u1));
    out[i] = (float)(m * cos(2.0 * M_PI * u2));
    if (i + 1 < n) out[i + 1] = (float)(m * sin(2.0 * M_PI * u2));
  }
}

void or_gen_zipf_i32(uint64_t seed, double s, int32_t nbins, int32_t* out, int64_t n) {
  double* cdf = (double*)malloc(sizeof(double) * (size_t)nbins);
  double acc = 0;
  for (int32_t k = 0; k < nbins; ++k) { acc += pow((double)(k + 1), -s); cdf[k] = acc; }
  for (int32_t k = 0; k < nbins; ++k) cdf[k] /= acc;
  xo r; xo_seed(&r, seed);
  for (int64_t i = 0; i < n; ++i) {
    const double u = xo_unit(&r);
    int32_t a = 0, b = nbins - 1;
    while (a < b) { int32_t m = (a + b) / 2; if (cdf[m] > u) b = m; else a = m + 1; }
    out[i] = a;
  }
  free(cdf);
}

void or_gen_segment_lengths(uint64_t seed, int64_t nseg, int64_t maxlen, int64_t total, int64_t* out) {
  xo r; xo_seed(&r, seed);
  double* l = (double*)malloc(sizeof(double) * (size_t)nseg);
  double sum = 0;
  for (int64_t i = 0; i < nseg; ++i) { l[i] = exp(xo_unit(&r) * log((double)maxlen)); sum += l[i]; }
  int64_t tot = 0, big = 0;
  for (int64_t i = 0; i < nseg; ++i) {
    int64_t v = (int64_t)llround(l[i] * (double)total / sum);
    if (v < 1) v = 1;
    out[i] = v;
    tot += v;
    if (out[i] > out[big]) big = i;
  }
  if (tot <= total) {
    out[big] += total - tot;
  } else {
    // rounding and the floor of 1 oversupplied: take the excess off the
    // segments in index order, never below 1 (needs nseg <= total)
    int64_t excess = tot - total;
    for (int64_t i = 0; i < nseg && excess > 0; ++i) {
      const int64_t d = out[i] - 1 < excess ? out[i] - 1 : excess;
      out[i] -= d;
      excess -= d;
    }
  }
  free(l);
}
