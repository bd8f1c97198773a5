/*
 * accelerate_hip.h - C ABI of the MI355X (gfx950) backend for Accelerate.
 *
 * This is the whole boundary between a host program (the Haskell backend
 * package, a test, the serving loop) and the GPU.  It replaces every
 * `Foreign.CUDA.*` call site of accelerate-llvm-ptx and adds the skeleton
 * entry points that the `instance Execute` methods call.  Paths in the
 * citations are relative to the accelerate-llvm repository:
 *   PTX/  = accelerate-llvm-ptx/src/Data/Array/Accelerate/LLVM/PTX/
 *   LLVM/ = accelerate-llvm/src/Data/Array/Accelerate/LLVM/
 *   NAT/  = accelerate-llvm-native/src/Data/Array/Accelerate/LLVM/Native/
 *
 * Conventions
 *   - Every function returns an acc_status (0 = success).  Nothing throws
 *     across the ABI.  acc_hip_last_error() returns a thread-local message for
 *     the last failure on the calling thread.
 *   - Out of memory is ACC_ERROR_OUT_OF_MEMORY, distinct from other errors, so
 *     a caller can evict and retry (PTX/Array/Remote.hs:63-74).
 *   - Every call names its device (or a stream that belongs to one).  There is
 *     no thread-bound "current context": calls are thread-safe and may come
 *     from any OS thread (unlike PTX/Context.hs:135-139).
 *   - Skeleton calls are asynchronous on the given stream.  Array arguments are
 *     device pointers to struct-of-arrays fields (one buffer per scalar field,
 *     LLVM/Execute/Marshal.hs:71-100); extents are 64-bit Accelerate `Int`s.
 *   - Calls on one stream execute in order.  Two skeleton calls that run
 *     concurrently must use different streams: each stream owns the small
 *     device workspace (tile status words, tickets) its skeletons use.
 */
#ifndef ACCELERATE_HIP_H
#define ACCELERATE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ACC_HIP_ABI_VERSION 1

typedef enum acc_status {
  ACC_SUCCESS = 0,
  ACC_ERROR_INVALID_VALUE = 1,
  ACC_ERROR_OUT_OF_MEMORY = 2,
  ACC_ERROR_HIP = 3,
  ACC_ERROR_UNSUPPORTED = 4,
  ACC_ERROR_EMPTY_ARRAY = 5, /* fold1 / scan-less reductions of an empty array (PTX/Execute.hs:232) */
  ACC_ERROR_NOT_READY = 6,
  ACC_ERROR_NO_DEVICE = 7
} acc_status;

/* Scalar element types (Accelerate SingleType tags). */
typedef enum acc_type {
  ACC_INT32 = 0,
  ACC_INT64 = 1,
  ACC_WORD32 = 2,
  ACC_WORD64 = 3,
  ACC_FLOAT32 = 4,
  ACC_FLOAT64 = 5
} acc_type;

/* Combine functions available without run-time compilation.  They are the
 * functions the reference recognises in permute (LLVM/CodeGen/Permute.hs:117-145)
 * and the usual fold/scan operators.  Floating min/max are fminf/fmaxf
 * (LLVM/CodeGen/Arithmetic.hs:456-466); integer arithmetic wraps. */
typedef enum acc_op {
  ACC_OP_ADD = 0,
  ACC_OP_MUL = 1,
  ACC_OP_MIN = 2,
  ACC_OP_MAX = 3,
  ACC_OP_AND = 4, /* integral types only */
  ACC_OP_OR = 5,
  ACC_OP_XOR = 6,
  ACC_OP_SUB = 7,     /* permute only: old - new   (atomicrmw sub) */
  ACC_OP_EXCHANGE = 8 /* permute only: new         (permute const)  */
} acc_op;

typedef enum acc_direction { ACC_LEFT_TO_RIGHT = 0, ACC_RIGHT_TO_LEFT = 1 } acc_direction;

/* Stencil boundary conditions (LLVM/CodeGen/Stencil.hs:44-49, index maps :224-253).
 * ACC_BOUNDARY_FUNCTION is the Function boundary (an index outside the array
 * reads f(index), :215-218) with the affine functions of the index as its
 * function set: the `constant` argument then points to three values c0, cy, cx
 * of the element type and a 2-D index (y, x) outside reads
 * (c0 + cy*y) + cx*x, computed in the element type (y and x converted to it;
 * integers wrap; floating operations round one by one, no fused multiply-add).
 * The 2-D stencils take it; the 3-D stencil does not (ACC_ERROR_UNSUPPORTED). */
typedef enum acc_boundary {
  ACC_BOUNDARY_CLAMP = 0,
  ACC_BOUNDARY_MIRROR = 1,
  ACC_BOUNDARY_WRAP = 2,
  ACC_BOUNDARY_CONSTANT = 3,
  ACC_BOUNDARY_FUNCTION = 4
} acc_boundary;

typedef struct acc_stream_s* acc_stream_t;
typedef struct acc_event_s* acc_event_t;
typedef struct acc_module_s* acc_module_t;
typedef struct acc_graph_s* acc_graph_t;
typedef void* acc_function_t;

typedef struct acc_device_props {
  char name[64];
  char arch[32];            /* e.g. "gfx950" */
  int compute_units;
  int wave_size;
  int max_threads_per_block;
  size_t lds_per_block;
  size_t total_global_mem;
  int clock_khz;
  int memory_clock_khz;
  int memory_bus_width;
  int l2_cache_bytes;
} acc_device_props;

/* ---------------------------------------------------------------- errors */
const char* acc_hip_last_error(void);
int acc_hip_abi_version(void);

/* ----------------------------------------------------------------- device
 * Replaces CUDA.count/device/props (PTX/State.hs:145-160) and deviceInfo
 * (PTX/Context.hs:189-213).  Initialisation is implicit and per device. */
int acc_hip_device_count(int* count);
int acc_hip_device_props(int device, acc_device_props* props);
int acc_hip_device_synchronize(int device);

/* ----------------------------------------------------------------- memory
 * Replaces CUDA.mallocArray/free/getMemInfo (PTX/Array/Remote.hs:63-100,
 * PTX/Array/Table.hs:45-49).  The allocation limit is a test knob: with a
 * limit set, an allocation that would take the library's live device bytes
 * above it fails with ACC_ERROR_OUT_OF_MEMORY exactly as a full device does. */
int acc_hip_malloc(int device, size_t bytes, void** ptr);
int acc_hip_free(int device, void* ptr);
int acc_hip_mem_get_info(int device, size_t* free_bytes, size_t* total_bytes);
int acc_hip_set_alloc_limit(int device, size_t limit_bytes); /* 0 = no limit */
int acc_hip_allocated_bytes(int device, size_t* bytes);
/* pinned host memory: mallocHostForeignPtr (accelerate-llvm-ptx PTX.hs:548-553) */
int acc_hip_host_alloc(size_t bytes, void** ptr);
int acc_hip_host_free(void* ptr);

/* Transfers: pokeArrayAsync, peekArrayAsync, copyArrayAsync, memsetAsync
 * (PTX/Array/Prim.hs:125,159,188,217,295). */
int acc_hip_memcpy_h2d_async(acc_stream_t s, void* dst, const void* src, size_t bytes);
int acc_hip_memcpy_d2h_async(acc_stream_t s, void* dst, const void* src, size_t bytes);
int acc_hip_memcpy_d2d_async(acc_stream_t s, void* dst, const void* src, size_t bytes);
/* fill `count` elements of `elem_bytes` (1, 2, 4 or 8) with the bit pattern at `value` */
int acc_hip_fill_async(acc_stream_t s, void* dst, const void* value, size_t elem_bytes, size_t count);

/* ---------------------------------------------------------------- streams
 * Stream.create / Stream.finished (PTX/Execute/Stream.hs:125,
 * PTX/Execute/Stream/Reservoir.hs:73).  Streams are created at the default
 * priority. */
int acc_hip_stream_create(int device, acc_stream_t* s);
int acc_hip_stream_destroy(acc_stream_t s);
int acc_hip_stream_synchronize(acc_stream_t s);
int acc_hip_stream_query(acc_stream_t s); /* ACC_SUCCESS = idle, ACC_ERROR_NOT_READY = busy */
int acc_hip_stream_device(acc_stream_t s, int* device);
/* The HIP stream under `s` (a hipStream_t), for foreign array functions
 * (ForeignAcc, PTX/Foreign.hs:74-84: a host function given the backend's
 * stream and its arrays' device pointers): work a foreign function enqueues
 * there is ordered with the skeletons on `s`, and an event recorded on `s`
 * after it is the foreign result's future. */
int acc_hip_stream_native(acc_stream_t s, void** hip_stream);
/* Pre-size the stream's skeleton workspace (bytes of tile-status words, and as
 * many bytes of untagged per-tile scratch), so that no allocation happens
 * later, e.g. while the stream is being captured. */
int acc_hip_stream_reserve(acc_stream_t s, size_t status_bytes);

/* ----------------------------------------------------------------- events
 * Event.create/record/wait/query/block/elapsedTime
 * (PTX/Execute/Event.hs:65,117,129,138,144; PTX/Debug.hs:62-83). */
int acc_hip_event_create(int device, int timing, acc_event_t* e);
int acc_hip_event_destroy(acc_event_t e);
int acc_hip_event_record(acc_event_t e, acc_stream_t s);
int acc_hip_stream_wait_event(acc_stream_t s, acc_event_t e);
int acc_hip_event_query(acc_event_t e); /* ACC_SUCCESS = complete, ACC_ERROR_NOT_READY */
int acc_hip_event_synchronize(acc_event_t e);
int acc_hip_event_elapsed_ms(acc_event_t start, acc_event_t stop, float* ms);

/* ---------------------------------------------------------- code objects
 * loadDataFromPtrEx, unload, getFun, requires (PTX/Link.hs:64,75,101-106;
 * PTX/Embed.hs:55,77) and CUDA.launchKernel (PTX/Execute.hs:864). */
int acc_hip_module_load_data(int device, const void* image, size_t bytes, acc_module_t* m);
int acc_hip_module_unload(acc_module_t m);
int acc_hip_module_get_function(acc_module_t m, const char* name, acc_function_t* f);
/* registers, static LDS bytes and max threads per block of a function */
int acc_hip_function_attributes(acc_function_t f, int* num_regs, int* static_lds_bytes,
                                int* max_threads);
int acc_hip_launch(acc_function_t f, unsigned gx, unsigned gy, unsigned gz, unsigned bx,
                   unsigned by, unsigned bz, unsigned dynamic_lds_bytes, acc_stream_t s,
                   void** args);

/* ----------------------------------------------------------------- graphs
 * Stream capture into a hipGraph (new: the batched form).  A graph captured on
 * stream s replays only on s (launch on another stream is
 * ACC_ERROR_INVALID_VALUE): its skeleton kernels use s's tickets and status
 * words.  The graph keeps that workspace alive until acc_hip_graph_destroy,
 * also when s is destroyed first. */
int acc_hip_graph_begin_capture(acc_stream_t s);
int acc_hip_graph_end_capture(acc_stream_t s, acc_graph_t* g);
int acc_hip_graph_launch(acc_graph_t g, acc_stream_t s);
int acc_hip_graph_destroy(acc_graph_t g);

/* ============================================================== skeletons
 * One entry point per `class Execute` method on the hot path
 * (LLVM/Execute.hs:58-163), reproducing the dispatch of `instance Execute PTX`
 * (PTX/Execute.hs:83-109): output sizes, empty-array cases and clone versus
 * in-place.  `seed` points to one host element of the array's type, or is NULL
 * for the seedless variant (fold1, scanl1, ...).  All array pointers are
 * device pointers.  An array of shape (sh :. n) is passed as outer = size sh,
 * inner = n, row-major. */

/* fold / fold1 (LLVM/Execute.hs:87-92; PTX/Execute.hs:222-340).
 * out[r] = seed (+) (in[r][0] (+) ... (+) in[r][inner-1]) for r < outer; the
 * seed is combined on the left.  inner == 0 with a seed fills `out` with the
 * seed; without a seed it is ACC_ERROR_EMPTY_ARRAY. */
int acc_hip_fold(acc_stream_t s, acc_type t, acc_op op, const void* seed, const void* in,
                 int64_t outer, int64_t inner, void* out);

/* foldSeg / fold1Seg (LLVM/Execute.hs:94-102; PTX/Execute.hs:343-374).
 * `offsets` (nseg + 1 values of type ACC_INT32 or ACC_INT64) are the
 * `scanl (+) 0` of the segment lengths, relative to the start of each row.
 * out[r][s] = seed (+) fold over in[r][offsets[s] .. offsets[s+1]).  An empty
 * segment yields the seed; fold1Seg requires non-empty segments. */
int acc_hip_fold_seg(acc_stream_t s, acc_type t, acc_op op, const void* seed, const void* in,
                     int64_t outer, int64_t inner, acc_type offset_type, const void* offsets,
                     int64_t nseg, void* out);

/* Segment tile plans: foldSeg over one offsets array, planned once.  No
 * counterpart in the reference's FFI: its foldSeg maps segments to blocks
 * again at every launch (PTX/Execute.hs:343-374).  Accelerate arrays are
 * immutable, so a program that folds several arrays over one segment
 * descriptor (the `scanl (+) 0` of the lengths shared by its foldSegs) can
 * build the plan when the offsets are produced and keep it with them.
 * acc_hip_seg_plan_build enqueues the plan's computation on `s` for `nseg`
 * segments over rows of `n` elements of type `t` whose first row starts at
 * `in` (the plan depends on the element type and on `in` modulo 16 bytes).
 * acc_hip_fold_seg_planned is acc_hip_fold_seg over `outer` such rows, with
 * the offsets, `n` and `nseg` of the plan; it refuses a type or an alignment
 * other than the plan's (with outer > 1, rows must be 16-byte multiples).
 * The offsets must not change while the plan is used, and a fold on another
 * stream than the build's must be ordered after it (an event).  A build that
 * needs a larger plan synchronises the device before freeing the old one.  A
 * plan is not locked: no thread may build it while another uses it. */
typedef struct acc_segplan_s* acc_segplan_t;
int acc_hip_seg_plan_create(int device, acc_segplan_t* plan);
int acc_hip_seg_plan_build(acc_stream_t s, acc_segplan_t plan, acc_type t, const void* in, int64_t n,
                           acc_type offset_type, const void* offsets, int64_t nseg);
/* The same plan from the segment lengths, with the offsets: writes
 * offsets[0 .. nseg] = scanl (+) 0 lengths (Int32 or Int64, wrapping like the
 * scan) and plans over them, in one launch: the frontend's
 * `foldSeg' f z xs (scanl (+) 0 seg)` producer and the plan together.  The
 * plan then refers to `offsets`. */
int acc_hip_seg_plan_build_lengths(acc_stream_t s, acc_segplan_t plan, acc_type t, const void* in, int64_t n,
                                   acc_type len_type, const void* lengths, int64_t nseg, void* offsets);
int acc_hip_fold_seg_planned(acc_stream_t s, acc_segplan_t plan, acc_type t, acc_op op, const void* seed,
                             const void* in, int64_t outer, void* out);
/* Several foldSegs (rows of one, outer = 1) over one plan in one call: fold i
 * combines ins[i] with ops[i] (seeds[i] or NULL; seeds itself may be NULL)
 * into outs[i], 1 <= count <= 8, every input of the plan's type and
 * alignment.  The tile kernels run one after the other (two Float32 folds,
 * each ACC_OP_ADD or ACC_OP_MAX: both folds' tiles in one launch) and the
 * fix-ups of all the folds share one launch; every result is bit-identical to
 * the fold's own acc_hip_fold_seg_planned call. */
int acc_hip_fold_seg_planned_multi(acc_stream_t s, acc_segplan_t plan, acc_type t, int count, const acc_op* ops,
                                   const void* const* seeds, const void* const* ins, void* const* outs);
int acc_hip_seg_plan_destroy(acc_segplan_t plan);

/* Allocation-free foldSeg / fold1Seg.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * semantics and kernels as acc_hip_fold_seg, but the skeleton workspace and
 * scratch are a fixed 16 MiB device array inside the library (one per device,
 * held until the process ends) instead of the stream's allocations, so the
 * call allocates and frees nothing.  A call whose records would not fit
 * (about 2^28.8 Float32 or 2^27.8 Float64 elements per row) is refused with
 * ACC_ERROR_INVALID_VALUE before anything is launched.  All calls on one
 * device share that device's array: calls that overlap on one device (other
 * streams, other threads) race and must be serialised by the caller. */
int acc_hip_fold_seg_noalloc(acc_stream_t s, acc_type t, acc_op op, const void* seed, const void* in,
                             int64_t outer, int64_t inner, acc_type offset_type, const void* offsets,
                             int64_t nseg, void* out);

/* scanl1/scanr1 (seed == NULL, out has inner elements per row) and
 * scanl/scanr (seed != NULL, out has inner + 1 elements per row)
 * (LLVM/Execute.hs:104-111; PTX/Execute.hs:377-488; element order
 * NAT/CodeGen/Scan.hs:136-210). */
int acc_hip_scan(acc_stream_t s, acc_type t, acc_op op, acc_direction dir, const void* seed,
                 const void* in, int64_t outer, int64_t inner, void* out);

/* scanl'/scanr' (LLVM/Execute.hs:113-119; PTX/Execute.hs:491-594;
 * NAT/CodeGen/Scan.hs:213-270): out has inner elements per row (the exclusive
 * prefixes), sums has outer elements (the totals). */
int acc_hip_scan_prime(acc_stream_t s, acc_type t, acc_op op, acc_direction dir,
                       const void* seed, const void* in, int64_t outer, int64_t inner,
                       void* out, void* sums);

/* permute (LLVM/Execute.hs:121-129; PTX/Execute.hs:597-649;
 * PTX/CodeGen/Permute.hs:117-258).
 *   out[index[i]] = combine(src[i], out[index[i]])  for i < n with
 *   0 <= index[i] < m; other indices are `Nothing` (the element is dropped).
 * If out != defaults, defaults is first copied to out (clone); if they are
 * equal the update is in place.  src == NULL means a delayed constant source
 * (`fill sh v`) whose value is *src_fill (the histogram case costs no source
 * bytes).  index_type is ACC_INT32 or ACC_INT64. */
int acc_hip_permute(acc_stream_t s, acc_type t, acc_op combine, const void* defaults, void* out,
                    int64_t m, const void* src, const void* src_fill, acc_type index_type,
                    const void* index, int64_t n);

/* stencil1 on a DIM2 array (LLVM/Execute.hs:131-138; PTX/Execute.hs:652-776)
 * with the stencil function f(tuple) = sum over taps of w[dy][dx] * x[y+dy][x+dx],
 * taps in row-major order (LLVM/CodeGen/Stencil.hs:126-180), dy in [-ry, ry],
 * dx in [-rx, rx]; `weights` is a host array of (2ry+1)*(2rx+1) elements of
 * type t.  Radii 1..4 (3/5/7/9-point patterns).  `constant` is used with
 * ACC_BOUNDARY_CONSTANT (one value) and ACC_BOUNDARY_FUNCTION (c0, cy, cx). */
int acc_hip_stencil2d(acc_stream_t s, acc_type t, int ry, int rx, const void* weights,
                      acc_boundary boundary, const void* constant, const void* in,
                      int64_t height, int64_t width, void* out);

/* stencil2 on DIM2 arrays (LLVM/Execute.hs:147-156, called at :326; stencil2Op,
 * PTX/Execute.hs:669-684; NAT/CodeGen/Stencil.hs:75-107) with the stencil function
 * f(s1, s2) = sum over taps of w1[dy][dx] * in1[y+dy][x+dx] + sum over taps of
 * w2[dy][dx] * in2[y+dy][x+dx], taps in row-major order, in1's before in2's.  The
 * output is height x width with height = min(height1, height2) and width =
 * min(width1, width2) (the intersection, PTX/Execute.hs:682); each input is read
 * through its own boundary against its own extents.  Both stencils have radii
 * ry, rx in 1..2 (write a smaller one with zero weights); weights1/weights2 are
 * host arrays of (2ry+1)*(2rx+1) elements of type t.  The output must not alias
 * an input. */
int acc_hip_stencil2_2d(acc_stream_t s, acc_type t, int ry, int rx, const void* weights1,
                        acc_boundary boundary1, const void* constant1, const void* in1,
                        int64_t height1, int64_t width1, const void* weights2,
                        acc_boundary boundary2, const void* constant2, const void* in2,
                        int64_t height2, int64_t width2, void* out);

/* stencil1 on a DIM3 array (LLVM/Execute.hs:131-138; PTX/Execute.hs:652-776) with
 * f = sum over taps of w[dz][dy][dx] * x[z+dz][y+dy][x+dx], taps in tuple order
 * (planes, rows, columns), dz in [-rz, rz], dy in [-ry, ry], dx in [-rx, rx];
 * `weights` is a host array of (2rz+1)*(2ry+1)*(2rx+1) elements of type t.  Radii
 * 1..2 (3- and 5-point extents per axis); the array is depth x height x width,
 * row-major.  Boundaries as for acc_hip_stencil2d, on every axis. */
int acc_hip_stencil3d(acc_stream_t s, acc_type t, int rz, int ry, int rx, const void* weights,
                      acc_boundary boundary, const void* constant, const void* in, int64_t depth,
                      int64_t height, int64_t width, void* out);

/* ================================================================ producers
 * map, zipWith, generate, transform and backpermute (LLVM/Execute.hs:310-313;
 * launched by mapOp, generateOp, transformOp and backpermuteOp,
 * PTX/Execute.hs:117-206; kernels PTX/CodeGen/{Map,Generate,Transform}.hs).
 * The reference compiles the program's own element and index functions into
 * these kernels; here they come from the fixed sets below.  Integer
 * arithmetic wraps; floating a*x + b and a + i*b round after each step (no
 * fused multiply-add).  All arrays are device pointers, linear (an array of
 * any shape is its row-major elements). */

/* Element functions of map and transform. */
typedef enum acc_unop {
  ACC_UNOP_ID = 0,     /* x */
  ACC_UNOP_NEG = 1,    /* negate x = x * (-1)             (LLVM/CodeGen/Arithmetic.hs:69-73) */
  ACC_UNOP_ABS = 2,    /* fabs / llvm.abs, abs minBound == minBound, identity if unsigned (:75-86) */
  ACC_UNOP_AFFINE = 3, /* a * x + b; params: two host elements {a, b} of the array's type */
  ACC_UNOP_CAST = 4    /* C conversion of the input type to the output type (map only):
                          integers wrap, floating to integer truncates (in-range values) */
} acc_unop;

/* Generators. */
typedef enum acc_genop {
  ACC_GEN_FILL = 0, /* out[i] = a;          params: {a} */
  ACC_GEN_IOTA = 1  /* out[i] = a + i * b;  params: {a, b} (enumFromStepN) */
} acc_genop;

/* Index maps of transform / backpermute: out[i] = f(in[p(i)]). */
typedef enum acc_ixmap {
  ACC_IX_GATHER = 0,   /* p(i) = index[i], a linear index into in; an index outside
                          [0, in_outer * in_inner) gives 0 and reads nothing */
  ACC_IX_REVERSE = 1,  /* p(i) = n - 1 - i */
  ACC_IX_TRANSPOSE = 2 /* in is in_outer x in_inner, out is in_inner x in_outer:
                          out[c][r] = f(in[r][c]) */
} acc_ixmap;

/* map (PTX/Execute.hs:140-162, PTX/CodeGen/Map.hs:37-63): out[i] = f(in[i]),
 * i < n.  t_in == t_out except for ACC_UNOP_CAST.  in == out maps in place
 * (the reference's in-place case, PTX/Execute.hs:151-158). */
int acc_hip_map(acc_stream_t s, acc_type t_in, acc_type t_out, acc_unop f, const void* params,
                const void* in, int64_t n, void* out);

/* zipWith f a b, which the frontend delivers as a generate reading both arrays
 * (PTX/Execute.hs:164-173, PTX/CodeGen/Generate.hs:37-62): out[i] = f(a[i], b[i])
 * for f in ADD MUL MIN MAX AND OR XOR (as in fold) and SUB (a[i] - b[i]). */
int acc_hip_zipwith(acc_stream_t s, acc_type t, acc_op f, const void* a, const void* b, int64_t n,
                    void* out);

/* generate (PTX/Execute.hs:164-173, PTX/CodeGen/Generate.hs:37-62). */
int acc_hip_generate(acc_stream_t s, acc_type t, acc_genop g, const void* params, int64_t n,
                     void* out);

/* transform and backpermute (PTX/Execute.hs:175-206, PTX/CodeGen/Transform.hs:37-67;
 * backpermute is f = ACC_UNOP_ID).  f in {ID, NEG, ABS, AFFINE}.  n is the
 * output's element count (= in_outer * in_inner for REVERSE and TRANSPOSE).
 * index_type (ACC_INT32 / ACC_INT64) and index are used by ACC_IX_GATHER. */
int acc_hip_transform(acc_stream_t s, acc_type t, acc_ixmap p, acc_unop f, const void* params,
                      const void* in, int64_t in_outer, int64_t in_inner, acc_type index_type,
                      const void* index, int64_t n, void* out);

/* ================================================== run-time compiled skeletons
 * Skeletons whose scalar functions come from the program itself, compiled
 * at run time: the reference's compile / link / execute of one skeleton node
 * (compileForTarget PTX/Compile.hs:84-165 with the functions generated by
 * llvmOfFun1/2, LLVM/CodeGen/Exp.hs:65-80; link PTX/Link.hs:57-121; the
 * Execute methods PTX/Execute.hs:222-776).  A node is an S-expression:
 *
 *   (fold     (elt T..) (fun (a b) E) [(seed E)] [(input (fun (i) E))] [(free T..)])
 *   (scan     (elt T..) (fun (a b) E) [(dir l2r|r2l)] [(seed E)] [prime] [(input ..)] [(free ..)])
 *   (foldseg  (elt T..) (fun (a b) E) [(offsets i32|i64)] [(seed E)] [(input ..)] [(free ..)])
 *   (permute  (elt T..) (fun (new old) E) (project (fun (i) E)) [(input ..)] [(free ..)])
 *   (stencil  (in T) (out T) (radius RY RX) (boundary B) (fun () E) [(free ..)])
 *   (generate (elt T..) (fun (i) E) | (fun (r c) E) [(free ..)])
 *
 * T is i8 i16 i32 i64 u8 u16 u32 u64 f32 f64 bool (Int = i64, Word = u64,
 * Bool stored as one byte); an element has 1..4 fields, each its own array
 * (struct of arrays).  E is an expression:
 *   x | (const T lit) | (let ((x E) ..) E) | (cond E E E) | (tuple E ..) | (prj k E)
 *   | (read k E)   element E (an i64) of free array k, clamped into the array
 *   | (nb dy dx)   stencil functions: the neighbour at offset (dy, dx)
 *   | (foreign "BODY" (fun (x ..) E) E ..)   a foreign scalar function
 *                  (ForeignExp, PTX/Foreign.hs:74-84): BODY is the C++ body of
 *                  a device function of the parameters x .. (typed by the
 *                  arguments), compiled into the kernel; the fallback (fun ..)
 *                  is its meaning where no foreign code runs and gives its
 *                  result type (numeric scalars; \" and \\ escape in BODY)
 *   | (op E ..)    op one of the PrimFun primitives (LLVM/CodeGen/Exp.hs:244-310):
 *       add sub mul neg abs signum quot rem quotrem div mod divmod band bor bxor
 *       bnot shiftl shiftr rotatel rotater popcount clz ctz fdiv recip sin cos tan
 *       asin acos atan sinh cosh tanh asinh acosh atanh exp sqrt log fpow logbase
 *       atan2 isnan isinfinite lt gt lte gte eq neq min max land lor lnot,
 *       (truncate T E) (round T E) (floor T E) (ceiling T E)
 *       (fromintegral T E) (tofloating T E) (bitcast T E)
 *   with the reference's semantics (LLVM/CodeGen/Arithmetic.hs: wrapping
 *   integers, fminf/fmaxf, ordered float comparisons, round half away from
 *   zero, no contraction of floating operations).
 * B is clamp | mirror | wrap | (constant E) | (function (fun (y x) E)).
 *
 * Semantics per skeleton, with acc_kernel_args (outer x inner input, row-major):
 *   fold:    out[r] = seed (+) in[r][0] (+) .. (+) in[r][inner-1]  (in element order:
 *            the combine need only be associative); fold1 of an empty row is
 *            ACC_ERROR_EMPTY_ARRAY, fold of one writes the seed.
 *   scan:    l2r without seed scanl1 (inner outputs per row); with seed scanl
 *            (inner + 1: exclusive values, the total last); r2l scanr1 / scanr
 *            (inclusive values, the seed last); prime: scanl'/scanr' (inner
 *            exclusive values, the totals in sums[r]).
 *   foldseg: out[r*nseg + s] = seed (+) fold of in[r][offsets[s] .. offsets[s+1]);
 *            empty segments get the seed (fold1Seg leaves them unwritten).
 *   permute: out = defaults (copied unless defaults[0] == out[0] or NULL), then
 *            out[j] = f(src[i], out[j]) for every source index i < outer*inner
 *            whose projection j = project(i) lies in [0, nseg); other
 *            projections are Nothing.  f must be associative and commutative.
 *            A one-field element of 4 or 8 bytes is updated by compare-and-swap;
 *            others by a lock per output slot (PTX/CodeGen/Permute.hs:219-258).
 *   stencil: out[y][x] = f(neighbourhood) on a height (outer) x width (inner) array.
 *   generate: out[i] = f(i) (or f(i / inner, i mod inner)) for i < outer*inner.
 * A manifest input is read from in[field]; (input (fun (i) E)) computes it
 * instead (a delayed producer fused into the consumer, LLVM/CodeGen/Sugar.hs:70-76).
 *
 * Code objects are cached on disk by a hash of the generated source, the
 * compiler options and the library version (LLVM/Compile/Cache.hs:57-91):
 * $ACC_HIP_CACHE_DIR, else $XDG_CACHE_HOME/accelerate-hip, else
 * ~/.cache/accelerate-hip.  acc_hip_kernel_image / _from_image carry a
 * compiled node inside a program, which then loads it without compiling
 * (runQ embedding, PTX/Embed.hs:50-79). */
typedef struct acc_kernel_s* acc_kernel_t;

typedef enum acc_kernel_origin {
  ACC_KERNEL_COMPILED = 0,         /* compiled by this call (and stored in the cache) */
  ACC_KERNEL_FROM_DISK_CACHE = 1,  /* code object read from the on-disk cache */
  ACC_KERNEL_FROM_IMAGE = 2,       /* from an image given by the caller */
  ACC_KERNEL_LOADED = 3            /* the module was already loaded in this process */
} acc_kernel_origin;

typedef struct acc_kernel_info {
  int origin;              /* acc_kernel_origin */
  int lock_free_permute;   /* permute: 1 compare-and-swap, 0 lock per slot; other skeletons -1 */
  double compile_ms;       /* time spent compiling (0 unless origin == ACC_KERNEL_COMPILED) */
  char key[40];            /* 32 hex digits identifying the code object */
} acc_kernel_info;

typedef struct acc_kernel_args {
  const void* in[4];        /* manifest input fields */
  void* out[4];             /* output fields */
  const void* defaults[4];  /* permute: default fields */
  void* sums[4];            /* scan': totals */
  const void* offsets;      /* foldseg: nseg + 1 offsets per row */
  int64_t outer, inner;     /* input extents */
  int64_t nseg;             /* foldseg: segments; permute: output size */
  int32_t nfree;            /* free arrays given below */
  int32_t reserved;
  const void* free_ptr[8];  /* free array k (device) */
  int64_t free_len[8];      /* and its element count */
} acc_kernel_args;

/* The HIP source a node translates to (*len = its length; copied when buf is given). */
int acc_hip_node_source(const char* node, char* buf, size_t cap, size_t* len);
/* Compile a node into the on-disk cache without a device (*origin: COMPILED or FROM_DISK_CACHE). */
int acc_hip_node_precompile(const char* node, int* origin);
/* Compile (or fetch from the cache) and load a node on a device. */
int acc_hip_kernel_compile(int device, const char* node, acc_kernel_t* k);
/* Load a node from an image made by acc_hip_kernel_image (no compilation; the
 * image must have been made for the same node and library version). */
int acc_hip_kernel_from_image(int device, const char* node, const void* image, size_t bytes, acc_kernel_t* k);
/* The node's image (*bytes = size; copied when buf is given and cap suffices). */
int acc_hip_kernel_image(acc_kernel_t k, void* buf, size_t cap, size_t* bytes);
int acc_hip_kernel_info(acc_kernel_t k, acc_kernel_info* info);
/* Execute the node on a stream (asynchronous, like every skeleton call). */
int acc_hip_kernel_run(acc_kernel_t k, acc_stream_t s, const acc_kernel_args* args);
int acc_hip_kernel_destroy(acc_kernel_t k);

/* ===================================================== remote memory table
 * Host arrays, named by a key the host chooses (its array's identity), mapped
 * to device buffers: the reference's per-context memory table
 * (PTX/Array/Table.hs:35-49 over the core's LRU table; instance Remote,
 * PTX/Array/Data.hs:58-152).
 *   use   : upload a host array (if not resident) and freeze it: the host copy
 *           stays authoritative, so the device copy may be evicted and
 *           uploaded again later (useRemote, PTX/Array/Prim.hs:89-100).  The
 *           host buffer must stay valid while the key is in the table.
 *   alloc : a device buffer for a result (allocateRemote); `host`, if not
 *           NULL, is where it is copied back (lazily, or before an eviction).
 *   get   : the device pointer, re-uploading an evicted entry.
 *   mark  : work queued on `s` uses the entry (writes != 0: and writes it);
 *           it is not evicted, copied back or recycled before that work ends.
 *   peek  : copy a result back to its host buffer if the device copy is newer
 *           (copyToHostLazy, PTX/Array/Data.hs:91-152); *copied says whether it was.
 *   free  : the host array is gone (a finaliser); its block waits for reuse.
 * Any allocation that does not fit evicts least recently used entries (copying
 * dirty ones back first) and retries (PTX/Array/Remote.hs:63-74); entries that
 * exist only on the device are never evicted, and when nothing is left to evict
 * the call returns ACC_ERROR_OUT_OF_MEMORY.  Blocks are 4096-byte multiples
 * (PTX/Array/Remote.hs:101). */
typedef struct acc_table_s* acc_table_t;
typedef struct acc_table_stats {
  uint64_t uploads, downloads, evictions, nursery_hits, oom_retries;
  uint64_t bytes_up, bytes_down, bytes_evicted;
  uint64_t resident_bytes, nursery_bytes, entries;
} acc_table_stats;
int acc_hip_table_create(int device, acc_table_t* t);
int acc_hip_table_use(acc_table_t t, uint64_t key, const void* host, size_t bytes, acc_stream_t s, void** dev);
int acc_hip_table_alloc(acc_table_t t, uint64_t key, void* host, size_t bytes, void** dev);
int acc_hip_table_get(acc_table_t t, uint64_t key, acc_stream_t s, void** dev);
int acc_hip_table_mark(acc_table_t t, uint64_t key, acc_stream_t s, int writes);
int acc_hip_table_peek(acc_table_t t, uint64_t key, int* copied);
int acc_hip_table_free(acc_table_t t, uint64_t key);
int acc_hip_table_stats(acc_table_t t, acc_table_stats* st);
int acc_hip_table_destroy(acc_table_t t);

/* ================================================================ batched
 * Many independent scans, folds, segmented folds and histograms of uneven
 * sizes in ONE persistent launch (no counterpart in the reference, whose
 * per-call path is PTX/Execute.hs:835-881).  A batch is built once from
 * problem descriptors (device pointers) and can be launched many times,
 * directly or inside a captured graph.  One launch of a batch may be in flight
 * at a time.  Problem semantics:
 *   ACC_BATCH_SCAN:      has_seed == 0 -> scanl1 (n outputs);
 *                        has_seed != 0 -> scanl seed (n + 1 outputs, total last)
 *   ACC_BATCH_FOLD:      out[0] = seed (+) fold of in (fold1 when has_seed == 0)
 *   ACC_BATCH_FOLD_SEG:  out[s] = seed (+) fold of in[aux[s] .. aux[s+1]) for
 *                        s < m (fold1Seg when has_seed == 0: empty segments are
 *                        left as they are); aux: m + 1 offsets of aux_type
 *   ACC_BATCH_HISTOGRAM: permute (+) of a constant source: out = m bins, each
 *                        first `seed`, then out[in[i]] (+)= fill for i < n with
 *                        0 <= in[i] < m (in: indices of aux_type); m <= 4096
 *   ACC_BATCH_STENCIL:   acc_hip_stencil2d on an n x m image (rows x columns):
 *                        params = ry | rx << 8 | boundary << 16 with
 *                        ry == rx == 1 (3 x 3) and a clamp, mirror, wrap or constant
 *                        boundary (its value in `fill`); aux = the weights
 *                        (device memory, (2ry+1)(2rx+1) of the element type)
 * Element types INT32, INT64, FLOAT32, FLOAT64; combine ADD or MAX (ADD for
 * histograms; none for stencils).  Scan and fold inputs (and scan outputs)
 * must be 16-byte aligned. */
typedef enum acc_batch_kind {
  ACC_BATCH_SCAN = 0,
  ACC_BATCH_FOLD = 1,
  ACC_BATCH_FOLD_SEG = 2,
  ACC_BATCH_HISTOGRAM = 3,
  ACC_BATCH_STENCIL = 4
} acc_batch_kind;

typedef struct acc_batch_problem {
  int32_t kind;     /* acc_batch_kind */
  int32_t type;     /* acc_type */
  int32_t op;       /* acc_op */
  int32_t has_seed;
  int64_t n;        /* elements (scan, fold, foldSeg values), indices (histogram) */
  const void* in;   /* values; histogram: the indices */
  void* out;
  uint64_t seed;    /* bit pattern of the seed in the low bytes (histogram: the initial bin value) */
  const void* aux;  /* foldSeg: offsets; stencil: weights */
  int64_t m;        /* foldSeg: segments; histogram: bins; stencil: columns */
  int32_t aux_type; /* foldSeg offsets, histogram indices: ACC_INT32 or ACC_INT64 */
  int32_t params;   /* stencil: ry | rx << 8 | boundary << 16 */
  uint64_t fill;    /* histogram: the constant source value; stencil: the constant
                     * boundary's value (bit patterns) */
} acc_batch_problem;

typedef struct acc_batch_s* acc_batch_t;

/* tile = vectors per thread (2, 4 or 8, with 256-thread workgroups and runs of
 * 4 tiles); 0 lets the autotuner time the candidates (workgroup size 256 or
 * 512, vectors per thread, tiles per ticket) on this batch and remember the
 * winner (file named by $ACC_HIP_TUNE_CACHE, else
 * ~/.cache/accelerate-hip-tune.txt).  Tuning RUNS the batch several times (a
 * cached choice does not): it first waits for all work queued on the device,
 * so inputs written by earlier calls on any stream are complete, and it writes
 * every problem's output, so the outputs must not be in use by other work while
 * acc_hip_batch_create runs. */
int acc_hip_batch_create(int device, const acc_batch_problem* problems, int64_t count, int tile,
                         acc_batch_t* b);
int acc_hip_batch_launch(acc_batch_t b, acc_stream_t s);
/* The same problems through the plain entry points, one call (or two for a
 * histogram: fill, permute) per problem, in order on s: the per-call path the
 * batch replaces, for comparison. */
int acc_hip_batch_launch_plain(acc_batch_t b, acc_stream_t s);
/* vectors per thread and the number of tickets (runs of work) of one launch */
int acc_hip_batch_tile(acc_batch_t b, int* tile, uint32_t* total_tiles);
/* the chosen shape: workgroup size, vectors per thread, tiles per ticket, tickets */
int acc_hip_batch_shape(acc_batch_t b, int* block, int* vectors, int* run, uint32_t* tickets);
int acc_hip_batch_destroy(acc_batch_t b);

/* ================================================================ serving
 * A request loop (no counterpart in the reference, which runs one program per
 * `run`, PTX.hs:140-180): requests are skeleton invocations on device-resident
 * inputs; a scheduler thread admits them as execution slots (streams) and
 * output memory (a device arena, admission by free HBM) become free, so work
 * of requests that arrive while others execute overlaps on the GPU.  Latency
 * is measured from a request's arrival to its result being complete on the
 * device.  Requests may be submitted ahead with a future arrival time (an
 * open-loop arrival trace). */
typedef enum acc_request_kind {
  ACC_REQ_FOLD = 0,      /* fold/fold1 of in[0..n)                           */
  ACC_REQ_SCAN = 1,      /* scanl1/scanr1, or scanl/scanr with a seed        */
  ACC_REQ_FOLD_SEG = 2,  /* foldSeg/fold1Seg: aux = offsets (m + 1 values)   */
  ACC_REQ_PERMUTE = 3,   /* permute into m slots initialised to `seed`;
                            aux = indices (n), in = source or NULL (`fill`)  */
  ACC_REQ_STENCIL2D = 4  /* n x m image, weighted-sum stencil              */
} acc_request_kind;

typedef struct acc_request {
  int32_t skeleton;  /* acc_request_kind */
  int32_t type;      /* acc_type */
  int32_t op;        /* acc_op */
  int32_t has_seed;
  uint64_t seed;     /* bit pattern; permute: the defaults value */
  uint64_t fill;     /* permute: constant source; stencil: constant boundary value */
  const void* in;
  int64_t n;
  int64_t m;
  const void* aux;
  int32_t aux_type;  /* ACC_INT32 or ACC_INT64 */
  int32_t direction; /* scan */
  int32_t ry, rx;    /* stencil radii */
  int32_t boundary;  /* stencil */
  int32_t reserved;
  const void* weights; /* stencil weights (host; copied at submit) */
} acc_request;

typedef struct acc_server_stats {
  uint64_t completed, failed;
  double p50_us, p99_us, mean_us, max_us;
  double req_per_s, bytes_per_s, span_s; /* span: first arrival to last completion */
  size_t arena_bytes, arena_used;
  uint64_t rounds; /* fused rounds run (one batched launch each) */
  uint64_t fused;  /* requests that ran in a fused round */
} acc_server_stats;

typedef struct acc_server_s* acc_server_t;

/* hbm_budget: arena bytes (0 = half of the free device memory) */
int acc_hip_server_create(int device, int slots, size_t hbm_budget, acc_server_t* s);
/* arrival = now + arrival_offset_ns; keep_output: hold the result until
 * acc_hip_server_release (otherwise its memory is reused on completion).
 * acc_hip_server_wait blocks until the request is complete.  For a kept
 * request it returns the output's device address; for one submitted with
 * keep_output == 0 it returns *out = NULL, and once such a request has
 * completed and been retired its id is unknown (ACC_ERROR_INVALID_VALUE). */
/* Fused rounds.  With fuse == 1 (the default) a slot that admits a request
 * with a batched form -- a left-to-right scan, a fold or a foldSeg with (+) or
 * max, or a (+) permute of a constant source (in == NULL) into at most 4096
 * bins, over INT32/INT64/FLOAT32/FLOAT64, scan and fold inputs 16-byte
 * aligned -- also admits every other arrived request with a batched form whose
 * output fits, up to round_requests requests and round_bytes algorithmic
 * bytes, and runs them as ONE launch of the batched kernel (acc_hip_batch_*)
 * with the given shape; other requests run through the plain entry points.  A
 * fused request's result depends only on the request and the shape, not on
 * what else shares its round.  fuse == 0 runs every request through the plain
 * entry points, one per slot at a time (the per-call path).  Zero fields take
 * the defaults: 256 requests, 1 GiB, 256 threads x 8 vectors, runs of 4 tiles.
 * Applies to rounds formed after the call. */
typedef struct acc_server_config {
  int32_t fuse;
  int32_t round_requests;
  int64_t round_bytes;
  int32_t block, vectors, run; /* the batched kernel's shape (acc_hip_batch_shape) */
  int32_t reserved;
} acc_server_config;
int acc_hip_server_configure(acc_server_t s, const acc_server_config* cfg);
int acc_hip_server_submit(acc_server_t s, const acc_request* r, int64_t arrival_offset_ns, int keep_output,
                          uint64_t* id);
int acc_hip_server_wait(acc_server_t s, uint64_t id, void** out, double* latency_us);
int acc_hip_server_release(acc_server_t s, uint64_t id);
int acc_hip_server_drain(acc_server_t s);
int acc_hip_server_stats(acc_server_t s, acc_server_stats* st);
int acc_hip_server_reset_stats(acc_server_t s);
int acc_hip_server_destroy(acc_server_t s);

/* ================================================================ HBM ceilings
 * No counterpart in the reference: it calibrates the roofline fractions the
 * skeletons are measured against.  Times `reps` back-to-back launches of a
 * read-only reduction (src), a copy (src -> dst) or a write-only fill (dst)
 * over `bytes`, with `vectors` (4, 8, 16) 16-byte loads or stores in flight
 * per lane, `blocks_per_cu` 256-thread workgroups per CU and, when
 * `nontemporal`, the non-temporal hint on the loads (read, copy) or the
 * stores (write); *us = mean microseconds per launch.  The kernels stride
 * over whole grid-wide steps of vectors: a tail shorter than one step is left
 * out (none when bytes is a multiple of 16 * 256 * vectors * workgroups). */
typedef enum acc_probe_kind { ACC_PROBE_READ = 0, ACC_PROBE_COPY = 1, ACC_PROBE_WRITE = 2 } acc_probe_kind;
int acc_hip_bandwidth_probe(acc_stream_t s, int kind, int vectors, int blocks_per_cu, int nontemporal,
                            const void* src, void* dst, size_t bytes, int reps, double* us);

#ifdef __cplusplus
}
#endif

#endif /* ACCELERATE_HIP_H */
