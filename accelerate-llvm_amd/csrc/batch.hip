// Batched execution: many independent scans and folds of uneven sizes in ONE
// persistent launch, capturable as a hipGraph.
//
// The plain path costs one kernel launch (and the reference's path 1-4
// launches plus host marshalling, PTX/Execute.hs:835-881) per problem, so
// small problems are dominated by launch latency.  Here every problem is cut
// into tiles; the tiles of all problems form one ticket space; persistent
// workgroups take tiles in ticket order and dispatch on the problem's class
// (kind x element type x combine function).  Each problem owns a slice of the
// status words, so its scan look-back never crosses into another problem.
// A fold's intermediate tiles only publish their aggregate; its last tile
// looks back over all of them and writes the result with the seed on the
// left.
//
// The tile size (dwordx4 vectors per thread) is chosen by an autotuner that
// times the candidates on the batch itself and remembers the winner per
// batch signature in a small text file.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include <vector>
#include <algorithm>
#include <chrono>
#include "scan_tile.hpp"

namespace acc {

enum : int32_t { BK_SCAN = 0, BK_FOLD = 1 };

struct BatchDesc {
  int32_t cls;        // kind * 8 + type index * 2 + op index
  uint32_t has_seed;
  int64_t n;
  const void* in;
  void* out;
  uint64_t seed;      // bit pattern
  uint64_t status_off;  // first status word of this problem
  uint32_t ntiles;
  uint32_t tile0;     // first global ticket of this problem
};

struct BatchArgs {
  const BatchDesc* desc;
  const uint32_t* tile_prob;  // global tile -> problem
  uint32_t total_tiles;
  SyncCtl* ctl;
  uint64_t* status;
};

template <typename T> __device__ __forceinline__ T bits_to(uint64_t b) {
  T v;
  __builtin_memcpy(&v, &b, sizeof(T));
  return v;
}

// Fold of one tile; the problem's last tile finishes the fold.
template <typename T, typename Op, int BLOCK, int K, bool MASKED>
__device__ __forceinline__ void fold_tile(const BatchDesc& d, uint64_t* status, uint32_t t, uint32_t epoch,
                                          T* s_wave) {
  constexpr int VEC = Vec<T>::N;
  constexpr int W = BLOCK / 64;
  constexpr int64_t TILE = (int64_t)BLOCK * K * VEC;
  constexpr int NW = Tagged<T>::NW;
  Op op{};
  const int lane = lane_id(), w = threadIdx.x / 64;
  const T* in = (const T*)d.in;
  const int64_t base = (int64_t)t * TILE;
  const T id = identity<T>(op);
  T acc = id;
  if (!MASKED) {
    Vec<T> v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = load16_stream(in + base + ((int64_t)(w * K + k) * 64 + lane) * VEC);
    T r = v[0].e[0];
#pragma unroll
    for (int j = 1; j < VEC; ++j) r = op(r, v[0].e[j]);
#pragma unroll
    for (int k = 1; k < K; ++k)
#pragma unroll
      for (int j = 0; j < VEC; ++j) r = op(r, v[k].e[j]);
    acc = r;
  } else {
    // the problem's last tile: a vector past the end re-reads the vector
    // holding the last element (an aligned 16-byte block with a valid element
    // never crosses a page), so every load is one unconditional dwordx4
    Vec<T> v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int64_t e0 = base + ((int64_t)(w * K + k) * 64 + lane) * VEC;
      v[k] = load16_stream(in + min(e0, (d.n - 1) / VEC * VEC));
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int64_t e0 = base + ((int64_t)(w * K + k) * 64 + lane) * VEC;
#pragma unroll
      for (int j = 0; j < VEC; ++j) acc = op(acc, e0 + j < d.n ? v[k].e[j] : id);
    }
  }
  const T wr = wave_reduce(acc, op);
  if (lane == 0) s_wave[w] = wr;
  __syncthreads();
  if (w == 0) {
    T agg = id;
#pragma unroll
    for (int i = 0; i < W; ++i) agg = op(agg, s_wave[i]);
    uint64_t* slot = status + (size_t)t * NW;
    const bool first = t == 0, last = t + 1 == d.ntiles;
    if (!last) {
      if (lane == 0) Tagged<T>::publish(slot, agg, make_tag(epoch, first ? ST_PREFIX : ST_AGGREGATE));
    } else {
      T r = agg;
      if (!first) r = op(wave_lookback<T>(status, t, epoch, op), r);
      if (d.has_seed) r = op(bits_to<T>(d.seed), r);  // seed on the left
      if (lane == 0) *(T*)d.out = r;
    }
  }
}

template <typename T, typename Op, int BLOCK, int K>
__device__ __forceinline__ void batch_tile(const BatchDesc& d, int kind, uint64_t* status, uint32_t t,
                                           uint32_t epoch, unsigned char* s_raw) {
  constexpr int64_t TILE = (int64_t)BLOCK * K * Vec<T>::N;
  T* s_wave = (T*)s_raw;
  T* s_prefix = s_wave + BLOCK / 64;
  const int64_t base = (int64_t)t * TILE;
  const bool full = base + TILE <= d.n;
  if (kind == BK_SCAN) {
    // scanl1 without a seed; with a seed the exclusive scanl (total at out[n])
    TileIO<T> io{(const T*)d.in, (T*)d.out, d.n, 0u, d.ntiles - 1, d.has_seed, d.has_seed, bits_to<T>(d.seed),
                 d.has_seed ? (T*)d.out + d.n : nullptr, nullptr, status};
    if (full) scan_tile<T, Op, BLOCK, K, false, true, false>(io, Op{}, t, base, epoch, s_wave, s_prefix);
    else scan_tile<T, Op, BLOCK, K, false, true, true>(io, Op{}, t, base, epoch, s_wave, s_prefix);
  } else {
    if (full) fold_tile<T, Op, BLOCK, K, false>(d, status, t, epoch, s_wave);
    else fold_tile<T, Op, BLOCK, K, true>(d, status, t, epoch, s_wave);
  }
}

template <int BLOCK, int K>
__global__ __launch_bounds__(BLOCK) void k_batch(BatchArgs a) {
  __shared__ uint64_t s_tick;
  __shared__ __align__(16) unsigned char s_raw[(BLOCK / 64 + 1) * sizeof(double)];
  if (threadIdx.x == 0) s_tick = take_ticket(a.ctl);
  __syncthreads();
  const uint64_t tk0 = s_tick;
  const uint32_t epoch = (uint32_t)(tk0 >> 32);
  uint32_t g = (uint32_t)tk0;
  while (g < a.total_tiles) {
    const BatchDesc d = a.desc[a.tile_prob[g]];
    const uint32_t t = g - d.tile0;
    uint64_t* st = a.status + d.status_off;
    const int kind = d.cls >> 3;
    switch (d.cls & 7) {
      case 0: batch_tile<int32_t, OpAdd, BLOCK, K>(d, kind, st, t, epoch, s_raw); break;
      case 1: batch_tile<int32_t, OpMax, BLOCK, K>(d, kind, st, t, epoch, s_raw); break;
      case 2: batch_tile<int64_t, OpAdd, BLOCK, K>(d, kind, st, t, epoch, s_raw); break;
      case 3: batch_tile<int64_t, OpMax, BLOCK, K>(d, kind, st, t, epoch, s_raw); break;
      case 4: batch_tile<float, OpAdd, BLOCK, K>(d, kind, st, t, epoch, s_raw); break;
      case 5: batch_tile<float, OpMax, BLOCK, K>(d, kind, st, t, epoch, s_raw); break;
      case 6: batch_tile<double, OpAdd, BLOCK, K>(d, kind, st, t, epoch, s_raw); break;
      default: batch_tile<double, OpMax, BLOCK, K>(d, kind, st, t, epoch, s_raw); break;
    }
    __syncthreads();
    if (threadIdx.x == 0) s_tick = take_ticket(a.ctl);
    __syncthreads();
    g = (uint32_t)s_tick;
  }
  block_retire(a.ctl, gridDim.x, epoch);
}

constexpr int BATCH_BLOCK = 256;
static const int kBatchK[] = {1, 2, 4};

}  // namespace acc

// ----------------------------------------------------------------- host side
using namespace acc;

struct acc_batch_s {
  int device = 0;
  std::vector<acc_batch_problem> problems;
  int k = 0;                 // vectors per thread (tile size)
  uint32_t total_tiles = 0;
  void* d_desc = nullptr;
  void* d_map = nullptr;
  void* d_ws = nullptr;      // SyncCtl + status words
  size_t ws_bytes = 0;
};

namespace acc {

static int type_index(acc_type t) {
  switch (t) {
    case ACC_INT32: return 0;
    case ACC_INT64: return 1;
    case ACC_FLOAT32: return 2;
    case ACC_FLOAT64: return 3;
    default: return -1;
  }
}

static void batch_free(acc_batch_t b) {
  if (b->d_desc) dev_free(b->device, b->d_desc);
  if (b->d_map) dev_free(b->device, b->d_map);
  if (b->d_ws) dev_free(b->device, b->d_ws);
  b->d_desc = b->d_map = b->d_ws = nullptr;
}

// Lay out tiles for vectors-per-thread k and upload descriptors, map and a
// zeroed workspace.
static int batch_layout(acc_batch_t b, int k) {
  batch_free(b);
  const size_t n = b->problems.size();
  std::vector<BatchDesc> desc(n);
  std::vector<uint32_t> map;
  uint64_t words = 0;
  uint64_t tiles = 0;
  for (size_t i = 0; i < n; ++i) {
    const acc_batch_problem& p = b->problems[i];
    const size_t B = type_size((acc_type)p.type);
    const int64_t tile = (int64_t)BATCH_BLOCK * k * (int64_t)(16 / B);
    const int64_t nt = p.n > 0 ? (p.n + tile - 1) / tile : 0;
    BatchDesc& d = desc[i];
    d.cls = p.kind * 8 + type_index((acc_type)p.type) * 2 + (p.op == ACC_OP_MAX ? 1 : 0);
    d.has_seed = p.has_seed ? 1u : 0u;
    d.n = p.n;
    d.in = p.in;
    d.out = p.out;
    d.seed = p.seed;
    d.status_off = words;
    d.ntiles = (uint32_t)nt;
    d.tile0 = (uint32_t)tiles;
    words += (uint64_t)nt * (B == 8 ? 2 : 1);
    tiles += (uint64_t)nt;
    if (tiles > 0xffffffffULL) return fail(ACC_ERROR_UNSUPPORTED, "batch: too many tiles");
    for (int64_t t = 0; t < nt; ++t) map.push_back((uint32_t)i);
  }
  b->k = k;
  b->total_tiles = (uint32_t)tiles;
  ACC_CHECK(dev_malloc(b->device, std::max<size_t>(n, 1) * sizeof(BatchDesc), &b->d_desc));
  ACC_CHECK(dev_malloc(b->device, std::max<size_t>(map.size(), 1) * sizeof(uint32_t), &b->d_map));
  b->ws_bytes = 256 + words * 8;
  ACC_CHECK(dev_malloc(b->device, b->ws_bytes, &b->d_ws));
  ACC_HIP_CHECK(hipMemcpy(b->d_desc, desc.data(), n * sizeof(BatchDesc), hipMemcpyHostToDevice));
  if (!map.empty())
    ACC_HIP_CHECK(hipMemcpy(b->d_map, map.data(), map.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  ACC_HIP_CHECK(hipMemset(b->d_ws, 0, b->ws_bytes));
  // warm the occupancy cache so a later launch inside a stream capture only
  // enqueues the kernel
  switch (k) {
    case 1: blocks_per_cu((const void*)k_batch<BATCH_BLOCK, 1>, BATCH_BLOCK); break;
    case 2: blocks_per_cu((const void*)k_batch<BATCH_BLOCK, 2>, BATCH_BLOCK); break;
    default: blocks_per_cu((const void*)k_batch<BATCH_BLOCK, 4>, BATCH_BLOCK); break;
  }
  const DeviceInfo* di;
  return device_info(b->device, &di);
}

static int batch_enqueue(acc_batch_t b, hipStream_t stream) {
  if (!b->total_tiles) return ACC_SUCCESS;
  const DeviceInfo* di;
  ACC_CHECK(device_info(b->device, &di));
  BatchArgs a{(const BatchDesc*)b->d_desc, (const uint32_t*)b->d_map, b->total_tiles, (SyncCtl*)b->d_ws,
              (uint64_t*)((char*)b->d_ws + 256)};
  auto launch = [&](auto kern) -> int {
    int64_t grid = (int64_t)di->compute_units * blocks_per_cu((const void*)kern, BATCH_BLOCK);
    if (grid > (int64_t)b->total_tiles) grid = b->total_tiles;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(BATCH_BLOCK), 0, stream, a);
    ACC_HIP_CHECK(hipGetLastError());
    return ACC_SUCCESS;
  };
  switch (b->k) {
    case 1: return launch(k_batch<BATCH_BLOCK, 1>);
    case 2: return launch(k_batch<BATCH_BLOCK, 2>);
    case 4: return launch(k_batch<BATCH_BLOCK, 4>);
  }
  return fail(ACC_ERROR_INVALID_VALUE, "batch: tile size %d", b->k);
}

// Signature of a batch for the tuning cache: problem count, element-size
// histogram (log2 buckets) and class mix.
static std::string batch_signature(const acc_batch_t b) {
  int hist[64] = {0};
  int cls[16] = {0};
  for (auto& p : b->problems) {
    int lg = 0;
    while (lg < 63 && ((int64_t)1 << (lg + 1)) <= p.n) ++lg;
    hist[lg / 2]++;
    cls[(p.kind * 8 + type_index((acc_type)p.type) * 2 + (p.op == ACC_OP_MAX)) & 15]++;
  }
  std::string s = "n" + std::to_string(b->problems.size());
  for (int i = 0; i < 32; ++i) if (hist[i]) s += "_h" + std::to_string(i) + ":" + std::to_string(hist[i] * 20 / (int)b->problems.size());
  for (int i = 0; i < 16; ++i) if (cls[i]) s += "_c" + std::to_string(i) + ":" + std::to_string(cls[i] * 20 / (int)b->problems.size());
  return s;
}

static std::string tune_path() {
  if (const char* p = getenv("ACC_HIP_TUNE_CACHE")) return p;
  std::string dir;
  if (const char* x = getenv("XDG_CACHE_HOME")) dir = x;
  else if (const char* h = getenv("HOME")) dir = std::string(h) + "/.cache";
  else dir = "/tmp";
  return dir + "/accelerate-hip-tune.txt";
}

static int tune_lookup(const std::string& sig) {
  FILE* f = fopen(tune_path().c_str(), "r");
  if (!f) return 0;
  char line[1024];
  int k = 0;
  while (fgets(line, sizeof line, f)) {
    char key[900];
    int v;
    if (sscanf(line, "batch %899s %d", key, &v) == 2 && sig == key) k = v;
  }
  fclose(f);
  return k;
}

static void tune_store(const std::string& sig, int k) {
  FILE* f = fopen(tune_path().c_str(), "a");
  if (!f) return;
  fprintf(f, "batch %s %d\n", sig.c_str(), k);
  fclose(f);
}

// Time each candidate tile size on the batch itself; keep the fastest.
static int batch_autotune(acc_batch_t b) {
  const std::string sig = batch_signature(b);
  int k = tune_lookup(sig);
  if (k == 1 || k == 2 || k == 4) return batch_layout(b, k);
  // the timing runs read the inputs and write the outputs: order them after
  // everything already queued on the device (inputs produced on any stream)
  ACC_HIP_CHECK(hipDeviceSynchronize());
  hipStream_t st;
  ACC_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  hipEvent_t e0, e1;
  ACC_HIP_CHECK(hipEventCreate(&e0));
  ACC_HIP_CHECK(hipEventCreate(&e1));
  float best = 1e30f;
  int best_k = 2;
  for (int cand : kBatchK) {
    int rc = batch_layout(b, cand);
    if (!rc) rc = batch_enqueue(b, st);  // warm-up
    if (!rc) {
      (void)hipEventRecord(e0, st);
      for (int r = 0; r < 3 && !rc; ++r) rc = batch_enqueue(b, st);
      (void)hipEventRecord(e1, st);
      (void)hipEventSynchronize(e1);
    }
    float ms = 0;
    if (!rc && hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms < best) {
      best = ms;
      best_k = cand;
    }
    if (rc) break;
  }
  (void)hipStreamSynchronize(st);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(st);
  tune_store(sig, best_k);
  return batch_layout(b, best_k);
}

}  // namespace acc

extern "C" {

int acc_hip_batch_create(int device, const acc_batch_problem* problems, int64_t count, int tile, acc_batch_t* out) {
  if (!out || (count > 0 && !problems) || count < 0) return fail(ACC_ERROR_INVALID_VALUE, "batch: bad arguments");
  if (tile != 0 && tile != 1 && tile != 2 && tile != 4)
    return fail(ACC_ERROR_INVALID_VALUE, "batch: tile must be 0 (autotune), 1, 2 or 4");
  for (int64_t i = 0; i < count; ++i) {
    const acc_batch_problem& p = problems[i];
    if (p.kind != ACC_BATCH_SCAN && p.kind != ACC_BATCH_FOLD)
      return fail(ACC_ERROR_UNSUPPORTED, "batch: problem %lld: kind %d", (long long)i, p.kind);
    if (type_index((acc_type)p.type) < 0)
      return fail(ACC_ERROR_UNSUPPORTED, "batch: problem %lld: element type %d", (long long)i, p.type);
    if (p.op != ACC_OP_ADD && p.op != ACC_OP_MAX)
      return fail(ACC_ERROR_UNSUPPORTED, "batch: problem %lld: combine %d", (long long)i, p.op);
    if (p.n < 0) return fail(ACC_ERROR_INVALID_VALUE, "batch: problem %lld: negative size", (long long)i);
    if (p.kind == ACC_BATCH_FOLD && p.n == 0 && !p.has_seed)
      return fail(ACC_ERROR_EMPTY_ARRAY, "batch: problem %lld: fold1 of an empty array", (long long)i);
    if (p.n > 0 && (!p.in || !p.out || (uintptr_t)p.in % 16 || (p.kind == ACC_BATCH_SCAN && (uintptr_t)p.out % 16)))
      return fail(ACC_ERROR_INVALID_VALUE, "batch: problem %lld: arrays must be non-null and 16-byte aligned",
                  (long long)i);
    if (p.n == 0)
      return fail(ACC_ERROR_UNSUPPORTED, "batch: problem %lld: empty problems are served by the plain calls",
                  (long long)i);
  }
  ACC_CHECK(set_device(device));
  auto b = new acc_batch_s();
  b->device = device;
  b->problems.assign(problems, problems + count);
  int rc = tile ? batch_layout(b, tile) : batch_autotune(b);
  if (rc) {
    batch_free(b);
    delete b;
    return rc;
  }
  *out = b;
  return ACC_SUCCESS;
}

int acc_hip_batch_launch(acc_batch_t b, acc_stream_t s) {
  if (!b || !s) return fail(ACC_ERROR_INVALID_VALUE, "batch: null argument");
  if (s->device != b->device) return fail(ACC_ERROR_INVALID_VALUE, "batch: stream on another device");
  ACC_CHECK(set_device(s->device));
  return batch_enqueue(b, s->stream);
}

int acc_hip_batch_tile(acc_batch_t b, int* tile, uint32_t* total_tiles) {
  if (!b) return fail(ACC_ERROR_INVALID_VALUE, "batch: null");
  if (tile) *tile = b->k;
  if (total_tiles) *total_tiles = b->total_tiles;
  return ACC_SUCCESS;
}

int acc_hip_batch_destroy(acc_batch_t b) {
  if (!b) return ACC_SUCCESS;
  (void)set_device(b->device);
  (void)hipDeviceSynchronize();
  batch_free(b);
  delete b;
  return ACC_SUCCESS;
}

}  // extern "C"
