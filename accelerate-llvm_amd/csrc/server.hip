// Serving loop: requests (one skeleton invocation each, inputs resident on
// the device) arrive over time; one worker thread per execution slot admits
// them as slots and device memory free up, runs admitted work concurrently on
// per-slot streams, and records each request's latency from arrival to its
// result being complete on the device.
//
// Continuous batching: there is no fixed batch boundary.  A slot that admits
// a scan, fold, foldSeg, histogram or stencil also admits every other such
// request that has arrived and whose output fits, up to a round limit, and
// runs the whole round as ONE launch of the batched form's persistent kernel
// (plus one of its stencil kernel when the round holds stencils: one upload
// of descriptors, run maps and stencil weights per round instead of one or
// more launches per request); requests that arrive
// while a round executes form the next round on another slot, so rounds
// overlap on the GPU.  The combinations the batched kernel does not take run
// through the plain entry points on their slot.  Admission is by free HBM: outputs are carved from a
// device arena sized at creation from hipMemGetInfo (or the caller's budget,
// whichever is smaller); a request whose output does not fit waits, and later
// small requests may pass it.
//
// There is no counterpart in the reference (accelerate-llvm runs one program
// per `run` call, PTX.hs:140-180).  A fused request runs the same per-problem
// computation whatever else shares its round (its tiles, look-back order and
// per-segment order depend only on the problem and the round's fixed shape),
// so its result is the same alone and in the mix.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>
#include <string.h>
#include "host.hpp"
#include "batch_internal.hpp"
#include "lookback.hpp"

namespace acc {

using clk = std::chrono::steady_clock;

// First-fit sub-allocator over one device allocation (offsets, 256-byte units).
class Arena {
 public:
  void init(size_t bytes) {
    cap_ = bytes;
    free_.clear();
    if (bytes) free_[0] = bytes;
    used_ = 0;
  }
  bool alloc(size_t bytes, size_t* off) {
    bytes = (std::max<size_t>(bytes, 1) + 255) / 256 * 256;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        *off = it->first;
        const size_t rest = it->second - bytes;
        const size_t at = it->first + bytes;
        free_.erase(it);
        if (rest) free_[at] = rest;
        used_ += bytes;
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = (std::max<size_t>(bytes, 1) + 255) / 256 * 256;
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t capacity() const { return cap_; }
  size_t used() const { return used_; }

 private:
  size_t cap_ = 0, used_ = 0;
  std::map<size_t, size_t> free_;
};

struct Req {
  uint64_t id;
  acc_request r;
  std::vector<uint8_t> weights;  // stencil weights (host copy)
  clk::time_point arrival;
  clk::time_point done;
  size_t out_bytes = 0, out_off = 0;
  int slot = -1;
  int status = ACC_SUCCESS;
  bool finished = false;
  bool keep = false;
};

}  // namespace acc

using namespace acc;

namespace acc {
// Per-slot buffers of the fused rounds: pinned staging and device copies of
// the descriptors and run map, and the batched kernel's workspace (SyncCtl +
// status words).  Grown on demand, only while the slot has nothing in flight.
struct RoundBufs {
  void* h = nullptr;
  size_t h_cap = 0;
  void* d = nullptr;
  size_t d_cap = 0;
  void* ws = nullptr;
  size_t ws_cap = 0;
};
}  // namespace acc

struct acc_server_s {
  int device = 0;
  acc_server_config cfg{1, 256, (int64_t)1 << 30, 256, 8, 4, 0};
  std::vector<acc::RoundBufs> rounds_buf;  // one per slot
  uint64_t rounds = 0, fused = 0;
  void* arena_base = nullptr;
  Arena arena;
  std::vector<acc_stream_t> slots;
  std::vector<hipEvent_t> slot_event;
  std::vector<uint64_t> slot_req;  // 0 = idle
  std::mutex mu;
  std::condition_variable cv;
  std::map<uint64_t, std::shared_ptr<Req>> reqs;  // a waiter holds its own reference
  std::deque<uint64_t> pending;  // arrival order (by arrival time)
  std::vector<double> latencies_us;
  clk::time_point first_arrival{}, last_done{};
  bool any = false;
  uint64_t next_id = 1;
  uint64_t completed = 0, failed = 0;
  double bytes_done = 0;
  bool stop = false;
  std::vector<std::thread> workers;  // one per slot
};

namespace acc {

static size_t request_out_bytes(const acc_request& r) {
  const size_t B = type_size((acc_type)r.type);
  switch (r.skeleton) {
    case ACC_REQ_FOLD: return B;
    case ACC_REQ_SCAN: return B * (size_t)(r.has_seed ? r.n + 1 : r.n);
    case ACC_REQ_FOLD_SEG: return B * (size_t)r.m;
    case ACC_REQ_PERMUTE: return B * (size_t)r.m;
    case ACC_REQ_STENCIL2D: return B * (size_t)r.n * (size_t)r.m;
  }
  return 0;
}

// algorithmic HBM bytes of a request (for the aggregate bytes/s figure)
static double request_bytes(const acc_request& r) {
  const double B = (double)type_size((acc_type)r.type);
  const double I = r.aux_type == ACC_INT64 ? 8.0 : 4.0;
  switch (r.skeleton) {
    case ACC_REQ_FOLD: return r.n * B + B;
    case ACC_REQ_SCAN: return r.n * B + (r.has_seed ? r.n + 1 : r.n) * B;
    case ACC_REQ_FOLD_SEG: return r.n * B + (r.m + 1) * I + r.m * B;
    case ACC_REQ_PERMUTE: return r.n * I + (r.in ? r.n * B : 0) + 2 * r.m * B;
    case ACC_REQ_STENCIL2D: return 2.0 * r.n * r.m * B;
  }
  return 0;
}

static int run_request(acc_server_t s, Req& q, acc_stream_t st, void* out) {
  const acc_request& r = q.r;
  const void* seed = r.has_seed ? (const void*)&r.seed : nullptr;
  switch (r.skeleton) {
    case ACC_REQ_FOLD:
      return acc_hip_fold(st, (acc_type)r.type, (acc_op)r.op, seed, r.in, 1, r.n, out);
    case ACC_REQ_SCAN:
      return acc_hip_scan(st, (acc_type)r.type, (acc_op)r.op, (acc_direction)r.direction, seed, r.in, 1, r.n, out);
    case ACC_REQ_FOLD_SEG:
      return acc_hip_fold_seg(st, (acc_type)r.type, (acc_op)r.op, seed, r.in, 1, r.n, (acc_type)r.aux_type, r.aux,
                              r.m, out);
    case ACC_REQ_PERMUTE: {
      // defaults: the seed value (0 for a histogram), updated in place
      uint64_t z = r.seed;
      int rc = acc_hip_fill_async(st, out, &z, type_size((acc_type)r.type), (size_t)r.m);
      if (rc) return rc;
      return acc_hip_permute(st, (acc_type)r.type, (acc_op)r.op, out, out, r.m, r.in, &r.fill,
                             (acc_type)r.aux_type, r.aux, r.n);
    }
    case ACC_REQ_STENCIL2D:
      return acc_hip_stencil2d(st, (acc_type)r.type, r.ry, r.rx, q.weights.data(), (acc_boundary)r.boundary,
                               &r.fill, r.in, r.n, r.m, out);
  }
  return fail(ACC_ERROR_UNSUPPORTED, "server: request kind %d", r.skeleton);
}

// The batched kernel's form of a request, if it has one: left-to-right
// scans, folds and foldSegs with (+) or max, (+) permutes of a constant
// source into at most 4096 bins, and 3 x 3 stencils with a clamp,
// mirror, wrap or constant boundary, over Int32/Int64/Float32/Float64, with
// the alignment the batched kernel needs.  A stencil's weights pointer is
// filled in by the round (they travel with its descriptors).
static bool batchable(const acc_request& r, void* out, acc_batch_problem* p) {
  if (r.type != ACC_INT32 && r.type != ACC_INT64 && r.type != ACC_FLOAT32 && r.type != ACC_FLOAT64) return false;
  const bool add_max = r.op == ACC_OP_ADD || r.op == ACC_OP_MAX;
  const bool idx_ok = (r.aux_type == ACC_INT32 || r.aux_type == ACC_INT64) && r.aux;
  memset(p, 0, sizeof *p);
  p->type = r.type;
  p->op = r.op;
  p->has_seed = r.has_seed ? 1 : 0;
  p->seed = r.seed;
  p->n = r.n;
  p->out = out;
  switch (r.skeleton) {
    case ACC_REQ_SCAN:
    case ACC_REQ_FOLD:
      if (!add_max || r.n <= 0 || !r.in || (uintptr_t)r.in % 16 || (uintptr_t)out % 16) return false;
      if (r.skeleton == ACC_REQ_SCAN && r.direction != ACC_LEFT_TO_RIGHT) return false;
      p->kind = r.skeleton == ACC_REQ_SCAN ? ACC_BATCH_SCAN : ACC_BATCH_FOLD;
      p->in = r.in;
      return true;
    case ACC_REQ_FOLD_SEG:
      if (!add_max || r.m <= 0 || !idx_ok || (r.n > 0 && !r.in) || r.n < 0) return false;
      p->kind = ACC_BATCH_FOLD_SEG;
      p->in = r.in;
      p->aux = r.aux;
      p->aux_type = r.aux_type;
      p->m = r.m;
      return true;
    case ACC_REQ_STENCIL2D: {
      const size_t B = type_size((acc_type)r.type);
      if (r.ry != 1 || r.rx != 1 || r.n <= 0 || r.m <= 0 || !r.in) return false;
      if (r.boundary != ACC_BOUNDARY_CLAMP && r.boundary != ACC_BOUNDARY_MIRROR && r.boundary != ACC_BOUNDARY_WRAP &&
          r.boundary != ACC_BOUNDARY_CONSTANT)
        return false;
      if ((uintptr_t)r.in % B || (uintptr_t)out % B || r.n > ((int64_t)1 << 26) || r.m > ((int64_t)1 << 30)) return false;
      p->kind = ACC_BATCH_STENCIL;
      p->in = r.in;
      p->m = r.m;
      p->params = r.ry | (r.rx << 8) | (r.boundary << 16);
      p->fill = r.fill;
      return true;
    }
    case ACC_REQ_PERMUTE:
      if (r.op != ACC_OP_ADD || r.in || r.m <= 0 || r.m > 4096 || !idx_ok || r.n < 0) return false;
      p->kind = ACC_BATCH_HISTOGRAM;
      p->in = r.aux;  // the indices
      p->aux_type = r.aux_type;
      p->m = r.m;
      p->fill = r.fill;
      p->seed = r.seed;  // every bin starts at the defaults value
      return true;
  }
  return false;
}

static int grow_host(void** p, size_t* cap, size_t need) {
  if (need <= *cap) return ACC_SUCCESS;
  if (*p) (void)hipHostFree(*p);
  *p = nullptr;
  *cap = 0;
  need = std::max<size_t>(need * 2, 64 << 10);
  ACC_HIP_CHECK(hipHostMalloc(p, need, hipHostMallocDefault));
  *cap = need;
  return ACC_SUCCESS;
}

static int grow_device(int device, void** p, size_t* cap, size_t need) {
  if (need <= *cap) return ACC_SUCCESS;
  if (*p) dev_free(device, *p);
  *p = nullptr;
  *cap = 0;
  need = std::max<size_t>(need * 2, 64 << 10);
  ACC_CHECK(dev_malloc(device, need, p));
  *cap = need;
  return ACC_SUCCESS;
}

// One fused round on the slot's stream: plan, one upload of stencil weights,
// descriptors and run maps, one launch of the persistent kernel (and one of
// the stencil kernel when the round holds stencils).  The slot's previous round has
// completed, so its buffers can be reused (or replaced) freely.
static int run_round(acc_server_t s, int slot, std::vector<acc_batch_problem>& probs,
                     const std::vector<std::shared_ptr<Req>>& round, BatchPlan& plan) {
  const acc_server_config& c = s->cfg;
  RoundBufs& b = s->rounds_buf[slot];
  hipStream_t stream = s->slots[slot]->stream;
  // the weights go first, at offsets known before planning
  std::vector<size_t> w_off(probs.size(), 0);
  size_t w_bytes = 0;
  for (size_t i = 0; i < probs.size(); ++i) {
    if (probs[i].kind != ACC_BATCH_STENCIL) continue;
    w_off[i] = w_bytes;
    w_bytes += (round[i]->weights.size() + 15) / 16 * 16;
  }
  w_bytes = (w_bytes + 255) / 256 * 256;
  const size_t desc_bytes = (probs.size() * sizeof(BatchDescBytes) + 255) / 256 * 256;
  size_t bytes = 0, map_bytes = 0, smap_bytes = 0;
  // plan against the device buffer's address; growing it moves the weights,
  // so the plan is made again (at most once)
  for (int pass = 0; pass < 3; ++pass) {
    for (size_t i = 0; i < probs.size(); ++i)
      if (probs[i].kind == ACC_BATCH_STENCIL) probs[i].aux = (char*)b.d + w_off[i];
    ACC_CHECK(batch_plan(probs.data(), (int64_t)probs.size(), c.block, c.vectors, c.run, &plan));
    map_bytes = (plan.run_prob.size() * sizeof(uint32_t) + 255) / 256 * 256;
    smap_bytes = plan.stencil_prob.size() * sizeof(uint32_t);
    bytes = w_bytes + desc_bytes + map_bytes + smap_bytes;
    if (b.d && bytes <= b.d_cap) break;
    ACC_CHECK(grow_device(s->device, &b.d, &b.d_cap, bytes));
  }
  ACC_CHECK(grow_host(&b.h, &b.h_cap, bytes));
  const size_t ws_bytes = kSyncCtlBytes + plan.status_words * 8;
  if (ws_bytes > b.ws_cap) {
    ACC_CHECK(grow_device(s->device, &b.ws, &b.ws_cap, ws_bytes));
    // fresh memory: no stale word may carry the current epoch's tag
    ACC_HIP_CHECK(hipMemsetAsync(b.ws, 0, b.ws_cap, stream));
  }
  char* h = (char*)b.h;
  for (size_t i = 0; i < probs.size(); ++i)
    if (probs[i].kind == ACC_BATCH_STENCIL) memcpy(h + w_off[i], round[i]->weights.data(), round[i]->weights.size());
  memcpy(h + w_bytes, plan.desc.data(), probs.size() * sizeof(BatchDescBytes));
  memcpy(h + w_bytes + desc_bytes, plan.run_prob.data(), plan.run_prob.size() * sizeof(uint32_t));
  memcpy(h + w_bytes + desc_bytes + map_bytes, plan.stencil_prob.data(), smap_bytes);
  ACC_HIP_CHECK(hipMemcpyAsync(b.d, b.h, bytes, hipMemcpyHostToDevice, stream));
  char* d = (char*)b.d;
  return batch_launch_plan(s->device, plan, d + w_bytes, d + w_bytes + desc_bytes, d + w_bytes + desc_bytes + map_bytes,
                           b.ws, stream);
}

// One worker thread per slot: it takes the earliest arrived request whose
// output fits in the free arena (a request that does not fit is passed by
// later ones).  If that request has a batched form (and fusion is on), the
// slot also takes every other arrived request with a batched form whose
// output fits, up to the round limits, and runs them as one fused round;
// otherwise it runs the request through the plain entry points.  It then
// waits for the slot's completion event and retires what it ran.  Workers
// launch in parallel, so the host-side cost of the launches is spread over
// the slots.
static void slot_worker(acc_server_t s, int slot) {
  (void)hipSetDevice(s->device);
  acc_stream_t st = s->slots[slot];
  hipEvent_t ev = s->slot_event[slot];
  std::vector<std::shared_ptr<Req>> round;
  std::vector<acc_batch_problem> probs;
  BatchPlan plan;
  std::unique_lock<std::mutex> lk(s->mu);
  while (true) {
    const auto now = clk::now();
    round.clear();
    probs.clear();
    acc_batch_problem bp;
    double round_bytes = 0;
    for (size_t i = 0; i < s->pending.size();) {
      const std::shared_ptr<Req>& q = s->reqs[s->pending[i]];
      if (q->arrival > now) break;  // pending is ordered by arrival
      const bool fusable = s->cfg.fuse != 0;
      if (!round.empty()) {  // filling a fused round: batched forms only
        if ((int)round.size() >= s->cfg.round_requests || round_bytes >= (double)s->cfg.round_bytes) break;
        if (!fusable || !batchable(q->r, (void*)(uintptr_t)256, &bp)) {
          ++i;
          continue;
        }
      }
      size_t off = 0;
      if (!s->arena.alloc(q->out_bytes, &off)) {
        ++i;
        continue;
      }
      void* out = (char*)s->arena_base + off;
      const bool fused = fusable && batchable(q->r, out, &bp);
      if (!fused && !round.empty()) {  // (arena offsets keep every alignment the check above assumed)
        s->arena.release(off, q->out_bytes);
        ++i;
        continue;
      }
      q->out_off = off;
      q->slot = slot;
      round.push_back(q);
      s->pending.erase(s->pending.begin() + (std::ptrdiff_t)i);
      if (!fused) break;  // a plain request runs alone
      probs.push_back(bp);
      round_bytes += request_bytes(q->r);
    }
    if (!round.empty()) {
      s->slot_req[slot] = round.front()->id;
      lk.unlock();
      int rc;
      if (probs.empty()) {
        Req& q = *round.front();
        rc = run_request(s, q, st, (char*)s->arena_base + q.out_off);
      } else {
        rc = run_round(s, slot, probs, round, plan);
      }
      if (!rc && hipEventRecord(ev, st->stream) != hipSuccess) rc = ACC_ERROR_HIP;
      while (!rc) {  // poll: a blocking wait adds the scheduler's wake-up latency
        const hipError_t e = hipEventQuery(ev);
        if (e == hipSuccess) break;
        if (e != hipErrorNotReady) rc = ACC_ERROR_HIP;
        else std::this_thread::yield();
      }
      if (rc) {
        // kernels queued before the failure may still write the outputs
        (void)hipStreamSynchronize(st->stream);
      }
      const auto done = clk::now();
      lk.lock();
      s->slot_req[slot] = 0;
      if (!probs.empty() && !rc) {
        s->rounds++;
        s->fused += probs.size();
      }
      for (auto& qp : round) {
        Req& q = *qp;
        q.done = done;
        q.finished = true;
        if (rc) {
          q.status = rc;
          s->failed++;
          s->arena.release(q.out_off, q.out_bytes);
          continue;
        }
        s->completed++;
        s->latencies_us.push_back(std::chrono::duration<double, std::micro>(q.done - q.arrival).count());
        s->bytes_done += request_bytes(q.r);
        if (!s->any || q.done > s->last_done) s->last_done = q.done;
        if (!q.keep) {
          s->arena.release(q.out_off, q.out_bytes);
          s->reqs.erase(q.id);  // the round's reference (and a waiter's) keep q alive
        }
      }
      round.clear();
      s->cv.notify_all();
      continue;
    }
    if (s->stop && s->pending.empty()) break;
    if (s->pending.empty()) {
      s->cv.wait_for(lk, std::chrono::milliseconds(5));
    } else if (s->reqs[s->pending.front()]->arrival - clk::now() > std::chrono::microseconds(300)) {
      // idle until shortly before the next arrival (timed waits overshoot by
      // tens of microseconds), then poll
      s->cv.wait_until(lk, s->reqs[s->pending.front()]->arrival - std::chrono::microseconds(200));
    } else {
      // the next request is due, or does not fit yet: poll again at once
      lk.unlock();
      std::this_thread::yield();
      lk.lock();
    }
  }
}

}  // namespace acc

extern "C" {

int acc_hip_server_create(int device, int slots, size_t hbm_budget, acc_server_t* out) {
  if (!out || slots < 1 || slots > 64) return fail(ACC_ERROR_INVALID_VALUE, "server: 1..64 slots");
  ACC_CHECK(set_device(device));
  size_t fr = 0, tot = 0;
  ACC_CHECK(acc_hip_mem_get_info(device, &fr, &tot));
  size_t cap = hbm_budget ? std::min(hbm_budget, fr) : fr / 2;
  cap = cap / 256 * 256;
  auto s = std::make_unique<acc_server_s>();
  s->device = device;
  s->rounds_buf.resize(slots);
  if (cap) ACC_CHECK(dev_malloc(device, cap, &s->arena_base));
  s->arena.init(cap);
  for (int k = 0; k < slots; ++k) {
    acc_stream_t st;
    int rc = acc_hip_stream_create(device, &st);
    if (rc) return rc;
    s->slots.push_back(st);
    // skeleton workspace up front (enough for foldSeg over ~2^29 elements),
    // so that no request grows it while others are in flight
    ACC_CHECK(acc_hip_stream_reserve(st, (size_t)16 << 20));
    // round buffers up front (host staging and device copy of a round's
    // weights, descriptors and run map; the batched kernel's workspace,
    // zeroed now): rounds of up to ~2 MiB of those allocate nothing, where a
    // pinned host allocation alone costs ~0.1-0.3 ms
    RoundBufs& rb = s->rounds_buf[k];
    ACC_CHECK(grow_host(&rb.h, &rb.h_cap, (size_t)1 << 20));
    ACC_CHECK(grow_device(device, &rb.d, &rb.d_cap, (size_t)1 << 20));
    ACC_CHECK(grow_device(device, &rb.ws, &rb.ws_cap, (size_t)1 << 20));
    ACC_HIP_CHECK(hipMemset(rb.ws, 0, rb.ws_cap));
    // (the zeroing runs on the null stream, which does not order the slot's
    // non-blocking stream: it completes before any round can launch)
    ACC_HIP_CHECK(hipStreamSynchronize(nullptr));
    hipEvent_t e;
    ACC_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    s->slot_event.push_back(e);
    s->slot_req.push_back(0);
  }
  acc_server_s* raw = s.release();
  for (int k = 0; k < slots; ++k) raw->workers.emplace_back(slot_worker, raw, k);
  *out = raw;
  return ACC_SUCCESS;
}

int acc_hip_server_configure(acc_server_t s, const acc_server_config* cfg) {
  if (!s || !cfg) return fail(ACC_ERROR_INVALID_VALUE, "server: null argument");
  acc_server_config c = *cfg;
  if (c.fuse != 0 && c.fuse != 1) return fail(ACC_ERROR_INVALID_VALUE, "server: fuse must be 0 or 1");
  if (!c.round_requests) c.round_requests = 256;
  if (!c.round_bytes) c.round_bytes = (int64_t)1 << 30;
  if (!c.block) c.block = 256;
  if (!c.vectors) c.vectors = 8;
  if (!c.run) c.run = 4;
  if (c.round_requests < 1 || c.round_bytes < 1 || c.run < 1 || c.run > 64)
    return fail(ACC_ERROR_INVALID_VALUE, "server: round limits and run length must be positive (run <= 64)");
  if (!batch_shape_supported(c.block, c.vectors))
    return fail(ACC_ERROR_INVALID_VALUE, "server: no batched kernel of %d threads x %d vectors", c.block, c.vectors);
  c.reserved = 0;
  std::lock_guard<std::mutex> lk(s->mu);
  s->cfg = c;  // rounds formed from now on use it
  return ACC_SUCCESS;
}

int acc_hip_server_submit(acc_server_t s, const acc_request* r, int64_t arrival_offset_ns, int keep_output,
                          uint64_t* id) {
  if (!s || !r || !id) return fail(ACC_ERROR_INVALID_VALUE, "server: null argument");
  if (r->skeleton < ACC_REQ_FOLD || r->skeleton > ACC_REQ_STENCIL2D)
    return fail(ACC_ERROR_UNSUPPORTED, "server: request kind %d", r->skeleton);
  if (!type_size((acc_type)r->type)) return fail(ACC_ERROR_UNSUPPORTED, "server: element type %d", r->type);
  auto q = std::make_shared<Req>();
  q->r = *r;
  q->keep = keep_output != 0;
  q->out_bytes = request_out_bytes(*r);
  if (q->out_bytes > s->arena.capacity())
    return fail(ACC_ERROR_OUT_OF_MEMORY, "server: output of %zu bytes exceeds the arena (%zu)", q->out_bytes,
                s->arena.capacity());
  if (r->skeleton == ACC_REQ_STENCIL2D) {
    if (!r->weights) return fail(ACC_ERROR_INVALID_VALUE, "server: stencil weights");
    const size_t nb = (size_t)(2 * r->ry + 1) * (2 * r->rx + 1) * type_size((acc_type)r->type);
    q->weights.assign((const uint8_t*)r->weights, (const uint8_t*)r->weights + nb);
  }
  q->arrival = clk::now() + std::chrono::nanoseconds(arrival_offset_ns > 0 ? arrival_offset_ns : 0);
  std::lock_guard<std::mutex> lk(s->mu);
  q->id = s->next_id++;
  *id = q->id;
  if (!s->any || q->arrival < s->first_arrival) s->first_arrival = q->arrival;
  s->any = true;
  // keep pending ordered by arrival time
  auto pos = s->pending.end();
  while (pos != s->pending.begin() && s->reqs[*std::prev(pos)]->arrival > q->arrival) --pos;
  s->pending.insert(pos, q->id);
  s->reqs[q->id] = std::move(q);
  s->cv.notify_all();
  return ACC_SUCCESS;
}

int acc_hip_server_wait(acc_server_t s, uint64_t id, void** out, double* latency_us) {
  if (!s) return fail(ACC_ERROR_INVALID_VALUE, "server: null");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return fail(ACC_ERROR_INVALID_VALUE, "server: unknown or released request %llu",
                                       (unsigned long long)id);
  // the waiter's own reference keeps the record alive when the worker
  // retires a request submitted with keep_output == 0
  std::shared_ptr<Req> q = it->second;
  s->cv.wait(lk, [&] { return q->finished; });
  if (q->status) return fail(q->status, "server: request %llu failed", (unsigned long long)id);
  // a result that was not kept is reused once complete: no pointer to it
  if (out) *out = q->keep ? (char*)s->arena_base + q->out_off : nullptr;
  if (latency_us) *latency_us = std::chrono::duration<double, std::micro>(q->done - q->arrival).count();
  return ACC_SUCCESS;
}

int acc_hip_server_release(acc_server_t s, uint64_t id) {
  if (!s) return fail(ACC_ERROR_INVALID_VALUE, "server: null");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end() || !it->second->finished)
    return fail(ACC_ERROR_INVALID_VALUE, "server: request %llu not finished", (unsigned long long)id);
  if (it->second->status == ACC_SUCCESS) s->arena.release(it->second->out_off, it->second->out_bytes);
  s->reqs.erase(it);
  s->cv.notify_all();
  return ACC_SUCCESS;
}

int acc_hip_server_drain(acc_server_t s) {
  if (!s) return fail(ACC_ERROR_INVALID_VALUE, "server: null");
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv.wait(lk, [&] {
    bool busy = false;
    for (auto id : s->slot_req) busy |= id != 0;
    return !busy && s->pending.empty();
  });
  return ACC_SUCCESS;
}

int acc_hip_server_stats(acc_server_t s, acc_server_stats* st) {
  if (!s || !st) return fail(ACC_ERROR_INVALID_VALUE, "server: null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  memset(st, 0, sizeof *st);
  st->completed = s->completed;
  st->failed = s->failed;
  std::vector<double> l = s->latencies_us;
  if (!l.empty()) {
    std::sort(l.begin(), l.end());
    auto pct = [&](double p) { return l[std::min(l.size() - 1, (size_t)(p * (l.size() - 1) + 0.5))]; };
    st->p50_us = pct(0.50);
    st->p99_us = pct(0.99);
    st->max_us = l.back();
    double sum = 0;
    for (double v : l) sum += v;
    st->mean_us = sum / l.size();
    const double span = std::chrono::duration<double>(s->last_done - s->first_arrival).count();
    st->span_s = span;
    if (span > 0) {
      st->req_per_s = l.size() / span;
      st->bytes_per_s = s->bytes_done / span;
    }
  }
  st->arena_bytes = s->arena.capacity();
  st->arena_used = s->arena.used();
  st->rounds = s->rounds;
  st->fused = s->fused;
  return ACC_SUCCESS;
}

int acc_hip_server_reset_stats(acc_server_t s) {
  if (!s) return fail(ACC_ERROR_INVALID_VALUE, "server: null");
  std::lock_guard<std::mutex> lk(s->mu);
  s->latencies_us.clear();
  s->completed = s->failed = 0;
  s->rounds = s->fused = 0;
  s->bytes_done = 0;
  s->any = false;
  return ACC_SUCCESS;
}

int acc_hip_server_destroy(acc_server_t s) {
  if (!s) return ACC_SUCCESS;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv.notify_all();
  }
  for (auto& w : s->workers)
    if (w.joinable()) w.join();
  for (auto st : s->slots) acc_hip_stream_destroy(st);
  for (auto e : s->slot_event) (void)hipEventDestroy(e);
  for (auto& b : s->rounds_buf) {
    if (b.h) (void)hipHostFree(b.h);
    if (b.d) dev_free(s->device, b.d);
    if (b.ws) dev_free(s->device, b.ws);
  }
  if (s->arena_base) dev_free(s->device, s->arena_base);
  delete s;
  return ACC_SUCCESS;
}

}  // extern "C"
