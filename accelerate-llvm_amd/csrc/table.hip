// The remote memory table: host arrays (by a key the host chooses, its
// identity) mapped to device buffers, in the host layer of the backend.
//
// The reference keeps one per context (PTX/Array/Table.hs:35-49, over the
// core's LRU memory table): `use` uploads an array and marks it frozen (the
// host copy stays valid, so the device copy may be evicted and uploaded again,
// PTX/Array/Prim.hs:89-100); `allocateRemote` makes a buffer for a result,
// whose device copy is the only one until it is copied back; when an
// allocation runs out of memory it evicts the least recently used arrays and
// retries (PTX/Array/Remote.hs:63-74); results are copied back lazily, only
// when the host asks and only if the device copy is newer
// (PTX/Array/Data.hs:91-152).  Allocation granularity is 4096 bytes
// (PTX/Array/Remote.hs:101); freed buffers wait in a nursery for reuse by an
// allocation of the same size, since a device free synchronises the device.
//
// An entry used by queued work carries an event (acc_hip_table_mark); it is
// not evicted, copied or recycled before that event has completed.
#include <hip/hip_runtime.h>
#include <list>
#include <map>
#include <mutex>
#include <unordered_map>
#include "host.hpp"

namespace acc {

constexpr size_t kGranule = 4096;

struct TEntry {
  void* dev = nullptr;
  size_t bytes = 0, alloc = 0;
  void* host = nullptr;   // host copy / copy-back target (may be null for a result)
  bool dirty = false;     // the device copy is newer than the host copy
  bool frozen = false;    // a `use`d array: the host copy is the original
  hipEvent_t ev = nullptr;
  bool ev_live = false;   // ev recorded and not yet seen complete
  bool in_lru = false;    // resident: `lru` points at its place in the list
  std::list<uint64_t>::iterator lru;
};

}  // namespace acc

using namespace acc;

struct acc_table_s {
  int device = 0;
  std::mutex mu;
  std::unordered_map<uint64_t, TEntry> map;
  std::list<uint64_t> lru;  // front: most recently used (resident entries only)
  std::multimap<size_t, std::pair<void*, hipEvent_t>> nursery;  // alloc size -> block, its last use
  size_t nursery_bytes = 0;
  acc_table_stats st{};
};

namespace acc {

static void wait_entry(TEntry& e) {
  if (e.ev_live) {
    (void)hipEventSynchronize(e.ev);
    e.ev_live = false;
  }
}

static void flush_nursery(acc_table_t t) {
  for (auto& kv : t->nursery) {
    if (kv.second.second) {
      (void)hipEventSynchronize(kv.second.second);
      (void)hipEventDestroy(kv.second.second);
    }
    (void)dev_free(t->device, kv.second.first);
  }
  t->nursery.clear();
  t->nursery_bytes = 0;
}

// Copy a dirty resident entry back to its host copy (synchronous).
static int copy_back(acc_table_t t, TEntry& e) {
  if (!e.dirty || !e.dev) return ACC_SUCCESS;
  if (!e.host) return fail(ACC_ERROR_INVALID_VALUE, "table: no host copy to write back to");
  wait_entry(e);
  ACC_HIP_CHECK(hipMemcpy(e.host, e.dev, e.bytes, hipMemcpyDeviceToHost));
  e.dirty = false;
  t->st.downloads++;
  t->st.bytes_down += e.bytes;
  return ACC_SUCCESS;
}

// Free the device copy of the least recently used evictable entry.  Returns
// false when nothing can be evicted.
static bool evict_one(acc_table_t t, uint64_t keep) {
  for (auto it = t->lru.rbegin(); it != t->lru.rend(); ++it) {
    const uint64_t k = *it;
    if (k == keep) continue;
    auto mit = t->map.find(k);  // never inserts: references into the map stay valid
    if (mit == t->map.end()) continue;
    TEntry& e = mit->second;
    if (e.dirty && !e.host) continue;  // the device copy is the only one
    if (copy_back(t, e) != ACC_SUCCESS) continue;
    wait_entry(e);
    (void)dev_free(t->device, e.dev);
    e.dev = nullptr;
    e.in_lru = false;
    t->lru.erase(std::next(it).base());
    t->st.evictions++;
    t->st.bytes_evicted += e.alloc;
    t->st.resident_bytes -= e.alloc;
    return true;
  }
  return false;
}

// A device block of `alloc` bytes: from the nursery, else allocated; on out of
// memory the nursery is released, then entries are evicted one at a time.
static int table_malloc(acc_table_t t, size_t alloc, uint64_t keep, void** out) {
  auto nit = t->nursery.find(alloc);
  if (nit != t->nursery.end()) {
    if (nit->second.second) {
      (void)hipEventSynchronize(nit->second.second);
      (void)hipEventDestroy(nit->second.second);
    }
    *out = nit->second.first;
    t->nursery_bytes -= alloc;
    t->nursery.erase(nit);
    t->st.nursery_hits++;
    return ACC_SUCCESS;
  }
  while (true) {
    const int rc = dev_malloc(t->device, alloc, out);
    if (rc != ACC_ERROR_OUT_OF_MEMORY) return rc;
    t->st.oom_retries++;
    if (!t->nursery.empty()) {
      flush_nursery(t);
      continue;
    }
    if (!evict_one(t, keep))
      return fail(ACC_ERROR_OUT_OF_MEMORY, "table: %zu bytes do not fit and nothing can be evicted", alloc);
  }
}

static void touch(acc_table_t t, uint64_t key, TEntry& e) {
  if (e.in_lru) t->lru.erase(e.lru);
  t->lru.push_front(key);
  e.lru = t->lru.begin();
  e.in_lru = true;
}

// Make the entry resident (uploading its host copy if it was evicted).
static int make_resident(acc_table_t t, uint64_t key, TEntry& e, acc_stream_t s) {
  if (e.dev) {
    touch(t, key, e);
    return ACC_SUCCESS;
  }
  if (!e.host) return fail(ACC_ERROR_INVALID_VALUE, "table: entry %llu has no copy anywhere", (unsigned long long)key);
  void* p = nullptr;
  ACC_CHECK(table_malloc(t, e.alloc, key, &p));
  e.dev = p;
  touch(t, key, e);
  t->st.resident_bytes += e.alloc;
  if (e.bytes) {
    ACC_CHECK(acc_hip_memcpy_h2d_async(s, e.dev, e.host, e.bytes));
    t->st.uploads++;
    t->st.bytes_up += e.bytes;
    if (!e.ev) ACC_HIP_CHECK(hipEventCreateWithFlags(&e.ev, hipEventDisableTiming));
    ACC_HIP_CHECK(hipEventRecord(e.ev, s->stream));
    e.ev_live = true;
  }
  return ACC_SUCCESS;
}

static size_t round_up(size_t b) { return (b ? (b + kGranule - 1) / kGranule : 1) * kGranule; }

}  // namespace acc

extern "C" {

int acc_hip_table_create(int device, acc_table_t* out) {
  if (!out) return fail(ACC_ERROR_INVALID_VALUE, "table: null argument");
  const DeviceInfo* di;
  ACC_CHECK(device_info(device, &di));
  auto t = new acc_table_s();
  t->device = device;
  *out = t;
  return ACC_SUCCESS;
}

int acc_hip_table_use(acc_table_t t, uint64_t key, const void* host, size_t bytes, acc_stream_t s, void** dev) {
  if (!t || !s || !dev || (bytes && !host)) return fail(ACC_ERROR_INVALID_VALUE, "table use: bad argument");
  ACC_CHECK(set_device(t->device));
  std::lock_guard<std::mutex> lk(t->mu);
  auto it = t->map.find(key);
  if (it == t->map.end()) {
    TEntry e;
    e.bytes = bytes;
    e.alloc = round_up(bytes);
    e.host = (void*)host;
    e.frozen = true;
    it = t->map.emplace(key, e).first;
  } else if (it->second.bytes != bytes || (it->second.host && it->second.host != host)) {
    return fail(ACC_ERROR_INVALID_VALUE, "table use: key %llu is registered with another array", (unsigned long long)key);
  }
  ACC_CHECK(make_resident(t, key, it->second, s));
  *dev = it->second.dev;
  return ACC_SUCCESS;
}

int acc_hip_table_alloc(acc_table_t t, uint64_t key, void* host, size_t bytes, void** dev) {
  if (!t || !dev) return fail(ACC_ERROR_INVALID_VALUE, "table alloc: bad argument");
  ACC_CHECK(set_device(t->device));
  std::lock_guard<std::mutex> lk(t->mu);
  if (t->map.count(key)) return fail(ACC_ERROR_INVALID_VALUE, "table alloc: key %llu exists", (unsigned long long)key);
  TEntry e;
  e.bytes = bytes;
  e.alloc = round_up(bytes);
  e.host = host;
  e.dirty = true;  // produced on the device
  void* p = nullptr;
  ACC_CHECK(table_malloc(t, e.alloc, key, &p));
  e.dev = p;
  auto it = t->map.emplace(key, e).first;
  touch(t, key, it->second);
  t->st.resident_bytes += e.alloc;
  *dev = p;
  return ACC_SUCCESS;
}

int acc_hip_table_get(acc_table_t t, uint64_t key, acc_stream_t s, void** dev) {
  if (!t || !s || !dev) return fail(ACC_ERROR_INVALID_VALUE, "table get: bad argument");
  ACC_CHECK(set_device(t->device));
  std::lock_guard<std::mutex> lk(t->mu);
  auto it = t->map.find(key);
  if (it == t->map.end()) return fail(ACC_ERROR_INVALID_VALUE, "table get: unknown key %llu", (unsigned long long)key);
  ACC_CHECK(make_resident(t, key, it->second, s));
  *dev = it->second.dev;
  return ACC_SUCCESS;
}

int acc_hip_table_mark(acc_table_t t, uint64_t key, acc_stream_t s, int writes) {
  if (!t || !s) return fail(ACC_ERROR_INVALID_VALUE, "table mark: bad argument");
  ACC_CHECK(set_device(t->device));
  std::lock_guard<std::mutex> lk(t->mu);
  auto it = t->map.find(key);
  if (it == t->map.end() || !it->second.dev)
    return fail(ACC_ERROR_INVALID_VALUE, "table mark: key %llu is not resident", (unsigned long long)key);
  TEntry& e = it->second;
  if (!e.ev) ACC_HIP_CHECK(hipEventCreateWithFlags(&e.ev, hipEventDisableTiming));
  ACC_HIP_CHECK(hipEventRecord(e.ev, s->stream));
  e.ev_live = true;
  if (writes) {
    if (e.frozen) return fail(ACC_ERROR_INVALID_VALUE, "table mark: a used array is read-only");
    e.dirty = true;
  }
  touch(t, key, e);
  return ACC_SUCCESS;
}

int acc_hip_table_peek(acc_table_t t, uint64_t key, int* copied) {
  if (!t) return fail(ACC_ERROR_INVALID_VALUE, "table peek: bad argument");
  ACC_CHECK(set_device(t->device));
  std::lock_guard<std::mutex> lk(t->mu);
  auto it = t->map.find(key);
  if (it == t->map.end()) return fail(ACC_ERROR_INVALID_VALUE, "table peek: unknown key %llu", (unsigned long long)key);
  TEntry& e = it->second;
  const bool need = e.dirty && e.dev;
  if (need) ACC_CHECK(copy_back(t, e));
  else wait_entry(e);
  if (copied) *copied = need ? 1 : 0;
  return ACC_SUCCESS;
}

int acc_hip_table_free(acc_table_t t, uint64_t key) {
  if (!t) return fail(ACC_ERROR_INVALID_VALUE, "table free: bad argument");
  std::lock_guard<std::mutex> lk(t->mu);
  auto it = t->map.find(key);
  if (it == t->map.end()) return ACC_SUCCESS;  // finalisers may run after the entry is gone
  TEntry& e = it->second;
  if (e.dev) {
    if (e.in_lru) t->lru.erase(e.lru);
    t->st.resident_bytes -= e.alloc;
    // recycled once the work that used it is complete (its event travels along)
    t->nursery.emplace(e.alloc, std::make_pair(e.dev, e.ev_live ? e.ev : nullptr));
    t->nursery_bytes += e.alloc;
    if (!e.ev_live && e.ev) (void)hipEventDestroy(e.ev);
  } else if (e.ev) {
    (void)hipEventDestroy(e.ev);
  }
  t->map.erase(it);
  return ACC_SUCCESS;
}

int acc_hip_table_stats(acc_table_t t, acc_table_stats* out) {
  if (!t || !out) return fail(ACC_ERROR_INVALID_VALUE, "table stats: bad argument");
  std::lock_guard<std::mutex> lk(t->mu);
  *out = t->st;
  out->entries = t->map.size();
  out->nursery_bytes = t->nursery_bytes;
  return ACC_SUCCESS;
}

int acc_hip_table_destroy(acc_table_t t) {
  if (!t) return ACC_SUCCESS;
  (void)set_device(t->device);
  {
    std::lock_guard<std::mutex> lk(t->mu);
    for (auto& kv : t->map) {
      TEntry& e = kv.second;
      wait_entry(e);
      if (e.dev) (void)dev_free(t->device, e.dev);
      if (e.ev) (void)hipEventDestroy(e.ev);
    }
    flush_nursery(t);
  }
  delete t;
  return ACC_SUCCESS;
}

}  // extern "C"
