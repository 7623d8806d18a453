// Caching allocators of the library: device memory, pinned host memory and
// streams.
//
// hipMalloc, hipFree, hipHostFree and hipStreamDestroy synchronise with the
// whole device, so a job server that starts and releases RVEs while others
// compute would stall every running job at each of them.  Memory the library
// releases is therefore kept, per rounded size, and handed out again; the
// runtime is called only when no cached block of the size exists, when the
// cache is trimmed, or when an allocation fails (then the cache is emptied
// and the allocation tried again).  Callers return a block only after the
// work that used it has completed (the owners synchronise their stream
// first), so a block handed out again is never still in use.
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "common.h"

namespace sca {
namespace {

constexpr size_t GRAIN = 4096;  // sizes are rounded up to this

struct Cache {
  std::mutex m;
  std::map<size_t, std::vector<void*>> free_blocks;  // rounded size -> blocks
  std::unordered_map<void*, size_t> live;            // block -> rounded size
  size_t cached = 0;                                  // bytes in free_blocks
};
Cache& dev_cache() {
  static Cache* c = new Cache();  // never destroyed: blocks live until the process ends
  return *c;
}
Cache& host_cache() {
  static Cache* c = new Cache();
  return *c;
}

size_t rounded(size_t bytes) { return ((bytes == 0 ? 1 : bytes) + GRAIN - 1) / GRAIN * GRAIN; }

// frees cached blocks (largest first) until at most keep bytes are cached;
// called with the lock held
void trim_locked(Cache& c, size_t keep, bool host) {
  for (auto it = c.free_blocks.rbegin(); it != c.free_blocks.rend() && c.cached > keep; ++it) {
    auto& v = it->second;
    while (!v.empty() && c.cached > keep) {
      if (host) (void)hipHostFree(v.back());
      else (void)hipFree(v.back());
      c.cached -= it->first;
      v.pop_back();
    }
  }
}

void* take(Cache& c, size_t bytes, bool host) {
  const size_t r = rounded(bytes);
  std::lock_guard<std::mutex> lk(c.m);
  auto it = c.free_blocks.find(r);
  if (it != c.free_blocks.end() && !it->second.empty()) {
    void* p = it->second.back();
    it->second.pop_back();
    c.cached -= r;
    c.live[p] = r;
    return p;
  }
  void* p = nullptr;
  hipError_t e = host ? hipHostMalloc(&p, r, hipHostMallocDefault) : hipMalloc(&p, r);
  if (e != hipSuccess) {  // give the cached memory back and try once more
    (void)hipGetLastError();
    trim_locked(c, 0, host);
    p = nullptr;
    e = host ? hipHostMalloc(&p, r, hipHostMallocDefault) : hipMalloc(&p, r);
  }
  SCA_HIP(e);
  c.live[p] = r;
  return p;
}

void give(Cache& c, void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(c.m);
  auto it = c.live.find(p);
  if (it == c.live.end()) throw std::runtime_error("block not from the library's allocator");
  c.free_blocks[it->second].push_back(p);
  c.cached += it->second;
  c.live.erase(it);
}

std::mutex g_stream_mu;
std::vector<hipStream_t>* g_streams = new std::vector<hipStream_t>();

}  // namespace

void* dev_alloc(size_t bytes) { return take(dev_cache(), bytes, false); }
void dev_free(void* p) { give(dev_cache(), p); }
void* host_alloc(size_t bytes) { return take(host_cache(), bytes, true); }
void host_free(void* p) { give(host_cache(), p); }

void pool_trim(size_t keep_bytes) {
  Cache& c = dev_cache();
  std::lock_guard<std::mutex> lk(c.m);
  trim_locked(c, keep_bytes, false);
}
size_t pool_cached_bytes() {
  Cache& c = dev_cache();
  std::lock_guard<std::mutex> lk(c.m);
  return c.cached;
}

hipStream_t stream_get() {
  {
    std::lock_guard<std::mutex> lk(g_stream_mu);
    if (!g_streams->empty()) {
      hipStream_t s = g_streams->back();
      g_streams->pop_back();
      return s;
    }
  }
  hipStream_t s = nullptr;
  SCA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));  // default priority
  return s;
}
void stream_put(hipStream_t s) {
  if (!s) return;
  std::lock_guard<std::mutex> lk(g_stream_mu);
  g_streams->push_back(s);
}

}  // namespace sca
