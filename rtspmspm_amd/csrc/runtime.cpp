// Device properties, pool allocator, library stream and the
// per-kernel hipEvent timer behind rt_prof_*.
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>

#include "common.hpp"

namespace rt {

const DeviceInfo& device() {
  static DeviceInfo info;
  static std::once_flag once;
  static int init_err = 0;
  std::call_once(once, [] {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
      init_err = 1;
      return;
    }
    hipDeviceProp_t p;
    if (hipGetDevice(&info.id) != hipSuccess || hipGetDeviceProperties(&p, info.id) != hipSuccess) {
      init_err = 1;
      return;
    }
    info.cus = p.multiProcessorCount;
    info.lds_per_block = p.sharedMemPerBlock;
    info.total_mem = p.totalGlobalMem;
    info.name = p.gcnArchName;
  });
  if (init_err) fail(RT_E_NODEVICE, "no HIP device visible");
  return info;
}

// ---- pool -----------------------------------------------------------------
namespace {
struct Pool {
  std::mutex mu;
  std::multimap<size_t, void*> free_blocks;  // size -> ptr
  std::map<void*, size_t> live;              // ptr -> size
  size_t cached = 0;
  size_t live_bytes = 0;
  i64 fail_after = 0;  // test hook: > 0 counts down to one injected failure
};
Pool& pool() {
  static Pool* p = new Pool();  // never destroyed: blocks outlive static dtors
  return *p;
}
size_t round_size(size_t b) {
  const size_t g = b < (size_t(1) << 20) ? 256 : (size_t(2) << 20);
  return (b + g - 1) / g * g;
}
}  // namespace

void* dmalloc(size_t bytes) {
  if (bytes == 0) bytes = 1;
  size_t sz = round_size(bytes);
  Pool& P = pool();
  {
    std::lock_guard<std::mutex> g(P.mu);
    if (P.fail_after > 0 && --P.fail_after == 0)
      fail(RT_E_NOMEM, "device allocation of " + std::to_string(sz) +
                           " bytes failed (injected by rt_debug_fail_alloc)");
    auto it = P.free_blocks.lower_bound(sz);
    if (it != P.free_blocks.end() && it->first <= sz + sz / 4) {
      void* ptr = it->second;
      size_t got = it->first;
      P.free_blocks.erase(it);
      P.cached -= got;
      P.live[ptr] = got;
      P.live_bytes += got;
      return ptr;
    }
  }
  void* ptr = nullptr;
  hipError_t e = hipMalloc(&ptr, sz);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    pool_trim();  // give cached blocks back and retry once
    e = hipMalloc(&ptr, sz);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      fail(RT_E_NOMEM, "device allocation of " + std::to_string(sz) + " bytes failed");
    }
  }
  std::lock_guard<std::mutex> g(P.mu);
  P.live[ptr] = sz;
  P.live_bytes += sz;
  return ptr;
}

void dfree(void* ptr) {
  if (!ptr) return;
  Pool& P = pool();
  std::lock_guard<std::mutex> g(P.mu);
  auto it = P.live.find(ptr);
  if (it == P.live.end()) return;  // not ours
  P.free_blocks.emplace(it->second, ptr);
  P.cached += it->second;
  P.live_bytes -= it->second;
  P.live.erase(it);
}

void pool_trim() {
  Pool& P = pool();
  std::lock_guard<std::mutex> g(P.mu);
  if (!P.free_blocks.empty()) (void)hipDeviceSynchronize();
  for (auto& kv : P.free_blocks) (void)hipFree(kv.second);
  P.free_blocks.clear();
  P.cached = 0;
}

size_t pool_cached_bytes() {
  std::lock_guard<std::mutex> g(pool().mu);
  return pool().cached;
}

void pool_stats(i64& live_bytes, i64& live_blocks, i64& cached_bytes) {
  Pool& P = pool();
  std::lock_guard<std::mutex> g(P.mu);
  live_bytes = (i64)P.live_bytes;
  live_blocks = (i64)P.live.size();
  cached_bytes = (i64)P.cached;
}

void pool_fail_after(i64 n) {
  Pool& P = pool();
  std::lock_guard<std::mutex> g(P.mu);
  P.fail_after = n > 0 ? n : 0;
}

// ---- stream ------------------------------------------------------------------
hipStream_t lib_stream() {
  static hipStream_t s = nullptr;
  static std::once_flag once;
  std::call_once(once, [] {
    device();
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;
  });
  return s;
}

// ---- profiling ---------------------------------------------------------------
thread_local bool Prof::enabled = false;
namespace {
struct ProfEntry {
  i64 launches = 0;
  double ms = 0;
};
struct Pending {
  std::string name;
  hipEvent_t a, b;
};
std::mutex g_prof_mu;
std::vector<std::pair<std::string, ProfEntry>> g_prof;
std::vector<Pending> g_pending;
std::vector<hipEvent_t> g_event_cache;
thread_local std::string g_cur_name;  // the open bracket of this thread
thread_local hipEvent_t g_cur_start;

hipEvent_t get_event() {
  if (!g_event_cache.empty()) {
    hipEvent_t e = g_event_cache.back();
    g_event_cache.pop_back();
    return e;
  }
  hipEvent_t e;
  RT_HIP(hipEventCreate(&e));
  return e;
}
ProfEntry& entry(const std::string& n) {
  for (auto& kv : g_prof)
    if (kv.first == n) return kv.second;
  g_prof.emplace_back(n, ProfEntry{});
  return g_prof.back().second;
}
}  // namespace

void Prof::begin(const char* name, hipStream_t s) {
  std::lock_guard<std::mutex> g(g_prof_mu);
  g_cur_name = name;
  g_cur_start = get_event();
  RT_HIP(hipEventRecord(g_cur_start, s));
}

void Prof::end(hipStream_t s) {
  std::lock_guard<std::mutex> g(g_prof_mu);
  hipEvent_t e = get_event();
  RT_HIP(hipEventRecord(e, s));
  g_pending.push_back({g_cur_name, g_cur_start, e});
}

void Prof::flush() {
  std::lock_guard<std::mutex> g(g_prof_mu);
  for (auto& p : g_pending) {
    RT_HIP(hipEventSynchronize(p.b));
    float ms = 0;
    RT_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    ProfEntry& e = entry(p.name);
    e.launches += 1;
    e.ms += ms;
    g_event_cache.push_back(p.a);
    g_event_cache.push_back(p.b);
  }
  g_pending.clear();
}

void Prof::record(const char* name, double ms) {
  std::lock_guard<std::mutex> g(g_prof_mu);
  ProfEntry& e = entry(name);
  e.launches += 1;
  e.ms += ms;
}

void prof_reset() {
  Prof::flush();
  std::lock_guard<std::mutex> g(g_prof_mu);
  g_prof.clear();
}
int prof_count() {
  Prof::flush();
  std::lock_guard<std::mutex> g(g_prof_mu);
  return (int)g_prof.size();
}
bool prof_get(int i, std::string& name, i64& launches, double& ms) {
  std::lock_guard<std::mutex> g(g_prof_mu);
  if (i < 0 || i >= (int)g_prof.size()) return false;
  name = g_prof[i].first;
  launches = g_prof[i].second.launches;
  ms = g_prof[i].second.ms;
  return true;
}

}  // namespace rt
