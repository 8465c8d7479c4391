// Host-side infrastructure shared by every translation unit of librtspmspm:
// error propagation (C++ exceptions inside the library, codes at the C ABI),
// the device memory pool, the library stream and per-kernel event timing.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rt_spmspm.h"

namespace rt {

using i64 = int64_t;
using i32 = int32_t;
using u32 = uint32_t;
using u64 = uint64_t;

// Exception carrying an RT_E* code; converted to a return code at the ABI.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(int code, const std::string& msg);
void set_last_error(const std::string& msg);

#define RT_HIP(call)                                                         \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      ::rt::fail(e_ == hipErrorOutOfMemory ? RT_E_NOMEM : RT_E_HIP,          \
                 std::string(#call) + ": " + hipGetErrorString(e_) + " at " + \
                     __FILE__ + ":" + std::to_string(__LINE__));             \
  } while (0)

#define RT_CHECK(cond, code, msg)            \
  do {                                       \
    if (!(cond)) ::rt::fail((code), (msg));  \
  } while (0)

// Host checks of an rt_csr argument (capi.cpp): the header fields, and the
// full structure of a host CSR (row_ptr from 0 to nnz, monotone; columns in
// range, sorted and unique within each row).  Both throw RT_E_INVALID.
void check_csr_header(const rt_csr* M, const char* name);
void validate_host(const rt_csr* M, const char* name);

// ---- device properties ---------------------------------------------------
struct DeviceInfo {
  int id = 0;
  int cus = 0;
  size_t lds_per_block = 0;
  size_t total_mem = 0;
  std::string name;
};
const DeviceInfo& device();

// ---- pool allocator -------------------------------------------------------
// Caching allocator over hipMalloc: freed blocks are kept and reused for a
// later request of at most 1.25x their size, so repeated products of the same
// shape (benchmark steps, the learned path's temp buffers) never go back to
// the driver.  Thread safe.
void* dmalloc(size_t bytes);
void dfree(void* p);
void pool_trim();
size_t pool_cached_bytes();
void pool_stats(i64& live_bytes, i64& live_blocks, i64& cached_bytes);
// Test hook: the n-th dmalloc from now fails with RT_E_NOMEM (0 disarms).
void pool_fail_after(i64 n);

// RAII device buffer from the pool.
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  explicit DBuf(size_t count) { alloc(count); }
  void alloc(size_t count) {
    release();
    n = count;
    p = count ? static_cast<T*>(dmalloc(count * sizeof(T))) : nullptr;
  }
  void release() {
    if (p) dfree(p);
    p = nullptr;
    n = 0;
  }
  T* release_ownership() {
    T* q = p;
    p = nullptr;
    n = 0;
    return q;
  }
  ~DBuf() { release(); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
};

// ---- stream and kernel timing --------------------------------------------
hipStream_t lib_stream();

// When profiling is on, launch() brackets the kernel with hipEvents recorded
// on `s` and accumulates its duration under `name`.  The switch and the open
// bracket are per host thread, so that concurrent calls on different streams
// time only their own launches; the accumulated table is shared.
struct Prof {
  static thread_local bool enabled;
  static void begin(const char* name, hipStream_t s);
  static void end(hipStream_t s);
  static void flush();  // resolve pending event pairs (synchronises)
  // Adds an already measured duration (e.g. a whole captured graph).
  static void record(const char* name, double ms);
};

struct ProfScope {
  bool on;
  hipStream_t s;
  ProfScope(const char* name, hipStream_t st) : on(Prof::enabled), s(st) {
    if (on) Prof::begin(name, s);
  }
  ~ProfScope() {
    if (on) Prof::end(s);
  }
};

// Copy helpers on the library stream (synchronous for D2H of small values).
template <class T>
T read_scalar(const T* dptr, hipStream_t s) {
  T v{};
  RT_HIP(hipMemcpyAsync(&v, dptr, sizeof(T), hipMemcpyDeviceToHost, s));
  RT_HIP(hipStreamSynchronize(s));
  return v;
}

}  // namespace rt
