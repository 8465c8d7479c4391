// extern "C" boundary of librtspmspm (declared in include/rt_spmspm.h).
// Every entry point catches the library's exceptions and returns an RT_E*
// code; rt_last_error() carries the message (the reference's equivalent is
// the CUDA_CHECK/OPTIX_CHECK exception caught in main, optixSpMSpM.cpp:654-658).
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "hostcsr.hpp"
#include "learned.hpp"
#include "spgemm.hpp"

namespace rt {
const char* last_error_cstr();
void prof_reset();
int prof_count();
bool prof_get(int i, std::string& name, i64& launches, double& ms);
double device_copy(void* dst, const void* src, i64 bytes, hipStream_t s);
}  // namespace rt

using namespace rt;

namespace {

template <class F>
int guard(F&& f) {
  try {
    set_last_error("");
    f();
    return RT_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return RT_E_NOMEM;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return RT_E_INVALID;
  }
}

void zero_csr(rt_csr* C) {
  if (C) std::memset(C, 0, sizeof(*C));
}

Csr view(const rt_csr* M) {
  Csr v;
  v.rows = M->rows;
  v.cols = M->cols;
  v.nnz = M->nnz;
  v.rp = M->row_ptr;
  v.col = M->col;
  v.val = M->val;
  return v;
}

struct DevCopy {
  DBuf<i64> rp;
  DBuf<i32> col;
  DBuf<float> val;
  Csr v;
  DevCopy(const rt_csr* M, hipStream_t s) {
    rp.alloc(M->rows + 1);
    col.alloc(std::max<i64>(M->nnz, 1));
    val.alloc(std::max<i64>(M->nnz, 1));
    RT_HIP(hipMemcpyAsync(rp.p, M->row_ptr, (M->rows + 1) * sizeof(i64), hipMemcpyHostToDevice, s));
    if (M->nnz) {
      RT_HIP(hipMemcpyAsync(col.p, M->col, M->nnz * sizeof(i32), hipMemcpyHostToDevice, s));
      RT_HIP(hipMemcpyAsync(val.p, M->val, M->nnz * sizeof(float), hipMemcpyHostToDevice, s));
    }
    v.rows = M->rows;
    v.cols = M->cols;
    v.nnz = M->nnz;
    v.rp = rp.p;
    v.col = col.p;
    v.val = val.p;
  }
};

void free_out(CsrOut& C) {
  dfree(C.rp);
  dfree(C.col);
  dfree(C.val);
  C = CsrOut{};
}

void export_device(CsrOut& Cd, rt_csr* C) {
  C->rows = Cd.rows;
  C->cols = Cd.cols;
  C->nnz = Cd.nnz;
  C->row_ptr = Cd.rp;
  C->col = Cd.col;
  C->val = Cd.val;
  C->mem = RT_MEM_DEVICE;
  C->owner = 1;
  Cd = CsrOut{};
}

void export_host(CsrOut& Cd, rt_csr* C, hipStream_t s) {
  HostCsr h;
  h.alloc(Cd.rows, Cd.cols, Cd.nnz);
  RT_HIP(hipMemcpyAsync(h.rp, Cd.rp, (Cd.rows + 1) * sizeof(i64), hipMemcpyDeviceToHost, s));
  if (Cd.nnz) {
    RT_HIP(hipMemcpyAsync(h.col, Cd.col, Cd.nnz * sizeof(i32), hipMemcpyDeviceToHost, s));
    RT_HIP(hipMemcpyAsync(h.val, Cd.val, Cd.nnz * sizeof(float), hipMemcpyDeviceToHost, s));
  }
  RT_HIP(hipStreamSynchronize(s));
  free_out(Cd);
  h.export_to(C);
}

struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
  EventPair() {
    RT_HIP(hipEventCreate(&a));
    RT_HIP(hipEventCreate(&b));
  }
  ~EventPair() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
  double ms() {
    RT_HIP(hipEventSynchronize(b));
    float t = 0;
    RT_HIP(hipEventElapsedTime(&t, a, b));
    return t;
  }
};

// Runs one product with explicit settings (autotuner); returns the device
// ms and leaves the result in C.
double timed_product(const Csr& A, const Csr& B, const Batch& bt, int mode, const Tuning& tu,
                     hipStream_t s, CsrOut& C) {
  RunStats st;
  EventPair ev;
  RT_HIP(hipEventRecord(ev.a, s));
  try {
    // batched calls run their stages as graphs, as rt_spmspm_batched does
    if (mode == RT_MODE_LEARNED && bt.b_base) spgemm_learned_graph(A, B, bt, C, tu, s, &st);
    else if (mode == RT_MODE_LEARNED) spgemm_learned(A, B, bt, C, tu, s, &st);
    else if (bt.b_base) spgemm_plain_graph(A, B, bt, C, tu, s, &st);
    else spgemm_plain(A, B, bt, C, tu, s, &st);
  } catch (...) {
    (void)hipStreamSynchronize(s);
    free_out(C);
    throw;
  }
  RT_HIP(hipEventRecord(ev.b, s));
  return ev.ms();
}

// Runs the selected method on device operands; C receives pool arrays.
void run_device(const Csr& A, const Csr& B, const Batch& bt, CsrOut& C, rt_opts& o,
                hipStream_t s) {
  RunStats st;
  EventPair ev;
  RT_CHECK(o.learned_long >= RT_LONG_DEFAULT && o.learned_long <= RT_LONG_EXACT, RT_E_INVALID,
           "learned_long must be one of RT_LONG_*");
  const bool batched = bt.b_base != nullptr;
  Tuning tu = active_tuning(A, B, o.use_tuned != 0, o.mode, batched);
  if (o.learned_long == RT_LONG_LEARNED) tu.learned_heavy = 1;
  if (o.learned_long == RT_LONG_EXACT) tu.learned_heavy = 0;
  Prof::enabled = o.profile != 0;
  RT_HIP(hipEventRecord(ev.a, s));
  try {
    const bool graph = o.graph == 1 || (o.graph == 0 && batched);
    switch (o.mode) {
      case RT_MODE_PLAIN:
        if (graph) spgemm_plain_graph(A, B, bt, C, tu, s, &st);
        else spgemm_plain(A, B, bt, C, tu, s, &st);
        break;
      case RT_MODE_RAY:
        RT_CHECK(!batched, RT_E_INVALID, "the ray method takes a single product");
        spgemm_ray(A, B, C, o.ablation, o.ray_output, s, &st);
        break;
      case RT_MODE_LEARNED:
        if (graph) spgemm_learned_graph(A, B, bt, C, tu, s, &st);
        else spgemm_learned(A, B, bt, C, tu, s, &st);
        break;
      default: fail(RT_E_INVALID, "unknown mode " + std::to_string(o.mode));
    }
  } catch (...) {
    (void)hipStreamSynchronize(s);  // drain before pool blocks are reused
    free_out(C);
    Prof::enabled = false;
    throw;
  }
  RT_HIP(hipEventRecord(ev.b, s));
  o.ms_total = ev.ms();
  if (Prof::enabled) Prof::flush();
  Prof::enabled = false;
  o.products = st.products;
  o.overflow_rows = st.overflow_rows;
  o.predicted_rows = st.predicted_rows;
}

// Stacked batch operands: per-row B bases and column limits, validated
// (unless trusted) and viewed as one product.
struct BatchView {
  DBuf<i64> base, alim, blim;
  Csr A, B;
  Batch bt;
};

void prepare_batch(int32_t count, const int64_t* a_rows, const int64_t* b_rows,
                   const int64_t* b_cols, const rt_csr* A, const rt_csr* B, bool trusted,
                   hipStream_t s, BatchView& v) {
  RT_CHECK(a_rows && b_rows && b_cols, RT_E_INVALID, "NULL argument");
  RT_CHECK(count >= 1, RT_E_INVALID, "count must be >= 1");
  check_csr_header(A, "A_stack");
  check_csr_header(B, "B_stack");
  RT_CHECK(A->mem == RT_MEM_DEVICE && B->mem == RT_MEM_DEVICE, RT_E_INVALID,
           "batched operands must be device memory");
  i64 ar = 0, br = 0, maxk = 0, maxc = 0;
  for (int b = 0; b < count; ++b) {
    RT_CHECK(a_rows[b] >= 0 && b_rows[b] >= 0 && b_cols[b] >= 0, RT_E_INVALID, "negative pair size");
    ar += a_rows[b];
    br += b_rows[b];
    maxk = std::max<i64>(maxk, b_rows[b]);
    maxc = std::max<i64>(maxc, b_cols[b]);
  }
  RT_CHECK(ar == A->rows && br == B->rows, RT_E_SHAPE, "stacked row counts do not match the pair sizes");
  RT_CHECK(A->cols >= maxk && B->cols >= maxc, RT_E_SHAPE, "stacked column counts too small");
  device();
  v.base.alloc(std::max<i64>(ar, 1));
  {
    std::vector<i64> bstart(count), av(count), bv(count);
    i64 acc = 0;
    for (int b = 0; b < count; ++b) {
      bstart[b] = acc;
      acc += b_rows[b];
      av[b] = b_rows[b];
      bv[b] = b_cols[b];
    }
    fill_segments(count, a_rows, bstart.data(), v.base.p, ar, s);
    if (!trusted) {
      v.alim.alloc(std::max<i64>(ar, 1));
      v.blim.alloc(std::max<i64>(br, 1));
      fill_segments(count, a_rows, av.data(), v.alim.p, ar, s);
      fill_segments(count, b_rows, bv.data(), v.blim.p, br, s);
    }
  }
  v.A = view(A);
  v.B = view(B);
  v.A.cols = maxk;
  if (!trusted) {
    validate_csr_dev(v.A, A->cols, v.alim.p, s, "A_stack");
    validate_csr_dev(v.B, B->cols, v.blim.p, s, "B_stack");
  }
  v.B.cols = maxc;
  v.bt.b_base = v.base.p;
}

// The tuner proper: times the candidates of autotune_search on (A, B, bt),
// rejects any whose C differs from the first candidate's (structure hashes
// exactly; value sums beyond 1e-3 of sum |v|, far above the rounding of a
// different accumulation order), records the fastest under `key`.
std::string tune_operands(const Csr& a, const Csr& b, const Batch& bt, int mode, int reps,
                          const std::string& key, hipStream_t s) {
  bool have_ref = false;
  Fingerprint ref;
  auto eval = [&](const Tuning& t) -> double {
    {
      CsrOut C;  // warm-up (allocations, caches), and the result check
      timed_product(a, b, bt, mode, t, s, C);
      Fingerprint f;
      try {
        f = fingerprint(C, s);
      } catch (...) {
        free_out(C);
        throw;
      }
      free_out(C);
      if (!have_ref) {
        ref = f;
        have_ref = true;
      } else {
        const double tol = 1e-3 * ref.vsum[0] + 1e-30;
        if (f.nnz != ref.nnz || f.hrp != ref.hrp || f.hcol != ref.hcol ||
            std::abs(f.vsum[0] - ref.vsum[0]) > tol || std::abs(f.vsum[1] - ref.vsum[1]) > tol)
          return -1.0;
      }
    }
    double t_ms = 0;
    for (int r = 0; r < reps; ++r) {
      CsrOut C;
      t_ms += timed_product(a, b, bt, mode, t, s, C);
      free_out(C);
    }
    return t_ms / reps;
  };
  std::string log;
  const Tuning best = autotune_search(mode == RT_MODE_LEARNED, eval, log);
  record_tuning(key, best);
  return "class " + key + " " + log + "; " + tuning_report(best);
}

void copy_desc(const std::string& line, char* desc, int64_t desc_len) {
  if (desc && desc_len > 0) {
    std::strncpy(desc, line.c_str(), desc_len - 1);
    desc[desc_len - 1] = '\0';
  }
}

}  // namespace

extern "C" {

const char* rt_last_error(void) { return last_error_cstr(); }

int rt_spmspm(const rt_csr* A, const rt_csr* B, rt_csr* C) {
  return rt_spmspm_ex(A, B, C, nullptr, nullptr);
}

int rt_spmspm_ex(const rt_csr* A, const rt_csr* B, rt_csr* C, rt_opts* opts, void* stream) {
  zero_csr(C);
  return guard([&] {
    RT_CHECK(C != nullptr, RT_E_INVALID, "C is NULL");
    check_csr_header(A, "A");
    check_csr_header(B, "B");
    RT_CHECK(A->cols == B->rows, RT_E_SHAPE,
             "shape mismatch: A is " + std::to_string(A->rows) + "x" + std::to_string(A->cols) +
                 ", B is " + std::to_string(B->rows) + "x" + std::to_string(B->cols));
    RT_CHECK(A->mem == B->mem, RT_E_INVALID, "A and B must live in the same memory space");
    rt_opts local{};
    rt_opts& o = opts ? *opts : local;
    device();
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : lib_stream();
    CsrOut Cd;
    if (A->mem == RT_MEM_HOST) {
      validate_host(A, "A");
      validate_host(B, "B");
      DevCopy dA(A, s), dB(B, s);
      run_device(dA.v, dB.v, Batch{}, Cd, o, s);
      export_host(Cd, C, s);
    } else {
      if (!(o.flags & RT_FLAG_TRUSTED)) {
        validate_csr_dev(view(A), A->cols, nullptr, s, "A");
        validate_csr_dev(view(B), B->cols, nullptr, s, "B");
      }
      run_device(view(A), view(B), Batch{}, Cd, o, s);
      export_device(Cd, C);
    }
  });
}

int rt_spmspm_batched(int32_t count, const int64_t* a_rows, const int64_t* b_rows,
                      const int64_t* b_cols, const rt_csr* A, const rt_csr* B, rt_csr* C,
                      rt_opts* opts, void* stream) {
  zero_csr(C);
  return guard([&] {
    RT_CHECK(C != nullptr, RT_E_INVALID, "NULL argument");
    rt_opts local{};
    rt_opts& o = opts ? *opts : local;
    RT_CHECK(o.mode == RT_MODE_PLAIN || o.mode == RT_MODE_LEARNED, RT_E_INVALID,
             "batched products support the plain and learned methods");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : nullptr;
    if (!s) {
      device();
      s = lib_stream();
    }
    BatchView v;
    prepare_batch(count, a_rows, b_rows, b_cols, A, B, (o.flags & RT_FLAG_TRUSTED) != 0, s, v);
    CsrOut Cd;
    run_device(v.A, v.B, v.bt, Cd, o, s);
    export_device(Cd, C);
  });
}

int rt_spmspm_noalloc(const rt_csr* A, const rt_csr* B, int64_t* c_row_ptr, int32_t* c_col,
                      float* c_val, int64_t c_capacity, int64_t* c_nnz, int32_t flags,
                      void* stream) {
  if (c_nnz) *c_nnz = -1;
  return guard([&] {
    check_csr_header(A, "A");
    check_csr_header(B, "B");
    RT_CHECK(A->mem == RT_MEM_DEVICE && B->mem == RT_MEM_DEVICE, RT_E_INVALID,
             "rt_spmspm_noalloc takes device operands");
    RT_CHECK(A->cols == B->rows, RT_E_SHAPE,
             "shape mismatch: A is " + std::to_string(A->rows) + "x" + std::to_string(A->cols) +
                 ", B is " + std::to_string(B->rows) + "x" + std::to_string(B->cols));
    RT_CHECK(c_row_ptr != nullptr && c_capacity >= 0, RT_E_INVALID,
             "c_row_ptr is NULL or c_capacity negative");
    RT_CHECK(c_capacity == 0 || (c_col && c_val), RT_E_INVALID, "NULL c_col / c_val");
    device();
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : lib_stream();
    i64 n = -1;
    try {
      spgemm_noalloc(view(A), view(B), c_row_ptr, c_col, c_val, c_capacity, &n,
                     (flags & RT_FLAG_TRUSTED) != 0, s);
    } catch (...) {
      if (c_nnz) *c_nnz = n;
      (void)hipStreamSynchronize(s);
      throw;
    }
    if (c_nnz) *c_nnz = n;
  });
}

int rt_autotune(const rt_csr* A, const rt_csr* B, int32_t mode, int32_t reps, char* desc,
                int64_t desc_len) {
  return guard([&] {
    check_csr_header(A, "A");
    check_csr_header(B, "B");
    RT_CHECK(A->mem == RT_MEM_DEVICE && B->mem == RT_MEM_DEVICE, RT_E_INVALID,
             "the autotuner takes device operands");
    RT_CHECK(A->cols == B->rows, RT_E_SHAPE, "shape mismatch");
    RT_CHECK(mode == RT_MODE_PLAIN || mode == RT_MODE_LEARNED, RT_E_INVALID,
             "tune the plain or the learned method");
    RT_CHECK(reps >= 1, RT_E_INVALID, "reps must be >= 1");
    device();
    hipStream_t s = lib_stream();
    const Csr a = view(A), b = view(B);
    validate_csr_dev(a, A->cols, nullptr, s, "A");
    validate_csr_dev(b, B->cols, nullptr, s, "B");
    copy_desc(tune_operands(a, b, Batch{}, mode, reps, tune_key(a, b, mode, false), s), desc,
              desc_len);
  });
}

int rt_autotune_batched(int32_t count, const int64_t* a_rows, const int64_t* b_rows,
                        const int64_t* b_cols, const rt_csr* A, const rt_csr* B, int32_t mode,
                        int32_t reps, char* desc, int64_t desc_len) {
  return guard([&] {
    RT_CHECK(mode == RT_MODE_PLAIN || mode == RT_MODE_LEARNED, RT_E_INVALID,
             "tune the plain or the learned method");
    RT_CHECK(reps >= 1, RT_E_INVALID, "reps must be >= 1");
    device();
    hipStream_t s = lib_stream();
    BatchView v;
    prepare_batch(count, a_rows, b_rows, b_cols, A, B, false, s, v);
    copy_desc(tune_operands(v.A, v.B, v.bt, mode, reps, tune_key(v.A, v.B, mode, true), s), desc,
              desc_len);
  });
}

int rt_tune_cache_check(const char* path, int32_t* valid, int32_t* rejected) {
  return guard([&] {
    RT_CHECK(path != nullptr, RT_E_INVALID, "NULL path");
    int v = 0, r = 0;
    tune_cache_check(path, v, r);
    if (valid) *valid = v;
    if (rejected) *rejected = r;
  });
}

void rt_csr_free(rt_csr* C) {
  if (!C) return;
  if (C->owner == 1) {
    if (C->mem == RT_MEM_DEVICE) {
      dfree(C->row_ptr);
      dfree(C->col);
      dfree(C->val);
    } else {
      std::free(C->row_ptr);
      std::free(C->col);
      std::free(C->val);
    }
  }
  zero_csr(C);
}

int rt_csr_to_device(const rt_csr* h, rt_csr* d) {
  zero_csr(d);
  return guard([&] {
    RT_CHECK(d != nullptr, RT_E_INVALID, "output is NULL");
    check_csr_header(h, "input");
    RT_CHECK(h->mem == RT_MEM_HOST, RT_E_INVALID, "input must be host memory");
    device();
    hipStream_t s = lib_stream();
    CsrOut o;
    o.rows = h->rows;
    o.cols = h->cols;
    o.nnz = h->nnz;
    o.rp = static_cast<i64*>(dmalloc((h->rows + 1) * sizeof(i64)));
    o.col = static_cast<i32*>(dmalloc(std::max<i64>(h->nnz, 1) * sizeof(i32)));
    o.val = static_cast<float*>(dmalloc(std::max<i64>(h->nnz, 1) * sizeof(float)));
    try {
      RT_HIP(hipMemcpyAsync(o.rp, h->row_ptr, (h->rows + 1) * sizeof(i64), hipMemcpyHostToDevice, s));
      if (h->nnz) {
        RT_HIP(hipMemcpyAsync(o.col, h->col, h->nnz * sizeof(i32), hipMemcpyHostToDevice, s));
        RT_HIP(hipMemcpyAsync(o.val, h->val, h->nnz * sizeof(float), hipMemcpyHostToDevice, s));
      }
      RT_HIP(hipStreamSynchronize(s));
    } catch (...) {
      free_out(o);
      throw;
    }
    export_device(o, d);
  });
}

int rt_csr_to_host(const rt_csr* d, rt_csr* h) {
  zero_csr(h);
  return guard([&] {
    RT_CHECK(h != nullptr, RT_E_INVALID, "output is NULL");
    check_csr_header(d, "input");
    RT_CHECK(d->mem == RT_MEM_DEVICE, RT_E_INVALID, "input must be device memory");
    device();
    hipStream_t s = lib_stream();
    HostCsr o;
    o.alloc(d->rows, d->cols, d->nnz);
    RT_HIP(hipMemcpyAsync(o.rp, d->row_ptr, (d->rows + 1) * sizeof(i64), hipMemcpyDeviceToHost, s));
    if (d->nnz) {
      RT_HIP(hipMemcpyAsync(o.col, d->col, d->nnz * sizeof(i32), hipMemcpyDeviceToHost, s));
      RT_HIP(hipMemcpyAsync(o.val, d->val, d->nnz * sizeof(float), hipMemcpyDeviceToHost, s));
    }
    RT_HIP(hipStreamSynchronize(s));
    o.export_to(h);
  });
}

int rt_memcpy_d2h(void* dst, const void* src, int64_t bytes) {
  return guard([&] {
    RT_CHECK(bytes >= 0 && (bytes == 0 || (dst && src)), RT_E_INVALID, "bad copy arguments");
    device();
    if (bytes) {
      hipStream_t s = lib_stream();
      RT_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, s));
      RT_HIP(hipStreamSynchronize(s));
    }
  });
}

int rt_device_copy(void* dst, const void* src, int64_t bytes, double* ms) {
  return guard([&] {
    RT_CHECK(bytes >= 0 && (bytes == 0 || (dst && src)), RT_E_INVALID, "bad copy arguments");
    device();
    const double t = bytes ? device_copy(dst, src, bytes, lib_stream()) : 0.0;
    if (ms) *ms = t;
  });
}

int rt_device_alloc(void** ptr, int64_t bytes) {
  return guard([&] {
    RT_CHECK(ptr && bytes >= 0, RT_E_INVALID, "bad allocation arguments");
    device();
    *ptr = dmalloc((size_t)bytes);
  });
}

int rt_device_free(void* ptr) {
  return guard([&] { dfree(ptr); });
}

int rt_device_synchronize(void) {
  return guard([&] {
    device();
    RT_HIP(hipDeviceSynchronize());
  });
}

int rt_pool_trim(void) {
  return guard([&] { pool_trim(); });
}

int rt_pool_stats(int64_t* live_bytes, int64_t* live_blocks, int64_t* cached_bytes) {
  return guard([&] {
    i64 lb = 0, lk = 0, cb = 0;
    pool_stats(lb, lk, cb);
    if (live_bytes) *live_bytes = lb;
    if (live_blocks) *live_blocks = lk;
    if (cached_bytes) *cached_bytes = cb;
  });
}

int rt_debug_fail_alloc(int64_t n) {
  return guard([&] {
    RT_CHECK(n >= 0, RT_E_INVALID, "n must be >= 0");
    pool_fail_after(n);
  });
}

int rt_stream_create(void** stream) {
  return guard([&] {
    RT_CHECK(stream != nullptr, RT_E_INVALID, "NULL argument");
    *stream = nullptr;
    device();
    hipStream_t st = nullptr;
    RT_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    *stream = st;
  });
}

int rt_stream_destroy(void* stream) {
  return guard([&] {
    if (stream) RT_HIP(hipStreamDestroy(static_cast<hipStream_t>(stream)));
  });
}

int rt_device_info(char* name, int64_t name_len, int64_t* total_mem, int32_t* cus) {
  return guard([&] {
    const DeviceInfo& d = device();
    if (name && name_len > 0) {
      std::strncpy(name, d.name.c_str(), name_len - 1);
      name[name_len - 1] = '\0';
    }
    if (total_mem) *total_mem = (int64_t)d.total_mem;
    if (cus) *cus = d.cus;
  });
}

void rt_prof_reset(void) {
  try {
    prof_reset();
  } catch (...) {
  }
}
int32_t rt_prof_count(void) {
  try {
    return prof_count();
  } catch (...) {
    return 0;
  }
}
int rt_prof_get(int32_t i, char* name, int64_t name_len, int64_t* launches, double* total_ms) {
  return guard([&] {
    std::string n;
    i64 l = 0;
    double ms = 0;
    RT_CHECK(prof_get(i, n, l, ms), RT_E_INVALID, "profile index out of range");
    if (name && name_len > 0) {
      std::strncpy(name, n.c_str(), name_len - 1);
      name[name_len - 1] = '\0';
    }
    if (launches) *launches = l;
    if (total_ms) *total_ms = ms;
  });
}

// ---- generators -------------------------------------------------------------
int rt_gen_uniform_k(int64_t n, int64_t k, uint64_t seed, rt_csr* out) {
  zero_csr(out);
  return guard([&] {
    HostCsr h;
    gen_uniform_k(n, k, seed, h);
    h.export_to(out);
  });
}
int rt_gen_rmat(int32_t scale, int32_t ef, double a, double b, double c, uint64_t seed, rt_csr* out) {
  zero_csr(out);
  return guard([&] {
    HostCsr h;
    gen_rmat(scale, ef, a, b, c, seed, h);
    h.export_to(out);
  });
}
int rt_gen_stencil27(int64_t n, uint64_t seed, rt_csr* out) {
  zero_csr(out);
  return guard([&] {
    HostCsr h;
    gen_stencil27(n, seed, h);
    h.export_to(out);
  });
}
int rt_gen_aggregation(int64_t n, rt_csr* P, rt_csr* R) {
  zero_csr(P);
  zero_csr(R);
  return guard([&] {
    HostCsr hp, hr;
    gen_aggregation(n, hp, hr);
    hp.export_to(P);
    hr.export_to(R);
  });
}
int rt_gen_lognormal_rows(int64_t rows, int64_t cols, double mu, double sigma, int64_t lo,
                          int64_t hi, uint64_t seed, uint64_t stream, rt_csr* out) {
  zero_csr(out);
  return guard([&] {
    HostCsr h;
    gen_lognormal_rows(rows, cols, mu, sigma, lo, hi, seed, stream, h);
    h.export_to(out);
  });
}
int rt_gen_fixed_normal(int64_t rows, int64_t cols, int64_t k, uint64_t seed, uint64_t stream,
                        rt_csr* out) {
  zero_csr(out);
  return guard([&] {
    HostCsr h;
    gen_fixed_normal(rows, cols, k, seed, stream, h);
    h.export_to(out);
  });
}
int rt_gen_poisson(int64_t n, double lambda, uint64_t seed, uint64_t stream, rt_csr* out) {
  zero_csr(out);
  return guard([&] {
    HostCsr h;
    gen_poisson(n, lambda, seed, stream, h);
    h.export_to(out);
  });
}
int rt_gen_batch_poisson(int32_t count, int64_t nmin, int64_t nmax, double lmin, double lmax,
                         uint64_t seed, int64_t* a_rows, int64_t* b_rows, int64_t* b_cols,
                         rt_csr* A, rt_csr* B) {
  zero_csr(A);
  zero_csr(B);
  return guard([&] {
    RT_CHECK(a_rows && b_rows && b_cols && A && B, RT_E_INVALID, "NULL argument");
    HostCsr ha, hb;
    gen_batch_poisson(count, nmin, nmax, lmin, lmax, seed, a_rows, b_rows, b_cols, ha, hb);
    ha.export_to(A);
    hb.export_to(B);
  });
}

// ---- Matrix Market -----------------------------------------------------------
int rt_mtx_read(const char* path, rt_csr* out) {
  zero_csr(out);
  return guard([&] {
    RT_CHECK(path && out, RT_E_INVALID, "NULL argument");
    HostCsr h;
    mtx_read(path, h);
    h.export_to(out);
  });
}
int rt_mtx_write(const char* path, const rt_csr* C, int32_t ref_compat) {
  return guard([&] {
    RT_CHECK(path && C, RT_E_INVALID, "NULL argument");
    check_csr_header(C, "C");
    RT_CHECK(C->mem == RT_MEM_HOST, RT_E_INVALID, "rt_mtx_write takes a host CSR");
    mtx_write(path, *C, ref_compat);
  });
}

}  // extern "C"
