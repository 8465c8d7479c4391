// Numeric-only re-run on a fixed structure (rt_plan_*).
//
// The reference's comparator can be built to use cuSPARSE's SpGEMMreuse API
// (cuSparse/src/main.cpp:98-363): symbolic phase once (:213-262), numeric
// phase twice with values re-uploaded in between (:264-296).  A plan keeps
// what the exact method learns about the structure -- exact row nnz (and so
// C's row pointer), and the long rows' per-tile counts -- and reruns only the
// numeric kernels for new values.
#include <memory>

#include "common.hpp"
#include "spgemm.hpp"

struct rt_plan {
  rt::Csr A, B;  // structure borrowed from the caller; values replaced per run
  rt::DBuf<rt::i64> nnz;
  rt::HeavyRows heavy;
  rt::Tuning tune;
  hipStream_t stream = nullptr;
  float* c_val = nullptr;  // identifies the C this plan filled
};

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
}  // namespace

extern "C" {

int rt_plan_create(const rt_csr* A, const rt_csr* B, rt_plan** plan, rt_csr* C, void* stream) {
  if (plan) *plan = nullptr;
  if (C) *C = rt_csr{};
  return guard([&] {
    RT_CHECK(A && B && plan && C, RT_E_INVALID, "NULL argument");
    check_csr_header(A, "A");
    check_csr_header(B, "B");
    RT_CHECK(A->mem == RT_MEM_DEVICE && B->mem == RT_MEM_DEVICE, RT_E_INVALID,
             "plans take device operands");
    RT_CHECK(A->cols == B->rows, RT_E_SHAPE, "shape mismatch");
    RT_CHECK(A->rows < (i64)INT32_MAX, RT_E_LIMIT, "A has too many rows");
    device();
    auto p = std::make_unique<rt_plan>();
    p->stream = stream ? static_cast<hipStream_t>(stream) : lib_stream();
    p->A = view(A);
    p->B = view(B);
    validate_csr_dev(p->A, A->cols, nullptr, p->stream, "A");
    validate_csr_dev(p->B, B->cols, nullptr, p->stream, "B");
    const i64 m = A->rows;
    p->nnz.alloc(std::max<i64>(m, 1));
    DBuf<i64> prod(std::max<i64>(m, 1));
    row_products(p->A, p->B, nullptr, prod.p, p->stream, nullptr);
    RT_HIP(hipMemsetAsync(p->nnz.p, 0, std::max<i64>(m, 1) * sizeof(i64), p->stream));
    symbolic_pass(p->A, p->B, nullptr, prod.p, p->nnz.p, p->tune, p->stream, p->heavy);
    CsrOut Cd;
    try {
      alloc_c(Cd, p->nnz.p, m, B->cols, p->stream);
      numeric_pass(p->A, p->B, nullptr, p->nnz.p, p->heavy, Cd, p->tune, p->stream);
    } catch (...) {
      (void)hipStreamSynchronize(p->stream);
      dfree(Cd.rp);
      dfree(Cd.col);
      dfree(Cd.val);
      throw;
    }
    C->rows = Cd.rows;
    C->cols = Cd.cols;
    C->nnz = Cd.nnz;
    C->row_ptr = Cd.rp;
    C->col = Cd.col;
    C->val = Cd.val;
    C->mem = RT_MEM_DEVICE;
    C->owner = 1;
    p->c_val = Cd.val;
    *plan = p.release();
  });
}

int rt_plan_execute(rt_plan* p, const float* a_val, const float* b_val, rt_csr* C, double* ms) {
  return guard([&] {
    RT_CHECK(p && C && a_val && b_val, RT_E_INVALID, "NULL argument");
    RT_CHECK(C->val == p->c_val && C->rows == p->A.rows && C->cols == p->B.cols, RT_E_INVALID,
             "C is not the result this plan created");
    Csr A = p->A, B = p->B;
    A.val = a_val;
    B.val = b_val;
    CsrOut Cd;
    Cd.rows = C->rows;
    Cd.cols = C->cols;
    Cd.nnz = C->nnz;
    Cd.rp = C->row_ptr;
    Cd.col = C->col;
    Cd.val = C->val;
    hipEvent_t e0, e1;
    RT_HIP(hipEventCreate(&e0));
    RT_HIP(hipEventCreate(&e1));
    RT_HIP(hipEventRecord(e0, p->stream));
    numeric_pass(A, B, nullptr, p->nnz.p, p->heavy, Cd, p->tune, p->stream);
    RT_HIP(hipEventRecord(e1, p->stream));
    RT_HIP(hipEventSynchronize(e1));
    float t = 0;
    RT_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms) *ms = t;
  });
}

void rt_plan_destroy(rt_plan* p) {
  if (!p) return;
  (void)hipStreamSynchronize(p->stream);
  delete p;
}

}  // extern "C"
