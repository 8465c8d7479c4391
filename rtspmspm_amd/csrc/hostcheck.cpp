// Host checks of the rt_csr arguments of the C ABI (declared in
// common.hpp).  Host code only: the checks run before any upload or launch,
// so a malformed operand is an RT_E_INVALID error and never reaches a kernel
// (the reference validates nothing; a bad MTX line is silently dropped,
// optixSpMSpM.cpp:113-116).
#include <string>

#include "common.hpp"

using namespace rt;

void rt::check_csr_header(const rt_csr* M, const char* name) {
  RT_CHECK(M != nullptr, RT_E_INVALID, std::string(name) + " is NULL");
  RT_CHECK(M->rows >= 0 && M->cols >= 0 && M->nnz >= 0, RT_E_INVALID,
           std::string(name) + " has negative dimensions");
  RT_CHECK(M->mem == RT_MEM_HOST || M->mem == RT_MEM_DEVICE, RT_E_INVALID,
           std::string(name) + " has an unknown memory space");
  RT_CHECK(M->row_ptr != nullptr, RT_E_INVALID, std::string(name) + ".row_ptr is NULL");
  RT_CHECK(M->nnz == 0 || (M->col != nullptr && M->val != nullptr), RT_E_INVALID,
           std::string(name) + " has NULL col/val");
  RT_CHECK(M->cols <= (i64)INT32_MAX, RT_E_LIMIT, std::string(name) + " has more than 2^31-1 columns");
}

// Full structural check of a host CSR: row_ptr monotone from 0 to nnz,
// columns in range and strictly increasing inside each row.
void rt::validate_host(const rt_csr* M, const char* name) {
  RT_CHECK(M->row_ptr[0] == 0 && M->row_ptr[M->rows] == M->nnz, RT_E_INVALID,
           std::string(name) + ": row_ptr must run from 0 to nnz");
  for (i64 i = 0; i < M->rows; ++i) {
    const i64 a = M->row_ptr[i], b = M->row_ptr[i + 1];
    RT_CHECK(a <= b, RT_E_INVALID, std::string(name) + ": row_ptr not monotone");
    for (i64 e = a; e < b; ++e) {
      RT_CHECK(M->col[e] >= 0 && M->col[e] < M->cols, RT_E_INVALID,
               std::string(name) + ": column index out of range");
      RT_CHECK(e == a || M->col[e] > M->col[e - 1], RT_E_INVALID,
               std::string(name) + ": columns must be sorted and unique within a row");
    }
  }
}
