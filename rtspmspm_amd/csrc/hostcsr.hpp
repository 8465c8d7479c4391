// Host CSR with malloc-owned arrays, handed out through the C ABI as an
// rt_csr with owner = 1 (released by rt_csr_free with free()).
#pragma once
#include <cstdlib>

#include "common.hpp"

namespace rt {

struct HostCsr {
  i64 rows = 0, cols = 0, nnz = 0;
  i64* rp = nullptr;
  i32* col = nullptr;
  float* val = nullptr;

  void alloc(i64 r, i64 c, i64 n) {
    release();
    rows = r;
    cols = c;
    rp = static_cast<i64*>(std::malloc((r + 1) * sizeof(i64)));
    if (!rp) fail(RT_E_NOMEM, "host allocation failed");
    set_nnz(n);
  }
  void set_nnz(i64 n) {
    std::free(col);
    std::free(val);
    nnz = n;
    col = static_cast<i32*>(std::malloc((n > 0 ? n : 1) * sizeof(i32)));
    val = static_cast<float*>(std::malloc((n > 0 ? n : 1) * sizeof(float)));
    if (!col || !val) fail(RT_E_NOMEM, "host allocation failed");
  }
  void release() {
    std::free(rp);
    std::free(col);
    std::free(val);
    rp = nullptr;
    col = nullptr;
    val = nullptr;
    rows = cols = nnz = 0;
  }
  // Hand the arrays to an rt_csr (owner = 1, host memory).
  void export_to(rt_csr* o) {
    o->rows = rows;
    o->cols = cols;
    o->nnz = nnz;
    o->row_ptr = rp;
    o->col = col;
    o->val = val;
    o->mem = RT_MEM_HOST;
    o->owner = 1;
    rp = nullptr;
    col = nullptr;
    val = nullptr;
  }
  ~HostCsr() { release(); }
};

void gen_uniform_k(i64 n, i64 k, u64 seed, HostCsr& out);
void gen_fixed_normal(i64 rows, i64 cols, i64 k, u64 seed, u64 stream, HostCsr& out);
void gen_lognormal_rows(i64 rows, i64 cols, double mu, double sigma, i64 lo, i64 hi, u64 seed,
                        u64 stream, HostCsr& out);
void gen_poisson(i64 n, double lambda, u64 seed, u64 stream, HostCsr& out);
void gen_rmat(int scale, int edge_factor, double a, double b, double c, u64 seed, HostCsr& out);
void gen_stencil27(i64 n, u64 seed, HostCsr& out);
void gen_aggregation(i64 n, HostCsr& P, HostCsr& R);
void gen_batch_poisson(i32 count, i64 nmin, i64 nmax, double lmin, double lmax, u64 seed,
                       i64* a_rows, i64* b_rows, i64* b_cols, HostCsr& A, HostCsr& B);

void mtx_read(const char* path, HostCsr& out);
void mtx_write(const char* path, const rt_csr& C, int ref_compat);

}  // namespace rt
