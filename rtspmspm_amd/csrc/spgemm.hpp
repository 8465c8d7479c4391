// Internal interface between the C ABI (capi.cpp) and the device pipelines.
#pragma once
#include <string>

#include "common.hpp"

namespace rt {

// Device CSR operand (borrowed).
struct Csr {
  i64 rows = 0, cols = 0, nnz = 0;
  const i64* rp = nullptr;
  const i32* col = nullptr;
  const float* val = nullptr;
};

// Device CSR result; arrays come from the pool (dmalloc) and are released
// with dfree (rt_csr_free does that for library-owned device results).
struct CsrOut {
  i64 rows = 0, cols = 0, nnz = 0;
  i64* rp = nullptr;
  i32* col = nullptr;
  float* val = nullptr;
};

// Batched problems: A row r belongs to the pair whose B rows start at
// b_base[r] in the stacked B.  nullptr for a single product.
struct Batch {
  const i64* b_base = nullptr;  // per A row
};

// Row-bin settings of the two-phase method.  Thresholds are inclusive upper
// bounds: symbolic bins use U = min(products, cols); numeric bins use exact
// nnz.  Bins 0..4 are LDS-hash kernels with tables of 32/128/512/2048/8192
// slots (symbolic: keys; numeric: keys + values), so a threshold sets the
// bin's load factor (at most 1/2: a numeric row's sorted entries are staged
// in half the table); bin 5 (above sym[4] / num[4]) is the LDS-bitmap
// symbolic kernel / the long-row numeric kernels.
// Invariant: sym[4] <= num[4] so every numeric bin-5 row has tile counts.
constexpr int HASH_T[5] = {32, 128, 512, 2048, 8192};
struct Tuning {
  i64 sym[5] = {16, 64, 256, 1024, 4096};
  i64 num[5] = {16, 64, 256, 1024, 4096};
  // Learned method: long rows (U > num[4]) sized by the predictor and run by
  // the learned dense-tile kernel (1), or sent to the exact path (0).
  int learned_heavy = 0;
  // Per hash bin, the workgroup shape: 0 the default lanes per row (8, 16,
  // 64, 128, 512 for bins 0..4), 1 the alternate (16, 32, 32, 256, 1024);
  // bin 4 also has shape 2, 256 lanes per row over a 4096-slot table (for a
  // long-row cut of at most 2048).
  int shape[5] = {0, 0, 0, 0, 0};
  // Learned method: numeric bins (bit b) whose rows are sized by the
  // predictor; rows predicted into other bins run the exact path.
  int learned_bins = 31;
};
// Lanes per row of hash bin b in shape `alt`.
inline int hash_group(int b, int alt) {
  static const int g[2][5] = {{8, 16, 64, 128, 512}, {16, 32, 32, 256, 1024}};
  if (b == 4 && alt == 2) return 256;
  return g[alt ? 1 : 0][b];
}
// Table slots of hash bin b in shape `alt`.
inline int hash_slots(int b, int alt) { return b == 4 && alt == 2 ? 4096 : HASH_T[b]; }
// Largest shape value of hash bin b.
inline int hash_shapes(int b) { return b == 4 ? 2 : 1; }
// Empty string when t is usable by the kernels, else what is wrong with it.
std::string tuning_problem(const Tuning& t);

struct RunStats {
  i64 products = 0;
  i64 max_row_products = 0;
  i64 overflow_rows = 0;
  i64 predicted_rows = 0;
};

// Rows the symbolic pass handled with the LDS-bitmap kernel, with their
// per-tile (TILE_W columns) counts, used by the long-row numeric kernels.
struct HeavyRows {
  DBuf<i32> list;  // all binned rows; the heavy ones start at `off`
  i32 off = 0, n = 0, nt = 1;
  DBuf<i32> tilecnt;  // n x nt: entries | nonzero bitmap words << 16 per tile
  // Each long row's column set (both may be empty): per tile 16 records of
  // its 32-word groups, the mask of the group's nonzero 32-column words with
  // the group's entry count in the high half (rowmask, n x nt x 16), and the
  // nonzero words packed in column order (rowbits, n x nt x TILE_W/32 words
  // of room).
  DBuf<u32> rowbits;
  DBuf<u64> rowmask;
  DBuf<i32> next;     // row-slot counter of the symbolic long-row kernel
  DBuf<i32> slab;     // B.rows x (nt+1) tile boundaries of B's rows
  DBuf<i32> colpad;   // B's columns with a spare entry past the end
};

// Pipeline pieces.  `key` arrays select and size rows: rows with key 0 are
// skipped.  Symbolic keys are upper bounds (products); numeric keys are the
// exact nnz; learned keys are predicted capacities.
void row_products(const Csr& A, const Csr& B, const i64* bbase, i64* prod, hipStream_t s,
                  RunStats* st);
// counters: bin_scratch_ints() device ints of scratch (nullptr: taken from
// the pool).
void bin_rows(const i64* x, i64 n, i64 cap, const i64 th[5], i32* list, int cnt[6], int off[7],
              hipStream_t s, i32* counters = nullptr);
i64 bin_scratch_ints();
// Reorders the long-row list `rows` (n entries) so that rows gathering the
// same long B rows run together (k_order.hip).
void order_long_rows(const Csr& A, const Csr& B, i32* rows, i32 n, hipStream_t s);
void symbolic_pass(const Csr& A, const Csr& B, const i64* bbase, const i64* key, i64* nnz,
                   const Tuning& tu, hipStream_t s, HeavyRows& heavy);
void numeric_pass(const Csr& A, const Csr& B, const i64* bbase, const i64* key,
                  const HeavyRows& heavy, CsrOut& C, const Tuning& tu, hipStream_t s);
void learned_numeric_pass(const Csr& A, const Csr& B, const i64* bbase, const i64* key,
                          const i64* cap, const i64* tmp_off, i32* tcol, float* tval, i64* nnz,
                          i32* ovf, const Tuning& tu, hipStream_t s);
// Long rows (key > 0) sized by the predictor: dense-tile kernel writing at
// tmp_off with capacity cap; nnz or ovf set per row.
void learned_heavy_pass(const Csr& A, const Csr& B, const i64* bbase, const i64* key,
                        const i64* cap, const i64* tmp_off, i32* tcol, float* tval, i64* nnz,
                        i32* ovf, hipStream_t s);
// C.rp = exclusive scan of nnz; allocates C.col / C.val.
void alloc_c(CsrOut& C, const i64* nnz, i64 m, i64 cols, hipStream_t s);

// Exact two-phase product.  C.rp/col/val allocated here.
void spgemm_plain(const Csr& A, const Csr& B, const Batch& batch, CsrOut& C,
                  const Tuning& tune, hipStream_t s, RunStats* st);

// The same method with each stage captured into a hipGraph (persistent row
// kernels, device-side binning; one host read for C's size).
void spgemm_plain_graph(const Csr& A, const Csr& B, const Batch& batch, CsrOut& C,
                        const Tuning& tune, hipStream_t s, RunStats* st);

// The learned method with its three stages captured in hipGraphs.
void spgemm_learned_graph(const Csr& A, const Csr& B, const Batch& batch, CsrOut& C,
                          const Tuning& tune, hipStream_t s, RunStats* st);

// Faithful one-work-item-per-A-nonzero launch (reference raygen + any-hit).
void spgemm_ray(const Csr& A, const Csr& B, CsrOut& C, int ablation, int output,
                hipStream_t s, RunStats* st);

// Learned row sizing: predicted capacities, exact rerun of overflow rows.
void spgemm_learned(const Csr& A, const Csr& B, const Batch& batch, CsrOut& C,
                    const Tuning& tune, hipStream_t s, RunStats* st);

// Per-row features (n x RT_NFEAT, device) and optionally exact nnz.
void row_features(const Csr& A, const Csr& B, const Batch& batch, float* feat,
                  const i64* prod, hipStream_t s);

// Structural check of a device CSR (k_validate); RT_E_INVALID when
// malformed.  flag: one device int of scratch (nullptr: from the pool).
void validate_csr_dev(const Csr& M, i64 col_limit, const i64* row_limit, hipStream_t s,
                      const char* what, int* flag = nullptr);

// C = A*B into caller-provided arrays, every scratch buffer taken from one
// file-scope device array (rt_spmspm_noalloc).  *c_nnz receives nnz(C) (or
// -1 when the scratch is too small); RT_E_LIMIT when C needs more than
// `capacity` entries (c_row_ptr is then filled, c_col / c_val untouched).
constexpr i64 NOALLOC_SCRATCH_BYTES = (i64)64 << 20;
void spgemm_noalloc(const Csr& A, const Csr& B, i64* c_row_ptr, i32* c_col, float* c_val,
                    i64 capacity, i64* c_nnz, bool trusted, hipStream_t s);

// Fingerprint of a result (k_fingerprint): structure hashes and value sums.
struct Fingerprint {
  i64 nnz = 0;
  u64 hrp = 0, hcol = 0;
  double vsum[2] = {0, 0};  // sum |v|, sum v * w(position)
};
Fingerprint fingerprint(const CsrOut& C, hipStream_t s);

// Device utilities.
// partials: scan_partials(n) device i64 of scratch (nullptr: from the pool).
void exclusive_scan_i64(const i64* in, i64* out /* n+1 */, i64 n, hipStream_t s,
                        i64* partials = nullptr);
i64 scan_partials(i64 n);
// out[r] = values[b] for every row r of segment b (segments of seg_rows[b]
// consecutive rows); used for the per-row B base and column limits of batches.
void fill_segments(i32 count, const i64* seg_rows, const i64* values, i64* out_dev,
                   i64 total_rows, hipStream_t s);

}  // namespace rt
