// Learned row sizing and autotuned row-bin settings (internal interface).
#pragma once
#include <functional>
#include <string>
#include <vector>

#include "spgemm.hpp"

namespace rt {

// Row-bin settings for this product: defaults, or the autotuner's cached
// choice for the operands' matrix class when use_tuned is set.
Tuning active_tuning(const Csr& A, const Csr& B, bool use_tuned, int mode, bool batched);

// Autotuner pieces (autotune.cpp).
std::string matrix_class(const Csr& A, const Csr& B, bool batched);
std::string tune_key(const Csr& A, const Csr& B, int mode, bool batched);
// Greedy search over candidate settings; eval returns a candidate's device
// ms, or a negative value when its result was rejected.  log receives the
// per-candidate times.
Tuning autotune_search(bool learned, const std::function<double(const Tuning&)>& eval,
                       std::string& log);
void record_tuning(const std::string& key, const Tuning& t);
// Cache-line fields of t, and a readable summary (table sizes, thresholds,
// lanes per row).
std::string tuning_str(const Tuning& t);
std::string tuning_report(const Tuning& t);
void tune_cache_check(const std::string& path, int& valid, int& rejected);

// Launch wrappers of the learned pipeline's kernels (k_learned.hip): kernels
// only, for the graph-captured learned path.
void learned_weights(std::vector<float>& params, float& margin);
void learned_predict_launch(const Csr& A, const Csr& B, const i64* bbase, const i64* prod,
                            const float* params, float* feat, float* pred, i64 skip_u,
                            i64 skip_lo, hipStream_t s);
// Rows whose bound min(P, cols) is at most this never land in a predicted
// bin (their capacity is below the lowest one learned_bins selects): the
// learned method skips their features (-1: none skipped).
i64 predicted_floor(const Tuning& tu);
void learned_capacity_launch(i64 m, const i64* prod, i64 ncols, const float* pred, float margin,
                             const Tuning& tu, i64* cap, i64* key_light, i64* key_heavy,
                             i64* key_exact, i32* ovf, hipStream_t s);
void learned_merge_launch(i64 m, const i64* prod, const i64* cap, const i32* ovf, i64* key_exact,
                          unsigned long long* counts, hipStream_t s);
void learned_finish_launch(i64 m, const i64* cap, const i32* ovf, const i64* tmp_off,
                           const i32* tcol, const float* tval, const i64* key_exact,
                           const i64* nnz, i64* key_num, const CsrOut& C, hipStream_t s);

}  // namespace rt
