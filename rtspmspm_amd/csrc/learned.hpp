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

}  // namespace rt
