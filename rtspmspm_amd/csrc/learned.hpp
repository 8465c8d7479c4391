// Learned row sizing and autotuned bin thresholds (internal interface).
#pragma once
#include <string>
#include <vector>

#include "spgemm.hpp"

namespace rt {

// Bin thresholds for this product: defaults, or the autotuner's cached
// choice for the operands' matrix class when use_tuned is set.
Tuning active_tuning(const Csr& A, const Csr& B, bool use_tuned, int mode);

// Autotuner pieces (autotune.cpp).
std::string matrix_class(const Csr& A, const Csr& B, bool batched);
std::string tune_key(const Csr& A, const Csr& B, int mode);
std::vector<Tuning> tune_candidates(bool learned);
void record_tuning(const std::string& key, const Tuning& t);

}  // namespace rt
