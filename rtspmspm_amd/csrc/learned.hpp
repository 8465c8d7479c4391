// Learned row sizing and autotuned bin thresholds (internal interface).
#pragma once
#include "spgemm.hpp"

namespace rt {

// Bin thresholds for this product: defaults, or the autotuner's cached
// choice for the operands' matrix class when use_tuned is set.
Tuning active_tuning(const Csr& A, const Csr& B, bool use_tuned);

}  // namespace rt
