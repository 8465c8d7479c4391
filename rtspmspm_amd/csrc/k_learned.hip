// Learned row sizing (placeholder until the predictor kernels land).
#include "learned.hpp"

namespace rt {

Tuning active_tuning(const Csr&, const Csr&, bool) { return Tuning{}; }

void spgemm_learned(const Csr&, const Csr&, const Batch&, CsrOut&, const Tuning&, hipStream_t,
                    RunStats*) {
  fail(RT_E_INVALID, "learned mode is not available in this build");
}

}  // namespace rt

extern "C" {
static int not_yet() {
  rt::set_last_error("learned predictor not available in this build");
  return RT_E_INVALID;
}
int rt_row_features(const rt_csr*, const rt_csr*, float*, int64_t*) { return not_yet(); }
int64_t rt_mlp_param_count(void) { return 0; }
int rt_mlp_set_weights(const float*, int64_t, float) { return not_yet(); }
int rt_mlp_get_weights(float*, int64_t, float*) { return not_yet(); }
int rt_mlp_train(const float*, const float*, int64_t, int32_t, int32_t, float, uint64_t, float*) {
  return not_yet();
}
int rt_mlp_loss_grad(const float*, const float*, const float*, int64_t, float*, float*) {
  return not_yet();
}
int rt_mlp_predict(const float*, const float*, int64_t, float*) { return not_yet(); }
}
