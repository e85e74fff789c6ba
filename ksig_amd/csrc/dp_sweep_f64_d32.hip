// DP forward (fused static kernel on the VALU), T = double, d <= 32.
#include "dp_sweep.h"

namespace ksig {

#define KSIG_CASE(L, O) \
  if (a.n_levels == L && a.order == O) return launch_dp_sweep<double, L, O, 32>(a, lds, s);

hipError_t dp_sweep_f64_d32(const DpFwdArgs& a, size_t lds, hipStream_t s) {
  KSIG_DP_LEVEL_ORDER_TABLE(KSIG_CASE)
  return hipErrorInvalidValue;
}

}  // namespace ksig
