// DP forward (fused static kernel on the VALU), T = float, d <= 16.
#include "dp_sweep.h"

namespace ksig {

#define KSIG_CASE(L, O) \
  if (a.n_levels == L && a.order == O) return launch_dp_sweep<float, L, O, 16>(a, lds, s);

hipError_t dp_sweep_f32_d16(const DpFwdArgs& a, size_t lds, hipStream_t s) {
  KSIG_DP_LEVEL_ORDER_TABLE(KSIG_CASE)
  return hipErrorInvalidValue;
}

}  // namespace ksig
