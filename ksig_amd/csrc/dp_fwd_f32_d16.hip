// DP forward instantiations: T = float, fused static kernel, d <= 16.
#include "dp_forward_impl.h"

namespace ksig {

#define KSIG_CASE(L, O) \
  if (a.n_levels == L && a.order == O) return launch_dp<float, L, O, 16, false>(a, lds, s);

hipError_t dp_fwd_f32_d16(const DpFwdArgs& a, size_t lds, hipStream_t s) {
  KSIG_DP_LEVEL_ORDER_TABLE(KSIG_CASE)
  return hipErrorInvalidValue;
}

}  // namespace ksig
