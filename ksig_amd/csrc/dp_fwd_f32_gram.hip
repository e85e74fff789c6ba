// DP forward instantiations: T = float, caller-supplied Gram input.
#include "dp_forward_impl.h"

namespace ksig {

#define KSIG_CASE(L, O) \
  if (a.n_levels == L && a.order == O) return launch_dp<float, L, O, 8, true>(a, lds, s);

hipError_t dp_fwd_f32_gram(const DpFwdArgs& a, size_t lds, hipStream_t s) {
  KSIG_DP_LEVEL_ORDER_TABLE(KSIG_CASE)
  return hipErrorInvalidValue;
}

}  // namespace ksig
