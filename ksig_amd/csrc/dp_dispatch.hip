// Dispatch of the DP forward sweep over (dtype, feature bucket, input kind).
#include "dp_forward_impl.h"

namespace ksig {

hipError_t dp_fwd_f32_d8(const DpFwdArgs& a, size_t lds, hipStream_t s);
hipError_t dp_fwd_f32_d16(const DpFwdArgs& a, size_t lds, hipStream_t s);
hipError_t dp_fwd_f32_d32(const DpFwdArgs& a, size_t lds, hipStream_t s);
hipError_t dp_fwd_f32_gram(const DpFwdArgs& a, size_t lds, hipStream_t s);
hipError_t dp_fwd_f64_d8(const DpFwdArgs& a, size_t lds, hipStream_t s);
hipError_t dp_fwd_f64_d16(const DpFwdArgs& a, size_t lds, hipStream_t s);
hipError_t dp_fwd_f64_d32(const DpFwdArgs& a, size_t lds, hipStream_t s);
hipError_t dp_fwd_f64_gram(const DpFwdArgs& a, size_t lds, hipStream_t s);

int dp_flow_count(int nlev, int order) { return 1 + nlev + dp_num_side(nlev, order); }

bool dp_supported(int n_levels, int order) {
#define KSIG_SUP(L, O) if (n_levels == L && order == O) return true;
  KSIG_DP_LEVEL_ORDER_TABLE(KSIG_SUP)
#undef KSIG_SUP
  return false;
}

hipError_t dp_forward_launch(const DpFwdArgs& a, int dtype, size_t lds, hipStream_t s) {
  const bool gram = a.kind == kGram;
  if (dtype == KSIG_F32) {
    if (gram) return dp_fwd_f32_gram(a, lds, s);
    if (a.d <= 8) return dp_fwd_f32_d8(a, lds, s);
    if (a.d <= 16) return dp_fwd_f32_d16(a, lds, s);
    return dp_fwd_f32_d32(a, lds, s);
  }
  if (gram) return dp_fwd_f64_gram(a, lds, s);
  if (a.d <= 8) return dp_fwd_f64_d8(a, lds, s);
  if (a.d <= 16) return dp_fwd_f64_d16(a, lds, s);
  return dp_fwd_f64_d32(a, lds, s);
}

}  // namespace ksig
