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

hipError_t dp_sweep_f32_d8(const DpFwdArgs& a, size_t lds, hipStream_t s);
hipError_t dp_sweep_f32_d16(const DpFwdArgs& a, size_t lds, hipStream_t s);
hipError_t dp_sweep_f32_d32(const DpFwdArgs& a, size_t lds, hipStream_t s);
hipError_t dp_sweep_f64_d8(const DpFwdArgs& a, size_t lds, hipStream_t s);
hipError_t dp_sweep_f64_d16(const DpFwdArgs& a, size_t lds, hipStream_t s);
hipError_t dp_sweep_f64_d32(const DpFwdArgs& a, size_t lds, hipStream_t s);

size_t dp_sweep_lds(int dtype, int d, int cols, int n_levels, int order) {
  const size_t esz = dtype == KSIG_F32 ? 4 : 8;
  const int dmax = d <= 8 ? 8 : (d <= 16 ? 16 : 32);
  const int quad = (int)(16 / esz);
  const int rs = ((dmax / quad) % 2 == 0) ? dmax + quad : dmax + 2 * quad;
  const int pc = cols > kWave ? cols : kWave + 1;
  const int rows_per_band = kWave * (dtype == KSIG_F32 ? 4 : 1);
  return ((size_t)rows_per_band * rs + (size_t)cols * rs +
          (size_t)pc * dp_flow_count(n_levels, order)) * esz;
}

hipError_t dp_sweep_launch(const DpFwdArgs& a, int dtype, size_t lds, hipStream_t s) {
  if (dtype == KSIG_F32) {
    if (a.d <= 8) return dp_sweep_f32_d8(a, lds, s);
    if (a.d <= 16) return dp_sweep_f32_d16(a, lds, s);
    return dp_sweep_f32_d32(a, lds, s);
  }
  if (a.d <= 8) return dp_sweep_f64_d8(a, lds, s);
  if (a.d <= 16) return dp_sweep_f64_d16(a, lds, s);
  return dp_sweep_f64_d32(a, lds, s);
}

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
