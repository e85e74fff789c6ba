// Dispatch of the MFMA-Gram DP forward over (precision mode, feature bucket).
#include "dp_mfma.h"

namespace ksig {

hipError_t dp_mfma_m0_d8(const DpFwdArgs& a, size_t lds, hipStream_t s);
hipError_t dp_mfma_m0_d16(const DpFwdArgs& a, size_t lds, hipStream_t s);
hipError_t dp_mfma_m0_d32(const DpFwdArgs& a, size_t lds, hipStream_t s);
hipError_t dp_mfma_m1_d8(const DpFwdArgs& a, size_t lds, hipStream_t s);
hipError_t dp_mfma_m1_d16(const DpFwdArgs& a, size_t lds, hipStream_t s);
hipError_t dp_mfma_m1_d32(const DpFwdArgs& a, size_t lds, hipStream_t s);

bool dp_mfma_supported(int n_levels, int order) {
  return (n_levels == 1 && order == 1) || (n_levels == 2 && order == 1) || (n_levels == 3 && order == 1) || (n_levels == 4 && order == 1) || (n_levels == 5 && order == 1) || (n_levels == 6 && order == 1) || (n_levels == 7 && order == 1) || (n_levels == 8 && order == 1) || (n_levels == 2 && order == 2) || (n_levels == 3 && order == 2) || (n_levels == 4 && order == 2) || (n_levels == 5 && order == 2) || (n_levels == 6 && order == 2) || (n_levels == 7 && order == 2) || (n_levels == 8 && order == 2) || (n_levels == 10 && order == 1) || (n_levels == 12 && order == 1);
}

hipError_t dp_mfma_launch(const DpFwdArgs& a, int mode, size_t lds, hipStream_t s) {
  if (mode == 1) {
    if (a.d <= 8) return dp_mfma_m0_d8(a, lds, s);
    if (a.d <= 16) return dp_mfma_m0_d16(a, lds, s);
    return dp_mfma_m0_d32(a, lds, s);
  }
  if (a.d <= 8) return dp_mfma_m1_d8(a, lds, s);
  if (a.d <= 16) return dp_mfma_m1_d16(a, lds, s);
  return dp_mfma_m1_d32(a, lds, s);
}

}  // namespace ksig
