// DP forward with MFMA static Gram: bf16 hi/lo split MFMA, d <= 32.
#include "dp_mfma.h"

namespace ksig {

hipError_t dp_mfma_m1_d32(const DpFwdArgs& a, size_t lds, hipStream_t s) {
  if (a.n_levels == 1 && a.order == 1) return launch_dp_mfma<1, 1, 32, 1>(a, lds, s);
  if (a.n_levels == 2 && a.order == 1) return launch_dp_mfma<2, 1, 32, 1>(a, lds, s);
  if (a.n_levels == 3 && a.order == 1) return launch_dp_mfma<3, 1, 32, 1>(a, lds, s);
  if (a.n_levels == 4 && a.order == 1) return launch_dp_mfma<4, 1, 32, 1>(a, lds, s);
  if (a.n_levels == 5 && a.order == 1) return launch_dp_mfma<5, 1, 32, 1>(a, lds, s);
  if (a.n_levels == 6 && a.order == 1) return launch_dp_mfma<6, 1, 32, 1>(a, lds, s);
  if (a.n_levels == 7 && a.order == 1) return launch_dp_mfma<7, 1, 32, 1>(a, lds, s);
  if (a.n_levels == 8 && a.order == 1) return launch_dp_mfma<8, 1, 32, 1>(a, lds, s);
  if (a.n_levels == 2 && a.order == 2) return launch_dp_mfma<2, 2, 32, 1>(a, lds, s);
  if (a.n_levels == 3 && a.order == 2) return launch_dp_mfma<3, 2, 32, 1>(a, lds, s);
  if (a.n_levels == 4 && a.order == 2) return launch_dp_mfma<4, 2, 32, 1>(a, lds, s);
  if (a.n_levels == 5 && a.order == 2) return launch_dp_mfma<5, 2, 32, 1>(a, lds, s);
  if (a.n_levels == 6 && a.order == 2) return launch_dp_mfma<6, 2, 32, 1>(a, lds, s);
  if (a.n_levels == 7 && a.order == 2) return launch_dp_mfma<7, 2, 32, 1>(a, lds, s);
  if (a.n_levels == 8 && a.order == 2) return launch_dp_mfma<8, 2, 32, 1>(a, lds, s);
  if (a.n_levels == 10 && a.order == 1) return launch_dp_mfma<10, 1, 32, 1>(a, lds, s);
  if (a.n_levels == 12 && a.order == 1) return launch_dp_mfma<12, 1, 32, 1>(a, lds, s);
  return hipErrorInvalidValue;
}

}  // namespace ksig
