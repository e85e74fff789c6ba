// Instantiations of the phase-aligned signature-PDE forward for double inputs.
#include "pde_fwd_aligned.h"

namespace ksig {

hipError_t pde_fwd_aligned_launch_f64(const PdeFwdArgs& a, int dmax, size_t lds, hipStream_t s) {
  return dispatch_pde_aligned<double>(a, dmax, lds, s);
}

}  // namespace ksig
