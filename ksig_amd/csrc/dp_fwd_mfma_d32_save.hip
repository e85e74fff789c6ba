// Instantiations of the MFMA-Gram first-order forward: features up to 32, with the adjoint's boundary state.
#include "dp_fwd_mfma.h"

namespace ksig {

hipError_t dp_fwd_mfma_d32_save(const DpFwdArgs& a, size_t lds, hipStream_t s) {
  return dispatch_fwd_mfma<32, true>(a, lds, s);
}

}  // namespace ksig
