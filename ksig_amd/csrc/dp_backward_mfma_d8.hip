// Instantiations of the MFMA-Gram DP adjoint for features up to 8.
#include "dp_backward_mfma.h"

namespace ksig {

hipError_t dp_backward_mfma_d8(const DpBwdArgs& a, size_t lds, hipStream_t s) {
  return dispatch_bwd_mfma<8>(a, lds, s);
}

}  // namespace ksig
