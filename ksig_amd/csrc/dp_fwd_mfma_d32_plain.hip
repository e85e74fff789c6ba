// Instantiations of the MFMA-Gram first-order forward: features up to 32, levels
// only; host-side sizing and dispatch of the MFMA-Gram forward.
#include "dp_fwd_mfma.h"

namespace ksig {

hipError_t dp_fwd_mfma_d32_plain(const DpFwdArgs& a, size_t lds, hipStream_t s) {
  return dispatch_fwd_mfma<32, false>(a, lds, s);
}

}  // namespace ksig

namespace ksig {

hipError_t dp_fwd_mfma_d32_save(const DpFwdArgs& a, size_t lds, hipStream_t s);

size_t dp_fwd_mfma_lds() { return (size_t)kBwdRingW * kRingS * sizeof(float); }

int dp_fwd_mfma_fifo_cols(int cols) { return fwd_mfma_pc(cols - 1); }  // Nc = cols - 1

size_t dp_fwd_mfma_fifo_bytes(int cols, int n_levels) {
  const int nfg = (1 + n_levels + 3) & ~3;
  return (size_t)dp_fwd_mfma_fifo_cols(cols) * nfg * sizeof(float);
}

bool dp_fwd_mfma_supported(int d, int n_levels) {
  // the 32-feature bucket: at d <= 16 the row stream is faster (its y rows are
  // short enough for the vector-memory path)
  if (d < 17 || d > 32) return false;
#define KSIG_SUP(L) if (n_levels == L) return true;
  KSIG_FWD_MFMA_LEVELS(KSIG_SUP)
#undef KSIG_SUP
  return false;
}

hipError_t dp_fwd_mfma_launch(const DpFwdArgs& a, bool save, size_t lds, hipStream_t s) {
  if (a.d <= 16) return hipErrorInvalidValue;
  return save ? dp_fwd_mfma_d32_save(a, lds, s) : dp_fwd_mfma_d32_plain(a, lds, s);
}

}  // namespace ksig
