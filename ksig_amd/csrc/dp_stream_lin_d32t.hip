// Row-stream DP training forward (order 1, saves the adjoint's boundary state),
// float32, linear static kernel, d == 32, 16-byte vector loads, LDS FIFO.
#include "dp_stream.h"

namespace ksig {

#define KSIG_CASE(L, O) \
  if (O == 1 && n_levels == L) return stream_entry<L, 1, 32, kLinear, true, false, true>();

StreamEntry dp_stream_lin_d32t(int n_levels) {
  KSIG_DP_LEVEL_ORDER_TABLE(KSIG_CASE)
  return StreamEntry{nullptr, 0};
}

}  // namespace ksig
