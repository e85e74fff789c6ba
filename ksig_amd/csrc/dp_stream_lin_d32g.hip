// Row-stream DP forward, float32, linear static kernel, d <= 32, LDS FIFO.
#include "dp_stream.h"

namespace ksig {

#define KSIG_CASE(L, O) \
  if (n_levels == L && order == O) \
    return stream_entry<L, O, 32, kLinear, false, false>();

StreamEntry dp_stream_lin_d32g(int n_levels, int order) {
  KSIG_DP_LEVEL_ORDER_TABLE(KSIG_CASE)
  return StreamEntry{nullptr, 0};
}

}  // namespace ksig
