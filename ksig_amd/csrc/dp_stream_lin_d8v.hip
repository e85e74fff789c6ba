// Row-stream DP forward, float32, linear static kernel, d == 8, 16-byte vector loads, LDS FIFO.
#include "dp_stream.h"

namespace ksig {

#define KSIG_CASE(L, O) \
  if (n_levels == L && order == O) \
    return stream_entry<L, O, 8, kLinear, true, false>();

StreamEntry dp_stream_lin_d8v(int n_levels, int order) {
  KSIG_DP_LEVEL_ORDER_TABLE(KSIG_CASE)
  return StreamEntry{nullptr, 0};
}

}  // namespace ksig
