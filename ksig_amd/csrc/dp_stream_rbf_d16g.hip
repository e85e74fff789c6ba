// Row-stream DP forward, float32, RBF static kernel, d <= 16, LDS FIFO.
#include "dp_stream.h"

namespace ksig {

#define KSIG_CASE(L, O) \
  if (n_levels == L && order == O) \
    return stream_entry<L, O, 16, kRBF, false, false>();

StreamEntry dp_stream_rbf_d16g(int n_levels, int order) {
  KSIG_DP_LEVEL_ORDER_TABLE(KSIG_CASE)
  return StreamEntry{nullptr, 0};
}

}  // namespace ksig
