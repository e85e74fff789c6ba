// Row-stream DP forward: kernel selection, occupancy query and launch.
#include "dp_stream.h"

namespace ksig {

StreamEntry dp_stream_rbf_d8v(int n_levels, int order);
StreamEntry dp_stream_rbf_d8g(int n_levels, int order);
StreamEntry dp_stream_rbf_d8w(int n_levels, int order);
StreamEntry dp_stream_rbf_d16v(int n_levels, int order);
StreamEntry dp_stream_rbf_d16g(int n_levels, int order);
StreamEntry dp_stream_rbf_d16w(int n_levels, int order);
StreamEntry dp_stream_rbf_d32v(int n_levels, int order);
StreamEntry dp_stream_rbf_d32g(int n_levels, int order);
StreamEntry dp_stream_rbf_d32w(int n_levels, int order);
StreamEntry dp_stream_lin_d8v(int n_levels, int order);
StreamEntry dp_stream_lin_d8g(int n_levels, int order);
StreamEntry dp_stream_lin_d8w(int n_levels, int order);
StreamEntry dp_stream_lin_d16v(int n_levels, int order);
StreamEntry dp_stream_lin_d16g(int n_levels, int order);
StreamEntry dp_stream_lin_d16w(int n_levels, int order);
StreamEntry dp_stream_lin_d32v(int n_levels, int order);
StreamEntry dp_stream_lin_d32g(int n_levels, int order);
StreamEntry dp_stream_lin_d32w(int n_levels, int order);

StreamEntry dp_stream_rbf_d8t(int n_levels);
StreamEntry dp_stream_rbf_d16t(int n_levels);
StreamEntry dp_stream_rbf_d32t(int n_levels);
StreamEntry dp_stream_lin_d8t(int n_levels);
StreamEntry dp_stream_lin_d16t(int n_levels);
StreamEntry dp_stream_lin_d32t(int n_levels);

StreamEntry dp_stream_save_entry(int kind, int dmax, int n_levels) {
  if (kind == kRBF) {
    if (dmax == 8) return dp_stream_rbf_d8t(n_levels);
    if (dmax == 16) return dp_stream_rbf_d16t(n_levels);
    if (dmax == 32) return dp_stream_rbf_d32t(n_levels);
  } else if (kind == kLinear) {
    if (dmax == 8) return dp_stream_lin_d8t(n_levels);
    if (dmax == 16) return dp_stream_lin_d16t(n_levels);
    if (dmax == 32) return dp_stream_lin_d32t(n_levels);
  }
  return StreamEntry{nullptr, 0};
}

StreamEntry dp_stream_entry(int kind, int dmax, bool vec, bool global_fifo, int n_levels,
                            int order) {
  if (global_fifo && !vec) return StreamEntry{nullptr, 0};
#define KSIG_PICK_D(K, D)                                                          \
  if (dmax == D)                                                                   \
    return global_fifo ? dp_stream_##K##_d##D##w(n_levels, order)                  \
                       : (vec ? dp_stream_##K##_d##D##v(n_levels, order)           \
                              : dp_stream_##K##_d##D##g(n_levels, order));
#define KSIG_PICK(K, KN)                                                           \
  if (kind == KN) {                                                                \
    KSIG_PICK_D(K, 8)                                                              \
    KSIG_PICK_D(K, 16)                                                             \
    KSIG_PICK_D(K, 32)                                                             \
  }
  KSIG_PICK(rbf, kRBF)
  KSIG_PICK(lin, kLinear)
#undef KSIG_PICK
#undef KSIG_PICK_D
  return StreamEntry{nullptr, 0};
}

hipError_t dp_stream_launch(const StreamEntry& e, const DpStreamArgs& a, size_t lds, int wpb,
                            hipStream_t s) {
  if (!e.fn) return hipErrorInvalidValue;
  if (lds > 64 * 1024) {
    hipError_t err = hipFuncSetAttribute(e.fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)lds);
    if (err != hipSuccess) return err;
  }
  DpStreamArgs copy = a;
  void* args[] = {&copy};
  const unsigned blocks = (unsigned)((a.n_waves + wpb - 1) / wpb);
  return hipLaunchKernel(e.fn, dim3(blocks), dim3(kWave * wpb), args, lds, s);
}

int dp_stream_occupancy(const StreamEntry& e, size_t lds, int wpb) {
  if (!e.fn) return 0;
  if (lds > 64 * 1024 &&
      hipFuncSetAttribute(e.fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
          hipSuccess)
    return 0;
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, e.fn, kWave * wpb, lds) != hipSuccess)
    return 0;
  return nb * wpb;
}

}  // namespace ksig
