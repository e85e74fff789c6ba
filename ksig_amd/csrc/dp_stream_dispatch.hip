// Row-stream DP forward: kernel selection, occupancy query and launch.
#include "dp_stream.h"

namespace ksig {

StreamEntry dp_stream_rbf_d8v(int n_levels, int order);
StreamEntry dp_stream_rbf_d8g(int n_levels, int order);
StreamEntry dp_stream_rbf_d16v(int n_levels, int order);
StreamEntry dp_stream_rbf_d16g(int n_levels, int order);
StreamEntry dp_stream_rbf_d32v(int n_levels, int order);
StreamEntry dp_stream_rbf_d32g(int n_levels, int order);
StreamEntry dp_stream_lin_d8v(int n_levels, int order);
StreamEntry dp_stream_lin_d8g(int n_levels, int order);
StreamEntry dp_stream_lin_d16v(int n_levels, int order);
StreamEntry dp_stream_lin_d16g(int n_levels, int order);
StreamEntry dp_stream_lin_d32v(int n_levels, int order);
StreamEntry dp_stream_lin_d32g(int n_levels, int order);

StreamEntry dp_stream_entry(int kind, int dmax, bool vec, int n_levels, int order) {
#define KSIG_PICK(K, KN)                                                           \
  if (kind == KN) {                                                                \
    if (dmax == 8) return vec ? dp_stream_##K##_d8v(n_levels, order)               \
                              : dp_stream_##K##_d8g(n_levels, order);              \
    if (dmax == 16) return vec ? dp_stream_##K##_d16v(n_levels, order)             \
                               : dp_stream_##K##_d16g(n_levels, order);            \
    if (dmax == 32) return vec ? dp_stream_##K##_d32v(n_levels, order)             \
                               : dp_stream_##K##_d32g(n_levels, order);            \
  }
  KSIG_PICK(rbf, kRBF)
  KSIG_PICK(lin, kLinear)
#undef KSIG_PICK
  return StreamEntry{nullptr, 0};
}

hipError_t dp_stream_launch(const StreamEntry& e, const DpStreamArgs& a, size_t lds,
                            hipStream_t s) {
  if (!e.fn) return hipErrorInvalidValue;
  if (lds > 64 * 1024) {
    hipError_t err = hipFuncSetAttribute(e.fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)lds);
    if (err != hipSuccess) return err;
  }
  DpStreamArgs copy = a;
  void* args[] = {&copy};
  return hipLaunchKernel(e.fn, dim3((unsigned)a.n_waves), dim3(kWave), args, lds, s);
}

int dp_stream_occupancy(const StreamEntry& e, size_t lds) {
  if (!e.fn) return 0;
  if (lds > 64 * 1024 &&
      hipFuncSetAttribute(e.fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
          hipSuccess)
    return 0;
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, e.fn, kWave, lds) != hipSuccess)
    return 0;
  return nb;
}

}  // namespace ksig
