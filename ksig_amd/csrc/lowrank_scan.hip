// Level recursion of the first-order low-rank signature features with an
// elementwise (tensorized random) projection (gfx950).
//
// Replaces the level loop of ksig/algorithms.py:218-246 for a
// TensorizedRandomProjection (ksig/projections.py:566-584): with z_m(t) the
// projected increments of level m (already scaled; level 0 projected alone),
//   R_1(t) = z_0(t),  R_m(t) = C_{m-1}(t) z_{m-1}(t),  C_m(t) = sum_{t' < t} R_m(t'),
//   P_m = sum_t R_m(t),
// i.e. the exclusive time cumsum, the elementwise product and the time sum
// of every level in one pass over z instead of three whole-tensor passes per
// level.  One thread per (sequence, component): the time walk is sequential
// in registers, consecutive threads read consecutive components (coalesced),
// and the level recursion keeps its running sums in float64.  HBM-bound: z
// is read once, the features written once.

#include "ksig_common.h"
#include "ksig_hip.h"

namespace ksig {

template <typename T, int NL>
__global__ void __launch_bounds__(256)
lowrank_scan_kernel(const T* __restrict__ z, int64_t n_seq, int steps, int n_comp,
                    double* __restrict__ out) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n_seq * n_comp) return;
  const int64_t a = idx / n_comp;
  const int p = (int)(idx - a * n_comp);
  const int64_t level_stride = n_seq * (int64_t)steps * n_comp;
  const T* zp = z + a * (int64_t)steps * n_comp + p;
  double C[NL], S[NL];
#pragma unroll
  for (int m = 0; m < NL; ++m) C[m] = S[m] = 0.0;
  for (int t = 0; t < steps; ++t) {
    double R[NL];
    R[0] = (double)zp[(int64_t)t * n_comp];
#pragma unroll
    for (int m = 1; m < NL; ++m) R[m] = C[m - 1] * (double)zp[m * level_stride + (int64_t)t * n_comp];
#pragma unroll
    for (int m = 0; m < NL; ++m) {
      C[m] += R[m];
      S[m] += R[m];
    }
  }
  double* o = out + a * (int64_t)NL * n_comp + p;
#pragma unroll
  for (int m = 0; m < NL; ++m) o[(int64_t)m * n_comp] = S[m];
}

template <typename T, int NL>
static hipError_t launch_scan(const void* z, int64_t n_seq, int steps, int n_comp,
                              double* out, hipStream_t s) {
  const int64_t threads = n_seq * (int64_t)n_comp;
  const unsigned blocks = (unsigned)((threads + 255) / 256);
  hipLaunchKernelGGL((lowrank_scan_kernel<T, NL>), dim3(blocks), dim3(256), 0, s,
                     reinterpret_cast<const T*>(z), n_seq, steps, n_comp, out);
  return hipGetLastError();
}

template <typename T>
static hipError_t dispatch_scan(const void* z, int n_levels, int64_t n_seq, int steps,
                                int n_comp, double* out, hipStream_t s) {
  switch (n_levels) {
    case 1: return launch_scan<T, 1>(z, n_seq, steps, n_comp, out, s);
    case 2: return launch_scan<T, 2>(z, n_seq, steps, n_comp, out, s);
    case 3: return launch_scan<T, 3>(z, n_seq, steps, n_comp, out, s);
    case 4: return launch_scan<T, 4>(z, n_seq, steps, n_comp, out, s);
    case 5: return launch_scan<T, 5>(z, n_seq, steps, n_comp, out, s);
    case 6: return launch_scan<T, 6>(z, n_seq, steps, n_comp, out, s);
    case 7: return launch_scan<T, 7>(z, n_seq, steps, n_comp, out, s);
    case 8: return launch_scan<T, 8>(z, n_seq, steps, n_comp, out, s);
  }
  return hipErrorInvalidValue;
}

hipError_t lowrank_scan_launch(int dtype, const void* z, int n_levels, int64_t n_seq,
                               int steps, int n_comp, double* out, hipStream_t s) {
  if (dtype == KSIG_F32) return dispatch_scan<float>(z, n_levels, n_seq, steps, n_comp, out, s);
  return dispatch_scan<double>(z, n_levels, n_seq, steps, n_comp, out, s);
}

}  // namespace ksig
