// Level recursion of the low-rank signature features with an elementwise
// (tensorized random) projection, first and higher order (gfx950).
//
// Replaces the level loops of ksig/algorithms.py:218-246 (order 1) and
// :296-337 (order p > 1) for a TensorizedRandomProjection
// (ksig/projections.py:566-584): with z_m(t) the projected increments of
// level m (already scaled; level 0 projected alone) and d_m = min(m+1, p)
// the embedding orders carried at level m,
//   R_0[0](t) = z_0(t),
//   R_m[0](t) = C_{m-1}(t) z_m(t),   C_m(t) = sum_{t' < t} sum_r R_m[r](t'),
//   R_m[r](t) = R_{m-1}[r-1](t) z_m(t) / (r+1),   1 <= r < d_m,
//   P_m = sum_t sum_r R_m[r](t),
// i.e. the exclusive time cumsum, the elementwise products and the time sums
// of every level and order in one pass over z instead of several
// whole-tensor passes per level (at p = 1 only r = 0 exists and this is the
// first-order recursion).  One thread per (sequence, component): the time
// walk is sequential in registers, consecutive threads read consecutive
// components (coalesced), and the recursion keeps its running sums and the
// current R values in float64.  HBM-bound: z is read once, the features
// written once.

#include "ksig_common.h"
#include "ksig_hip.h"

namespace ksig {

// The diagonal scan below keeps the loads of kAheadDiag time steps in flight
// in a register ring, refilled as each step is consumed (one thread per
// sequence and component pair: about one wave per SIMD at 1000 sequences, so
// the walk is bound by load latency unless several steps are outstanding;
// 0.428 -> 0.323 ms at 1000 x 199 steps, 50 pairs, 5 levels).  The
// tensorized scan has twice the threads and is left to the compiler's load
// scheduling: rings of depth 1 and 4 measured 0.221 and 0.138 ms against its
// 0.130.
constexpr int kAheadDiag = 4;

template <typename T, int NL, int ORD>
__global__ void __launch_bounds__(256)
lowrank_scan_kernel(const T* __restrict__ z, int64_t n_seq, int steps, int n_comp,
                    double* __restrict__ out) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n_seq * n_comp) return;
  const int64_t a = idx / n_comp;
  const int p = (int)(idx - a * n_comp);
  const int64_t level_stride = n_seq * (int64_t)steps * n_comp;
  const T* zp = z + a * (int64_t)steps * n_comp + p;
  // C[m]: sum over r and over the steps walked so far of R_m[r] -- the
  // exclusive running sum while a step is formed, the level feature at the end
  double C[NL];
#pragma unroll
  for (int m = 0; m < NL; ++m) C[m] = 0.0;
  for (int t = 0; t < steps; ++t) {
    // R[m][r] of this time step; entries with r >= min(m+1, ORD) stay unused
    double R[NL][ORD];
    R[0][0] = (double)zp[(int64_t)t * n_comp];
#pragma unroll
    for (int m = 1; m < NL; ++m) {
      const double zm = (double)zp[m * level_stride + (int64_t)t * n_comp];
      R[m][0] = C[m - 1] * zm;
#pragma unroll
      for (int r = 1; r < ORD; ++r)
        if (r < m + 1) R[m][r] = R[m - 1][r - 1] * zm * (1.0 / (r + 1));
    }
#pragma unroll
    for (int m = 0; m < NL; ++m) {
      double sum = R[m][0];
#pragma unroll
      for (int r = 1; r < ORD; ++r)
        if (r < m + 1) sum += R[m][r];
      C[m] += sum;
    }
  }
  double* o = out + a * (int64_t)NL * n_comp + p;
#pragma unroll
  for (int m = 0; m < NL; ++m) o[(int64_t)m * n_comp] = C[m];
}

template <typename T, int NL, int ORD>
static hipError_t launch_scan(const void* z, int64_t n_seq, int steps, int n_comp,
                              double* out, hipStream_t s) {
  const int64_t threads = n_seq * (int64_t)n_comp;
  const unsigned blocks = (unsigned)((threads + 255) / 256);
  hipLaunchKernelGGL((lowrank_scan_kernel<T, NL, ORD>), dim3(blocks), dim3(256), 0, s,
                     reinterpret_cast<const T*>(z), n_seq, steps, n_comp, out);
  return hipGetLastError();
}

// order in [1, NL] (order >= NL carries every order: the full recursion)
template <typename T, int NL>
static hipError_t dispatch_order(const void* z, int order, int64_t n_seq, int steps,
                                 int n_comp, double* out, hipStream_t s) {
  switch (order < NL ? order : NL) {
    case 1: return launch_scan<T, NL, 1>(z, n_seq, steps, n_comp, out, s);
    case 2: if constexpr (NL >= 2) return launch_scan<T, NL, 2>(z, n_seq, steps, n_comp, out, s); break;
    case 3: if constexpr (NL >= 3) return launch_scan<T, NL, 3>(z, n_seq, steps, n_comp, out, s); break;
    case 4: if constexpr (NL >= 4) return launch_scan<T, NL, 4>(z, n_seq, steps, n_comp, out, s); break;
    case 5: if constexpr (NL >= 5) return launch_scan<T, NL, 5>(z, n_seq, steps, n_comp, out, s); break;
    case 6: if constexpr (NL >= 6) return launch_scan<T, NL, 6>(z, n_seq, steps, n_comp, out, s); break;
    case 7: if constexpr (NL >= 7) return launch_scan<T, NL, 7>(z, n_seq, steps, n_comp, out, s); break;
    case 8: if constexpr (NL >= 8) return launch_scan<T, NL, 8>(z, n_seq, steps, n_comp, out, s); break;
  }
  return hipErrorInvalidValue;
}

template <typename T>
static hipError_t dispatch_scan(const void* z, int n_levels, int order, int64_t n_seq,
                                int steps, int n_comp, double* out, hipStream_t s) {
  switch (n_levels) {
    case 1: return dispatch_order<T, 1>(z, order, n_seq, steps, n_comp, out, s);
    case 2: return dispatch_order<T, 2>(z, order, n_seq, steps, n_comp, out, s);
    case 3: return dispatch_order<T, 3>(z, order, n_seq, steps, n_comp, out, s);
    case 4: return dispatch_order<T, 4>(z, order, n_seq, steps, n_comp, out, s);
    case 5: return dispatch_order<T, 5>(z, order, n_seq, steps, n_comp, out, s);
    case 6: return dispatch_order<T, 6>(z, order, n_seq, steps, n_comp, out, s);
    case 7: return dispatch_order<T, 7>(z, order, n_seq, steps, n_comp, out, s);
    case 8: return dispatch_order<T, 8>(z, order, n_seq, steps, n_comp, out, s);
  }
  return hipErrorInvalidValue;
}

// The same recursion with a diagonal projection (ksig/projections.py:587-674)
// of internal size 2 on random Fourier features: level i's increments are
// the (sin, cos) pair u_i(t) = (dU_i[t, k], dU_i[t, H + k]) of component k,
// and R_i(t) = sqrt(H) C_{i-1}(t) (x) u_i(t) has 2^(i+1) entries per
// component (the q-index of level 0 most significant), so the level features
// are [2^(i+1), H] per sequence.  One thread per (sequence, component k)
// keeps every level's running sums and time sums in float64 registers.
template <typename T, int NL>
__global__ void __launch_bounds__(256)
lowrank_diag_scan_kernel(const T* __restrict__ du, int64_t n_seq, int steps, int half,
                         double* __restrict__ out) {
  constexpr int kTot = (2 << NL) - 2;  // sum over levels of 2^(i+1)
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n_seq * half) return;
  const int64_t a = idx / half;
  const int k = (int)(idx - a * half);
  const int width = 2 * half;
  const int64_t level_stride = n_seq * (int64_t)steps * width;
  const T* up = du + a * (int64_t)steps * width + k;
  const double sc = sqrt((double)half);
  // running sums over the steps walked so far; the level features at the end
  double C[kTot];
#pragma unroll
  for (int j = 0; j < kTot; ++j) C[j] = 0.0;
  T ring[kAheadDiag][2 * NL];
  auto load = [&](T* dst, int t) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      const T* ut = up + i * level_stride + (int64_t)t * width;
      dst[2 * i] = t < steps ? ut[0] : T(0);
      dst[2 * i + 1] = t < steps ? ut[half] : T(0);
    }
  };
#pragma unroll
  for (int u = 0; u < kAheadDiag; ++u) load(ring[u], u);
  for (int t0 = 0; t0 < steps; t0 += kAheadDiag) {
#pragma unroll
    for (int u = 0; u < kAheadDiag; ++u) {
      if (t0 + u < steps) {  // uniform over the grid
        double uv[2 * NL];
#pragma unroll
        for (int j = 0; j < 2 * NL; ++j) uv[j] = (double)ring[u][j];
        load(ring[u], t0 + u + kAheadDiag);
        // highest level first, so that level i reads level i-1's running
        // sums before this step's terms are added to them (exclusive sum)
#pragma unroll
        for (int i = NL - 1; i >= 1; --i) {
          const int off = (2 << i) - 2;  // first entry of level i
          const int prev = (2 << (i - 1)) - 2;
#pragma unroll
          for (int j = 0; j < (1 << i); ++j) {
            const double cs = sc * C[prev + j];
            C[off + 2 * j] += cs * uv[2 * i];
            C[off + 2 * j + 1] += cs * uv[2 * i + 1];
          }
        }
        C[0] += uv[0];
        C[1] += uv[1];
      }
    }
  }
  double* o = out + a * (int64_t)kTot * half + k;
#pragma unroll
  for (int j = 0; j < kTot; ++j) o[(int64_t)j * half] = C[j];
}

template <typename T, int NL>
static hipError_t launch_diag(const void* du, int64_t n_seq, int steps, int half,
                              double* out, hipStream_t s) {
  const int64_t threads = n_seq * (int64_t)half;
  const unsigned blocks = (unsigned)((threads + 255) / 256);
  hipLaunchKernelGGL((lowrank_diag_scan_kernel<T, NL>), dim3(blocks), dim3(256), 0, s,
                     reinterpret_cast<const T*>(du), n_seq, steps, half, out);
  return hipGetLastError();
}

hipError_t lowrank_diag_scan_launch(int dtype, const void* du, int n_levels, int64_t n_seq,
                                    int steps, int half, double* out, hipStream_t s) {
  switch (n_levels) {
#define KSIG_DIAG_CASE(L)                                                          \
    case L:                                                                      \
      return dtype == KSIG_F32 ? launch_diag<float, L>(du, n_seq, steps, half, out, s) \
                               : launch_diag<double, L>(du, n_seq, steps, half, out, s);
    KSIG_DIAG_CASE(1)
    KSIG_DIAG_CASE(2)
    KSIG_DIAG_CASE(3)
    KSIG_DIAG_CASE(4)
    KSIG_DIAG_CASE(5)
#undef KSIG_DIAG_CASE
  }
  return hipErrorInvalidValue;
}

hipError_t lowrank_scan_launch(int dtype, const void* z, int n_levels, int order,
                               int64_t n_seq, int steps, int n_comp, double* out,
                               hipStream_t s) {
  if (dtype == KSIG_F32)
    return dispatch_scan<float>(z, n_levels, order, n_seq, steps, n_comp, out, s);
  return dispatch_scan<double>(z, n_levels, order, n_seq, steps, n_comp, out, s);
}

}  // namespace ksig

namespace ksig {

// Random Fourier features and their time increments in one pass: from the
// projections p = X W [n_seq, L, half] with rows ld elements apart (several
// levels' weights may share one GEMM) (ksig/static/features.py:322-342),
// out[a, t, k] = (f(p[a, t+1, k]) - f(p[a, t, k])) / sqrt(half) with
// f = sin for k < half and cos for k >= half (difference), or f(p[a, t, k])
// / sqrt(half) without it — the concatenation, scaling and time difference
// of ksig/algorithms.py:196 as one HBM read and one write.  One thread per
// (sequence, chunk of kTimeChunk time steps, component): it walks its chunk
// in time, so each point's sin and cos are evaluated once (the difference
// reuses the previous point's), the chunk's loads are independent and in
// flight together, and consecutive threads take consecutive components
// (coalesced rows).  The index is split once per thread, not per output.
constexpr int kTimeChunk = 8;

__device__ __forceinline__ void sin_cos(float x, float* s, float* c) { sincosf(x, s, c); }
__device__ __forceinline__ void sin_cos(double x, double* s, double* c) { sincos(x, s, c); }

template <typename T>
__global__ void __launch_bounds__(256)
rff_diff_kernel(const T* __restrict__ p, int64_t n_seq, int L, int half, int ld,
                int difference, T* __restrict__ out) {
  const int steps = difference ? L - 1 : L;
  const int nchunk = (steps + kTimeChunk - 1) / kTimeChunk;
  const int64_t total = n_seq * (int64_t)nchunk * half;
  const T sc = T(1) / sqrt(T(half));
  const int64_t w = 2 * (int64_t)half;  // output row
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(i % half);
    const int64_t r = i / half;
    const int t0 = (int)(r % nchunk) * kTimeChunk;
    const int64_t a = r / nchunk;
    const int nt = min(kTimeChunk, steps - t0);
    const int lag = difference ? 1 : 0;
    const T* pa = p + (a * L + t0) * ld + k;
    T* o = out + (a * steps + t0) * w + k;
    T x[kTimeChunk];
#pragma unroll
    for (int u = 0; u < kTimeChunk; ++u) x[u] = u < nt ? pa[(int64_t)(u + lag) * ld] : T(0);
    T sp = T(0), cp = T(0);
    if (difference) sin_cos(pa[0], &sp, &cp);
#pragma unroll
    for (int u = 0; u < kTimeChunk; ++u) {
      if (u < nt) {
        T sv, cv;
        sin_cos(x[u], &sv, &cv);
        o[u * w] = (sv - sp) * sc;
        o[u * w + half] = (cv - cp) * sc;
        if (difference) {
          sp = sv;
          cp = cv;
        }
      }
    }
  }
}

hipError_t rff_diff_launch(int dtype, const void* p, int64_t n_seq, int L, int half, int ld,
                           int difference, void* out, hipStream_t s) {
  const int steps = difference ? L - 1 : L;
  const int64_t total =
      n_seq * (int64_t)((steps + kTimeChunk - 1) / kTimeChunk) * (int64_t)half;
  int64_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  if (dtype == KSIG_F32)
    hipLaunchKernelGGL(rff_diff_kernel<float>, dim3((unsigned)blocks), dim3(256), 0, s,
                       reinterpret_cast<const float*>(p), n_seq, L, half, ld, difference,
                       reinterpret_cast<float*>(out));
  else
    hipLaunchKernelGGL(rff_diff_kernel<double>, dim3((unsigned)blocks), dim3(256), 0, s,
                       reinterpret_cast<const double*>(p), n_seq, L, half, ld, difference,
                       reinterpret_cast<double*>(out));
  return hipGetLastError();
}

}  // namespace ksig
