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
#include "lowrank.h"

namespace ksig {


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
// of internal size q on random Fourier features, at any embedding order
// (ksig/algorithms.py:296-337 keeps [d, n, L, q^(i+1), H] terms per level for
// this projection): level i's increments are the q values u_i(t)[b] =
// dU_i[t, b H + h] of component h (b < q), and
//   R_0[0](t) = u_0(t),
//   R_i[0](t) = sqrt(H) C_{i-1}(t) (x) u_i(t),   C_i(t) = sum_{t' < t} sum_r R_i[r](t'),
//   R_i[r](t) = sqrt(H) R_{i-1}[r-1](t) (x) u_i(t) / (r+1),   1 <= r < min(i+1, p),
// each with q^(i+1) entries per component (the internal index of level 0
// most significant), so the level features are [q^(i+1), H] per sequence.
// One thread per (sequence, component h) keeps every level's running sums in
// float64 registers (the projection is diagonal over h: no thread reads
// another's state); shapes whose state does not fit a register budget are
// not built (diag_state_doubles below, shared with the host).
template <typename T, int Q, int NL, int ORD>
__global__ void __launch_bounds__(256)
lowrank_diag_scan_kernel(const T* __restrict__ du, int64_t n_seq, int steps, int half,
                         double* __restrict__ out) {
  constexpr int kTot = diag_total(Q, NL);
  constexpr int kR = diag_rbuf(Q, NL, ORD);
  constexpr int kAhead = diag_ahead(Q, NL, ORD, (int)sizeof(T));
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n_seq * half) return;
  const int64_t a = idx / half;
  const int k = (int)(idx - a * half);
  const int width = Q * half;
  const int64_t level_stride = n_seq * (int64_t)steps * width;
  const T* up = du + a * (int64_t)steps * width + k;
  const double sc = sqrt((double)half);
  // running sums over the steps walked so far; the level features at the end
  double C[kTot];
#pragma unroll
  for (int j = 0; j < kTot; ++j) C[j] = 0.0;
  T ring[kAhead][Q * NL];
  auto load = [&](T* dst, int t) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      const T* ut = up + i * level_stride + (int64_t)t * width;
#pragma unroll
      for (int b = 0; b < Q; ++b) dst[Q * i + b] = t < steps ? ut[b * half] : T(0);
    }
  };
#pragma unroll
  for (int u = 0; u < kAhead; ++u) load(ring[u], u);
  for (int t0 = 0; t0 < steps; t0 += kAhead) {
#pragma unroll
    for (int u = 0; u < kAhead; ++u) {
      if (t0 + u < steps) {  // uniform over the grid
        double uv[Q * NL];
#pragma unroll
        for (int j = 0; j < Q * NL; ++j) uv[j] = (double)ring[u][j];
        load(ring[u], t0 + u + kAhead);
        if constexpr (ORD == 1) {
          // highest level first, so that level i reads level i-1's running
          // sums before this step's terms are added to them (exclusive sum)
#pragma unroll
          for (int i = NL - 1; i >= 1; --i) {
            const int off = diag_total(Q, i);  // first entry of level i
            const int prev = diag_total(Q, i - 1);
#pragma unroll
            for (int j = 0; j < ipow(Q, i); ++j) {
              const double cs = sc * C[prev + j];
#pragma unroll
              for (int b = 0; b < Q; ++b) C[off + Q * j + b] += cs * uv[Q * i + b];
            }
          }
#pragma unroll
          for (int b = 0; b < Q; ++b) C[b] += uv[b];
        } else {
          // lowest level first: level i needs this step's terms of level
          // i-1 (Rp, its d_{i-1} orders stacked) and C_{i-1} before they are
          // added to it; the top level's terms go straight into its sums
          double Rp[kR], Rn[kR];
#pragma unroll
          for (int b = 0; b < Q; ++b) Rp[b] = uv[b];
          if constexpr (NL == 1) {
#pragma unroll
            for (int b = 0; b < Q; ++b) C[b] += Rp[b];
          }
#pragma unroll
          for (int i = 1; i < NL; ++i) {
            const int off = diag_total(Q, i);
            const int prev = diag_total(Q, i - 1);
            const int np = ipow(Q, i);                   // entries per order, level i-1
            const int dp = i < ORD ? i : ORD;            // orders of level i-1
            const int dn = i + 1 < ORD ? i + 1 : ORD;    // orders of level i
#pragma unroll
            for (int j = 0; j < np; ++j) {
#pragma unroll
              for (int b = 0; b < Q; ++b) {
                const double ub = sc * uv[Q * i + b];
                double v = C[prev + j] * ub;
                if (i < NL - 1) Rn[Q * j + b] = v;
                else C[off + Q * j + b] += v;
#pragma unroll
                for (int r = 1; r < dn; ++r) {
                  v = Rp[(r - 1) * np + j] * ub * (1.0 / (r + 1));
                  if (i < NL - 1) Rn[r * np * Q + Q * j + b] = v;
                  else C[off + Q * j + b] += v;
                }
              }
            }
#pragma unroll
            for (int j = 0; j < np; ++j) {
#pragma unroll
              for (int r = 0; r < dp; ++r) C[prev + j] += Rp[r * np + j];
            }
            if (i < NL - 1) {
#pragma unroll
              for (int j = 0; j < dn * np * Q; ++j) Rp[j] = Rn[j];
            }
          }
        }
      }
    }
  }
  double* o = out + a * (int64_t)kTot * half + k;
#pragma unroll
  for (int j = 0; j < kTot; ++j) o[(int64_t)j * half] = C[j];
}

template <typename T, int Q, int NL, int ORD>
static hipError_t launch_diag(const void* du, int64_t n_seq, int steps, int half,
                              double* out, hipStream_t s) {
  if constexpr (!diag_fits(Q, NL, ORD)) {
    return hipErrorInvalidValue;
  } else {
    const int64_t threads = n_seq * (int64_t)half;
    const unsigned blocks = (unsigned)((threads + 255) / 256);
    hipLaunchKernelGGL((lowrank_diag_scan_kernel<T, Q, NL, ORD>), dim3(blocks), dim3(256), 0, s,
                       reinterpret_cast<const T*>(du), n_seq, steps, half, out);
    return hipGetLastError();
  }
}

template <typename T, int Q, int NL>
static hipError_t diag_order(int order, const void* du, int64_t n_seq, int steps, int half,
                             double* out, hipStream_t s) {
  switch (order) {
    case 1: return launch_diag<T, Q, NL, 1>(du, n_seq, steps, half, out, s);
    case 2: if constexpr (NL >= 2) return launch_diag<T, Q, NL, 2>(du, n_seq, steps, half, out, s); break;
    case 3: if constexpr (NL >= 3) return launch_diag<T, Q, NL, 3>(du, n_seq, steps, half, out, s); break;
    case 4: if constexpr (NL >= 4) return launch_diag<T, Q, NL, 4>(du, n_seq, steps, half, out, s); break;
    case 5: if constexpr (NL >= 5) return launch_diag<T, Q, NL, 5>(du, n_seq, steps, half, out, s); break;
  }
  return hipErrorInvalidValue;
}

template <typename T, int Q>
static hipError_t diag_levels(int n_levels, int order, const void* du, int64_t n_seq, int steps,
                              int half, double* out, hipStream_t s) {
  switch (n_levels) {
    case 1: return diag_order<T, Q, 1>(order, du, n_seq, steps, half, out, s);
    case 2: return diag_order<T, Q, 2>(order, du, n_seq, steps, half, out, s);
    case 3: return diag_order<T, Q, 3>(order, du, n_seq, steps, half, out, s);
    case 4: return diag_order<T, Q, 4>(order, du, n_seq, steps, half, out, s);
    case 5: return diag_order<T, Q, 5>(order, du, n_seq, steps, half, out, s);
  }
  return hipErrorInvalidValue;
}

template <typename T>
static hipError_t diag_q(int q, int n_levels, int order, const void* du, int64_t n_seq,
                         int steps, int half, double* out, hipStream_t s) {
  switch (q) {
    case 1: return diag_levels<T, 1>(n_levels, order, du, n_seq, steps, half, out, s);
    case 2: return diag_levels<T, 2>(n_levels, order, du, n_seq, steps, half, out, s);
    case 3: return diag_levels<T, 3>(n_levels, order, du, n_seq, steps, half, out, s);
    case 4: return diag_levels<T, 4>(n_levels, order, du, n_seq, steps, half, out, s);
  }
  return hipErrorInvalidValue;
}

// order: 1 <= order <= n_levels (the caller folds the full recursion into
// order = n_levels); diag_fits(q, n_levels, order) must hold
hipError_t lowrank_diag_scan_launch(int dtype, const void* du, int q, int n_levels, int order,
                                    int64_t n_seq, int steps, int half, double* out,
                                    hipStream_t s) {
  if (!diag_fits(q, n_levels, order)) return hipErrorInvalidValue;
  return dtype == KSIG_F32 ? diag_q<float>(q, n_levels, order, du, n_seq, steps, half, out, s)
                           : diag_q<double>(q, n_levels, order, du, n_seq, steps, half, out, s);
}

// The level recursion with TensorSketch projections (ksig/projections.py:
// 379-458): level i's projection keeps a hash h_i[j] in [0, m) and a sign
// s_i[j] per input feature j < width; _project is the count sketch
// CS_i(u)[k] = sum_{j: h_i[j] = k} s_i[j] u[j] (ksig/utils.py:279-298) and the
// outer-product step is the circular convolution A (*) CS_i(u) of
// ksig/utils.py:301-313, which the sparse sketch turns into
//   (A (*) CS_i(u))[k] = sum_j s_i[j] u[j] A[(k - h_i[j]) mod m]
// (width terms per component instead of three length-m FFTs), with level 0's
// count sketch the same sum over the unit vector A = e_0.  The recursion is
// that of lowrank_scan_kernel with (*) in place of the elementwise product:
//   R_0[0](t) = CS_0(u_0(t)),  R_i[0](t) = C_{i-1}(t) (*) CS_i(u_i(t)),
//   R_i[r](t) = R_{i-1}[r-1](t) (*) CS_i(u_i(t)) / (r+1),   P_i = C_i(end).
// A convolution reads every component of its source, so one workgroup walks
// one sequence's time steps.  Its threads form ng groups of w: thread l of a
// group owns components l + c w (c < kSketchKpt), and group g convolves the
// g-th slice of the input features (several waves per sequence hide the LDS
// latency that one wave per sequence exposes); the groups' partial sums meet
// in LDS and group 0 adds them to the running sums (float64, in the LDS rows
// the next level reads; the top level's in registers).  The sources a convolution reads (e_0, the running sums
// C_0..C_{NL-2} and, at order p > 1, this step's terms of the levels below
// the top) sit in LDS as rows of 2m doubles holding the vector twice, so that
// A[(k - h) mod m] is row[k + m - h] with no wrap test.  The hashes, the signs
// and each step's signed increments s_i[j] u_i(t)[j] are staged in LDS too
// (one coalesced load of the step's rows by the whole workgroup), so the
// convolution loop reads only LDS: broadcast reads of the features' hashes
// and signed increments and one gather per feature and component.
// Order 1: every level reads only the previous step's sums, so one pass of
// convolutions, a barrier, the reduction with the sum updates and the next
// step's staging, and a second barrier per step.  Order p > 1: level i reads
// level i-1's terms of the same step, so each level's convolutions, a
// barrier, its reduction into the term rows and a barrier, then the sum
// updates.  A hash value outside [0, m) selects no component (the
// reference's 0/1 mask has no column for it): its sign is staged as zero.
constexpr int kSketchKpt = 2;

// sum over features j in [j0, j1) of cs[j] row[k + m - hs[j]] for the
// thread's components (two accumulator chains per component); ID: hs[j] = j
// (cs is then a count sketch, indexed by component)
template <bool ID>
__device__ __forceinline__ void sketch_conv(const double* __restrict__ row,
                                            const double* __restrict__ cs,
                                            const int* __restrict__ hs, int j0, int j1, int m,
                                            const int* kr, double* acc) {
  double e[kSketchKpt], o[kSketchKpt];
#pragma unroll
  for (int c = 0; c < kSketchKpt; ++c) e[c] = o[c] = 0.0;
  const double* rm = row + m;
  int j = j0;
  for (; j + 4 <= j1; j += 4) {
    double cj[4];
    int hj[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      cj[u] = cs[j + u];
      hj[u] = ID ? j + u : hs[j + u];
    }
#pragma unroll
    for (int c = 0; c < kSketchKpt; ++c) {
      e[c] = __builtin_fma(cj[0], rm[kr[c] - hj[0]], e[c]);
      o[c] = __builtin_fma(cj[1], rm[kr[c] - hj[1]], o[c]);
      e[c] = __builtin_fma(cj[2], rm[kr[c] - hj[2]], e[c]);
      o[c] = __builtin_fma(cj[3], rm[kr[c] - hj[3]], o[c]);
    }
  }
  for (; j < j1; ++j)
#pragma unroll
    for (int c = 0; c < kSketchKpt; ++c)
      e[c] = __builtin_fma(cs[j], rm[kr[c] - (ID ? j : hs[j])], e[c]);
#pragma unroll
  for (int c = 0; c < kSketchKpt; ++c) acc[c] = e[c] + o[c];
}

// per-level stride of the staged tables: rows of 4-feature groups
__host__ __device__ inline int sketch_wpad(int width) { return (width + 3) & ~3; }

// DENSE (m < width): each step's increments are first summed into their count
// sketch b_i = CS_i(u_i) (m values, by buckets of features listed once per
// launch), and the convolution runs over the m components of b instead of
// the width features: (A (*) b)[k] = sum_p b[p] A[(k - p) mod m].
template <typename T, int NL, bool DENSE>
__global__ void __launch_bounds__(1024)
lowrank_sketch_kernel(const T* __restrict__ du, int64_t n_seq, int steps, int width,
                      const int* __restrict__ hidx, const T* __restrict__ hbit, int m, int order,
                      int ng, double* __restrict__ out) {
  extern __shared__ double sk_lds[];
  const int64_t a = blockIdx.x;
  const int tid = threadIdx.x, bs = blockDim.x;
  const int w = bs / ng;
  const int g = __builtin_amdgcn_readfirstlane(tid / w);  // w is a multiple of 64
  const int l = tid - g * w;
  const int m2 = 2 * m;
  const int rows = sketch_rows(NL, order);
  const int wpad = sketch_wpad(width);
  double* sv = sk_lds + rows * m2;   // [NL][wpad] signs, zero for a hash outside [0, m)
  double* cst = sv + NL * wpad;      // [NL][wpad] this step's signed increments
  double* red = cst + NL * wpad;     // [ng][NL][m] the groups' partial sums (ng > 1)
  const int mpad = sketch_wpad(m);
  double* bsk = red + (ng > 1 ? ng * NL * m : 0);  // DENSE: [NL][mpad] the step's count sketches
  int* hs = reinterpret_cast<int*>(bsk + (DENSE ? NL * mpad : 0));  // [NL][wpad]
  int* boff = hs + NL * wpad;        // DENSE: [NL][m+1] bucket offsets
  int* bidx = boff + (DENSE ? NL * (m + 1) : 0);  // DENSE: [NL][width] features by bucket
  // group g's slice of the terms (features, or DENSE components), in whole
  // groups of 4
  const int nterm = DENSE ? m : width;
  const int chunk = ((nterm + 4 * ng - 1) / (4 * ng)) * 4;
  const int j0 = min(nterm, g * chunk), j1 = min(nterm, j0 + chunk);
  for (int j = tid; j < rows * m2; j += bs)
    sk_lds[j] = (j == 0 || j == m) ? 1.0 : 0.0;  // row 0: e_0 twice; the rest zero
  for (int e = tid; e < NL * wpad; e += bs) {
    const int i = e / wpad, j = e - i * wpad;
    const int h = j < width ? hidx[i * width + j] : -1;
    const bool ok = (unsigned)h < (unsigned)m;
    hs[e] = ok ? h : 0;
    sv[e] = ok ? (double)hbit[i * width + j] : 0.0;
    cst[e] = 0.0;
  }
  if constexpr (DENSE) {
    __syncthreads();  // the hashes are staged
    if (tid < NL) {   // level tid's features by bucket, in feature order
      const int i = tid;
      int* bo = boff + i * (m + 1);
      int* cur = reinterpret_cast<int*>(bsk + i * mpad);  // scratch until the first sketch
      for (int p = 0; p <= m; ++p) bo[p] = 0;
      for (int j = 0; j < width; ++j)
        if ((unsigned)hidx[i * width + j] < (unsigned)m) ++bo[hs[i * wpad + j] + 1];
      for (int p = 0; p < m; ++p) bo[p + 1] += bo[p];
      for (int p = 0; p < m; ++p) cur[p] = bo[p];
      for (int j = 0; j < width; ++j)
        if ((unsigned)hidx[i * width + j] < (unsigned)m) bidx[i * width + cur[hs[i * wpad + j]]++] = j;
    }
  }
  int kc[kSketchKpt], kr[kSketchKpt];  // components; the ones read (0 past m)
#pragma unroll
  for (int c = 0; c < kSketchKpt; ++c) {
    kc[c] = l + c * w;
    kr[c] = kc[c] < m ? kc[c] : 0;
  }
  const int64_t level_stride = n_seq * (int64_t)steps * width;
  const T* ua = du + a * (int64_t)steps * width;
  auto stage = [&](int t) {  // signed increments of step t (after the signs)
    for (int e = tid; e < NL * width; e += bs) {
      const int i = e / width, j = e - i * width;
      cst[i * wpad + j] = sv[i * wpad + j] * (double)ua[i * level_stride + (int64_t)t * width + j];
    }
  };
  auto bucket = [&]() {  // DENSE: count sketches of the staged increments
    for (int e = tid; e < NL * m; e += bs) {
      const int i = e / m, p = e - i * m;
      const int* bo = boff + i * (m + 1);
      double sum = 0.0;
      for (int k = bo[p]; k < bo[p + 1]; ++k) sum += cst[i * wpad + bidx[i * width + k]];
      bsk[i * mpad + p] = sum;
    }
  };
  // the coefficients a convolution of level i reads, and their hashes
  auto coef = [&](int i) { return DENSE ? bsk + i * mpad : cst + i * wpad; };
  auto crow = [&](int i) { return sk_lds + (1 + i) * m2; };  // C_i, i < NL-1
  // at order p > 1: level i's order-r terms of this step, i < NL-1
  auto rrow = [&](int i, int r) {
    int base = NL;
    for (int q = 0; q < i; ++q) base += q + 1 < order ? q + 1 : order;
    return sk_lds + (base + r) * m2;
  };
  auto put = [&](double* row, int c, double v) {
    if (kc[c] < m) {
      row[kc[c]] = v;
      row[kc[c] + m] = v;
    }
  };
  // the partial sums v of slot s (a level, or an order term) of this group,
  // and group 0's total over the groups
  auto part = [&](int slot, const double* v) {
    if (ng > 1)
#pragma unroll
      for (int c = 0; c < kSketchKpt; ++c)
        if (kc[c] < m) red[(g * NL + slot) * m + kc[c]] = v[c];
  };
  auto total = [&](int slot, double* v) {
    if (ng > 1)
#pragma unroll
      for (int c = 0; c < kSketchKpt; ++c) {
        double sum = 0.0;
        for (int q = 0; q < ng; ++q) sum += red[(q * NL + slot) * m + kr[c]];
        v[c] = sum;
      }
  };
  // the top level's running sums (the lower levels' live in their rows)
  double Ctop[kSketchKpt];
#pragma unroll
  for (int c = 0; c < kSketchKpt; ++c) Ctop[c] = 0.0;
  __syncthreads();  // the signs (and buckets) are staged
  if (steps > 0) stage(0);
  __syncthreads();
  if constexpr (DENSE) {
    if (steps > 0) bucket();
    __syncthreads();
  }
  for (int t = 0; t < steps; ++t) {
    // DENSE: the convolutions read the sketches, so the next step's
    // increments are staged now and sketched after the mid-step barrier
    if constexpr (DENSE) {
      if (t + 1 < steps) stage(t + 1);
    }
    double S[NL][kSketchKpt];  // this step's terms summed over the orders
    if (order == 1) {
#pragma unroll
      for (int i = 0; i < NL; ++i) {
        sketch_conv<DENSE>(i == 0 ? sk_lds : crow(i - 1), coef(i), hs + i * wpad, j0, j1, m, kr,
                           S[i]);
        part(i, S[i]);
      }
      __syncthreads();  // every convolution has read the sums and the step
      if (g == 0)
#pragma unroll
        for (int i = 0; i < NL; ++i) total(i, S[i]);
    } else {
#pragma unroll
      for (int i = 0; i < NL; ++i) {
        const int d = i + 1 < order ? i + 1 : order;
        auto src = [&](int r) { return i == 0 ? sk_lds : (r == 0 ? crow(i - 1) : rrow(i - 1, r - 1)); };
        if (ng > 1) {
          for (int r = 0; r < d; ++r) {
            double v[kSketchKpt];
            sketch_conv<DENSE>(src(r), coef(i), hs + i * wpad, j0, j1, m, kr, v);
            part(r, v);
          }
          __syncthreads();  // the partials of every group are in
        }
        if (g == 0) {
#pragma unroll
          for (int c = 0; c < kSketchKpt; ++c) S[i][c] = 0.0;
          for (int r = 0; r < d; ++r) {
            double v[kSketchKpt];
            if (ng > 1) total(r, v);
            else sketch_conv<DENSE>(src(r), coef(i), hs + i * wpad, j0, j1, m, kr, v);
            const double wt = 1.0 / (r + 1);
#pragma unroll
            for (int c = 0; c < kSketchKpt; ++c) {
              v[c] *= wt;
              S[i][c] += v[c];
              if (i + 1 < NL) put(rrow(i, r), c, v[c]);
            }
          }
        }
        __syncthreads();  // level i's terms are visible to level i+1
      }
    }
    if (g == 0) {
#pragma unroll
      for (int i = 0; i + 1 < NL; ++i)
#pragma unroll
        for (int c = 0; c < kSketchKpt; ++c) put(crow(i), c, crow(i)[kr[c]] + S[i][c]);
#pragma unroll
      for (int c = 0; c < kSketchKpt; ++c) Ctop[c] += S[NL - 1][c];
    }
    if (t + 1 < steps) {
      if constexpr (DENSE) bucket();
      else stage(t + 1);
    }
    __syncthreads();  // sums and the next step staged before they are read
  }
  if (g != 0) return;
  double* o = out + a * (int64_t)NL * m;
#pragma unroll
  for (int c = 0; c < kSketchKpt; ++c) {
    if (kc[c] < m) {
#pragma unroll
      for (int i = 0; i + 1 < NL; ++i) o[(int64_t)i * m + kc[c]] = crow(i)[kc[c]];
      o[(int64_t)(NL - 1) * m + kc[c]] = Ctop[c];
    }
  }
}

// threads per group for m components, the groups per workgroup and the LDS
// bytes of one workgroup
__host__ inline int sketch_threads(int m) {
  const int per = (m + kSketchKpt - 1) / kSketchKpt;
  return ((per + kWave - 1) / kWave) * kWave;
}
// the count sketch first when it has fewer terms than the features
bool lowrank_sketch_dense(int m, int width) { return m < width; }
size_t lowrank_sketch_lds(int n_levels, int order, int m, int width, int ng) {
  const bool dense = lowrank_sketch_dense(m, width);
  return (size_t)sketch_rows(n_levels, order) * 2 * (size_t)m * sizeof(double) +
         (size_t)n_levels * sketch_wpad(width) * (2 * sizeof(double) + sizeof(int)) +
         (ng > 1 ? (size_t)ng * n_levels * m * sizeof(double) : 0) +
         (dense ? (size_t)n_levels * (sketch_wpad(m) * sizeof(double) +
                                      (size_t)(m + 1 + width) * sizeof(int))
                : 0);
}
// up to 8 waves per sequence where the components leave room, no more groups
// than 4-term slices, and fewer where the LDS would not fit
int lowrank_sketch_groups(int n_levels, int order, int m, int width, size_t lds_limit) {
  const int w = sketch_threads(m);
  int ng = 512 / w;
  if (ng < 1) ng = 1;
  const int nterm = lowrank_sketch_dense(m, width) ? m : width;
  const int slices = (nterm + 3) / 4;
  if (ng > slices) ng = slices > 0 ? slices : 1;
  while (ng > 1 && lowrank_sketch_lds(n_levels, order, m, width, ng) > lds_limit) --ng;
  return ng;
}
int lowrank_sketch_max_comp() { return 1024 * kSketchKpt; }

template <typename T, int NL, bool DENSE>
static hipError_t launch_sketch(const void* du, int64_t n_seq, int steps, int width,
                                const int* hidx, const void* hbit, int m, int order, int ng,
                                double* out, hipStream_t s) {
  auto k = lowrank_sketch_kernel<T, NL, DENSE>;
  const size_t lds = lowrank_sketch_lds(NL, order, m, width, ng);
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)k,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k, dim3((unsigned)n_seq), dim3(ng * sketch_threads(m)), lds, s,
                     reinterpret_cast<const T*>(du), n_seq, steps, width, hidx,
                     reinterpret_cast<const T*>(hbit), m, order, ng, out);
  return hipGetLastError();
}

template <typename T>
static hipError_t sketch_levels(int n_levels, const void* du, int64_t n_seq, int steps, int width,
                                const int* hidx, const void* hbit, int m, int order, int ng,
                                double* out, hipStream_t s) {
  const bool dense = lowrank_sketch_dense(m, width);
  switch (n_levels) {
#define KSIG_SKETCH_CASE(L)                                                                   \
    case L:                                                                                   \
      return dense ? launch_sketch<T, L, true>(du, n_seq, steps, width, hidx, hbit, m, order, \
                                               ng, out, s)                                    \
                   : launch_sketch<T, L, false>(du, n_seq, steps, width, hidx, hbit, m,       \
                                                order, ng, out, s);
    KSIG_SKETCH_CASE(1)
    KSIG_SKETCH_CASE(2)
    KSIG_SKETCH_CASE(3)
    KSIG_SKETCH_CASE(4)
    KSIG_SKETCH_CASE(5)
    KSIG_SKETCH_CASE(6)
    KSIG_SKETCH_CASE(7)
    KSIG_SKETCH_CASE(8)
#undef KSIG_SKETCH_CASE
  }
  return hipErrorInvalidValue;
}

// 1 <= order <= n_levels <= 8, 1 <= m <= lowrank_sketch_max_comp(), ng from
// lowrank_sketch_groups and the LDS of lowrank_sketch_lds within the
// device's limit (checked by the caller)
hipError_t lowrank_sketch_launch(int dtype, const void* du, int n_levels, int order,
                                 int64_t n_seq, int steps, int width, const int* hidx,
                                 const void* hbit, int m, int ng, double* out, hipStream_t s) {
  if (dtype == KSIG_F32)
    return sketch_levels<float>(n_levels, du, n_seq, steps, width, hidx, hbit, m, order, ng, out,
                                s);
  return sketch_levels<double>(n_levels, du, n_seq, steps, width, hidx, hbit, m, order, ng, out,
                               s);
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
