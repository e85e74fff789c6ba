// Global alignment kernel in log space, forward pass (gfx950).
//
// Computes log GA(x, y) of ksig/algorithms.py:514-559 (launcher) and
// :453-511 (device recursion): with the transformed static kernel
// M'(i,j) = M/(2 - M) (made "infinitely divisible", :532) and
// l(i,j) = log(max(M'(i,j), 1e-12)) (:534),
//   logK(i,j) = l(i,j) + LSE(logK(i-1,j), logK(i,j-1), logK(i-1,j-1)),
// logK(-1,-1) = 0, logK(-1,j) = logK(i,-1) = -inf, result logK(Lx-1, Ly-1),
// LSE by the max-shift trick (:492-496).
//
// The reference sweeps antidiagonals with two block barriers per step and a
// three-slot LDS ring per antidiagonal element.  Here the same recurrence is
// a skewed systolic wave (as the signature-PDE forward): lane i owns row
// 64*band + i, walks the columns one step behind lane i-1 and receives
// logK(r-1, c) from it through DPP wave_shr:1; lane 63's values go to an LDS
// FIFO (one slot per column) that lane 0 reads in the next band.  No barrier
// inside the sweep.  The longer sequence is the row side (GA is symmetric
// under transposition: the alignment paths and M' transpose together).
//
// Precision: the log values grow like the path length, so they are carried
// in float64 (full rate on the vector ALUs of gfx950); for float32 inputs the
// transcendentals (exp / log of the shifted LSE terms, of the static kernel)
// run on the float32 hardware units (v_exp_f32 / v_log_f32), whose results
// are O(1) and enter the float64 sum with absolute error ~1e-7 per cell.
// float64 inputs use float64 exp / log throughout.
//
// log(0) is represented by kGakNeg = -1e300: every real cell has at least one
// finite predecessor, and exp(kGakNeg - m) underflows to exactly 0 for any
// finite m, while cells outside the grid (all three inputs kGakNeg) stay
// finite instead of producing NaN.

#include <math.h>

#include "ksig_common.h"
#include "ksig_hip.h"
#include "gak.h"

namespace ksig {

constexpr double kGakNeg = -1e300;
constexpr double kLn2 = 0.69314718055994530942;
constexpr double kLog2e = 1.44269504088896340736;

// log(max(M/(2-M), 1e-12)) from the static-kernel value M (linear, Gram).
__device__ __forceinline__ double gak_logm_direct(float m) {
  const float mp = m / (2.0f - m);
  return (double)(__builtin_amdgcn_logf(fmaxf(mp, 1e-12f)) * (float)kLn2);
}
__device__ __forceinline__ double gak_logm_direct(double m) {
  return log(fmax(m / (2.0 - m), 1e-12));
}

// The same for RBF from q = |x' - y'|^2 with M = 2^-q (scales folded in):
// log M' = -q ln 2 - log(2 - M), clamped at log(1e-12).
__device__ __forceinline__ double gak_logm_rbf(float q) {
  const float m = __builtin_amdgcn_exp2f(-q);
  const float lg = -q * (float)kLn2 - __builtin_amdgcn_logf(2.0f - m) * (float)kLn2;
  return (double)fmaxf(lg, -27.631021115928547f);
}
__device__ __forceinline__ double gak_logm_rbf(double q) {
  const double m = exp2(-q);
  return fmax(-q * kLn2 - log(2.0 - m), -27.631021115928547);
}

// log(e^a + e^b + e^c) with m = max(a, b, c) shifted out.
template <typename T>
__device__ __forceinline__ double gak_lse3(double a, double b, double c) {
  const double m = fmax(fmax(a, b), c);
  if constexpr (sizeof(T) == 4) {
    const float s = __builtin_amdgcn_exp2f((float)((a - m) * kLog2e)) +
                    __builtin_amdgcn_exp2f((float)((b - m) * kLog2e)) +
                    __builtin_amdgcn_exp2f((float)((c - m) * kLog2e));
    return m + (double)(__builtin_amdgcn_logf(s) * (float)kLn2);
  } else {
    return m + log(exp(a - m) + exp(b - m) + exp(c - m));
  }
}

template <typename T, int DMAX>
__global__ void __launch_bounds__(kWave)
gak_forward_kernel(const GakArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];

  const int lane = threadIdx.x;
  const int64_t pr = blockIdx.x;
  const int kind = a.kind;
  const int d = a.d;

  const T* xs = nullptr;
  const T* ys = nullptr;
  const T* gm = nullptr;
  int Lx, Ly;
  if (kind == kGram) {
    Lx = a.gram_lx;
    Ly = a.gram_ly;
    gm = reinterpret_cast<const T*>(a.gram) + pr * (int64_t)Lx * Ly;
  } else {
    const int32_t ia = a.pairs[2 * pr];
    const int32_t ib = a.pairs[2 * pr + 1];
    xs = reinterpret_cast<const T*>(a.x.data) + a.x.off[ia] * d;
    Lx = a.x.len[ia];
    ys = reinterpret_cast<const T*>(a.y.data) + a.y.off[ib] * d;
    Ly = a.y.len[ib];
    if (Ly > Lx) {
      const T* tp = xs; xs = ys; ys = tp;
      const int tl = Lx; Lx = Ly; Ly = tl;
    }
  }
  if (Lx <= 0 || Ly <= 0) {
    if (lane == 0) a.out[pr] = -INFINITY;  // empty sum of alignments
    return;
  }
  const int Pc = Ly < kWave ? kWave : Ly;
  const int nb = (Lx + kWave - 1) / kWave;

  // LDS: FIFO [Pc] doubles, then y rows [Ly][DMAX] (scaled) and norms [Ly].
  double* fifo = reinterpret_cast<double*>(smem_raw);
  T* ylds = reinterpret_cast<T*>(fifo + Pc);
  T* ynrm = ylds + (kind == kGram ? 0 : Ly * DMAX);
  const T xscale = T(a.xscale), yscale = T(a.yscale);
  if (kind != kGram) {
    for (int j = lane; j < Ly; j += kWave) {
      T n2 = T(0);
#pragma unroll
      for (int k = 0; k < DMAX; ++k) {
        const T v = k < d ? ys[(int64_t)j * d + k] * yscale : T(0);
        ylds[j * DMAX + k] = v;
        n2 += v * v;
      }
      ynrm[j] = n2;
    }
  }
  __syncthreads();

  auto load_x = [&](T* xv, T& xnrm, int xr) {
    const bool ok = xr < Lx && kind != kGram;
    T n2 = T(0);
#pragma unroll
    for (int k = 0; k < DMAX; ++k) {
      const T v = (ok && k < d) ? xs[(int64_t)xr * d + k] * xscale : T(0);
      xv[k] = v;
      n2 += v * v;
    }
    xnrm = n2;
  };
  // log M'(r, c) of this lane's row r (x row in registers) and column c.
  auto logm = [&](const T* xv, T xnrm, int r, int c) -> double {
    if (kind == kGram) return gak_logm_direct(gm[(int64_t)r * Ly + c]);
    const T* yr = ylds + c * DMAX;
    T dot = T(0);
#pragma unroll
    for (int k = 0; k < DMAX; ++k) dot += xv[k] * yr[k];
    if (kind == kRBF) return gak_logm_rbf(xnrm + ynrm[c] - T(2) * dot);
    return gak_logm_direct(dot);
  };

  int band = 0;
  int r = lane;
  T xv[DMAX], xnrm;
  load_x(xv, xnrm, r);
  T xv2[DMAX], xnrm2;  // next band's row, prefetched
  load_x(xv2, xnrm2, r + kWave);

  double up = kGakNeg;     // logK(r, c) flowing down to row r+1
  double left = kGakNeg;   // logK(r, c-1)
  double diag = kGakNeg;   // logK(r-1, c-1)
  int cc = -lane;

  const int total = nb * Pc + kWave - 1;
  for (int step = 0; step < total; ++step) {
    const bool active = cc >= 0 && band < nb;
    double fin = kGakNeg;  // row above the band: row -1 (band 0) or the FIFO
    if (lane == 0 && active && band > 0) fin = fifo[cc];
    const double in_up = shift_down(up, fin);

    const bool cell = active && cc < Ly && r < Lx;
    double val = kGakNeg;
    if (cell) {
      const double dg = cc == 0 ? (r == 0 ? 0.0 : kGakNeg) : diag;
      const double lf = cc == 0 ? kGakNeg : left;
      val = logm(xv, xnrm, r, cc) + gak_lse3<T>(in_up, lf, dg);
      if (r == Lx - 1 && cc == Ly - 1) a.out[pr] = val;
    }
    if (lane == kWave - 1 && active) fifo[cc] = val;
    up = val;
    left = val;
    diag = in_up;

    ++cc;
    if (cc == Pc) {  // wrap to the next band
      cc = 0;
      ++band;
      r += kWave;
#pragma unroll
      for (int k = 0; k < DMAX; ++k) xv[k] = xv2[k];
      xnrm = xnrm2;
      load_x(xv2, xnrm2, r + kWave);
      left = kGakNeg;
    }
  }
}

template <typename T, int DMAX>
static hipError_t launch_gak(const GakArgs& a, size_t lds, hipStream_t s) {
  auto k = gak_forward_kernel<T, DMAX>;
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)k,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k, dim3((unsigned)a.n_pairs), dim3(kWave), lds, s, a);
  return hipGetLastError();
}

template <typename T>
static hipError_t dispatch_gak(const GakArgs& a, size_t lds, hipStream_t s) {
  if (a.kind == kGram || a.d <= 8) return launch_gak<T, 8>(a, lds, s);
  if (a.d <= 16) return launch_gak<T, 16>(a, lds, s);
  return launch_gak<T, 32>(a, lds, s);
}

size_t gak_lds_bytes(int dtype, int kind, int d, int cols) {
  const size_t esz = dtype == KSIG_F32 ? 4 : 8;
  const int dmax = d <= 8 ? 8 : (d <= 16 ? 16 : 32);
  const size_t pc = cols < kWave ? kWave : cols;
  const size_t ycols = kind == kGram ? 0 : (size_t)cols;
  return pc * 8 + (ycols * dmax + ycols) * esz;
}

hipError_t gak_forward_launch(const GakArgs& a, int dtype, size_t lds, hipStream_t s) {
  if (dtype == KSIG_F32) return dispatch_gak<float>(a, lds, s);
  return dispatch_gak<double>(a, lds, s);
}

}  // namespace ksig
