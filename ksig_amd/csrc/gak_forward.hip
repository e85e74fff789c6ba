// Global alignment kernel (log of the value), forward pass (gfx950).
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
// a skewed systolic wave (as the DP row stream): lane i owns R consecutive
// rows of each 64R-row band, walks the columns one step behind lane i-1 and
// receives K(r-1, c) of the row above its first row from it through DPP
// wave_shr:1; the R cells of a step share the column's y row and the step's
// control work.  Lane 63's last row goes to an LDS FIFO (one slot per
// column) that lane 0 reads in the next band.  No barrier
// inside the sweep.  The longer sequence is the row side (GA is symmetric
// under transposition: the alignment paths and M' transpose together).
//
// Representation: the sum over alignments is carried in exp space, each
// lane's values as float64 multiples of 2^E with one int32 exponent E per
// lane, so the cell update is the exp-space form of the same recurrence,
//   K(i,j) = M'(i,j) (K(i-1,j) + K(i,j-1) + K(i-1,j-1)),
// two adds and a multiply per cell.  The value the lane above hands down
// (with its own exponent) is scaled into this lane's by one ldexp per step
// (exact), and every kGakRenorm steps a lane rescales its values so that the
// largest has exponent 0 (exact), so nothing overflows or underflows at any
// length: between rescalings a value grows at most by (3 M'max)^kGakRenorm
// and adjacent rows differ by bounded factors.  Power-of-two scalings are
// exact, so each add and multiply rounds as it would on the unscaled values:
// one float64 rounding per add and per multiply, no exp / log per cell (the
// log-sum-exp form needs three exp and a log).  A term more than 1074
// binades below the lane's scale flushes to zero, as exp(a - m) underflows
// in the log-sum-exp.  The result is log v + E ln 2 at the last cell.  The
// static-kernel value and M' = max(M / (2 - M), 1e-12) are computed in the
// input precision (float32 inputs: v_exp_f32 and v_rcp_f32, ~1 ulp each).
// An empty sum (log 0) is v = 0; a handed-down empty sum carries exponent
// kGakZeroExp.

#include <math.h>

#include "ksig_common.h"
#include "ksig_hip.h"
#include "gak.h"

namespace ksig {

constexpr int kGakZeroExp = -(1 << 29);
constexpr double kLn2 = 0.69314718055994530942;

// M' = max(M / (2 - M), 1e-12) from the static-kernel value M.
__device__ __forceinline__ double gak_mprime(float m) {
  const float mp = m * __builtin_amdgcn_rcpf(2.0f - m);
  return (double)fmaxf(mp, 1e-12f);
}
__device__ __forceinline__ double gak_mprime(double m) {
  return fmax(m / (2.0 - m), 1e-12);
}

// Steps between a lane's rescalings, and how far (in binades) the row above
// may outgrow the lane's scale before the lane rescales to it at once: with
// M' < 2^54 (float64 inputs next to the pole M = 2) values grow at most
// 2^(55 kGakRenorm) between rescalings, so 2^(55 kGakRenorm) * 2^kGakRescale *
// 2^(55 kGakRenorm) stays far below the float64 range.
constexpr int kGakRenorm = 4;
constexpr int kGakRescale = 128;

// MB: pairs may span several row bands; the next band's x rows are then
// staged by the whole wave into a second register set at one step a band
// ahead (lane l changes band at its own step b*Pc - 1 + l; loading there ran
// the loads and norms under one lane's mask at 64 steps of every band).
// Launches whose rows all fit one band use MB = false and keep the
// registers.
template <typename T, int DMAX, int KIND, int R, bool MB>
__global__ void __launch_bounds__(kWave)
gak_forward_kernel(const GakArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr bool kGramIn = KIND == kGram;
  constexpr int BR = kWave * R;  // rows per band

  const int lane = threadIdx.x;
  const int64_t pr = blockIdx.x;
  const int d = a.d;

  const T* xs = nullptr;
  const T* ys = nullptr;
  const T* gm = nullptr;
  int Lx, Ly;
  if constexpr (kGramIn) {
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
  // Columns per band pass; more than 64 so that lane 0's FIFO read, issued
  // one step ahead, comes at least a step after lane 63's write of the slot.
  const int Pc = Ly <= kWave ? kWave + 1 : Ly;
  const int nb = (Lx + BR - 1) / BR;
  if (!MB && nb > 1) {  // a row side longer than the launch's stated maximum
    if (lane == 0) a.out[pr] = NAN;
    return;
  }

  // LDS: FIFO mantissas [Pc] (double) and exponents [Pc] (int), then y rows
  // [Ly][DMAX] (scaled) and norms [Ly].
  double* fifo_v = reinterpret_cast<double*>(smem_raw);
  int* fifo_e = reinterpret_cast<int*>(fifo_v + Pc);
  T* ylds = reinterpret_cast<T*>(fifo_e + ((Pc + 3) & ~3));
  T* ynrm = ylds + (kGramIn ? 0 : Ly * DMAX);
  const T xscale = T(a.xscale), yscale = T(a.yscale);
  if constexpr (!kGramIn) {
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

  // this lane's R rows of band `b`: x rows scaled, with their norms (and,
  // MB, the next band's)
  T xv[R][DMAX], xn[R];
  T pxv[MB ? R : 1][DMAX], pxn[MB ? R : 1];
  auto load_rows = [&](int b, T (&dv)[MB ? R : 1][DMAX], T (&dn)[MB ? R : 1]) __attribute__((always_inline)) {
#pragma unroll
    for (int q = 0; q < (MB ? R : 1); ++q) {
      const int xr = b * BR + lane * R + q;
      const bool ok = xr < Lx && !kGramIn;
      T n2 = T(0);
#pragma unroll
      for (int k = 0; k < DMAX; ++k) {
        const T v = (ok && k < d) ? xs[(int64_t)xr * d + k] * xscale : T(0);
        dv[q][k] = v;
        n2 += v * v;
      }
      dn[q] = n2;
    }
  };
  auto load_first = [&]() __attribute__((always_inline)) {  // band 0 into the current set
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int xr = lane * R + q;
      const bool ok = xr < Lx && !kGramIn;
      T n2 = T(0);
#pragma unroll
      for (int k = 0; k < DMAX; ++k) {
        const T v = (ok && k < d) ? xs[(int64_t)xr * d + k] * xscale : T(0);
        xv[q][k] = v;
        n2 += v * v;
      }
      xn[q] = n2;
    }
  };

  // Inputs of one step, read a step ahead: the column's y row and norm (or
  // the caller's matrix values of the lane's rows), and lane 0's FIFO entry.
  struct StepIn {
    T y[kGramIn ? R : DMAX];
    T yn;
    double fv;
    int fe;
  };
  auto fetch = [&](StepIn& in, int c, int b) __attribute__((always_inline)) {
    const int cy = c < 0 ? 0 : (c >= Ly ? Ly - 1 : c);
    if constexpr (kGramIn) {
#pragma unroll
      for (int q = 0; q < R; ++q) {
        int rr = b * BR + lane * R + q;
        rr = rr < Lx ? rr : Lx - 1;
        in.y[q] = gm[(int64_t)rr * Ly + cy];
      }
    } else {
#pragma unroll
      for (int k = 0; k < DMAX; ++k) in.y[k] = ylds[cy * DMAX + k];
      in.yn = ynrm[cy];
    }
    in.fv = 0.0;
    in.fe = kGakZeroExp;
    if (lane == 0 && b > 0 && b < nb && c >= 0) {
      in.fv = fifo_v[c];
      in.fe = fifo_e[c];
    }
  };
  // M' of row q of this lane and the fetched column.
  auto mprime = [&](int q, const StepIn& in) __attribute__((always_inline)) -> double {
    if constexpr (kGramIn) {
      return gak_mprime(in.y[q]);
    } else {
      T dot = T(0);
#pragma unroll
      for (int k = 0; k < DMAX; ++k) dot += xv[q][k] * in.y[k];
      if constexpr (KIND == kRBF) return gak_mprime(fast_exp2(-(xn[q] + in.yn - T(2) * dot)));
      else return gak_mprime(dot);
    }
  };

  int band = 0;
  load_first();
  // MB: band b+1 is staged at step b*Pc + 64, when every lane has taken band
  // b and before lane 0 takes band b+1 (Pc >= 65)
  int next_stage = kWave;
  int stage_band = 1;
  // K of this lane's rows at column c-1 and of the row above at c-1 (the
  // first row's diagonal), as multiples of 2^E; the last row's value handed
  // down with its exponent
  double lv[R];
#pragma unroll
  for (int q = 0; q < R; ++q) lv[q] = 0.0;
  int E = 0;
  double pin_v = 0.0, up_v = 0.0;
  int up_e = kGakZeroExp;
  int cc = -lane;

  StepIn cur;
  fetch(cur, cc, band);
  const int total = nb * Pc + kWave - 1;
  for (int step = 0; step < total; ++step) {
    if constexpr (MB) {
      if (step == next_stage) {  // wave-uniform
        if (stage_band < nb) load_rows(stage_band, pxv, pxn);
        ++stage_band;
        next_stage += Pc;
      }
    }
    const bool active = cc >= 0 && band < nb;
    const bool colok = active && cc < Ly;
    const bool c0 = cc == 0;
    // row above the lane's first row: lane i-1's last row, or the FIFO
    // (row -1 in band 0: empty sums)
    const double in_v = shift_down(up_v, cur.fv);
    const int in_e = __builtin_amdgcn_update_dpp(cur.fe, up_e, 0x138, 0xf, 0xf, false);
    double mp[R];
#pragma unroll
    for (int q = 0; q < R; ++q) mp[q] = mprime(q, cur);
    // next step's inputs, in flight while this step computes
    {
      const bool wrap = cc + 1 == Pc;
      fetch(cur, wrap ? 0 : cc + 1, band + (wrap ? 1 : 0));
    }

    // the first column of a band pass takes the exponent handed down (the
    // lane's own values are empty sums there); later the row above may
    // outgrow the lane's scale, and the lane rescales to it
    if (c0) E = in_v != 0.0 ? in_e : 0;
    int al = in_e - E;
    if (al > kGakRescale) {
#pragma unroll
      for (int q = 0; q < R; ++q) lv[q] = __builtin_amdgcn_ldexp(lv[q], -al);
      pin_v = __builtin_amdgcn_ldexp(pin_v, -al);
      E = in_e;
      al = 0;
    }
    const double in_s = __builtin_amdgcn_ldexp(in_v, al < -2000 ? -2000 : al);

    const int row0 = band * BR + lane * R;
    double uv = in_s, dv = pin_v;   // K(row-1, c), K(row-1, c-1)
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int row = row0 + q;
      // column -1 holds empty sums (the left values are reset at each band
      // start, and rows q >= 1 take their diagonal from them), except
      // K(-1, -1) = 1 above-left of (0, 0) (exponent 0: E = 0 there); the
      // first row's diagonal at column 0 comes from the row above, which is
      // not yet in the band
      double dq = dv;
      if (q == 0) dq = c0 ? (row == 0 ? 1.0 : 0.0) : dv;
      double v = ((uv + lv[q]) + dq) * mp[q];
      const bool cell = colok && row < Lx;
      v = cell ? v : 0.0;
      if (cell && row == Lx - 1 && cc == Ly - 1) a.out[pr] = log(v) + (double)E * kLn2;
      dv = lv[q];
      lv[q] = v;
      uv = v;
    }
    pin_v = in_s;
    if ((step & (kGakRenorm - 1)) == kGakRenorm - 1) {
      double mx = pin_v;
#pragma unroll
      for (int q = 0; q < R; ++q) mx = fmax(mx, lv[q]);
      if (mx > 0.0) {
        const int sh = __builtin_amdgcn_frexp_exp(mx);
#pragma unroll
        for (int q = 0; q < R; ++q) lv[q] = __builtin_amdgcn_ldexp(lv[q], -sh);
        pin_v = __builtin_amdgcn_ldexp(pin_v, -sh);
        E += sh;
      }
    }
    uv = lv[R - 1];
    if (lane == kWave - 1 && active) {
      fifo_v[cc] = uv;
      fifo_e[cc] = E;
    }
    up_v = uv;
    up_e = E;

    ++cc;
    if (cc == Pc) {  // wrap to the next band: column -1 of the new rows
      cc = 0;
      ++band;
      if constexpr (MB) {
        if (band < nb) {
#pragma unroll
          for (int q = 0; q < R; ++q) {
#pragma unroll
            for (int k = 0; k < DMAX; ++k) xv[q][k] = pxv[q][k];
            xn[q] = pxn[q];
          }
        }
      }
#pragma unroll
      for (int q = 0; q < R; ++q) lv[q] = 0.0;
    }
  }
}

template <typename T, int DMAX, int KIND, int R, bool MB>
static hipError_t launch_gak_mb(const GakArgs& a, size_t lds, hipStream_t s) {
  auto k = gak_forward_kernel<T, DMAX, KIND, R, MB>;
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)k,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k, dim3((unsigned)a.n_pairs), dim3(kWave), lds, s, a);
  return hipGetLastError();
}

// one band (every pair's row side within 64R rows) or several
template <typename T, int DMAX, int KIND, int R>
static hipError_t launch_gak(const GakArgs& a, size_t lds, hipStream_t s) {
  if (a.rows_max > kWave * R) return launch_gak_mb<T, DMAX, KIND, R, true>(a, lds, s);
  return launch_gak_mb<T, DMAX, KIND, R, false>(a, lds, s);
}

// Rows per lane from register pressure: float32 4 (d <= 16) or 2, float64 2
// (d <= 8) or 1.
template <typename T, int KIND>
static hipError_t dispatch_gak_d(const GakArgs& a, size_t lds, hipStream_t s) {
  constexpr bool f32 = sizeof(T) == 4;
  if (a.d <= 8) return launch_gak<T, 8, KIND, f32 ? 4 : 2>(a, lds, s);
  if (a.d <= 16) return launch_gak<T, 16, KIND, f32 ? 4 : 1>(a, lds, s);
  return launch_gak<T, 32, KIND, f32 ? 2 : 1>(a, lds, s);
}

template <typename T>
static hipError_t dispatch_gak(const GakArgs& a, size_t lds, hipStream_t s) {
  if (a.kind == kGram) return launch_gak<T, 8, kGram, sizeof(T) == 4 ? 4 : 2>(a, lds, s);
  if (a.kind == kRBF) return dispatch_gak_d<T, kRBF>(a, lds, s);
  return dispatch_gak_d<T, kLinear>(a, lds, s);
}

size_t gak_lds_bytes(int dtype, int kind, int d, int cols) {
  const size_t esz = dtype == KSIG_F32 ? 4 : 8;
  const int dmax = d <= 8 ? 8 : (d <= 16 ? 16 : 32);
  const size_t pc = cols <= kWave ? kWave + 1 : cols;
  const size_t ycols = kind == kGram ? 0 : (size_t)cols;
  return pc * 8 + ((pc + 3) & ~(size_t)3) * 4 + (ycols * dmax + ycols) * esz;
}

hipError_t gak_forward_launch(const GakArgs& a, int dtype, size_t lds, hipStream_t s) {
  if (dtype == KSIG_F32) return dispatch_gak<float>(a, lds, s);
  return dispatch_gak<double>(a, lds, s);
}

}  // namespace ksig
