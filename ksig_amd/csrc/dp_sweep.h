// Truncated signature kernel by dynamic programming, forward pass, fused
// static kernel evaluated on the vector ALUs (gfx950).  Template
// implementation; instantiated by dp_sweep_*.hip.  This is the production
// path for difference=True with the RBF / linear static kernels; the
// caller-supplied-M path, difference=False and the training forward (which
// also saves adjoint boundary state) use dp_forward_impl.h.
//
// Recursion: ksig/algorithms.py:47-85 (first order) and :88-140 (higher
// order) on M = ΔΔ k(x_i, y_j) (ksig/algorithms.py:64), k from
// ksig/static/kernels.py:248-260 (RBF) or :145-155 (linear).
//
// Systolic wave as in dp_forward_impl.h (lane i owns grid row 64*band + i and
// walks the columns one step behind lane i-1; column sums flow down the lanes
// by DPP, lane 63 hands over to the next band through an LDS FIFO), with the
// per-step work cut to what the recursion needs:
//   * the lane's x row lives in registers; it is reloaded from an LDS staging
//     buffer right after the last dot product of a band, so the load hides
//     behind the next step's y-row read;
//   * y rows sit in LDS with a row stride of an odd number of 16-byte quads,
//     so the skewed per-lane row reads are bank-conflict free;
//   * all lanes read lane 0's FIFO slot (one broadcast address);
//   * inputs are replicated past the sequence ends (rows clamped to Lx-1,
//     columns to Ly-1), which makes ΔΔ exactly zero there: no bounds masks;
//   * the row-start reset of the running row sums is a multiply by a 0/1
//     factor fused into the update (the column inputs are zero in the
//     priming column).
#pragma once

#include "ksig_common.h"
#include "ksig_hip.h"
#include "dp_forward.h"
#include "dp_forward_impl.h"

namespace ksig {

// One cell of the level recursion with the row-start reset folded in:
// keepf = 0 in the priming column of a row (where M = 0 and the column inputs
// are 0), 1 elsewhere.
template <typename T, int NLEV, int ORDER>
__device__ __forceinline__ void dp_cell_keep(const T M, const T* in, T* out, T* PS, T* HB,
                                             const T keepf) {
  dp_cell_core<T, NLEV, ORDER>(M, in, out, PS, HB,
                               [&](T old, T inc) { return dp_fma(old, keepf, inc); });
}

template <typename T, int DMAX>
struct SweepLayout {
  // LDS row strides in elements: an odd number of 16-byte quads.
  static constexpr int QUAD = 16 / sizeof(T);
  static constexpr int RS = ((DMAX / QUAD) % 2 == 0) ? DMAX + QUAD : DMAX + 2 * QUAD;
};

// R = rows per lane: lane i owns grid rows 64R*band + R*i + k, k < R, and
// evaluates them top to bottom in the same step (the column values pass from
// row k to row k+1 in registers; only the last row's go to lane i+1).  The y
// row, the FIFO read and the lane-to-lane shifts are shared by the R rows.
template <typename T, int NLEV, int ORDER, int DMAX, int R>
__global__ void __launch_bounds__(kWave)
dp_sweep_kernel(const DpFwdArgs a) {
  constexpr int NSIDE = dp_num_side(NLEV, ORDER);
  constexpr int NF = 1 + NLEV + NSIDE;
  constexpr int NPS = NLEV > 1 ? NLEV - 1 : 1;
  constexpr int NHB = NSIDE > 0 ? NSIDE : 1;
  constexpr int RS = SweepLayout<T, DMAX>::RS;  // features + norm at [DMAX]
  constexpr int RB = kWave * R;                   // rows per band

  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* smem = reinterpret_cast<T*>(smem_raw);

  const int lane = threadIdx.x;
  const int64_t pr = blockIdx.x;
  double* out = a.out + pr * (NLEV + 1);
  const int d = a.d;
  const bool rbf = a.kind == kRBF;

  const int32_t ia = a.pairs[2 * pr];
  const int32_t ib = a.pairs[2 * pr + 1];
  const T* xs = reinterpret_cast<const T*>(a.x.data) + a.x.off[ia] * d;
  int Lx = a.x.len[ia];
  const T* ys = reinterpret_cast<const T*>(a.y.data) + a.y.off[ib] * d;
  int Ly = a.y.len[ib];
  if (Ly > Lx) {
    const T* tp = xs; xs = ys; ys = tp;
    int tl = Lx; Lx = Ly; Ly = tl;
  }
  const int Nr = Lx - 1, Nc = Ly - 1;
  if (Nr <= 0 || Nc <= 0) {
    if (lane <= NLEV) out[lane] = lane == 0 ? 1.0 : 0.0;
    return;
  }
  // Pc >= 65: lane 0's FIFO slot for step T+1 (loaded during step T) was
  // written by lane 63 at step T+1-Pc+63 < T.
  const int Pc = Nc + 1 > kWave ? Nc + 1 : kWave + 1;
  const int nb = (Nr + RB - 1) / RB;
  const T xscale = T(a.xscale), yscale = T(a.yscale);

  // LDS: x staging [RB][RS] (one band), y rows [Ly][RS], FIFO [Pc][NF].
  T* xl = smem;
  T* yl = xl + RB * RS;
  T* fifo = yl + Ly * RS;

  auto stage_x = [&](int band) {
#pragma unroll
    for (int k = 0; k < R; ++k) {
      int xr = band * RB + R * lane + k + 1;
      xr = xr < Lx - 1 ? xr : Lx - 1;
      T n2 = T(0);
      T* dst = xl + (R * lane + k) * RS;
#pragma unroll
      for (int f = 0; f < DMAX; ++f) {
        const T v = f < d ? xs[(int64_t)xr * d + f] * xscale : T(0);
        dst[f] = v;
        n2 += v * v;
      }
      dst[DMAX] = n2;
    }
  };
  T xr[R][DMAX];
  T xn[R];
  auto load_xregs = [&]() {
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const T* src = xl + (R * lane + k) * RS;
#pragma unroll
      for (int f = 0; f < DMAX; ++f) xr[k][f] = src[f];
      xn[k] = src[DMAX];
    }
  };

  for (int j = lane; j < Ly; j += kWave) {
    T n2 = T(0);
#pragma unroll
    for (int f = 0; f < DMAX; ++f) {
      const T v = f < d ? ys[(int64_t)j * d + f] * yscale : T(0);
      yl[j * RS + f] = v;
      n2 += v * v;
    }
    yl[j * RS + DMAX] = n2;
  }
  stage_x(0);
  __syncthreads();

  // FIFO: G(0, c') of the row above the grid (columns clamped), zero sums.
  {
    T x0[DMAX];
    T x0n = T(0);
#pragma unroll
    for (int f = 0; f < DMAX; ++f) {
      x0[f] = f < d ? xs[f] * xscale : T(0);
      x0n += x0[f] * x0[f];
    }
    for (int s = lane; s < Pc; s += kWave) {
      const int yc = s < Ly - 1 ? s : Ly - 1;
      const T* yr = yl + yc * RS;
      T dot = T(0);
#pragma unroll
      for (int f = 0; f < DMAX; ++f) dot += x0[f] * yr[f];
      fifo[s * NF] = rbf ? fast_exp2(-(x0n + yr[DMAX] - T(2) * dot)) : dot;
#pragma unroll
      for (int k = 1; k < NF; ++k) fifo[s * NF + k] = T(0);
    }
  }
  load_xregs();
  __syncthreads();
  if (nb > 1) stage_x(1);  // every lane holds band 0 in registers

  T fl[NF];
#pragma unroll
  for (int k = 0; k < NF; ++k) fl[k] = T(0);
  T PS[R][NPS];
  T HB[R][NHB];
  T gleft[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
#pragma unroll
    for (int m = 0; m < NPS; ++m) PS[k][m] = T(0);
#pragma unroll
    for (int m = 0; m < NHB; ++m) HB[k][m] = T(0);
    gleft[k] = T(0);
  }
  T gdiag = T(0);

  // Column c' of G this lane evaluates; lane i starts i steps late (negative
  // c' = not started: such steps are treated as priming columns, so the lane
  // carries zeros until its first band begins).
  int cp = -lane;
  const int ncols = nb * Pc;
  const int total = ncols + kWave - 1;
  int s0 = 0;                  // lane 0's FIFO slot: T mod Pc
  int s63 = Pc - (kWave - 1);  // lane 63's FIFO slot: (T - 63) mod Pc
  // Band b >= 2 is staged at the start of step (b-1)*Pc + 64: lane 63 loaded
  // band b-1 from the buffer at the end of step (b-1)*Pc + 62, and lane 0
  // loads band b at the end of step b*Pc - 1 (Pc >= 65).
  int next_stage = Pc + kWave;
  int stage_band = 2;

  // Software pipeline: the y row (features + norm) and lane 0's FIFO slot of
  // step T+1 are loaded while step T computes; two register sets alternate
  // (the loop body holds two steps).
  T yA[DMAX + 1], yB[DMAX + 1], fA[NF], fB[NF];
  auto load_next = [&](T* yv, T* fv, int c, int slot) {
    const unsigned yc = c < 0 ? 0u : (unsigned)(c < Ly - 1 ? c : Ly - 1);
    const T* yr = yl + __umul24(yc, (unsigned)RS);
#pragma unroll
    for (int f = 0; f <= DMAX; ++f) yv[f] = yr[f];
#pragma unroll
    for (int k = 0; k < NF; ++k) fv[k] = fifo[slot * NF + k];
  };
  auto step = [&](int T_, const T* yv, const T* fv, T* ynext, T* fnext) {
    if (T_ == next_stage) {
      if (stage_band < nb) stage_x(stage_band);
      next_stage += Pc;
      ++stage_band;
    }
    {
      int cn = cp + 1;
      cn = cn == Pc ? 0 : cn;
      int sn = s0 + 1;
      sn = sn == Pc ? 0 : sn;
      load_next(ynext, fnext, cn, sn);
    }
    T in[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) in[k] = shift_down(fl[k], fv[k]);

    T g[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      T dot = T(0);
#pragma unroll
      for (int f = 0; f < DMAX; ++f) dot += xr[k][f] * yv[f];
      g[k] = rbf ? fast_exp2(-(xn[k] + yv[DMAX] - T(2) * dot)) : dot;
    }
    const bool keep = cp > 0;
    const T keepf = keep ? T(1) : T(0);
    T Mk[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const T up = k == 0 ? in[0] : g[k - 1];
      const T dg = k == 0 ? gdiag : gleft[k - 1];
      const T m = (g[k] - gleft[k]) - (up - dg);
      Mk[k] = keep ? m : T(0);
    }
#pragma unroll
    for (int k = 0; k < R; ++k) gleft[k] = g[k];
    gdiag = in[0];

    T cur[NF];
#pragma unroll
    for (int j = 0; j < NF; ++j) cur[j] = in[j];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      T of[NF];
      of[0] = g[k];
      dp_cell_keep<T, NLEV, ORDER>(Mk[k], cur, of, PS[k], HB[k], keepf);
#pragma unroll
      for (int j = 0; j < NF; ++j) cur[j] = of[j];
    }
#pragma unroll
    for (int k = 0; k < NF; ++k) fl[k] = cur[k];
    if (lane == kWave - 1 && T_ >= kWave - 1 && T_ < total) {
#pragma unroll
      for (int k = 0; k < NF; ++k) fifo[s63 * NF + k] = cur[k];
    }

    ++cp;
    if (cp == Pc) {  // this lane enters its next band: load that band's x rows
      cp = 0;
      load_xregs();
    }
    ++s0;
    s0 = s0 == Pc ? 0 : s0;
    ++s63;
    s63 = s63 == Pc ? 0 : s63;
  };

  load_next(yA, fA, cp, s0);
  for (int T_ = 0; T_ < total; T_ += 2) {
    step(T_, yA, fA, yB, fB);
    step(T_ + 1, yB, fB, yA, fA);
  }
  __syncthreads();

  double acc[NLEV];
#pragma unroll
  for (int m = 0; m < NLEV; ++m) acc[m] = 0.0;
  for (int s = lane + 1; s <= Nc; s += kWave) {
#pragma unroll
    for (int m = 0; m < NLEV; ++m) acc[m] += (double)fifo[s * NF + 1 + m];
  }
#pragma unroll
  for (int m = 0; m < NLEV; ++m) acc[m] = wave_sum(acc[m]);
  if (lane == 0) out[0] = 1.0;
#pragma unroll
  for (int m = 0; m < NLEV; ++m)
    if (lane == m + 1) out[m + 1] = acc[m];
}

template <typename T, int NLEV, int ORDER, int DMAX>
hipError_t launch_dp_sweep(const DpFwdArgs& a, size_t lds, hipStream_t s) {
  constexpr int R = sizeof(T) == 4 ? 4 : 1;
  auto k = dp_sweep_kernel<T, NLEV, ORDER, DMAX, R>;
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)k,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k, dim3((unsigned)a.n_pairs), dim3(kWave), lds, s, a);
  return hipGetLastError();
}

}  // namespace ksig
