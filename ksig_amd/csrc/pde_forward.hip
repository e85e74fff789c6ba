// Signature-PDE kernel, forward pass (gfx950).
//
// Solves the explicit Goursat scheme of ksig/algorithms.py:380-385
//   U(i,j) = (U(i-1,j) + U(i,j-1)) (1 + M/2 + M^2/12) - U(i-1,j-1) (1 - M^2/12)
// with U = 1 on the row/column before the grid, returning U at the last cell
// (ksig/algorithms.py:344-399, host launcher :402-446).  The reference sweeps
// antidiagonals with two block barriers per step; here the same recurrence is
// evaluated as a skewed systolic wave: lane i owns R consecutive refined rows
// of each 64R-row band (R = 2 or 4 where registers allow; at dyadic order 1
// each pair of rows is one original row, whose static-kernel value the two
// rows share), walks the
// refined columns one step behind lane i-1 and receives U of the row above
// its first row from it through DPP wave_shr:1 (lane 0 from an LDS FIFO per
// column), so a step costs no barrier at all.  A step's inputs (the y row of
// the column, lane 0's FIFO entry) are read a step ahead.
//
// Precision: the update is evaluated in increment form.  With
// V(i,j) = U(i,j) - U(i,j-1), the scheme reads
//   V(i,j) = V(i-1,j) + δ(i,j),   U(i,j) = U(i,j-1) + V(i,j),
//   δ = (U(i-1,j) + U(i,j-1)) (M/2 + M^2/12) + U(i-1,j-1) M^2/12,
// which is algebraically identical but never subtracts two O(U) terms, so
// float32 arithmetic keeps ~1e-7 relative accuracy where the direct form
// loses ~1e-4 (V flows down the column like the DP's column sums).
//
// Dyadic refinement (no reference counterpart; dyadic_order = 0 is the
// reference scheme): each ΔΔ cell becomes a 2^λ × 2^λ block of cells carrying
// M / 4^λ.  The first refined row of each original row ("leader" lane)
// evaluates one static-kernel value G(r+1, c+1) per original column, every
// 2^λ steps (the leaders of a wave do this in lock step); the other lanes of
// the block receive the two Gram values they need from the lane above.

#include <type_traits>

#include "ksig_common.h"
#include "ksig_hip.h"
#include "dp_forward.h"
#include "pde_fwd_aligned.h"

namespace ksig {

template <typename T, int DMAX, int KIND, int R>
__global__ void __launch_bounds__(kWave)
pde_forward_kernel(const PdeFwdArgs a) {
  constexpr int NF = 4;  // flow: [G(r, c+1), G(r+1, c+1), U, V]
  constexpr bool kGramIn = KIND == kGram;
  constexpr int BR = kWave * R;  // refined rows per band
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* smem = reinterpret_cast<T*>(smem_raw);

  const int lane = threadIdx.x;
  const int64_t pr = blockIdx.x;
  const int d = a.d;

  const T* xs = nullptr;
  const T* ys = nullptr;
  int Lx, Ly;
  const T* gm = nullptr;
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
    if (Ly > Lx) {  // the Goursat problem is symmetric under transposition
      const T* tp = xs; xs = ys; ys = tp;
      int tl = Lx; Lx = Ly; Ly = tl;
    }
  }
  const int diff = a.difference;
  const int Nr = diff ? Lx - 1 : Lx;
  const int Nc = diff ? Ly - 1 : Ly;
  if (Nr <= 0 || Nc <= 0) {
    if (lane == 0) a.out[pr] = 1.0;
    return;
  }
  const int lam = a.dyadic;
  const int F = 1 << lam;
  const int Nrr = Nr << lam, Ncr = Nc << lam;
  // priming column (F steps) + refined columns; more than 64 so that lane 0's
  // FIFO read, issued a step ahead, follows lane 63's write of the slot
  int Pc = (Nc + 1) << lam;
  if (Pc <= kWave) Pc = kWave + 1;
  const int nb = (Nrr + BR - 1) / BR;
  const T inv4 = T(1) / T(F * F);
  const int rshift = diff ? 1 : 0;

  const T xscale = T(a.xscale), yscale = T(a.yscale);
  T* ylds = smem;
  T* ynrm = ylds + (kGramIn ? 0 : Ly * DMAX);
  T* fifo = ynrm + (kGramIn ? 0 : Ly);

  if constexpr (!kGramIn) {
    for (int j = lane; j < Ly; j += kWave) {
      T n2 = T(0);
#pragma unroll
      for (int k = 0; k < DMAX; ++k) {
        T v = k < d ? ys[(int64_t)j * d + k] * yscale : T(0);
        ylds[j * DMAX + k] = v;
        n2 += v * v;
      }
      ynrm[j] = n2;
    }
  }
  __syncthreads();

  auto load_x = [&](T* xv, T& xnrm, int xr) {
    const bool ok = xr >= 0 && xr < Lx && !kGramIn;
    T n2 = T(0);
#pragma unroll
    for (int k = 0; k < DMAX; ++k) {
      T v = (ok && k < d) ? xs[(int64_t)xr * d + k] * xscale : T(0);
      xv[k] = v;
      n2 += v * v;
    }
    xnrm = n2;
  };

  // FIFO initial contents: the row above the grid, G(0, c+1) as "bottom"
  // value and U = 1.  Slot s covers refined column s - F (original column
  // (s >> lam) - 1, whose right edge is G(0, s >> lam)).
  {
    T x0[DMAX], x0n;
    load_x(x0, x0n, 0);
    for (int s = lane; s < Pc; s += kWave) {
      const int yc = (s >> lam);
      T g = T(0);
      if (diff && yc < Ly) {
        if constexpr (kGramIn) {
          g = gm[yc];
        } else {
          const T* yr = ylds + yc * DMAX;
          T dot = T(0);
#pragma unroll
          for (int k = 0; k < DMAX; ++k) dot += x0[k] * yr[k];
          g = KIND == kRBF ? fast_exp2(-(x0n + ynrm[yc] - T(2) * dot)) : dot;
        }
      }
      fifo[s * NF + 0] = T(0);
      fifo[s * NF + 1] = g;
      fifo[s * NF + 2] = T(1);
      fifo[s * NF + 3] = T(0);
    }
  }
  __syncthreads();

  // Per-lane state: R consecutive refined rows of the band.
  int band = 0;
  int rr0 = lane * R;            // first refined row of the lane
  T xv[R][DMAX], xnrm[R];
  T xv2[R][DMAX], xnrm2[R];      // next band's rows, prefetched
#pragma unroll
  for (int q = 0; q < R; ++q) {
    load_x(xv[q], xnrm[q], ((rr0 + q) >> lam) + rshift);
    load_x(xv2[q], xnrm2[q], ((rr0 + q + BR) >> lam) + rshift);
  }

  double* edges = a.edges ? reinterpret_cast<double*>(a.edges) + a.e_off[pr] : nullptr;

  T fl[NF] = {T(0), T(0), T(1), T(0)};
  T gt_old[R], gb_old[R], gt[R], gb[R], M[R], uleft[R], udiag[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    gt_old[q] = gb_old[q] = gt[q] = gb[q] = M[q] = T(0);
    uleft[q] = udiag[q] = T(1);
  }
  int cc = -F - lane;            // refined column (priming columns are -F..-1)

  // Inputs of one step, read a step ahead: the y row and norm of the
  // original column whose right edge the step's leader rows evaluate (or the
  // caller's values), and lane 0's FIFO entry.
  struct StepIn {
    T y[kGramIn ? R : DMAX];
    T yn;
    T f[NF];
  };
  auto fetch = [&](StepIn& in, int c, int b, int r0) __attribute__((always_inline)) {
    int yc = (c >> lam) + rshift;
    if constexpr (kGramIn) {
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int xr = ((r0 + q) >> lam) + rshift;
        in.y[q] = (xr < Lx && yc >= 0 && yc < Ly) ? gm[(int64_t)xr * Ly + yc] : T(0);
      }
    } else {
      yc = yc < 0 ? 0 : (yc >= Ly ? Ly - 1 : yc);
#pragma unroll
      for (int k = 0; k < DMAX; ++k) in.y[k] = ylds[yc * DMAX + k];
      in.yn = ynrm[yc];
    }
    in.f[0] = T(0);
    in.f[1] = T(0);
    in.f[2] = T(1);
    in.f[3] = T(0);
    if (lane == 0 && c >= -F && b < nb) {
      const int s = c + F;
#pragma unroll
      for (int k = 0; k < NF; ++k) in.f[k] = fifo[s * NF + k];
    }
  };

  StepIn cur;
  fetch(cur, cc, band, rr0);
  const int total = nb * Pc + kWave - 1;
  for (int step = 0; step < total; ++step) {
    const bool active = cc >= -F && band < nb;
    const T in_gt = shift_down(fl[0], cur.f[0]);
    const T in_gb = shift_down(fl[1], cur.f[1]);
    const T in_u = shift_down(fl[2], cur.f[2]);
    const T in_v = shift_down(fl[3], cur.f[3]);
    // static-kernel values of the leader rows at the right edge of the
    // original column (used at the column's first refined step)
    T gnew[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      if constexpr (kGramIn) {
        gnew[q] = cur.y[q];
      } else {
        T dot = T(0);
#pragma unroll
        for (int k = 0; k < DMAX; ++k) dot += xv[q][k] * cur.y[k];
        gnew[q] = KIND == kRBF ? fast_exp2(-(xnrm[q] + cur.yn - T(2) * dot)) : dot;
      }
    }
    {
      const bool wrap = cc + 1 == Pc - F;
      fetch(cur, wrap ? -F : cc + 1, band + (wrap ? 1 : 0), wrap ? rr0 + BR : rr0);
    }

    const bool newcol = (cc & (F - 1)) == 0;  // first refined step of a column
    const bool colok = active && cc >= 0 && cc < Ncr;
    T igt = in_gt, igb = in_gb, iu = in_u, iv = in_v;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int rr = rr0 + q;
      const bool leader = (rr & (F - 1)) == 0;
      // Gram values of this row's original cell (right edge, column ocol+1)
      const T ngt = leader ? igb : igt;
      const T ngb = leader ? (newcol ? gnew[q] : gb[q]) : igb;
      if (newcol) {
        gt_old[q] = gt[q];
        gb_old[q] = gb[q];
        const T m = diff ? ((ngb - gb_old[q]) - (ngt - gt_old[q])) : ngb;
        M[q] = m * inv4;
      }
      gt[q] = ngt;
      gb[q] = ngb;

      const bool cell = colok && rr < Nrr;
      const T b = M[q] * M[q] * T(1.0 / 12.0);
      const T aa = T(0.5) * M[q] + b;
      const T delta = (iu + uleft[q]) * aa + udiag[q] * b;
      T v = iv + delta;
      T u = uleft[q] + v;
      u = cell ? u : T(1);  // priming / padding columns carry the boundary value
      v = cell ? v : iv;
      if (cell && rr == Nrr - 1 && cc == Ncr - 1) a.out[pr] = (double)u;
      if (edges != nullptr && cell) {  // last refined row, then last refined column
        if (rr == Nrr - 1) edges[cc] = u;
        if (cc == Ncr - 1) edges[Ncr + rr] = u;
      }
      udiag[q] = iu;
      uleft[q] = u;
      igt = ngt;
      igb = ngb;
      iu = u;
      iv = v;
    }

    fl[0] = igt;
    fl[1] = igb;
    fl[2] = iu;
    fl[3] = iv;
    if (lane == kWave - 1 && active) {
      const int s = cc + F;
      fifo[s * NF + 0] = igt;
      fifo[s * NF + 1] = igb;
      fifo[s * NF + 2] = iu;
      fifo[s * NF + 3] = iv;
    }

    ++cc;
    if (cc == Pc - F) {  // wrap to the next band
      cc = -F;
      ++band;
      rr0 += BR;
#pragma unroll
      for (int q = 0; q < R; ++q) {
#pragma unroll
        for (int k = 0; k < DMAX; ++k) xv[q][k] = xv2[q][k];
        xnrm[q] = xnrm2[q];
        load_x(xv2[q], xnrm2[q], ((rr0 + q + BR) >> lam) + rshift);
        uleft[q] = T(1);
        gt[q] = gb[q] = T(0);
      }
    }
  }
}

// Refined rows per lane from register pressure (x rows of this and the next
// band stay in registers): 4 up to 4 float32 / 2 float64 features, 2 up to
// 8 float64 / 16 float32 features, else 1.
template <typename T, int DMAX>
constexpr int pde_rows() { return DMAX * (int)sizeof(T) <= 16 ? 4 : (DMAX * (int)sizeof(T) <= 64 ? 2 : 1); }

template <typename T, int DMAX, int KIND>
static hipError_t launch_pde(const PdeFwdArgs& a, size_t lds, hipStream_t s) {
  if (a.dyadic <= 1)
    return sizeof(T) == 4 ? pde_fwd_aligned_launch_f32(a, KIND == kGram ? 8 : DMAX, lds, s)
                          : pde_fwd_aligned_launch_f64(a, KIND == kGram ? 8 : DMAX, lds, s);
  auto k = pde_forward_kernel<T, DMAX, KIND, pde_rows<T, DMAX>()>;
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)k,
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k, dim3((unsigned)a.n_pairs), dim3(kWave), lds, s, a);
  return hipGetLastError();
}

template <typename T, int KIND>
static hipError_t dispatch_pde_d(const PdeFwdArgs& a, size_t lds, hipStream_t s) {
  if (a.d <= 4) return launch_pde<T, 4, KIND>(a, lds, s);
  if (a.d <= 8) return launch_pde<T, 8, KIND>(a, lds, s);
  if (a.d <= 16) return launch_pde<T, 16, KIND>(a, lds, s);
  return launch_pde<T, 32, KIND>(a, lds, s);
}

template <typename T>
static hipError_t dispatch_pde(const PdeFwdArgs& a, size_t lds, hipStream_t s) {
  if (a.kind == kGram) return launch_pde<T, 8, kGram>(a, lds, s);
  if (a.kind == kRBF) return dispatch_pde_d<T, kRBF>(a, lds, s);
  return dispatch_pde_d<T, kLinear>(a, lds, s);
}

hipError_t pde_forward_launch(const PdeFwdArgs& a, int dtype, size_t lds, hipStream_t s) {
  if (dtype == KSIG_F32) return dispatch_pde<float>(a, lds, s);
  return dispatch_pde<double>(a, lds, s);
}

}  // namespace ksig
