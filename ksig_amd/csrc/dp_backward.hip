// Adjoint of the first-order DP signature kernel (gfx950).
//
// The reference has no gradients (its recursion is CuPy-only); this is the
// reverse-mode derivative of the recursion of ksig/algorithms.py:47-85 on
// M = ΔΔ k(x_i, y_j) (ksig/algorithms.py:64, ksig/static/kernels.py:145-155,
// 248-260).  For L = sum_m w_m K_m with K_m = sum R_m, R_1 = M,
// R_m = M * E(R_{m-1}):
//   λ_n = w_n,  λ_m = w_m + E'(M λ_{m+1})    (E' = exclusive 2-D sum over the
//                                              cells below and to the right)
//   dL/dM(i,j) = λ_1 + sum_{m>=2} λ_m(i,j) P_{m-1}(i,j),   P_a = E(R_a).
//
// The wave sweeps the grid in reverse (bottom-up, right-to-left) with the same
// systolic structure as the forward: lane i owns one row, column sums flow
// from lane to lane (DPP), lane 63 hands over to the next band through an LDS
// FIFO.  The forward prefixes P_a are not stored per cell: the forward leaves
// the final column sums C_a(Nr, j) and each row's end prefix P_a(i, Nc), and the
// reverse sweep recovers every P_a(i, j) on the fly from
//   C_a(i, j) = C_a(i+1, j) - R_a(i, j),   P_a(i, j) = P_a(i, j+1) - C_a(i, j).
// The reverse-exclusive sums of the adjoint use the mirrored decomposition
// D_m (sum below, flows up) and S_m (row suffix, lane-local).
//
// Chain to the static kernel: dL/dG = ΔΔ^T(dL/dM) is finalised one step late
// (the lane for M row i owns G point (i+1, j+1)); the kernel writes
// omega = dL/dG * G (RBF) or dL/dG (linear) per G point, in the caller's
// (x, y) orientation, plus per-pair sum omega * D' (RBF, D' = scaled squared
// distance) or omega * G (linear) for the bandwidth / scale gradient.  The
// contraction of omega with the inputs (d-dimensional GEMMs) is left to the
// caller.  An extra virtual row (M row -1) and an extra column step (column
// -1) finalise G row 0 and G column 0.

#include "ksig_common.h"
#include "ksig_hip.h"
#include "dp_forward.h"

namespace ksig {

template <typename T, int NLEV, int DMAX>
__global__ void __launch_bounds__(kWave)
dp_backward_kernel(const DpBwdArgs a) {
  constexpr int NL1 = NLEV > 1 ? NLEV - 1 : 1;
  constexpr int NC = NLEV - 1;                  // prefix levels 1..NLEV-1
  constexpr int NF = 3 + 2 * NC;                // [G, D', gM, C_1.., Dq_1..]
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* smem = reinterpret_cast<T*>(smem_raw);

  const int lane = threadIdx.x;
  const int64_t pr = blockIdx.x;
  const int d = a.d;
  const int kind = a.kind;
  const int32_t ia = a.pairs[2 * pr];
  const int32_t ib = a.pairs[2 * pr + 1];
  const T* xs = reinterpret_cast<const T*>(a.x.data) + a.x.off[ia] * d;
  const T* ys = reinterpret_cast<const T*>(a.y.data) + a.y.off[ib] * d;
  int Lx = a.x.len[ia], Ly = a.y.len[ib];
  const bool swapped = Ly > Lx;  // same orientation rule as the forward
  if (swapped) {
    const T* tp = xs; xs = ys; ys = tp;
    int tl = Lx; Lx = Ly; Ly = tl;
  }
  const int Nr = Lx - 1, Nc = Ly - 1;
  if (Nr <= 0 || Nc <= 0) {
    if (lane == 0) a.gparam[pr] = 0.0;
    return;
  }
  const int nrows = Nr + 1;  // + virtual row above the grid
  const int Pc = Nc + 2 > kWave ? Nc + 2 : kWave;  // priming, Nc columns, column -1
  const int nb = (nrows + kWave - 1) / kWave;

  T w[NLEV + 1];
#pragma unroll
  for (int m = 1; m <= NLEV; ++m) w[m] = T(a.gout[pr * (NLEV + 1) + m]);
  const T* rowend = reinterpret_cast<const T*>(a.save) + a.save_off[pr];  // [NC][Nr]
  const T* colsum = rowend + NC * Nr;                                       // [NC][Nc]
  T* omega = reinterpret_cast<T*>(a.omega) + a.omega_off[pr];

  const T xscale = T(a.xscale), yscale = T(a.yscale);
  // LDS: x rows [2][64][DMAX], x norms [2][64], row-end prefixes [2][64][NL1],
  // y rows [Ly][DMAX], y norms [Ly], FIFO [Pc][NF].
  T* xl = smem;
  T* xnl = xl + 2 * kWave * DMAX;
  T* rel = xnl + 2 * kWave;
  T* yl = rel + 2 * kWave * NL1;
  T* ynl = yl + Ly * DMAX;
  T* fifo = ynl + Ly;

  // Stage flipped row fr (M row i = Nr-1-fr, whose top edge G row is i) into buffer b.
  auto stage = [&](int b, int fr) {
    const int i = Nr - 1 - fr;
    const bool ok = i >= 0 && fr < nrows;
    T n2 = T(0);
#pragma unroll
    for (int k = 0; k < DMAX; ++k) {
      const T v = (ok && k < d) ? xs[(int64_t)i * d + k] * xscale : T(0);
      xl[(b * kWave + lane) * DMAX + k] = v;
      n2 += v * v;
    }
    xnl[b * kWave + lane] = n2;
#pragma unroll
    for (int m = 0; m < NC; ++m)
      rel[(b * kWave + lane) * NL1 + m] = ok ? rowend[m * Nr + i] : T(0);
  };

  for (int j = lane; j < Ly; j += kWave) {
    T n2 = T(0);
#pragma unroll
    for (int k = 0; k < DMAX; ++k) {
      const T v = k < d ? ys[(int64_t)j * d + k] * yscale : T(0);
      yl[j * DMAX + k] = v;
      n2 += v * v;
    }
    ynl[j] = n2;
  }
  stage(0, lane);
  stage(1, kWave + lane);
  __syncthreads();

  // G value and its "distance" companion (D' for RBF, G for linear) for x row
  // held in buffer b and y row yc.
  auto gpoint = [&](int b, int yc, T& dv) -> T {
    const T* xr = xl + (b * kWave + lane) * DMAX;
    const T* yr = yl + yc * DMAX;
    T dot = T(0);
#pragma unroll
    for (int k = 0; k < DMAX; ++k) dot += xr[k] * yr[k];
    if (kind == kRBF) {
      dv = xnl[b * kWave + lane] + ynl[yc] - T(2) * dot;
      return fast_exp2(-dv);
    }
    dv = dot;
    return dot;
  };

  // FIFO: the row below the grid (G row Nr = Lx-1) with the forward's final
  // column sums; slot s covers column j = Nc - s.
  {
    T xb[DMAX];
    T xbn = T(0);
#pragma unroll
    for (int k = 0; k < DMAX; ++k) {
      xb[k] = k < d ? xs[(int64_t)Nr * d + k] * xscale : T(0);
      xbn += xb[k] * xb[k];
    }
    for (int s = lane; s < Pc; s += kWave) {
      const int j = Nc - s;
      T g = T(0), dv = T(0);
      if (j >= 0) {
        const T* yr = yl + j * DMAX;
        T dot = T(0);
#pragma unroll
        for (int k = 0; k < DMAX; ++k) dot += xb[k] * yr[k];
        if (kind == kRBF) {
          dv = xbn + ynl[j] - T(2) * dot;
          g = fast_exp2(-dv);
        } else {
          dv = g = dot;
        }
      }
      fifo[s * NF + 0] = g;
      fifo[s * NF + 1] = dv;
      fifo[s * NF + 2] = T(0);
#pragma unroll
      for (int m = 0; m < NC; ++m) {
        fifo[s * NF + 3 + m] = (j >= 0 && j < Nc) ? colsum[m * Nc + j] : T(0);
        fifo[s * NF + 3 + NC + m] = T(0);
      }
    }
  }
  __syncthreads();

  T fl[NF];
#pragma unroll
  for (int k = 0; k < NF; ++k) fl[k] = T(0);
  T P[NL1], S[NL1];
#pragma unroll
  for (int m = 0; m < NL1; ++m) P[m] = S[m] = T(0);
  T g_right = T(0), gup_right = T(0), dup_right = T(0), gm_right = T(0), gmup_right = T(0);
  T gsig_row = T(0);
  double gsig = 0.0;

  int s = -lane;  // step within the band: column j = Nc - s
  int band = 0;
  int fr = lane;  // flipped row
  const int total = nb * Pc + kWave - 1;
  int next_stage = Pc + kWave;
  int stage_band = 2;
  for (int step = 0; step < total; ++step) {
    if (step == next_stage) {
      stage(stage_band & 1, stage_band * kWave + lane);
      next_stage += Pc;
      ++stage_band;
    }
    const bool active = s >= 0 && band < nb;
    T fin[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) fin[k] = T(0);
    if (lane == 0 && active) {
#pragma unroll
      for (int k = 0; k < NF; ++k) fin[k] = fifo[s * NF + k];
    }
    T in[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) in[k] = shift_down(fl[k], fin[k]);

    const int j = Nc - s;
    const int i = Nr - 1 - fr;
    const int b = band & 1;
    const int yc = j < 0 ? 0 : (j > Ly - 1 ? Ly - 1 : j);
    T dv;
    T g = gpoint(b, yc, dv);
    const bool gvalid = active && i >= 0 && j >= 0;
    g = gvalid ? g : T(0);
    dv = gvalid ? dv : T(0);
    const bool prim = s == 0;
    const bool cell = active && i >= 0 && j >= 0 && j < Nc;
    const T M = cell ? (gup_right - in[0]) - (g_right - g) : T(0);

    // forward reconstruction: C_a(i,j), P_a(i,j)
    T Pn[NL1], Cout[NL1];
    {
      T R = M;
#pragma unroll
      for (int m = 0; m < NC; ++m) {
        const T C = in[3 + m] - R;
        Pn[m] = P[m] - C;
        Cout[m] = C;
        R = M * Pn[m];
      }
    }
    // adjoint
    T lam[NLEV + 1];
    lam[NLEV] = w[NLEV];
#pragma unroll
    for (int m = NLEV - 1; m >= 1; --m) lam[m] = w[m] + S[m - 1];
    T gM = lam[1];
#pragma unroll
    for (int m = 2; m <= NLEV; ++m) gM += lam[m] * Pn[m - 2];
    gM = cell ? gM : T(0);
    T Dout[NL1];
#pragma unroll
    for (int m = 0; m < NC; ++m) Dout[m] = in[3 + NC + m] + M * lam[m + 2];

    // finalise dL/dG at G point (i+1, j+1)
    const bool fin_ok = active && !prim && i >= -1 && j >= -1;
    const T gG = (gM - gm_right) - (in[2] - gmup_right);
    const T om = kind == kRBF ? gG * gup_right : gG;
    if (fin_ok) {
      const int p = i + 1, q = j + 1;
      const int64_t idx = swapped ? (int64_t)q * Lx + p : (int64_t)p * Ly + q;
      omega[idx] = om;
      gsig_row += om * dup_right;
    }

    // state
    g_right = g;
    gup_right = in[0];
    dup_right = in[1];
    gm_right = gM;
    gmup_right = in[2];
#pragma unroll
    for (int m = 0; m < NC; ++m) {
      P[m] = prim ? rel[(b * kWave + lane) * NL1 + m] : Pn[m];
      S[m] = prim ? T(0) : S[m] + in[3 + NC + m];
    }
    T of[NF];
    of[0] = g;
    of[1] = dv;
    of[2] = gM;
#pragma unroll
    for (int m = 0; m < NC; ++m) {
      of[3 + m] = Cout[m];
      of[3 + NC + m] = Dout[m];
    }
#pragma unroll
    for (int k = 0; k < NF; ++k) fl[k] = of[k];
    if (lane == kWave - 1 && active) {
#pragma unroll
      for (int k = 0; k < NF; ++k) fifo[s * NF + k] = of[k];
    }

    ++s;
    if (s == Pc) {
      s = 0;
      ++band;
      fr += kWave;
      gsig += (double)gsig_row;
      gsig_row = T(0);
    }
  }
  gsig += (double)gsig_row;
  gsig = wave_sum(gsig);
  if (lane == 0) a.gparam[pr] = gsig;
}

template <typename T, int NLEV, int DMAX>
static hipError_t launch_bwd(const DpBwdArgs& a, size_t lds, hipStream_t s) {
  auto k = dp_backward_kernel<T, NLEV, DMAX>;
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)k,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k, dim3((unsigned)a.n_pairs), dim3(kWave), lds, s, a);
  return hipGetLastError();
}

#define KSIG_BWD_LEVELS(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(12)

template <typename T, int DMAX>
static hipError_t dispatch_bwd_l(const DpBwdArgs& a, size_t lds, hipStream_t s) {
#define KSIG_CASE(L) if (a.n_levels == L) return launch_bwd<T, L, DMAX>(a, lds, s);
  KSIG_BWD_LEVELS(KSIG_CASE)
#undef KSIG_CASE
  return hipErrorInvalidValue;
}

template <typename T>
static hipError_t dispatch_bwd(const DpBwdArgs& a, size_t lds, hipStream_t s) {
  if (a.d <= 8) return dispatch_bwd_l<T, 8>(a, lds, s);
  if (a.d <= 16) return dispatch_bwd_l<T, 16>(a, lds, s);
  return dispatch_bwd_l<T, 32>(a, lds, s);
}

bool dp_backward_supported(int n_levels) {
#define KSIG_SUP(L) if (n_levels == L) return true;
  KSIG_BWD_LEVELS(KSIG_SUP)
#undef KSIG_SUP
  return false;
}

hipError_t dp_backward_launch(const DpBwdArgs& a, int dtype, size_t lds, hipStream_t s) {
  if (dtype == KSIG_F32) return dispatch_bwd<float>(a, lds, s);
  return dispatch_bwd<double>(a, lds, s);
}

}  // namespace ksig
