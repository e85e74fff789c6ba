// Adjoint of the first-order DP signature kernel (gfx950), float64 (the
// float32 adjoint, with the static Gram on MFMA, is dp_backward_mfma.h).
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
#include "dp_backward_mfma.h"

namespace ksig {

// LDS of the adjoint: lane-private staging of the next band's x row and row-end
// prefixes, the y norms, the FIFO.  The y rows themselves stream from L2.
template <typename T, int NLEV, int DMAX>
struct BwdLayout {
  static constexpr int NL1 = NLEV > 1 ? NLEV - 1 : 1;
  static constexpr int NF = 3 + 2 * (NLEV - 1);
  static constexpr int XS = DMAX + 1 + NL1;  // x row, norm, row-end prefixes
};

// One sequence point into registers: 16-byte vector loads when VEC (d == DMAX,
// aligned data), else one load per slot with the address clamped into the row
// and the slots past d zeroed.  Both forms are straight-line code, which keeps
// the compiler's wait counts exact so the two-step prefetch stays in flight.
template <typename T, int DMAX, bool VEC>
__device__ __forceinline__ void bwd_load_row(T* v, const T* p, int d) {
  if (VEC) {
    if constexpr (sizeof(T) == 4) {
#pragma unroll
      for (int f = 0; f < DMAX; f += 4) {
        const float4 t = *reinterpret_cast<const float4*>(p + f);
        v[f] = t.x; v[f + 1] = t.y; v[f + 2] = t.z; v[f + 3] = t.w;
      }
    } else {
#pragma unroll
      for (int f = 0; f < DMAX; f += 2) {
        const double2 t = *reinterpret_cast<const double2*>(p + f);
        v[f] = t.x; v[f + 1] = t.y;
      }
    }
  } else {
#pragma unroll
    for (int f = 0; f < DMAX; ++f) {
      const T t = p[f < d ? f : d - 1];
      v[f] = f < d ? t : T(0);
    }
  }
}

template <typename T, int NLEV, int DMAX, bool VEC>
__global__ void __launch_bounds__(kWave)
dp_backward_kernel(const DpBwdArgs a) {
  using L = BwdLayout<T, NLEV, DMAX>;
  constexpr int NL1 = L::NL1;
  constexpr int NC = NLEV - 1;                  // prefix levels 1..NLEV-1
  constexpr int NF = L::NF;                     // [G, D', gM, C_1.., Dq_1..]
  constexpr int XS = L::XS;
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
    if (lane == 0 && a.gparam) a.gparam[pr] = 0.0;
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

  // Scale factors are folded into the x side: staged x rows carry
  // xscale * yscale (so y rows are used raw), x norms xscale^2 |x|^2 and the y
  // norms yscale^2 |y|^2.
  const T xscale = T(a.xscale), yscale = T(a.yscale);
  const T xyscale = xscale * yscale, ys2 = yscale * yscale;
  T* xst = smem;                      // [64][XS], lane-private slots
  T* ynl = xst + kWave * XS;          // [Ly] scaled y norms
  T* fifo = ynl + Ly;                 // [Pc][NF]

  // Stage flipped row fr (M row i = Nr-1-fr, whose top edge G row is i) into
  // this lane's slot.
  auto stage = [&](int fr) {
    const int i = Nr - 1 - fr;
    const bool ok = i >= 0 && fr < nrows;
    T* dst = xst + lane * XS;
    T n2 = T(0);
#pragma unroll
    for (int k = 0; k < DMAX; ++k) {
      const T v = (ok && k < d) ? xs[(int64_t)i * d + k] : T(0);
      dst[k] = v * xyscale;
      n2 += v * v;
    }
    dst[DMAX] = n2 * xscale * xscale;
#pragma unroll
    for (int m = 0; m < NC; ++m) dst[DMAX + 1 + m] = ok ? rowend[m * Nr + i] : T(0);
  };
  T xr[DMAX], xn, rel[NL1];
  auto take = [&]() {
    const T* src = xst + lane * XS;
#pragma unroll
    for (int k = 0; k < DMAX; ++k) xr[k] = src[k];
    xn = src[DMAX];
#pragma unroll
    for (int m = 0; m < NL1; ++m) rel[m] = src[DMAX + 1 + m];
  };
  // raw y row yc into registers
  auto yrow = [&](T* dst, int yc) { bwd_load_row<T, DMAX, VEC>(dst, ys + (int64_t)yc * d, d); };

  for (int j = lane; j < Ly; j += kWave) {
    T yv[DMAX];
    yrow(yv, j);
    T n2 = T(0);
#pragma unroll
    for (int k = 0; k < DMAX; ++k) n2 += yv[k] * yv[k];
    ynl[j] = n2 * ys2;
  }

  // FIFO: the row below the grid (G row Nr = Lx-1) with the forward's final
  // column sums; slot s covers column j = Nc - s.
  {
    T xb[DMAX];
    T xbn = T(0);
#pragma unroll
    for (int k = 0; k < DMAX; ++k) {
      const T v = k < d ? xs[(int64_t)Nr * d + k] : T(0);
      xb[k] = v * xyscale;
      xbn += v * v;
    }
    xbn *= xscale * xscale;
    for (int s = lane; s < Pc; s += kWave) {
      const int j = Nc - s;
      T g = T(0), dv = T(0);
      if (j >= 0) {
        T yv[DMAX];
        yrow(yv, j);
        T dot = T(0), yn = T(0);
#pragma unroll
        for (int k = 0; k < DMAX; ++k) {
          dot += xb[k] * yv[k];
          yn += yv[k] * yv[k];
        }
        yn *= ys2;
        if (kind == kRBF) {
          dv = xbn + yn - T(2) * dot;
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
  stage(lane);
  __syncthreads();
  take();

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
  // band b+1 is staged at step b*Pc + 63: every lane has taken band b by then
  int next_stage = kWave - 1;
  int stage_band = 1;

  // y rows prefetched two steps ahead (column of step t+2 of this lane)
  int a_s = s;
  auto ycol = [&](int st) {
    int j = Nc - (st < 0 ? 0 : st);
    return j < 0 ? 0 : (j > Ly - 1 ? Ly - 1 : j);
  };
  auto yload = [&](T* dst) {
    yrow(dst, ycol(a_s));
    ++a_s;
    if (a_s == Pc) a_s = 0;
  };
  T y0[DMAX], y1[DMAX];
  yload(y0);
  yload(y1);

  auto step_fn = [&](int step, T* yv) {
    if (step == next_stage) {  // wave-uniform
      if (stage_band < nb) stage(stage_band * kWave + lane);
      ++stage_band;
      next_stage += Pc;
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
    const int yc = j < 0 ? 0 : (j > Ly - 1 ? Ly - 1 : j);
    T dv, g;
    {
      T dot = T(0);
#pragma unroll
      for (int k = 0; k < DMAX; ++k) dot += xr[k] * yv[k];
      if (kind == kRBF) {
        dv = xn + ynl[yc] - T(2) * dot;
        g = fast_exp2(-dv);
      } else {
        dv = g = dot;
      }
    }
    yload(yv);
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
      P[m] = prim ? rel[m] : Pn[m];
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
      take();
    }
  };

  int step = 0;
  for (; step + 2 <= total; step += 2) {
    step_fn(step, y0);
    step_fn(step + 1, y1);
  }
  if (step < total) step_fn(step, y0);
  gsig += (double)gsig_row;
  gsig = wave_sum(gsig);
  if (lane == 0 && a.gparam) a.gparam[pr] = gsig;
}

template <typename T, int NLEV, int DMAX, bool VEC>
static hipError_t launch_bwd(const DpBwdArgs& a, size_t lds, hipStream_t s) {
  auto k = dp_backward_kernel<T, NLEV, DMAX, VEC>;
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
  const bool vec = a.d == DMAX && ((uintptr_t)a.x.data & 15) == 0 &&
                   ((uintptr_t)a.y.data & 15) == 0;
#define KSIG_CASE(L)                                                              \
  if (a.n_levels == L)                                                            \
    return vec ? launch_bwd<T, L, DMAX, true>(a, lds, s) : launch_bwd<T, L, DMAX, false>(a, lds, s);
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

size_t dp_backward_lds(int dtype, int d, int cols, int n_levels) {
  const size_t esz = dtype == KSIG_F32 ? 4 : 8;
  const int dmax = d <= 8 ? 8 : (d <= 16 ? 16 : 32);
  const int nl1 = n_levels > 1 ? n_levels - 1 : 1;
  const int nf = 3 + 2 * (n_levels - 1);
  const int pc = cols + 1 > kWave ? cols + 1 : kWave;
  return ((size_t)kWave * (dmax + 1 + nl1) + (size_t)cols + (size_t)pc * nf) * esz;
}

hipError_t dp_backward_mfma_d8(const DpBwdArgs& a, size_t lds, hipStream_t s);
hipError_t dp_backward_mfma_d16(const DpBwdArgs& a, size_t lds, hipStream_t s);
hipError_t dp_backward_mfma_d32(const DpBwdArgs& a, size_t lds, hipStream_t s);

size_t dp_backward_mfma_lds(int d, int cols, int n_levels) {
  (void)d;
  (void)cols;
  (void)n_levels;
  return (size_t)kBwdRingW * kRingS * 4;  // the ring only: 20 KB, 8 waves per CU
}

int dp_backward_mfma_fifo_cols(int cols) { return bwd_mfma_pc(cols - 1); }  // Nc = cols - 1

size_t dp_backward_mfma_fifo_bytes(int cols, int n_levels) {
  const int nf = 3 + 2 * (n_levels - 1);
  return (size_t)dp_backward_mfma_fifo_cols(cols) * ((nf + 3) & ~3) * 4;
}

hipError_t dp_backward_mfma_launch(const DpBwdArgs& a, size_t lds, hipStream_t s) {
  if (a.d <= 8) return dp_backward_mfma_d8(a, lds, s);
  if (a.d <= 16) return dp_backward_mfma_d16(a, lds, s);
  return dp_backward_mfma_d32(a, lds, s);
}

hipError_t dp_backward_launch(const DpBwdArgs& a, int dtype, size_t lds, hipStream_t s) {
  // float32 runs the MFMA variant (dp_backward_mfma.h)
  if (dtype == KSIG_F32) return hipErrorInvalidValue;
  return dispatch_bwd<double>(a, lds, s);
}

}  // namespace ksig
