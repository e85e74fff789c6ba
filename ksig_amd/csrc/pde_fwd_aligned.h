// Phase-aligned signature-PDE forward for dyadic order 0 and 1 (gfx950):
// kernel template, instantiated by pde_fwd_aligned_f32.hip / _f64.hip.
// The scheme is ksig/algorithms.py:380-385 (host launcher :402-446).
#pragma once

#include <type_traits>

#include "ksig_common.h"
#include "ksig_hip.h"
#include "dp_forward.h"
#include "pde_common.h"

namespace ksig {

// ---------------------------------------------------------------------------
// Phase-aligned sweep for dyadic order 0 and 1 (F = 2^LAM = 1 or 2 refined
// cells per original cell and direction), R = 4 refined rows per lane.
//
// Lane i processes refined column c at step c + F*i (+ F: the priming
// columns), i.e. F steps behind lane i-1 rather than one: every lane is then
// at the same position within its original column, so the work that belongs
// to an original cell (one static-kernel value per original row and column,
// its ΔΔ, the coefficients a = m/2 + m^2/12 and b = m^2/12 of the scheme)
// runs once per original column, on wave-uniform steps, and a refined cell
// costs only the increment update
//   δ = (U(r-1,c) + U(r,c-1)) a + U(r-1,c-1) b,  V = V(r-1,c) + δ,  U = U(r,c-1) + V.
// The lane-to-lane hand-over reads what lane i-1 produced F steps earlier:
// at F = 2 the loop is unrolled by phase and each phase keeps its own output
// registers, so the delay costs no moves.  Lane 63's outputs go to an LDS
// FIFO slot per refined column, which lane 0 of the next band reads a step
// ahead (a band pass is at least 64F + 2 steps, so the slot is written
// first).  The lane's R rows hold R / F original rows; the static kernel at
// the bottom edge of its last original row flows down to the next lane as
// the top edge of that lane's first.
//
// Everything that is not per-cell arithmetic is scheduled on wave-uniform
// step counters: a lane leaves its band pass at a step known to the whole
// wave, so the band change (the left U values back to 1, the next band's x
// rows into the lane's LDS slot), the reset of its y-row fetch pointer, lane
// 63's FIFO store, lane 0's FIFO read and the writes of the result and the
// saved edges run under scalar branches for the one lane concerned, and no
// lane evaluates per-step predicates.  The y rows of a lane's next-but-one
// original column are fetched from global memory (L2-resident) into one of
// two register sets; their norm terms come from an LDS table.
// ---------------------------------------------------------------------------
template <typename T, int DMAX, int KIND, int LAM, bool EDGES, bool VEC>
__global__ void __launch_bounds__(kWave)
pde_fwd_aligned_kernel(const PdeFwdArgs a) {
  // U and V in float64 when the edges are saved for the adjoint, whose
  // backward rebuild amplifies any rounding of the values it starts from
  using TU = std::conditional_t<EDGES, double, T>;
  constexpr int F = 1 << LAM;
  constexpr int R = 4;
  constexpr int NO = R / F;       // original rows per lane
  constexpr int BR = kWave * R;   // refined rows per band
  constexpr bool kGramIn = KIND == kGram;
  constexpr bool kRbf = KIND == kRBF;
  constexpr int NY = kGramIn ? NO : DMAX;  // fetched per column: y row or the caller's values
  extern __shared__ __align__(16) unsigned char smem_raw[];

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
  const int Nrr = Nr * F, Ncr = Nc * F;
  const int Pc = pde_aligned_pc(Nc, LAM);
  const int ncols = Pc / F;  // original columns of a band pass, priming included
  const int nb = (Nrr + BR - 1) / BR;
  const int rshift = diff;
  const T sx = T(a.xscale), sy = T(a.yscale);
  const PdeAlignedLds lay = pde_aligned_lds((int)sizeof(T), (int)sizeof(TU), DMAX, KIND, Pc, LAM);
  TU* fifo_uv = reinterpret_cast<TU*>(smem_raw + lay.uv);  // [Pc][2]: U, V
  T* fifo_g = reinterpret_cast<T*>(smem_raw + lay.g);      // [Pc]: G
  T* xslot = reinterpret_cast<T*>(smem_raw + lay.x) + lane * NO * DMAX;  // [NO][DMAX]
  T* xnslot = reinterpret_cast<T*>(smem_raw + lay.xn) + lane * NO;      // [NO]
  T* ynorm = reinterpret_cast<T*>(smem_raw + lay.yn) + kWave;  // [j + 1] for original column j

  // ---- static-kernel inputs ----------------------------------------------
  // y rows through a descriptor over the sequence (kGram: the pair's matrix)
  const int ystride = d * (int)sizeof(T);
  const __amdgpu_buffer_rsrc_t yrs =
      kGramIn ? raw_rsrc(gm, unsigned((int64_t)Lx * Ly * sizeof(T)))
              : raw_rsrc(ys, unsigned((int64_t)Ly * ystride));
  // this lane's next band's rows (prepared), stored into its LDS slot at its
  // own band change
  T xnx[NO][DMAX], xnn[NO];
  auto load_rows = [&](int b) __attribute__((always_inline)) {
    if constexpr (!kGramIn) {
#pragma unroll
      for (int k = 0; k < NO; ++k) {
        const int row = (b * kWave + lane) * NO + k + rshift;
        const bool ok = row < Lx;
        T raw[DMAX];
#pragma unroll
        for (int e = 0; e < DMAX; ++e) raw[e] = (ok && e < d) ? xs[(int64_t)row * d + e] : T(0);
        pde_xrow_prep<T, DMAX, kRbf>(raw, xnx[k], xnn[k], sx, sy, LAM);
      }
    }
  };
  auto store_rows = [&]() __attribute__((always_inline)) {
    if constexpr (!kGramIn) {
#pragma unroll
      for (int k = 0; k < NO; ++k)
#pragma unroll
        for (int e = 0; e < DMAX; ++e) xslot[k * DMAX + e] = xnx[k][e];
      if constexpr (kRbf) {
#pragma unroll
        for (int k = 0; k < NO; ++k) xnslot[k] = xnn[k];
      }
    }
  };
  load_rows(0);
  store_rows();
  load_rows(1);

  // y norm terms of the band pass's columns j = -1 .. ncols (zero outside the
  // sequence) and 64 zeros before them
  if constexpr (kRbf) {
    for (int n = lane - kWave; n < ncols + 2; n += kWave) {
      const int yc = n - 1 + rshift;
      T n2 = T(0);
      if (n >= 0 && yc >= 0 && yc < Ly) {
#pragma unroll
        for (int e = 0; e < DMAX; ++e) {
          const T v = e < d ? ys[(int64_t)yc * d + e] * sy : T(0);
          n2 = pde_fma(v, v, n2);
        }
      }
      ynorm[n] = n2;
    }
  }

  // FIFO initial contents: the row above the grid.  Slot s holds refined
  // column s - F, of original column ((s - F) >> LAM), whose right edge is y
  // column ((s - F) >> LAM) + 1 (with differencing): G(x_0, y) and U = 1.
  {
    T x0[DMAX], nx0 = T(0);
    if constexpr (!kGramIn) {
      T raw[DMAX];
#pragma unroll
      for (int e = 0; e < DMAX; ++e) raw[e] = e < d ? xs[e] : T(0);
      pde_xrow_prep<T, DMAX, kRbf>(raw, x0, nx0, sx, sy, LAM);
    }
    for (int s = lane; s < Pc; s += kWave) {
      const int yc = ((s - F) >> LAM) + rshift;
      T g = T(0);
      if (diff && yc >= 0 && yc < Ly) {
        if constexpr (kGramIn) {
          g = gm[yc];
        } else {
          T yv[DMAX];
#pragma unroll
          for (int e = 0; e < DMAX; ++e) yv[e] = e < d ? ys[(int64_t)yc * d + e] : T(0);
          T yn = T(0);
          if constexpr (kRbf) {
#pragma unroll
            for (int e = 0; e < DMAX; ++e) {
              const T v = yv[e] * sy;
              yn = pde_fma(v, v, yn);
            }
          }
          g = pde_gval<T, DMAX, kRbf>(x0, nx0, yv, yn);
        }
      }
      fifo_g[s] = g;
      fifo_uv[2 * s + 0] = TU(1);
      fifo_uv[2 * s + 1] = TU(0);
    }
  }
  __syncthreads();

  // ---- per-lane state ------------------------------------------------------
  TU uleft[R], udiag[R];
#pragma unroll
  for (int q = 0; q < R; ++q) uleft[q] = udiag[q] = TU(1);
  T gbl[NO], gtl[NO];
  TU ca[NO], cb[NO];
#pragma unroll
  for (int k = 0; k < NO; ++k) {
    gbl[k] = gtl[k] = T(0);
    ca[k] = cb[k] = TU(0);
  }
  T pmask = T(0);  // 0 on a lane's priming column (and before its start)
  T flg = T(0);    // G at the bottom edge of the lowest original row, last column
  TU fu[F], fv[F]; // U, V of the last step of each phase
#pragma unroll
  for (int p = 0; p < F; ++p) {
    fu[p] = TU(1);
    fv[p] = TU(0);
  }

  // The y row of a lane's original column j is fetched two columns ahead
  // into register set (column parity).  Lane i starts at step F*i: its
  // phase-0 steps before that see columns -1-i, -i, ... (zero rows and norms).
  // fetch pointer: the column after the two prefetched ones
  const int jpri = -1;  // the priming column
  int ycol = jpri - lane + 2;
  auto yoff = [&](int j) __attribute__((always_inline)) -> int {
    if constexpr (kGramIn) return j + rshift;  // element column of the caller's matrix
    return __mul24(j + rshift, ystride);  // full-rate 24-bit multiply (|j| < 2^23)
  };
  // kGram: this lane's first original row of its band, as a matrix row offset
  int grow = ((lane * NO) + rshift) * Ly;
  T yA[NY], yB[NY], ynA = T(0), ynB = T(0);
  auto fetch = [&](T* yv, T& yn, int j) __attribute__((always_inline)) {
    if constexpr (kGramIn) {
      const bool ok = (unsigned)(j + rshift) < (unsigned)Ly;
#pragma unroll
      for (int k = 0; k < NO; ++k) {
        if constexpr (sizeof(T) == 4) {
          const float t = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
              yrs, (grow + j + rshift) * 4, k * Ly * 4, 0));
          yv[k] = ok ? T(t) : T(0);
        } else {
          typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
          const u32x2 t = __builtin_amdgcn_raw_buffer_load_b64(yrs, (grow + j + rshift) * 8,
                                                               k * Ly * 8, 0);
          yv[k] = ok ? T(__hiloint2double((int)t.y, (int)t.x)) : T(0);
        }
      }
    } else {
      pde_load_y<T, DMAX, VEC>(yv, yrs, yoff(j), d);
      if constexpr (kRbf) yn = ynorm[j + 1];
    }
  };
  fetch(yA, ynA, jpri - lane);
  fetch(yB, ynB, jpri - lane + 1);

  double* edges = EDGES ? reinterpret_cast<double*>(a.edges) + a.e_off[pr] : nullptr;
  // the lane holding the last refined row (band nb-1), and its row within
  const int lrow = Nrr - 1 - (nb - 1) * BR;
  const int Ls = lrow / R, qs = lrow % R;
  const int tlast = F * Ls + (nb - 1) * Pc + F;  // its step at refined column 0

  // lane 0's FIFO inputs of the next step, read a step ahead
  TU nfu = TU(1), nfv = TU(0);
  T nfg = fifo_g[0];
  nfu = fifo_uv[0];
  nfv = fifo_uv[1];

  // every lane's last band pass, rounded up to whole iterations of two
  // original columns (the extra steps store nothing)
  const int total = (nb * Pc + F * (kWave - 1) + 2 * F - 1) / (2 * F) * (2 * F);
  int slot0 = 1 % Pc;               // (t + 1) mod Pc
  int slot63 = Pc - F * (kWave - 1);  // (t - 63F) mod Pc
  int load_band = 2;
  // last refined column events (EDGES): (t - F - Ncr + 1) mod Pc and its wraps
  int qe = Pc - F - Ncr + 1;
  int be = -1;
  if (qe >= Pc) { qe -= Pc; be = 0; }

  auto step_fn = [&](auto phase_tag, auto set_tag, int t, int qq) __attribute__((always_inline)) {
    constexpr int P = decltype(phase_tag)::value;
    constexpr int S = decltype(set_tag)::value;
    T* yv = S == 0 ? yA : yB;
    T& yn = S == 0 ? ynA : ynB;
    TU iu = shift_down(fu[P], nfu);
    TU iv = shift_down(fv[P], nfv);
    // lanes with a band change this column: the one leaving its band pass
    // (its last refined column is this column's), and the one whose fetch
    // pointer goes back to the priming column
    int sw = qq + F;
    sw = sw >= Pc ? sw - Pc : sw;
    int sf = qq + 2 * F;
    sf = sf >= Pc ? sf - Pc : sf;
    sf = sf >= Pc ? sf - Pc : sf;
    if constexpr (P == 0) {
      T gt = shift_down(flg, nfg);
#pragma unroll
      for (int k = 0; k < NO; ++k) {
        T gb;
        if constexpr (kGramIn) {
          gb = yv[k];
        } else {
          T xv[DMAX];
#pragma unroll
          for (int e = 0; e < DMAX; ++e) xv[e] = xslot[k * DMAX + e];
          gb = pde_gval<T, DMAX, kRbf>(xv, kRbf ? xnslot[k] : T(0), yv, yn);
        }
        T m = diff ? (gb - gbl[k]) - (gt - gtl[k]) : gb;
        if constexpr (kGramIn) m *= T(1) / T(F * F);
        m *= pmask;
        gbl[k] = gb;
        gtl[k] = gt;
        T aq, bq;
        pde_coef(m, aq, bq);
        ca[k] = TU(aq);
        cb[k] = TU(bq);
        gt = gb;
      }
      flg = gt;
      // the mask of the lane's next phase-0 step
      if (t < kWave * F) pmask = t >= F * lane ? T(1) : T(0);
      else pmask = T(1);
      // fetch the y row of the column after the next into this set
      // (before a lane starts these resets fall on columns that are already
      // the priming column: no guard needed)
      if (sf < kWave * F) {  // wave-uniform
        if (lane == (sf >> LAM)) ycol = jpri;
      }
      fetch(yv, yn, ycol);
      ++ycol;
    }
    // lane 0's inputs of the next step (its slot is wave-uniform)
    {
      const int s0 = slot0;
      nfu = fifo_uv[2 * s0 + 0];
      nfv = fifo_uv[2 * s0 + 1];
      if constexpr (P == F - 1) nfg = fifo_g[s0];
    }
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int k = q / F;
      const TU delta = (iu + uleft[q]) * ca[k] + udiag[q] * cb[k];
      const TU v = iv + delta;
      const TU u = uleft[q] + v;
      udiag[q] = iu;
      uleft[q] = u;
      iu = u;
      iv = v;
    }
    fu[P] = iu;
    fv[P] = iv;
    if (t >= F * (kWave - 1)) {  // lane 63 has started: its outputs into its slot
      if (lane == kWave - 1) {
        fifo_uv[2 * slot63 + 0] = iu;
        fifo_uv[2 * slot63 + 1] = iv;
        if constexpr (P == 0) fifo_g[slot63] = flg;
      }
    }
    // the result and the last refined row (wave-uniform tests)
    {
      const int ce = t - tlast;
      if ((unsigned)ce < (unsigned)Ncr && (EDGES || ce == Ncr - 1)) {
        if (lane == Ls) {
          TU ul = uleft[0];
#pragma unroll
          for (int q = 1; q < R; ++q)
            if (qs == q) ul = uleft[q];
          if constexpr (EDGES) edges[ce] = (double)ul;
          if (ce == Ncr - 1) a.out[pr] = (double)ul;
        }
      }
    }
    if constexpr (P == F - 1) {
      // the last refined column: every lane reaching it saves its rows
      if constexpr (EDGES) {
        if (qe < kWave * F && be >= 0 && be < nb) {
          const int le = qe >> LAM;
          const int r0 = be * BR + le * R;
          if (lane == le) {
#pragma unroll
            for (int q = 0; q < R; ++q)
              if (r0 + q < Nrr) edges[Ncr + r0 + q] = (double)uleft[q];
          }
        }
      }
      // band change of the lane whose last refined column this was
      if (t + 1 >= Pc && sw < kWave * F) {  // wave-uniform
        const int lw = sw >> LAM;
        if (lane == lw) {
#pragma unroll
          for (int q = 0; q < R; ++q) uleft[q] = TU(1);
          pmask = T(0);
          store_rows();
          if constexpr (kGramIn) grow += kWave * NO * Ly;
        }
        if (lw == kWave - 1) {  // every lane has taken its next band's rows
          load_rows(load_band);
          ++load_band;
        }
      }
    }
    slot0 = slot0 + 1 == Pc ? 0 : slot0 + 1;
    slot63 = slot63 + 1 == Pc ? 0 : slot63 + 1;
    if (++qe == Pc) {
      qe = 0;
      ++be;
    }
  };

  using Z = std::integral_constant<int, 0>;
  using O = std::integral_constant<int, 1>;
  int qq = 0;  // t mod Pc at an original column's first step
  auto two_columns = [&](int t) __attribute__((always_inline)) {
    step_fn(Z{}, Z{}, t, qq);
    if constexpr (F == 2) step_fn(O{}, Z{}, t + 1, qq);
    qq += F;
    qq = qq >= Pc ? qq - Pc : qq;
    step_fn(Z{}, O{}, t + F, qq);
    if constexpr (F == 2) step_fn(O{}, O{}, t + F + 1, qq);
    qq += F;
    qq = qq >= Pc ? qq - Pc : qq;
  };
  // The first two columns run outside the loop: the loop is then entered
  // with the same loads in flight as at its back edge, and the compiler's
  // wait for a column's y row leaves the next column's fetch in flight
  // (entered from the prologue, it waited for every load at the loop head).
  two_columns(0);
  for (int t = 2 * F; t < total; t += 2 * F) two_columns(t);
}


hipError_t pde_fwd_aligned_launch_f32(const PdeFwdArgs& a, int dmax, size_t lds, hipStream_t s);
hipError_t pde_fwd_aligned_launch_f64(const PdeFwdArgs& a, int dmax, size_t lds, hipStream_t s);

// One phase-aligned instantiation set for element type T.
template <typename T, int DMAX, int KIND>
hipError_t launch_pde_aligned(const PdeFwdArgs& a, size_t lds, hipStream_t s) {
  const bool e = a.edges != nullptr;
  // 16-byte y-row loads: d fills the bucket and both sides are aligned (the
  // kernel's y side is the shorter sequence, from either set)
  const bool vec = KIND != kGram && a.d == DMAX && ((uintptr_t)a.x.data & 15) == 0 &&
                   ((uintptr_t)a.y.data & 15) == 0;
  auto pick = [&](auto lam_tag) {
    constexpr int L = decltype(lam_tag)::value;
    if constexpr (KIND == kGram) {  // caller values: no y rows, no saved edges
      return pde_fwd_aligned_kernel<T, DMAX, KIND, L, false, false>;
    } else {
      return e ? (vec ? pde_fwd_aligned_kernel<T, DMAX, KIND, L, true, true>
                      : pde_fwd_aligned_kernel<T, DMAX, KIND, L, true, false>)
               : (vec ? pde_fwd_aligned_kernel<T, DMAX, KIND, L, false, true>
                      : pde_fwd_aligned_kernel<T, DMAX, KIND, L, false, false>);
    }
  };
  auto k = a.dyadic == 0 ? pick(std::integral_constant<int, 0>{})
                         : pick(std::integral_constant<int, 1>{});
  if (lds > 64 * 1024) {
    hipError_t r = hipFuncSetAttribute((const void*)k,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (r != hipSuccess) return r;
  }
  hipLaunchKernelGGL(k, dim3((unsigned)a.n_pairs), dim3(kWave), lds, s, a);
  return hipGetLastError();
}

template <typename T>
hipError_t dispatch_pde_aligned(const PdeFwdArgs& a, int dmax, size_t lds, hipStream_t s) {
  if (a.kind == kGram) return launch_pde_aligned<T, 8, kGram>(a, lds, s);
  if (a.kind == kRBF) {
    if (dmax == 4) return launch_pde_aligned<T, 4, kRBF>(a, lds, s);
    if (dmax == 8) return launch_pde_aligned<T, 8, kRBF>(a, lds, s);
    if (dmax == 16) return launch_pde_aligned<T, 16, kRBF>(a, lds, s);
    return launch_pde_aligned<T, 32, kRBF>(a, lds, s);
  }
  if (dmax == 4) return launch_pde_aligned<T, 4, kLinear>(a, lds, s);
  if (dmax == 8) return launch_pde_aligned<T, 8, kLinear>(a, lds, s);
  if (dmax == 16) return launch_pde_aligned<T, 16, kLinear>(a, lds, s);
  return launch_pde_aligned<T, 32, kLinear>(a, lds, s);
}

}  // namespace ksig
