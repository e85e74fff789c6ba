// Signature-PDE kernel, forward pass (gfx950).
//
// Solves the explicit Goursat scheme of ksig/algorithms.py:380-385
//   U(i,j) = (U(i-1,j) + U(i,j-1)) (1 + M/2 + M^2/12) - U(i-1,j-1) (1 - M^2/12)
// with U = 1 on the row/column before the grid, returning U at the last cell
// (ksig/algorithms.py:344-399, host launcher :402-446).  The reference sweeps
// antidiagonals with two block barriers per step; here the same recurrence is
// evaluated as a skewed systolic wave: lane i owns refined row r' = 64*band+i,
// walks the refined columns one step behind lane i-1 and receives U(r'-1, c')
// from it through DPP wave_shr:1 (lane 0 from an LDS FIFO per column), so a
// step costs no barrier at all.
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

#include "ksig_common.h"
#include "ksig_hip.h"
#include "dp_forward.h"

namespace ksig {

template <typename T, int DMAX>
__global__ void __launch_bounds__(kWave)
pde_forward_kernel(const PdeFwdArgs a) {
  constexpr int NF = 4;  // flow: [G(r, c+1), G(r+1, c+1), U, V]
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* smem = reinterpret_cast<T*>(smem_raw);

  const int lane = threadIdx.x;
  const int64_t pr = blockIdx.x;
  const int kind = a.kind;
  const int d = a.d;

  const T* xs;
  const T* ys;
  int Lx, Ly;
  const T* gm = nullptr;
  if (kind == kGram) {
    Lx = a.gram_lx;
    Ly = a.gram_ly;
    gm = reinterpret_cast<const T*>(a.gram) + pr * (int64_t)Lx * Ly;
    xs = ys = nullptr;
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
  int Pc = (Nc + 1) << lam;  // priming column (F steps) + refined columns
  if (Pc < kWave) Pc = kWave;
  const int nb = (Nrr + kWave - 1) / kWave;
  const T inv4 = T(1) / T(F * F);
  const int rshift = diff ? 1 : 0;

  const T xscale = T(a.xscale), yscale = T(a.yscale);
  T* ylds = smem;
  T* ynrm = ylds + (kind == kGram ? 0 : Ly * DMAX);
  T* fifo = ynrm + (kind == kGram ? 0 : Ly);

  if (kind != kGram) {
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

  // Static-kernel value at (x row xr, y row yc), given x row registers.
  auto gval = [&](const T* xv, T xnrm, int xr, int yc) -> T {
    if (kind == kGram) {
      if (xr < 0 || xr >= Lx || yc < 0 || yc >= Ly) return T(0);
      return gm[(int64_t)xr * Ly + yc];
    }
    yc = yc < 0 ? 0 : (yc >= Ly ? Ly - 1 : yc);
    const T* yr = ylds + yc * DMAX;
    T dot = T(0);
#pragma unroll
    for (int k = 0; k < DMAX; ++k) dot += xv[k] * yr[k];
    return kind == kRBF ? fast_exp2(-(xnrm + ynrm[yc] - T(2) * dot)) : dot;
  };
  auto load_x = [&](T* xv, T& xnrm, int xr) {
    const bool ok = xr >= 0 && xr < Lx && kind != kGram;
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
      if (diff && yc < Ly) g = gval(x0, x0n, 0, yc);
      fifo[s * NF + 0] = T(0);
      fifo[s * NF + 1] = g;
      fifo[s * NF + 2] = T(1);
      fifo[s * NF + 3] = T(0);
    }
  }
  __syncthreads();

  // Per-lane state.
  int band = 0;
  int rr = lane;                 // refined row
  int orow = rr >> lam;          // original grid row
  bool leader = (rr & (F - 1)) == 0;
  T xv[DMAX], xnrm;
  load_x(xv, xnrm, orow + rshift);
  T xv2[DMAX], xnrm2;            // next band's row, prefetched
  load_x(xv2, xnrm2, ((rr + kWave) >> lam) + rshift);

  T* edges = a.edges ? reinterpret_cast<T*>(a.edges) + a.e_off[pr] : nullptr;

  T fl[NF] = {T(0), T(0), T(1), T(0)};
  T gt_old = T(0), gb_old = T(0), gt = T(0), gb = T(0);
  T M = T(0);
  T uleft = T(1), udiag = T(1);
  int cc = -F - lane;            // refined column (priming columns are -F..-1)

  const int total = nb * Pc + kWave - 1;
  for (int step = 0; step < total; ++step) {
    const bool active = cc >= -F && band < nb;
    T fin0 = T(0), fin1 = T(0), fin2 = T(1), fin3 = T(0);
    if (lane == 0 && active) {
      const int s = cc + F;
      fin0 = fifo[s * NF + 0];
      fin1 = fifo[s * NF + 1];
      fin2 = fifo[s * NF + 2];
      fin3 = fifo[s * NF + 3];
    }
    const T in_gt = shift_down(fl[0], fin0);
    const T in_gb = shift_down(fl[1], fin1);
    const T in_u = shift_down(fl[2], fin2);
    const T in_v = shift_down(fl[3], fin3);

    const bool newcol = (cc & (F - 1)) == 0;  // first refined step of a column
    const int ocol = cc >> lam;               // original column (-1 = priming)
    // Gram values of this lane's original cell (right edge, column ocol+1).
    T ngt = leader ? in_gb : in_gt;
    T ngb = in_gb;
    if (leader && newcol) ngb = gval(xv, xnrm, orow + rshift, ocol + rshift);
    else if (leader) ngb = gb;
    if (newcol) {
      gt_old = gt;
      gb_old = gb;
      const T m = diff ? ((ngb - gb_old) - (ngt - gt_old)) : ngb;
      M = m * inv4;
    }
    gt = ngt;
    gb = ngb;

    const bool cell = active && cc >= 0 && cc < Ncr && rr < Nrr;
    const T b = M * M * T(1.0 / 12.0);
    const T aa = T(0.5) * M + b;
    const T delta = (in_u + uleft) * aa + udiag * b;
    T v = in_v + delta;
    T u = uleft + v;
    u = cell ? u : T(1);  // priming / padding columns carry the boundary value
    v = cell ? v : in_v;
    if (cell && rr == Nrr - 1 && cc == Ncr - 1) a.out[pr] = (double)u;
    if (edges != nullptr && cell) {  // last refined row, then last refined column
      if (rr == Nrr - 1) edges[cc] = u;
      if (cc == Ncr - 1) edges[Ncr + rr] = u;
    }
    uleft = u;
    udiag = in_u;

    fl[0] = gt;
    fl[1] = gb;
    fl[2] = u;
    fl[3] = v;
    if (lane == kWave - 1 && active) {
      const int s = cc + F;
      fifo[s * NF + 0] = gt;
      fifo[s * NF + 1] = gb;
      fifo[s * NF + 2] = u;
      fifo[s * NF + 3] = v;
    }

    ++cc;
    if (cc == Pc - F) {  // wrap to the next band
      cc = -F;
      ++band;
      rr += kWave;
      orow = rr >> lam;
      leader = (rr & (F - 1)) == 0;
#pragma unroll
      for (int k = 0; k < DMAX; ++k) xv[k] = xv2[k];
      xnrm = xnrm2;
      load_x(xv2, xnrm2, ((rr + kWave) >> lam) + rshift);
      uleft = T(1);
      gt = gb = T(0);
    }
  }
}

template <typename T, int DMAX>
static hipError_t launch_pde(const PdeFwdArgs& a, size_t lds, hipStream_t s) {
  auto k = pde_forward_kernel<T, DMAX>;
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)k,
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k, dim3((unsigned)a.n_pairs), dim3(kWave), lds, s, a);
  return hipGetLastError();
}

template <typename T>
static hipError_t dispatch_pde(const PdeFwdArgs& a, size_t lds, hipStream_t s) {
  if (a.kind == kGram || a.d <= 8) return launch_pde<T, 8>(a, lds, s);
  if (a.d <= 16) return launch_pde<T, 16>(a, lds, s);
  return launch_pde<T, 32>(a, lds, s);
}

hipError_t pde_forward_launch(const PdeFwdArgs& a, int dtype, size_t lds, hipStream_t s) {
  if (dtype == KSIG_F32) return dispatch_pde<float>(a, lds, s);
  return dispatch_pde<double>(a, lds, s);
}

}  // namespace ksig
