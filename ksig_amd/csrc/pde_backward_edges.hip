// Adjoint of the signature-PDE Goursat sweep with the solution rebuilt in the
// reverse sweep (gfx950).  No reference counterpart: ksig/algorithms.py:344-446
// is forward-only; the scheme differentiated is ksig/algorithms.py:380-385.
//
// Refined cell (r, c), M' = M / 4^λ, A = 1 + a, a = M'/2 + M'^2/12,
// B = 1 - b, b = M'^2/12:
//   U(r,c) = (U(r-1,c) + U(r,c-1)) A(r,c) - U(r-1,c-1) B(r,c),   K = U(last cell).
// The adjoint λ = dK/dU is swept bottom-up, right-to-left, in increment form
// (μ(r,c) = λ(r,c) - λ(r,c+1)):
//   μ(r,c) = μ(r+1,c) + a(r+1,c) λ(r+1,c) + a(r,c+1) λ(r,c+1) + b(r+1,c+1) λ(r+1,c+1),
//   λ(r,c) = λ(r,c+1) + μ(r,c),
// and dL/dM'(r,c) = λ(r,c) [(U(r-1,c) + U(r,c-1)) (1/2 + M'/6) + U(r-1,c-1) M'/6].
//
// The update above is exactly invertible in its diagonal corner,
//   U(r-1,c-1) = ((U(r-1,c) + U(r,c-1)) A(r,c) - U(r,c)) / B(r,c),
// so the same reverse sweep rebuilds every U from the forward's last refined
// row and last refined column (O(Lx + Ly) values per pair, written by
// ksig_signature_pde_fwd_edges), with no per-cell storage.  The
// rebuild runs in float64 (full rate on the vector ALUs of gfx950): it is a
// 2-D difference recursion whose rounding errors add up along the sweep.
//
// Systolic layout as in the forward: lane k of band b owns flipped refined row
// r = Nrr-1-(64b+k) and walks the refined columns right-to-left one step behind
// lane k-1; the lane of row r passes λ, μ, aλ, bλ, M'(r,c) and U(r,c-1) up to
// the lane of row r-1 (DPP wave_shr:1; lane 63 -> lane 0 of the next band
// through an LDS FIFO).  The lane of row r-1 completes cell (r, c): it knows
// λ(r,c), U(r,c-1) from below and U(r-1,c), U(r-1,c-1) from its own rebuild.
// A virtual row r = -1 (U = 1) completes row 0.  The 2^λ x 2^λ refined cells
// of an original cell are summed along each lane's 2^λ columns and across the
// 2^λ lanes by a flowing partial; the last lane writes dL/dM of the original
// cell once (no atomics).  Static-kernel values are recomputed (two per lane
// per original column), not read back.

#include "ksig_common.h"
#include "ksig_hip.h"
#include "dp_forward.h"

namespace ksig {

template <typename T>
struct PdeSlot {
  T v[6];    // λ, μ, aλ, bλ, M', block partial
  double u;  // U(r, c-1)
};

template <typename T, int DMAX>
__global__ void __launch_bounds__(kWave)
pde_bwd_edges_kernel(const PdeBwdEdgeArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];

  const int lane = threadIdx.x;
  const int64_t pr = blockIdx.x;
  const int kind = a.kind;
  const int d = a.d;
  const int32_t ia = a.pairs[2 * pr];
  const int32_t ib = a.pairs[2 * pr + 1];
  const T* xs = reinterpret_cast<const T*>(a.x.data) + a.x.off[ia] * d;
  int Lx = a.x.len[ia];
  const T* ys = reinterpret_cast<const T*>(a.y.data) + a.y.off[ib] * d;
  int Ly = a.y.len[ib];
  if (Ly > Lx) {  // orientation of the forward (rows = longer sequence)
    const T* tp = xs; xs = ys; ys = tp;
    int tl = Lx; Lx = Ly; Ly = tl;
  }
  const int diff = a.difference;
  const int Nr = diff ? Lx - 1 : Lx;
  const int Nc = diff ? Ly - 1 : Ly;
  if (Nr <= 0 || Nc <= 0) return;
  const int lam = a.dyadic;
  const int F = 1 << lam;
  const int Nrr = Nr << lam, Ncr = Nc << lam;
  const int Pc = Ncr > kWave ? Ncr : kWave;
  const int nrows = Nrr + 1;  // + virtual row above the grid
  const int nb = (nrows + kWave - 1) / kWave;
  const T inv4 = T(1) / T(F * F);
  const T w = T(a.gout[pr]);
  const T* edge_row = reinterpret_cast<const T*>(a.edges) + a.e_off[pr];  // [Ncr]
  const T* edge_col = edge_row + Ncr;                                       // [Nrr]
  T* gm = reinterpret_cast<T*>(a.gm) + a.m_off[pr];                         // [Nr][Nc]

  // LDS: y rows [Ly][DMAX] (scaled), y norms [Ly], FIFO [Pc] slots
  T* ylds = reinterpret_cast<T*>(smem_raw);
  T* ynrm = ylds + Ly * DMAX;
  PdeSlot<T>* fifo = reinterpret_cast<PdeSlot<T>*>(
      smem_raw + (((size_t)Ly * (DMAX + 1) * sizeof(T) + 15) & ~(size_t)15));
  const T xscale = T(a.xscale), yscale = T(a.yscale);
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
  for (int s = lane; s < Pc; s += kWave) {
    PdeSlot<T> z;
#pragma unroll
    for (int k = 0; k < 6; ++k) z.v[k] = T(0);
    z.u = 1.0;
    fifo[s] = z;
  }
  __syncthreads();

  auto gval = [&](const T* xv, T xn, int yc) -> T {
    const T* yr = ylds + yc * DMAX;
    T dot = T(0);
#pragma unroll
    for (int k = 0; k < DMAX; ++k) dot += xv[k] * yr[k];
    return kind == kRBF ? fast_exp2(-(xn + ynrm[yc] - T(2) * dot)) : dot;
  };
  auto load_x = [&](T* xv, T& xn, int p) {
    T n2 = T(0);
#pragma unroll
    for (int k = 0; k < DMAX; ++k) {
      const T v = k < d ? xs[(int64_t)p * d + k] * xscale : T(0);
      xv[k] = v;
      n2 += v * v;
    }
    xn = n2;
  };

  int band = 0;
  int r = Nrr - 1 - lane;  // refined row of this lane (flipped order)
  T xt[DMAX], xb[DMAX], xtn = T(0), xbn = T(0);
  auto enter_row = [&]() {
    if (r >= 0) {
      const int i = r >> lam;
      load_x(xt, xtn, i);
      load_x(xb, xbn, diff ? i + 1 : i);
    }
  };
  enter_row();

  T fl[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) fl[k] = T(0);
  double flu = 1.0;
  T lam_r = T(0), a_r = T(0), beta_prev = T(0), gacc = T(0), M = T(0);
  T gtL = T(0), gbL = T(0);
  double u_self = 1.0, u_bp = 1.0;
  int s = -lane;
  const int total = nb * Pc + kWave - 1;
  for (int step = 0; step < total; ++step) {
    const bool active = s >= 0 && s < Ncr && band < nb && r >= -1;
    PdeSlot<T> fin;
#pragma unroll
    for (int k = 0; k < 6; ++k) fin.v[k] = T(0);
    fin.u = 1.0;
    if (lane == 0 && active) fin = fifo[s];
    T in[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) in[k] = shift_down(fl[k], fin.v[k]);
    const double in_u = shift_down(flu, fin.u);

    const int c = Ncr - 1 - s;
    const bool cell = active && r >= 0;
    if (active && s == 0) {  // right edge: saved last refined column
      u_self = r >= 0 ? (double)edge_col[r] : 1.0;
      u_bp = (r + 1 <= Nrr - 1) ? (double)edge_col[r + 1] : 1.0;
    }
    // this lane's cell value M'(r, c), recomputed once per original column
    if (cell && (c & (F - 1)) == F - 1) {
      const int j = c >> lam;
      if (diff) {
        T gtR, gbR;
        if (s == 0) {
          gtR = gval(xt, xtn, j + 1);
          gbR = gval(xb, xbn, j + 1);
        } else {
          gtR = gtL;
          gbR = gbL;
        }
        gtL = gval(xt, xtn, j);
        gbL = gval(xb, xbn, j);
        M = ((gbR - gbL) - (gtR - gtL)) * inv4;
      } else {
        M = gval(xt, xtn, j) * inv4;
      }
    }

    // adjoint λ of this lane's cell (r, c)
    const T bb = M * M * T(1.0 / 12.0);
    const T aa = T(0.5) * M + bb;
    const T eps = in[2] + a_r * lam_r + beta_prev;
    const T seed = (cell && r == Nrr - 1 && c == Ncr - 1) ? w : T(0);
    T mu = in[1] + eps + seed;
    T lv = lam_r + mu;
    mu = cell ? mu : in[1];
    lv = cell ? lv : T(0);

    // rebuild U(r, c-1)
    double u_new = 1.0;
    if (active && c >= 1) {
      if (r == Nrr - 1) {
        u_new = (double)edge_row[c - 1];
      } else if (r >= 0) {
        const double mb = (double)in[4];
        const double bq = mb * mb * (1.0 / 12.0);
        const double A = 1.0 + 0.5 * mb + bq;
        const double B = 1.0 - bq;
        u_new = ((u_self + in_u) * A - u_bp) / B;
      }
    }

    // cell (r+1, c), completed here: λ and M' from below, U from both rows
    const int rb = r + 1;
    const bool below = active && rb <= Nrr - 1;
    T g = T(0);
    if (below) {
      const double mb = (double)in[4];
      g = in[0] * (T)((u_self + in_u) * (0.5 + mb * (1.0 / 6.0)) + u_new * mb * (1.0 / 6.0));
    }
    gacc = ((c & (F - 1)) == F - 1 ? T(0) : gacc) + g;
    const T part = (((rb & (F - 1)) == F - 1) ? T(0) : in[5]) + gacc;
    if (below && (c & (F - 1)) == 0 && (rb & (F - 1)) == 0)
      gm[(int64_t)(rb >> lam) * Nc + (c >> lam)] = part * inv4;

    T of[6] = {lv, mu, cell ? aa * lv : T(0), cell ? bb * lv : T(0), cell ? M : T(0), part};
#pragma unroll
    for (int k = 0; k < 6; ++k) fl[k] = of[k];
    flu = u_new;
    if (lane == kWave - 1 && active) {
      PdeSlot<T> o;
#pragma unroll
      for (int k = 0; k < 6; ++k) o.v[k] = of[k];
      o.u = u_new;
      fifo[s] = o;
    }
    lam_r = lv;
    a_r = cell ? aa : T(0);
    beta_prev = in[3];
    u_bp = in_u;
    u_self = u_new;

    ++s;
    if (s == Pc) {
      s = 0;
      ++band;
      r -= kWave;
      enter_row();
      lam_r = a_r = beta_prev = gacc = M = T(0);
      u_self = u_bp = 1.0;
    }
  }
}

template <typename T, int DMAX>
static hipError_t launch_bwd_edges(const PdeBwdEdgeArgs& a, size_t lds, hipStream_t s) {
  auto k = pde_bwd_edges_kernel<T, DMAX>;
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)k,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k, dim3((unsigned)a.n_pairs), dim3(kWave), lds, s, a);
  return hipGetLastError();
}

template <typename T>
static hipError_t dispatch_bwd_edges(const PdeBwdEdgeArgs& a, size_t lds, hipStream_t s) {
  if (a.d <= 8) return launch_bwd_edges<T, 8>(a, lds, s);
  if (a.d <= 16) return launch_bwd_edges<T, 16>(a, lds, s);
  return launch_bwd_edges<T, 32>(a, lds, s);
}

size_t pde_bwd_edges_lds(int dtype, int d, int ly, int pc) {
  const size_t esz = dtype == KSIG_F32 ? 4 : 8;
  const int dmax = d <= 8 ? 8 : (d <= 16 ? 16 : 32);
  const size_t head = (((size_t)ly * (dmax + 1) * esz) + 15) & ~(size_t)15;
  const size_t slot = dtype == KSIG_F32 ? sizeof(PdeSlot<float>) : sizeof(PdeSlot<double>);
  return head + (size_t)pc * slot;
}

hipError_t pde_bwd_edges_launch(const PdeBwdEdgeArgs& a, int dtype, size_t lds, hipStream_t s) {
  if (dtype == KSIG_F32) return dispatch_bwd_edges<float>(a, lds, s);
  return dispatch_bwd_edges<double>(a, lds, s);
}

}  // namespace ksig
