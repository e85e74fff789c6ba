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
// through a FIFO in a global-memory workspace, read two steps ahead).  The
// saved edges and the next band's x rows are staged in LDS, so no lane waits
// on global memory at a band change.  The lane of row r-1 completes cell (r, c): it knows
// λ(r,c), U(r,c-1) from below and U(r-1,c), U(r-1,c-1) from its own rebuild.
// A virtual row r = -1 (U = 1) completes row 0.  The 2^λ x 2^λ refined cells
// of an original cell are summed along each lane's 2^λ columns and across the
// 2^λ lanes by a flowing partial; the last lane writes dL/dM of the original
// cell once (no atomics).  Static-kernel values are recomputed (two per lane
// per original column), not read back.

#include <type_traits>

#include "ksig_common.h"
#include "ksig_hip.h"
#include "dp_forward.h"
#include "pde_common.h"

namespace ksig {

// Band hand-over slot in the global workspace: λ, μ, aλ, bλ, M', block
// partial and U(r, c-1), padded to whole 16-byte quads.
template <typename T>
struct PdeSlot {
  T v[6];
  double u;
};
template <typename T>
constexpr int pde_slot_quads() { return (int)((sizeof(PdeSlot<T>) + 15) / 16); }

__host__ __device__ constexpr int pde_bwd_pc(int ncr) { return ncr > 80 ? ncr : 80; }

template <typename T, int DMAX, int R>
__global__ void __launch_bounds__(kWave)
pde_bwd_edges_kernel(const PdeBwdEdgeArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int XS = 2 * DMAX + 2;  // staged x rows of the next band: top, bottom, norms
  constexpr int SQ = pde_slot_quads<T>();
  constexpr int BR = kWave * R;     // refined rows per band

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
  // with an even number of rows per lane and dyadic order >= 1, each odd row
  // r0 - 2k and the row r0 - 2k - 1 after it lie in one original row and
  // share its static-kernel values (r0 = Nrr - 1 - R * (...) is odd)
  const bool shared = (R % 2 == 0) && lam >= 1;
  const int Nrr = Nr << lam, Ncr = Nc << lam;
  const int Pc = pde_bwd_pc(Ncr);
  const int nrows = Nrr + 1;  // + virtual row above the grid
  const int nb = (nrows + BR - 1) / BR;
  const T inv4 = T(1) / T(F * F);
  const T w = T(a.gout[pr]);
  const double* edge_row = reinterpret_cast<const double*>(a.edges) + a.e_off[pr];  // [Ncr]
  const double* edge_col = edge_row + Ncr;                                            // [Nrr]
  T* gm = reinterpret_cast<T*>(a.gm) + a.m_off[pr];                         // [Nr][Nc]
  float4* fifo = reinterpret_cast<float4*>(a.fifo) + (size_t)pr * a.fifo_cols * SQ;

  // LDS: y rows [Ly][DMAX] (scaled), y norms [Ly], the saved last refined row
  // [Ncr] and column [Nrr], lane-private staging of the next band's x rows
  T* ylds = reinterpret_cast<T*>(smem_raw);
  T* ynrm = ylds + Ly * DMAX;
  T* erow = ynrm + Ly;
  T* ecol = erow + Ncr;
  T* xst = ecol + Nrr;
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
  for (int c = lane; c < Ncr; c += kWave) erow[c] = T(edge_row[c]);
  for (int c = lane; c < Nrr; c += kWave) ecol[c] = T(edge_col[c]);

  auto gval = [&](const T* xv, T xn, int yc) __attribute__((always_inline)) -> T {
    const T* yr = ylds + yc * DMAX;
    T dot = T(0);
#pragma unroll
    for (int k = 0; k < DMAX; ++k) dot += xv[k] * yr[k];
    return kind == kRBF ? fast_exp2(-(xn + ynrm[yc] - T(2) * dot)) : dot;
  };
  // The y row (and norm) of the original column a step needs, read a step
  // ahead: column c of step s is Ncr - 1 - s, its original column c >> lam.
  struct YRow {
    T y[DMAX];
    T yn;
  };
  auto yfetch = [&](YRow& yr, int s_at) __attribute__((always_inline)) {
    int j = (Ncr - 1 - s_at) >> lam;
    j = j < 0 ? 0 : (j >= Ly ? Ly - 1 : j);
#pragma unroll
    for (int k = 0; k < DMAX; ++k) yr.y[k] = ylds[j * DMAX + k];
    yr.yn = ynrm[j];
  };
  auto gval_y = [&](const T* xv, T xn, const YRow& yr) __attribute__((always_inline)) -> T {
    T dot = T(0);
#pragma unroll
    for (int k = 0; k < DMAX; ++k) dot += xv[k] * yr.y[k];
    return kind == kRBF ? fast_exp2(-(xn + yr.yn - T(2) * dot)) : dot;
  };
  // stage the x rows of the lane's refined rows r0, r0-1, ... into its slots
  auto stage = [&](int r0) __attribute__((always_inline)) {
#pragma unroll
    for (int q = 0; q < R; ++q) {
      T* dst = xst + (lane * R + q) * XS;
      const int rr = r0 - q;
      if (rr >= 0) {
        const int i = rr >> lam;
        const int i2 = diff ? i + 1 : i;
        T n0 = T(0), n1 = T(0);
#pragma unroll
        for (int k = 0; k < DMAX; ++k) {
          const T v0 = k < d ? xs[(int64_t)i * d + k] * xscale : T(0);
          const T v1 = k < d ? xs[(int64_t)i2 * d + k] * xscale : T(0);
          dst[k] = v0;
          dst[DMAX + k] = v1;
          n0 += v0 * v0;
          n1 += v1 * v1;
        }
        dst[2 * DMAX] = n0;
        dst[2 * DMAX + 1] = n1;
      }
    }
  };
  int band = 0;
  int r0 = Nrr - 1 - lane * R;  // the lane's first refined row (flipped order)
  T xt[R][DMAX], xb[R][DMAX], xtn[R], xbn[R];
  auto take = [&]() __attribute__((always_inline)) {
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const T* src = xst + (lane * R + q) * XS;
#pragma unroll
      for (int k = 0; k < DMAX; ++k) {
        xt[q][k] = src[k];
        xb[q][k] = src[DMAX + k];
      }
      xtn[q] = src[2 * DMAX];
      xbn[q] = src[2 * DMAX + 1];
    }
  };
  stage(r0);
  __syncthreads();
  take();

  // Lane 0's input of step st, read two steps ahead by every lane (a
  // broadcast; only lane 0 keeps it): in bands 1 .. nb-1 the slot lane 63
  // wrote Pc - 65 >= 15 steps earlier; otherwise (band 0: the row below the
  // grid; steps where lane 0 is idle) a zero adjoint with U = 1, so that no
  // unwritten workspace is read.
  int fslot = 0, fband = 0;  // lane 0's slot and band at the next prefetched step
  auto fread = [&](T* v, double& u) __attribute__((always_inline)) {
    const int b0 = fband, f0 = fslot;
    if (++fslot == Pc) {
      fslot = 0;
      ++fband;
    }
    const bool live = b0 >= 1 && b0 < nb && f0 < Ncr;
    const int fs = live ? f0 : 0;
    float4 q[SQ];
#pragma unroll
    for (int k = 0; k < SQ; ++k) q[k] = fifo[fs * SQ + k];
    PdeSlot<T> sl;
    __builtin_memcpy(&sl, q, sizeof(PdeSlot<T>));
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] = live ? sl.v[k] : T(0);
    u = live ? sl.u : 1.0;
  };
  T fv0[6], fv1[6];
  double fu0, fu1;
  fread(fv0, fu0);
  fread(fv1, fu1);

  T fl[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) fl[k] = T(0);
  double flu = 1.0;
  // per-row state
  T lam_r[R], a_r[R], beta_prev[R], gacc[R], M[R], gtL[R], gbL[R];
  double u_self[R], u_bp[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    lam_r[q] = a_r[q] = beta_prev[q] = gacc[q] = M[q] = gtL[q] = gbL[q] = T(0);
    u_self[q] = u_bp[q] = 1.0;
  }
  int s = -lane;
  YRow ycur;
  yfetch(ycur, s);
  const int total = nb * Pc + kWave - 1;
  // band b+1 is staged at step b*Pc + 63: every lane has taken band b by then
  int next_stage = kWave - 1;
  int stage_band = 1;

  auto step_fn = [&](int step, T* fv, double& fu) __attribute__((always_inline)) {
    if (step == next_stage) {  // wave-uniform
      if (stage_band < nb) stage(r0 - BR);
      ++stage_band;
      next_stage += Pc;
    }
    const bool colok = s >= 0 && s < Ncr && band < nb;
    T in[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) in[k] = shift_down(fl[k], fv[k]);
    double in_u = shift_down(flu, fu);
    // this step's inputs are consumed: refill the set for the step after next
    fread(fv, fu);

    const int c = Ncr - 1 - s;
    const bool newblk = (c & (F - 1)) == F - 1;  // first refined column of an original column
    const int j = c >> lam;
    bool last_active = false;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int r = r0 - q;
      const bool active = colok && r >= -1;
      const bool cell = active && r >= 0;
      last_active = active;
      if (active && s == 0) {  // right edge: saved last refined column
        u_self[q] = r >= 0 ? (double)ecol[r] : 1.0;
        u_bp[q] = (r + 1 <= Nrr - 1) ? (double)ecol[r + 1] : 1.0;
      }
      // this row's cell value M'(r, c), recomputed once per original column
      if (cell && newblk) {
        if (shared && (q & 1)) {
          M[q] = M[q - 1];
        } else if (diff) {
          T gtR, gbR;
          if (s == 0) {
            gtR = gval(xt[q], xtn[q], j + 1);
            gbR = gval(xb[q], xbn[q], j + 1);
          } else {
            gtR = gtL[q];
            gbR = gbL[q];
          }
          gtL[q] = gval_y(xt[q], xtn[q], ycur);
          gbL[q] = gval_y(xb[q], xbn[q], ycur);
          M[q] = ((gbR - gbL[q]) - (gtR - gtL[q])) * inv4;
        } else {
          M[q] = gval_y(xt[q], xtn[q], ycur) * inv4;
        }
      }

      // adjoint λ of this row's cell (r, c)
      const T Mq = M[q];
      const T bb = Mq * Mq * T(1.0 / 12.0);
      const T aa = T(0.5) * Mq + bb;
      const T eps = in[2] + a_r[q] * lam_r[q] + beta_prev[q];
      const T seed = (cell && r == Nrr - 1 && c == Ncr - 1) ? w : T(0);
      T mu = in[1] + eps + seed;
      T lv = lam_r[q] + mu;
      mu = cell ? mu : in[1];
      lv = cell ? lv : T(0);

      // rebuild U(r, c-1)
      double u_new = 1.0;
      if (active && c >= 1) {
        if (r == Nrr - 1) {
          u_new = (double)erow[c - 1];
        } else if (r >= 0) {
          const double mb = (double)in[4];
          const double bq = mb * mb * (1.0 / 12.0);
          const double A = 1.0 + 0.5 * mb + bq;
          const double B = 1.0 - bq;
          // 1/B from the float32 reciprocal and two Newton steps (relative
          // error 2^-23 -> 2^-46 -> below double rounding), cheaper than the
          // float64 division sequence
          double rB = (double)__builtin_amdgcn_rcpf((float)B);
          rB = fma(rB, fma(-B, rB, 1.0), rB);
          rB = fma(rB, fma(-B, rB, 1.0), rB);
          u_new = ((u_self[q] + in_u) * A - u_bp[q]) * rB;
        }
      }

      // cell (r+1, c), completed here: λ and M' from below, U from both rows
      const int rb = r + 1;
      const bool below = active && rb <= Nrr - 1;
      T g = T(0);
      if (below) {
        const double mb = (double)in[4];
        g = in[0] * (T)((u_self[q] + in_u) * (0.5 + mb * (1.0 / 6.0)) +
                        u_new * mb * (1.0 / 6.0));
      }
      gacc[q] = (newblk ? T(0) : gacc[q]) + g;
      const T part = (((rb & (F - 1)) == F - 1) ? T(0) : in[5]) + gacc[q];
      if (below && (c & (F - 1)) == 0 && (rb & (F - 1)) == 0)
        gm[(int64_t)(rb >> lam) * Nc + (c >> lam)] = part * inv4;

      // this row's outputs are the next row's inputs
      const T o3 = in[3];
      in[0] = lv;
      in[1] = mu;
      in[2] = cell ? aa * lv : T(0);
      in[3] = cell ? bb * lv : T(0);
      in[4] = cell ? Mq : T(0);
      in[5] = part;
      lam_r[q] = lv;
      a_r[q] = cell ? aa : T(0);
      beta_prev[q] = o3;
      u_bp[q] = in_u;
      u_self[q] = u_new;
      in_u = u_new;
    }

    // the next step's y row, in flight while the flows are handed over
    yfetch(ycur, s + 1 == Pc ? 0 : s + 1);
#pragma unroll
    for (int k = 0; k < 6; ++k) fl[k] = in[k];
    flu = in_u;
    if (lane == kWave - 1 && last_active) {
      PdeSlot<T> o;
#pragma unroll
      for (int k = 0; k < 6; ++k) o.v[k] = fl[k];
      o.u = flu;
      float4 q4[SQ];
      __builtin_memset(q4, 0, sizeof(q4));
      __builtin_memcpy(q4, &o, sizeof(PdeSlot<T>));
#pragma unroll
      for (int k = 0; k < SQ; ++k) fifo[s * SQ + k] = q4[k];
    }

    ++s;
    if (s == Pc) {
      s = 0;
      ++band;
      r0 -= BR;
      take();
#pragma unroll
      for (int q = 0; q < R; ++q) {
        lam_r[q] = a_r[q] = beta_prev[q] = gacc[q] = M[q] = T(0);
        u_self[q] = u_bp[q] = 1.0;
      }
    }
  };

  int step = 0;
  for (; step + 2 <= total; step += 2) {
    step_fn(step, fv0, fu0);
    step_fn(step + 1, fv1, fu1);
  }
  if (step < total) step_fn(step, fv0, fu0);
}

// ---------------------------------------------------------------------------
// Phase-aligned adjoint for dyadic order 0 and 1 (F = 2^LAM), R = 4 refined
// rows per lane; the adjoint in the working precision, U rebuilt in float64.
//
// On the flipped grid (row index r counted up from the last refined row,
// column c counted left from the last refined column) the adjoint and the
// rebuild of U are both forward recursions over one stencil, so the sweep
// has the shape of the forward: lane i takes flipped refined column k at
// step k + F*(i+1), F steps behind lane i-1, and every lane is at the same
// position within its original column.  The row of U index r (1-based,
// U(0,.) = U(.,0) = 1) at column c receives from the row below (r + 1, the
// lane below or the previous row of the lane) mu(r+1,c), a(r+1,c) lam(r+1,c),
// b(r+1,c) lam(r+1,c) and its own V(r,c), U(r,c-1), which the row below
// rebuilt, and computes
//   mu(r,c)    = mu(r+1,c) + a(r+1,c) lam(r+1,c) + a(r,c+1) lam(r,c+1) + b(r+1,c+1) lam(r+1,c+1)
//   lam(r,c)   = lam(r,c+1) + mu(r,c)
//   S          = U(r-1,c) + U(r,c-1)
//   V(r-1,c)   = (V(r,c) - S a(r,c) - U(r-1,c) b(r,c)) / (1 - b(r,c))
//   U(r-1,c-1) = U(r-1,c) - V(r-1,c)
//   dL/dM'(r,c) = lam(r,c) (S (1/2 + m/6) + U(r-1,c-1) m/6)
// and hands mu, a lam, b lam, V(r-1,c), U(r-1,c-1) to the row above.  The
// rebuild is the forward's increment update (V(r,c) = V(r-1,c) + S a +
// U(r-1,c-1) b, U(r,c) = U(r,c-1) + V(r,c)) solved for the corner it did not
// know.  It runs in float64 whatever the working precision: solving the
// Goursat problem backwards from the saved last row and column amplifies
// rounding by a factor that grows quickly with U (at K ~ 2 a float32
// rebuild is 1e-5 off, at K ~ 100 it is meaningless; in float64 a pair with
// K ~ 1e5 still loses about five digits), and a float64 fused multiply-add
// issues at the rate of an unpacked float32 one on gfx950.  The row below
// the grid is the saved last refined row (V and U from it, mu = dL/dK at
// the last column, which seeds lam(Nrr,Ncr)); the top row r = 1 keeps
// U(0,.) = 1 through a zero rebuild factor.  The static kernel, its ΔΔ and the coefficients a = m/2 +
// m^2/12, b = m^2/12, 1/(1-b), 1/2 + m/6, m/6 are evaluated once per original
// cell on wave-uniform steps; they are zero on the priming, padding and
// out-of-grid cells, which therefore need no per-cell mask.  dL/dM of an
// original cell is the sum over its F x F refined cells, which sit in one
// lane, times 1/F^2, written once per original cell.
// ---------------------------------------------------------------------------
template <typename T, int DMAX, int KIND, int LAM>
__global__ void __launch_bounds__(kWave)
pde_bwd_aligned_kernel(const PdeBwdEdgeArgs a) {
  constexpr int F = 1 << LAM;
  constexpr int R = 4;
  constexpr int NO = R / F;       // original rows per lane
  constexpr int BR = kWave * R;   // refined rows per band
  constexpr int NF = 4;           // FIFO slot: [mu, a lam, b lam, G] (T), then [V, U] (double)
  constexpr int SB = NF * (int)sizeof(T) + 16;  // slot bytes
  constexpr int XS = DMAX + 1;    // staged x row: DMAX scaled features and the norm
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* smem = reinterpret_cast<T*>(smem_raw);

  const int lane = threadIdx.x;
  const int64_t pr = blockIdx.x;
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
  const int Nrr = Nr * F, Ncr = Nc * F;
  const int Pc = pde_bwd_aligned_pc(Nc, LAM);
  const int nb = (Nrr + BR - 1) / BR;
  const T inv4 = T(1) / T(F * F);
  const T xscale = T(a.xscale), yscale = T(a.yscale);
  const double* erow = reinterpret_cast<const double*>(a.edges) + a.e_off[pr];  // U(Nrr, c), c = 1..Ncr
  const double* ecol = erow + Ncr;                                                // U(r, Ncr), r = 1..Nrr
  T* gm = reinterpret_cast<T*>(a.gm) + a.m_off[pr];                    // [Nr][Nc]
  // band hand-over FIFO: this pair's [fifo_cols] slots of the workspace
  unsigned char* fifo = reinterpret_cast<unsigned char*>(a.fifo) + (size_t)pr * a.fifo_cols * SB;
  auto slot_t = [&](int sl) -> T* { return reinterpret_cast<T*>(fifo + (size_t)sl * SB); };
  auto slot_d = [&](int sl) -> double* {
    return reinterpret_cast<double*>(fifo + (size_t)sl * SB + NF * sizeof(T));
  };
  T* xslot = smem + lane * 2 * NO * XS;  // this lane's [2][NO][XS] in LDS, by band parity

  // Band b of this lane: original rows (64 b + lane) * NO + k counted up from
  // the last (cell row Nr - 1 - that), whose upper x row is the cell row's
  // index, prepared as the forward prepares it (pde_xrow_prep) into slot
  // b & 1 with its norm term; and U(r - 1, Ncr) of its refined rows r (the
  // rebuild's start) into registers.
  double ecn[R];
  auto load_rows = [&](int b) __attribute__((always_inline)) {
    T* dst = xslot + (b & 1) * NO * XS;
#pragma unroll
    for (int k = 0; k < NO; ++k) {
      const int i = Nr - 1 - ((b * kWave + lane) * NO + k);
      const bool ok = i >= 0;
      T raw[DMAX], xm[DMAX], nx;
#pragma unroll
      for (int e = 0; e < DMAX; ++e) raw[e] = (ok && e < d) ? xs[(int64_t)i * d + e] : T(0);
      pde_xrow_prep<T, DMAX, KIND == kRBF>(raw, xm, nx, xscale, yscale, LAM);
#pragma unroll
      for (int e = 0; e < DMAX; ++e) dst[k * XS + e] = xm[e];
      dst[k * XS + DMAX] = nx;
    }
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int rq = Nrr - ((b * kWave + lane) * R + q);
      ecn[q] = rq >= 2 ? ecol[rq - 2] : 1.0;
    }
  };
  // y row as stored and its norm term, as the forward forms them
  auto load_y = [&](T* yv, T& yn, int yc) __attribute__((always_inline)) {
    yc = yc < 0 ? 0 : (yc >= Ly ? Ly - 1 : yc);
    T n2 = T(0);
#pragma unroll
    for (int e = 0; e < DMAX; ++e) {
      yv[e] = e < d ? ys[(int64_t)yc * d + e] : T(0);
      const T v = yv[e] * yscale;
      n2 = pde_fma(v, v, n2);
    }
    yn = KIND == kRBF ? n2 : T(0);
  };
  auto kfused = [&](const T* xv, T nx, const T* yv, T yn) __attribute__((always_inline)) -> T {
    return pde_gval<T, DMAX, KIND == kRBF>(xv, nx, yv, yn);
  };
  load_rows(0);
  double up[R];  // U(r-1, c) of the lane's rows: the rebuild's right neighbour
#pragma unroll
  for (int q = 0; q < R; ++q) up[q] = ecn[q];
  load_rows(1);

  // FIFO initial contents: the row below the grid.  Slot s holds flipped
  // column k = s - F (U column c = Ncr - k): V(Nrr,c), U(Nrr,c-1), the seed
  // mu = dL/dK at k = 0, and the static kernel of the last x row at the
  // original column's left y column (the bottom edge of the lowest cells).
  {
    T xl[DMAX], xln;
    {
      T raw[DMAX];
#pragma unroll
      for (int e = 0; e < DMAX; ++e) raw[e] = (diff && e < d) ? xs[(int64_t)Nr * d + e] : T(0);
      pde_xrow_prep<T, DMAX, KIND == kRBF>(raw, xl, xln, xscale, yscale, LAM);
    }
    const T w = T(a.gout[pr]);
    for (int s = lane; s < Pc; s += kWave) {
      const int k = s - F;
      T mu = T(0), G = T(0);
      double V = 0.0, U = 1.0;
      if (k >= 0 && k < Ncr) {
        const int c = Ncr - k;
        const double uc = erow[c - 1];
        const double ul = c >= 2 ? erow[c - 2] : 1.0;
        V = uc - ul;
        U = ul;
        mu = k == 0 ? w : T(0);
      }
      const int yc = Nc - 1 - (k >> LAM);
      if (diff && yc >= 0 && yc <= Nc) {
        T yv[DMAX], yn;
        load_y(yv, yn, yc);
        G = kfused(xl, xln, yv, yn);
      }
      T* st = slot_t(s);
      st[0] = mu;
      st[1] = T(0);
      st[2] = T(0);
      st[3] = G;
      double* sd = slot_d(s);
      sd[0] = V;
      sd[1] = U;
    }
  }
  __syncthreads();

  // per-row state
  T lamp[R], alp[R], blb[R];  // lam(r,c+1), a(r,c+1) lam(r,c+1), the row below's b lam of the last step
#pragma unroll
  for (int q = 0; q < R; ++q) lamp[q] = alp[q] = blb[q] = T(0);
  // per original row of the lane: gradient partial, static kernel at the right
  // y column (lower and upper x row), the scheme's coefficients
  // (the rebuild's in double: the backward Goursat solve amplifies rounding
  // by about the growth of U, which float32 cannot absorb beyond K ~ 10)
  // original columns per written row segment (32 bytes; 16 at F = 1, where
  // the lane holds four original rows)
  constexpr int kSeg = F == 2 ? 8 : 4;
  T gseg[NO][kSeg];
#pragma unroll
  for (int k = 0; k < NO; ++k)
#pragma unroll
    for (int w = 0; w < kSeg; ++w) gseg[k][w] = T(0);
  T gacc[NO], gbr[NO], gtr[NO], ca[NO], cb[NO];
  double cad[NO], cbd[NO], cgd[NO], chd[NO], rbq[R];
#pragma unroll
  for (int k = 0; k < NO; ++k) {
    gacc[k] = gbr[k] = gtr[k] = ca[k] = cb[k] = T(0);
    cad[k] = cbd[k] = cgd[k] = chd[k] = 0.0;
  }
#pragma unroll
  for (int q = 0; q < R; ++q) rbq[q] = 1.0;
  T fl[F][NF];        // outputs of the last F steps, by phase
  double fd[F][2];    // V, U of the last F steps, by phase
#pragma unroll
  for (int p = 0; p < F; ++p) {
#pragma unroll
    for (int k = 0; k < NF; ++k) fl[p][k] = T(0);
    fd[p][0] = 0.0;
    fd[p][1] = 1.0;
  }
  int cc = -F - F * lane;  // flipped refined column of the current step
  int band = 0;
  int rr0 = lane * R;      // flipped first refined row of the lane in its band
  // 1 for the refined rows whose upper neighbour is rebuilt, 0 for r = 1
  // (U(0, .) = 1 stays exact) and the padding rows above the grid
  double rmask[R];
  auto set_rmask = [&]() __attribute__((always_inline)) {
#pragma unroll
    for (int q = 0; q < R; ++q) rmask[q] = Nrr - (rr0 + q) >= 2 ? 1.0 : 0.0;
  };
  set_rmask();

  // lane 0's FIFO slot, read two steps ahead into the register set of the
  // step's parity (every lane reads it at one address; only lane 0 uses it)
  T ny[DMAX], nyn = T(0), nf[2][NF];
  double nd[2][2];
  auto fetch_col = [&](int c) __attribute__((always_inline)) {
    load_y(ny, nyn, Nc - 1 - (c >> LAM));
  };
  // lane 63 wrote a slot Pc - 63F - 2 >= F + 30 steps before this read,
  // and every step issues the two loads of this read at least
  constexpr int kFifoWait = fifo_wait_count(F + 30, 2);
  auto fetch_fifo = [&](int Q, int c) __attribute__((always_inline)) {
    const int s0 = __builtin_amdgcn_readfirstlane(c + F);  // lane 0's slot (in [0, Pc))
    wait_vmcnt<kFifoWait>();
    const T* st = slot_t(s0);
    const double* sd = slot_d(s0);
#pragma unroll
    for (int k = 0; k < NF; ++k) nf[Q][k] = st[k];
    nd[Q][0] = sd[0];
    nd[Q][1] = sd[1];
  };
  fetch_col(cc);
  fetch_fifo(0, cc);
  fetch_fifo(1, cc + 1);

  auto step_fn = [&](auto phase_tag, auto parity_tag) __attribute__((always_inline)) {
    constexpr int P = decltype(phase_tag)::value;
    constexpr int Q = decltype(parity_tag)::value;  // register set of the FIFO prefetch
    const bool active = cc >= -F && band < nb;
    T imu = shift_down(fl[P][0], nf[Q][0]);
    T ial = shift_down(fl[P][1], nf[Q][1]);
    T ibl = shift_down(fl[P][2], nf[Q][2]);
    double iv = shift_down(fd[P][0], nd[Q][0]);
    double iu = shift_down(fd[P][1], nd[Q][1]);
    T oG = T(0);
    // a lane leaves its band after the last refined column of an original
    // column: only on phase F-1 steps
    const bool wrap = P == F - 1 && cc + 1 == Pc - F;
    const int orow = Nr - 1 - (rr0 >> LAM);  // cell row of the lane's lowest original row
    if constexpr (P == 0) {  // a new original column: once per original cell
      const bool colok = cc >= 0 && cc < Ncr;
      T gb = shift_down(fl[P][3], nf[Q][3]);  // lower x row of the lowest original row
      const T* xb = xslot + (band & 1) * NO * XS;
#pragma unroll
      for (int k = 0; k < NO; ++k) {
        T xv[DMAX];
#pragma unroll
        for (int e = 0; e < DMAX; ++e) xv[e] = xb[k * XS + e];
        const T gt = kfused(xv, xb[k * XS + DMAX], ny, nyn);
        T m = diff ? (gbr[k] - gb) - (gtr[k] - gt) : gt;
        const bool ok = colok && orow - k >= 0;
        m = ok ? m : T(0);  // 1/4^lam is folded into the static-kernel values
        gbr[k] = gb;
        gtr[k] = gt;
        // a and b as the forward forms them, so that the rebuild inverts the
        // very coefficients the forward used
        pde_coef(m, ca[k], cb[k]);
        cad[k] = (double)ca[k];
        cbd[k] = (double)cb[k];
        chd[k] = (double)m * (1.0 / 6.0);
        cgd[k] = ok ? 0.5 + chd[k] : 0.0;
        // 1 / (1 - b): the float32 reciprocal and two Newton steps
        // (relative error 2^-23 -> 2^-46 -> below float64 rounding)
        const double bb = 1.0 - cbd[k];
        double rinv = (double)__builtin_amdgcn_rcpf((float)bb);
        rinv = fma(rinv, fma(-bb, rinv, 1.0), rinv);
        rinv = fma(rinv, fma(-bb, rinv, 1.0), rinv);
#pragma unroll
        for (int f = 0; f < F; ++f) rbq[k * F + f] = rinv * rmask[k * F + f];
        gb = gt;
      }
      oG = gb;
      // the next original column's y row, F steps ahead (the next band's
      // priming column after the band's last original column)
      const bool wrapn = cc + F >= Pc - F;
      fetch_col(wrapn ? -F : cc + F);
    }
    const int cc1 = wrap ? -F : cc + 1;
    // this lane's column two steps ahead (only lane 0's is used)
    fetch_fifo(Q, cc + 2 >= Pc - F ? cc + 2 - Pc : cc + 2);

#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int k = q / F;
      const T mu = imu + ial + alp[q] + blb[q];
      const T lam = lamp[q] + mu;
      const double S = up[q] + iu;
      const double vn = (iv - S * cad[k] - up[q] * cbd[k]) * rbq[q];
      const double un = up[q] - vn;
      gacc[k] += lam * T(S * cgd[k] + un * chd[k]);
      const T al = ca[k] * lam, bl = cb[k] * lam;
      blb[q] = ibl;
      lamp[q] = lam;
      alp[q] = al;
      up[q] = un;
      imu = mu;
      ial = al;
      ibl = bl;
      iv = vn;
      iu = un;
    }
    if constexpr (P == F - 1) {  // the original column's last refined column
      // dL/dM of the lane's original cells goes through a shift register of
      // the last kSeg original columns, written as one contiguous segment of
      // each row: per-column stores land on 64 rows at once (the lanes are
      // skewed), which the memory system turns into a partial-line write
      // each, about ten times the bytes stored
      const int p = cc >> LAM;  // flipped original column
#pragma unroll
      for (int k = 0; k < NO; ++k) {
#pragma unroll
        for (int w = kSeg - 1; w > 0; --w) gseg[k][w] = gseg[k][w - 1];
        gseg[k][0] = gacc[k] * inv4;
        gacc[k] = T(0);
      }
      const bool flush = cc >= 0 && cc < Ncr && ((p & (kSeg - 1)) == kSeg - 1 || p == Nc - 1);
      if (flush) {
        const int j0 = Nc - 1 - p;             // the segment's first column
        const int cnt = (p & (kSeg - 1)) + 1;  // kSeg, or the grid's last partial segment
#pragma unroll
        for (int k = 0; k < NO; ++k) {
          if (orow - k >= 0) {
            T* g = gm + (int64_t)(orow - k) * Nc + j0;
#pragma unroll
            for (int w = 0; w < kSeg; ++w)
              if (w < cnt) g[w] = gseg[k][w];
          }
        }
      }
    }
    fl[P][0] = imu;
    fl[P][1] = ial;
    fl[P][2] = ibl;
    fl[P][3] = oG;
    fd[P][0] = iv;
    fd[P][1] = iu;
    if (lane == kWave - 1 && active) {
      const int s = cc + F;
      T* st = slot_t(s);
      st[0] = imu;
      st[1] = ial;
      st[2] = ibl;
      st[3] = oG;
      double* sd = slot_d(s);
      sd[0] = iv;
      sd[1] = iu;
    }
    cc = cc1;
    if constexpr (P == F - 1) {
      // a lane entering its next band starts from the right edge with a
      // zero adjoint; wave-uniform test, since few steps have such a lane
      if (__builtin_amdgcn_ballot_w64(wrap)) {
#pragma unroll
        for (int q = 0; q < R; ++q) {
          blb[q] = wrap ? T(0) : blb[q];
          lamp[q] = wrap ? T(0) : lamp[q];
          alp[q] = wrap ? T(0) : alp[q];
          up[q] = wrap ? ecn[q] : up[q];
        }
        band += wrap ? 1 : 0;
        rr0 += wrap ? BR : 0;
        if (wrap) set_rmask();
      }
    }
  };

  // The x rows of band b+1 go into the slot of band b-1 once lane 63 has
  // entered band b (every lane has left band b-1); U(r - 1, Ncr) of band
  // b+1 waits in registers until each lane's own entry into that band.
  const int total = nb * Pc + F * (kWave - 1);  // a multiple of F
  int next_load = F * (kWave - 1) + Pc;
  int load_band = 2;
  // two steps per iteration (both phases at F = 2): the FIFO prefetch
  // alternates between two register sets; band changes fall on a multiple of
  // F, and Pc is even at F = 2, so the loads of the next band's rows stay on
  // an iteration boundary at F = 2 and are checked on both steps at F = 1
  using Z = std::integral_constant<int, 0>;
  using O = std::integral_constant<int, 1>;
  int step = 0;
  for (; step + 2 <= total; step += 2) {
    if (step == next_load || (F == 1 && step + 1 == next_load)) {  // wave-uniform
      if (F == 1 && step + 1 == next_load) {
        step_fn(Z{}, Z{});
        load_rows(load_band);
        ++load_band;
        next_load += Pc;
        step_fn(Z{}, O{});
        continue;
      }
      load_rows(load_band);
      ++load_band;
      next_load += Pc;
    }
    step_fn(Z{}, Z{});
    if constexpr (F == 2) step_fn(O{}, O{});
    else step_fn(Z{}, O{});
  }
  if (step < total) step_fn(Z{}, Z{});  // F = 1 with an odd step count
}

// Refined rows per lane: 2 where the staged and register x rows fit (up to
// 16 float32 / 8 float64 features), else 1 (4 rows measured no faster at
// C2: the adjoint's float64 rebuild keeps it issue bound per row).
__host__ __device__ constexpr int pde_bwd_rows_for(int dmax, int esz) {
  return dmax * esz <= 64 ? 2 : 1;
}
template <typename T, int DMAX>
constexpr int pde_bwd_rows() { return pde_bwd_rows_for(DMAX, (int)sizeof(T)); }

template <typename T, int DMAX, int KIND>
static hipError_t launch_bwd_aligned(const PdeBwdEdgeArgs& a, size_t lds, hipStream_t s) {
  auto k = a.dyadic == 0 ? pde_bwd_aligned_kernel<T, DMAX, KIND, 0>
                         : pde_bwd_aligned_kernel<T, DMAX, KIND, 1>;
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)k,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k, dim3((unsigned)a.n_pairs), dim3(kWave), lds, s, a);
  return hipGetLastError();
}

template <typename T, int DMAX>
static hipError_t launch_bwd_edges(const PdeBwdEdgeArgs& a, size_t lds, hipStream_t s) {
  if (a.dyadic <= 1)
    return a.kind == kRBF ? launch_bwd_aligned<T, DMAX, kRBF>(a, lds, s)
                          : launch_bwd_aligned<T, DMAX, kLinear>(a, lds, s);
  auto k = pde_bwd_edges_kernel<T, DMAX, pde_bwd_rows<T, DMAX>()>;
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
  if (a.d <= 4) return launch_bwd_edges<T, 4>(a, lds, s);
  if (a.d <= 8) return launch_bwd_edges<T, 8>(a, lds, s);
  if (a.d <= 16) return launch_bwd_edges<T, 16>(a, lds, s);
  return launch_bwd_edges<T, 32>(a, lds, s);
}

size_t pde_bwd_edges_lds(int dtype, int d, int ly, int ncr, int nrr, int dyadic) {
  const size_t esz = dtype == KSIG_F32 ? 4 : 8;
  const int dmax = d <= 4 ? 4 : (d <= 8 ? 8 : (d <= 16 ? 16 : 32));
  if (dyadic <= 1)  // phase-aligned kernel: two bands of the lanes' x rows
    return (size_t)kWave * 2 * (4 >> dyadic) * (dmax + 1) * esz;
  const int rows = pde_bwd_rows_for(dmax, (int)esz);
  return ((size_t)ly * (dmax + 1) + (size_t)ncr + (size_t)nrr +
          (size_t)kWave * rows * (2 * dmax + 2)) * esz;
}

int pde_bwd_fifo_cols(int ncr, int dyadic) {
  return dyadic <= 1 ? pde_bwd_aligned_pc(ncr >> dyadic, dyadic) : pde_bwd_pc(ncr);
}

size_t pde_bwd_fifo_bytes(int dtype, int ncr, int dyadic) {
  if (dyadic <= 1)  // phase-aligned kernel: [mu, a lam, b lam, G] and [V, U] (float64) per slot
    return (size_t)pde_bwd_fifo_cols(ncr, dyadic) * (4 * (dtype == KSIG_F32 ? 4 : 8) + 16);
  const int quads = dtype == KSIG_F32 ? pde_slot_quads<float>() : pde_slot_quads<double>();
  return (size_t)pde_bwd_pc(ncr) * quads * 16;
}

hipError_t pde_bwd_edges_launch(const PdeBwdEdgeArgs& a, int dtype, size_t lds, hipStream_t s) {
  if (dtype == KSIG_F32) return dispatch_bwd_edges<float>(a, lds, s);
  return dispatch_bwd_edges<double>(a, lds, s);
}

}  // namespace ksig
