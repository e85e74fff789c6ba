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

#include "ksig_common.h"
#include "ksig_hip.h"
#include "dp_forward.h"

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
  const T* edge_row = reinterpret_cast<const T*>(a.edges) + a.e_off[pr];  // [Ncr]
  const T* edge_col = edge_row + Ncr;                                       // [Nrr]
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
  for (int c = lane; c < Ncr; c += kWave) erow[c] = edge_row[c];
  for (int c = lane; c < Nrr; c += kWave) ecol[c] = edge_col[c];

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

// Refined rows per lane: 2 where the staged and register x rows fit (up to
// 16 float32 / 8 float64 features), else 1 (4 rows measured no faster at
// C2: the adjoint's float64 rebuild keeps it issue bound per row).
__host__ __device__ constexpr int pde_bwd_rows_for(int dmax, int esz) {
  return dmax * esz <= 64 ? 2 : 1;
}
template <typename T, int DMAX>
constexpr int pde_bwd_rows() { return pde_bwd_rows_for(DMAX, (int)sizeof(T)); }

template <typename T, int DMAX>
static hipError_t launch_bwd_edges(const PdeBwdEdgeArgs& a, size_t lds, hipStream_t s) {
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

size_t pde_bwd_edges_lds(int dtype, int d, int ly, int ncr, int nrr) {
  const size_t esz = dtype == KSIG_F32 ? 4 : 8;
  const int dmax = d <= 4 ? 4 : (d <= 8 ? 8 : (d <= 16 ? 16 : 32));
  const int rows = pde_bwd_rows_for(dmax, (int)esz);
  return ((size_t)ly * (dmax + 1) + (size_t)ncr + (size_t)nrr +
          (size_t)kWave * rows * (2 * dmax + 2)) * esz;
}

int pde_bwd_fifo_cols(int ncr) { return pde_bwd_pc(ncr); }

size_t pde_bwd_fifo_bytes(int dtype, int ncr) {
  const int quads = dtype == KSIG_F32 ? pde_slot_quads<float>() : pde_slot_quads<double>();
  return (size_t)pde_bwd_pc(ncr) * quads * 16;
}

hipError_t pde_bwd_edges_launch(const PdeBwdEdgeArgs& a, int dtype, size_t lds, hipStream_t s) {
  if (dtype == KSIG_F32) return dispatch_bwd_edges<float>(a, lds, s);
  return dispatch_bwd_edges<double>(a, lds, s);
}

}  // namespace ksig
