// Adjoint of the phase-aligned signature-PDE sweep (dyadic order 0 and 1,
// difference = True) fused with the chain rule through the static kernel
// (gfx950).  No reference counterpart: ksig/algorithms.py:344-446 is
// forward-only; the scheme differentiated is ksig/algorithms.py:380-385 on
// M = ΔΔ k(x_i, y_j) (ksig/algorithms.py:409-413, ksig/static/kernels.py:248-260).
//
// The sweep is that of the two-pass adjoint (pde_backward_edges.hip): on the
// flipped grid the adjoint λ and the rebuild of U from the forward's last
// refined row and column are both forward recursions, lane i takes flipped
// refined column k at step k + F*(i+1), F = 2^λ steps behind lane i-1, U is
// rebuilt in float64, and every lane is at the same position within its
// original column.  Instead of writing dL/dM for a second kernel, each lane
// finishes the chain rule itself, once per original column:
//   dL/dG(a, b) = gM(a-1, b-1) - gM(a-1, b) - gM(a, b-1) + gM(a, b)
//   (gM = dL/dM of an original cell, zero outside the grid), and
//   ω(a, b)     = dL/dG(a, b) G(a, b)          (RBF; G evaluated in the sweep)
//               = dL/dG(a, b)                  (linear),
// for the static-kernel points a on the bottom edges of the lane's original
// rows (the cell row below belongs to the lane below, which finished that
// column F steps earlier and passes its top row's gM up with the other
// flows) and b = the column right of the one just finished.  The sums the
// gradients need are accumulated in registers:
//   per point row a (this lane):   Σ_b ω,  Σ_b ω y_b,
//   per point column b (a flow):   Σ_a ω,  Σ_a ω x_a, passed up the lanes and
//                                  through the band FIFO like the adjoint's flows,
// and turned into gradients when a lane leaves its band pass (rows) and when
// the last band's lane 63 finishes a column (columns):
//   RBF    G = exp(-s |x - y|^2):  dL/dx_a = -2s (x_a Σ_b ω - Σ_b ω y_b),
//                                  dL/dy_b = -2s (y_b Σ_a ω - Σ_a ω x_a),
//                                  dL/dσ  = Σ ω |x_a - y_b|^2 / σ^3,
//   linear G = c <x, y>:           dL/dx_a = c Σ_b ω y_b,  dL/dy_b = c Σ_a ω x_a,
//                                  dL/dc  = Σ ω <x_a, y_b>.
// |x_a - y_b|^2 is expanded as |x_a|^2 + |y_b|^2 - 2 <x_a, y_b> over the same
// sums, combined in float64.  Row and column gradients go to the caller's
// arrays with float atomics (many pairs share a sequence); the parameter part
// is one float64 per pair.  An extra original row above the grid (the
// virtual row, no cells) gives point row 0 its owner.
#include <type_traits>

#include "ksig_common.h"
#include "ksig_hip.h"
#include "dp_forward.h"
#include "pde_common.h"

namespace ksig {

namespace {

// float words per FIFO slot after its two float64 (V, U): mu, a lam, b lam,
// G, gM and the column sums (DMAX of Σ ω x, then Σ ω), whole 16-byte quads
template <int DMAX>
constexpr int fused_nf() { return ((6 + DMAX) + 3) & ~3; }

// LDS per wave: each lane's staged x rows (two bands of 2^(2-lam) + 1 rows of
// dmax + 1 floats), then each lane's two bands of 4 float64 rebuild starts
__host__ __device__ constexpr int fused_x_bytes(int dmax, int lam) {
  return kWave * 2 * ((4 >> lam) + 1) * (dmax + 1) * 4;
}
__host__ __device__ constexpr int fused_lds_bytes(int dmax, int lam) {
  return fused_x_bytes(dmax, lam) + kWave * 2 * 4 * 8;
}

template <typename T, int DMAX, int KIND, int LAM, bool VEC>
__global__ void __launch_bounds__(kWave)
pde_bwd_fused_kernel(const PdeBwdFusedArgs a) {
  constexpr int F = 1 << LAM;
  constexpr int R = 4;
  constexpr int NO = R / F;        // original rows per lane
  constexpr int BR = kWave * R;    // refined rows per band
  constexpr int NX = NO + 1;       // staged x rows: upper rows of the original rows, lower row of the first
  constexpr int XS = DMAX + 1;     // prepared row and its norm term
  constexpr bool kRbf = KIND == kRBF;
  constexpr int NTF = fused_nf<DMAX>();   // float words of a slot
  constexpr int NQ = 1 + NTF / 4;         // 16-byte quads of a slot
  constexpr int SB = 16 * NQ;             // slot bytes
  extern __shared__ __align__(16) unsigned char smem_raw[];

  const int lane = threadIdx.x;
  const int64_t pr = blockIdx.x;
  const int d = a.d;
  const int32_t ia = a.pairs[2 * pr];
  const int32_t ib = a.pairs[2 * pr + 1];
  const T* xs = reinterpret_cast<const T*>(a.x.data) + a.x.off[ia] * d;
  int Lx = a.x.len[ia];
  const T* ys = reinterpret_cast<const T*>(a.y.data) + a.y.off[ib] * d;
  int Ly = a.y.len[ib];
  // gradient rows of the kernel's x (rows) and y (columns) side
  T* gxs = a.gx ? reinterpret_cast<T*>(a.gx) + a.x.off[ia] * d : nullptr;
  T* gys = a.gy ? reinterpret_cast<T*>(a.gy) + a.y.off[ib] * d : nullptr;
  if (Ly > Lx) {  // orientation of the forward (rows = longer sequence)
    const T* tp = xs; xs = ys; ys = tp;
    int tl = Lx; Lx = Ly; Ly = tl;
    T* tg = gxs; gxs = gys; gys = tg;
  }
  const int Nr = Lx - 1, Nc = Ly - 1;
  if (Nr <= 0 || Nc <= 0) {
    if (lane == 0 && a.gp) a.gp[pr] = 0.0;
    return;
  }
  const int Nrr = Nr * F, Ncr = Nc * F;
  const int Pc = pde_fused_pc(Nc, LAM);
  const int nb = ((Nr + 1) * F + BR - 1) / BR;  // + the virtual row above the grid
  const T sx = T(a.xscale), sy = T(a.yscale);
  const T xmul = kRbf ? T(2) * sx * sy : sx * sy / T(F * F);  // prepared row = xmul * x
  const T rxmul = T(1) / xmul;
  const T inv4 = T(1) / T(F * F);
  const double* erow = reinterpret_cast<const double*>(a.edges) + a.e_off[pr];  // U(Nrr, c), c = 1..Ncr
  const double* ecol = erow + Ncr;                                                // U(r, Ncr), r = 1..Nrr
  // band FIFO: this pair's [fifo_cols] slots of the workspace through a
  // descriptor; lane 0's loads and lane 63's stores are the only ones in range
  unsigned char* fifo = reinterpret_cast<unsigned char*>(a.fifo) + (size_t)pr * a.fifo_cols * SB;
  const __amdgpu_buffer_rsrc_t fifo_rs = raw_rsrc(fifo, unsigned(Pc * SB));
  const int fin_vo = lane == 0 ? 0 : int(kBufOob);
  const int fout_vo = lane == kWave - 1 ? 0 : int(kBufOob);
  // y rows through a descriptor (columns past the sequence read zero)
  const int ystride = d * (int)sizeof(T);
  const __amdgpu_buffer_rsrc_t yrs = raw_rsrc(ys, unsigned((int64_t)Ly * ystride));
  T* xslot = reinterpret_cast<T*>(smem_raw) + lane * 2 * NX * XS;  // [2][NX][XS] by band parity
  // U(r - 1, Ncr) of the lane's refined rows for its next band, by band parity
  double* eslot = reinterpret_cast<double*>(smem_raw + fused_x_bytes(DMAX, LAM)) + lane * 2 * R;

  // ---- band staging --------------------------------------------------------
  // Band b: original row k of this lane is flipped row o = (64 b + lane) NO + k,
  // cell row i_k = Nr - 1 - o, upper x row i_k; slot row NO holds the lower
  // x row of row 0 (x_{i_0 + 1}).  U(r - 1, Ncr) of the refined rows (the
  // rebuild's start) waits in LDS for the lane's own band change.
  auto load_rows = [&](int b) __attribute__((always_inline)) {
    T* dst = xslot + (b & 1) * NX * XS;
    const int o0 = (b * kWave + lane) * NO;
#pragma unroll
    for (int k = 0; k < NX; ++k) {
      const int xr = k < NO ? Nr - 1 - (o0 + k) : Nr - o0;  // x row index
      const bool ok = xr >= 0 && xr < Lx;
      T raw[DMAX], xm[DMAX], nx;
#pragma unroll
      for (int e = 0; e < DMAX; ++e) raw[e] = (ok && e < d) ? xs[(int64_t)xr * d + e] : T(0);
      pde_xrow_prep<T, DMAX, kRbf>(raw, xm, nx, sx, sy, LAM);
#pragma unroll
      for (int e = 0; e < DMAX; ++e) dst[k * XS + e] = xm[e];
      dst[k * XS + DMAX] = nx;
    }
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int rq = Nrr - ((b * kWave + lane) * R + q);
      eslot[(b & 1) * R + q] = rq >= 2 ? ecol[rq - 2] : 1.0;
    }
  };
  load_rows(0);
  double up[R];  // U(r-1, c) of the lane's refined rows: the rebuild's right neighbour
#pragma unroll
  for (int q = 0; q < R; ++q) up[q] = eslot[q];
  load_rows(1);

  // FIFO initial contents: the row below the grid.  Slot s holds flipped
  // column k = s - F (U column c = Ncr - k): V(Nrr,c), U(Nrr,c-1), the seed
  // mu = dL/dK at k = 0, and the static kernel of the last x row at the
  // original column's left y column (the bottom edge of the lowest cells);
  // no cells below the grid, so gM and the column sums start at zero.
  {
    T xl[DMAX], nxl;
    {
      T raw[DMAX];
#pragma unroll
      for (int e = 0; e < DMAX; ++e) raw[e] = e < d ? xs[(int64_t)Nr * d + e] : T(0);
      pde_xrow_prep<T, DMAX, kRbf>(raw, xl, nxl, sx, sy, LAM);
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
      if (yc >= 0 && yc <= Nc) {
        T yv[DMAX], yn = T(0);
#pragma unroll
        for (int e = 0; e < DMAX; ++e) {
          yv[e] = e < d ? ys[(int64_t)yc * d + e] : T(0);
          const T v = yv[e] * sy;
          yn = pde_fma(v, v, yn);
        }
        G = pde_gval<T, DMAX, kRbf>(xl, nxl, yv, kRbf ? yn : T(0));
      }
      unsigned char* sl = fifo + (size_t)s * SB;
      *reinterpret_cast<double2*>(sl) = double2{V, U};
      T* sf = reinterpret_cast<T*>(sl + 16);
      sf[0] = mu;
#pragma unroll
      for (int f = 1; f < NTF; ++f) sf[f] = f == 3 ? G : T(0);
    }
  }
  __syncthreads();

  // ---- per-lane state ------------------------------------------------------
  T lamp[R], alp[R], blb[R];  // lam(r,c+1), a(r,c+1) lam(r,c+1), the row below's b lam of the last step
#pragma unroll
  for (int q = 0; q < R; ++q) lamp[q] = alp[q] = blb[q] = T(0);
  // per original row: dL/dM partial of the current cell, static kernel at
  // the right y column (lower and upper x row) and one column further right
  // (the point the chain rule finishes), the scheme's coefficients
  T gacc[NO], gbr[NO], gtr[NO], gbo[NO], ca[NO], cb[NO];
  double cad[NO], cbd[NO], cgd[NO], chd[NO], rbq[R];
  float rmask[R];  // 1 for the refined rows whose upper neighbour is rebuilt
  T rowm[NO];  // 1 for rows inside the grid
#pragma unroll
  for (int k = 0; k < NO; ++k) {
    gacc[k] = gbr[k] = gtr[k] = gbo[k] = ca[k] = cb[k] = T(0);
    cad[k] = cbd[k] = cgd[k] = chd[k] = 0.0;
  }
  // chain rule: gM of the last finished column (this lane's rows, the row
  // below), per point row Σ ω and Σ ω y, the column-sum flow
  T gmp[NO], gmbp = T(0), sw[NO], swy[NO][DMAX];
#pragma unroll
  for (int k = 0; k < NO; ++k) {
    gmp[k] = sw[k] = T(0);
#pragma unroll
    for (int e = 0; e < DMAX; ++e) swy[k][e] = T(0);
  }
  double gsig = 0.0;  // parameter-gradient part (before its scale)
  int band = 0;
  int xpar = 0;  // LDS offset of the current band's staged rows (band parity)
  int pcol = -1 - lane;  // flipped original column of the lane's next phase-0 step
  auto set_masks = [&]() __attribute__((always_inline)) {
    const int o0 = (band * kWave + lane) * NO;
#pragma unroll
    for (int k = 0; k < NO; ++k) rowm[k] = Nr - 1 - (o0 + k) >= 0 ? T(1) : T(0);
    const int r0 = (band * kWave + lane) * R;
#pragma unroll
    for (int q = 0; q < R; ++q) rmask[q] = Nrr - (r0 + q) >= 2 ? 1.f : 0.f;
  };
  set_masks();
#pragma unroll
  for (int q = 0; q < R; ++q) rbq[q] = 1.0;
  T fl[F][3];        // mu, a lam, b lam of the last step of each phase
  double fd[F][2];   // V, U
  T flg = T(0);      // G at the bottom edge of the lowest original row (phase 0)
  T flm = T(0);      // gM of the top original row (last phase)
  T fcs[DMAX + 1];   // column sums (last phase)
#pragma unroll
  for (int p = 0; p < F; ++p) {
#pragma unroll
    for (int k = 0; k < 3; ++k) fl[p][k] = T(0);
    fd[p][0] = 0.0;
    fd[p][1] = 1.0;
  }
#pragma unroll
  for (int e = 0; e <= DMAX; ++e) fcs[e] = T(0);

  // y rows: the row of a lane's flipped column p (its left y column
  // Nc - 1 - p) is fetched two columns ahead into the set of the loop's
  // column parity; the chain rule of column p needs the row of column p - 1,
  // kept in the history set of that parity when its column began
  auto yoff = [&](int p) __attribute__((always_inline)) {
    return __mul24(Nc - 1 - p, ystride);  // full-rate 24-bit multiply
  };
  T yA[DMAX], yB[DMAX], hA[DMAX], hB[DMAX], ynA = T(0), ynB = T(0);
#pragma unroll
  for (int e = 0; e < DMAX; ++e) hA[e] = hB[e] = T(0);
  auto fetch = [&](T* yv, T& yn, int p) __attribute__((always_inline)) {
    pde_load_y<T, DMAX, VEC>(yv, yrs, yoff(p), d);
    if constexpr (kRbf) {  // as the forward forms it (features past d count zero)
      T n2 = T(0);
#pragma unroll
      for (int e = 0; e < DMAX; ++e) {
        const T v = (VEC || e < d ? yv[e] : T(0)) * sy;
        n2 = pde_fma(v, v, n2);
      }
      yn = n2;
    }
  };
  // lane i's first two phase-0 steps see columns -1 - i and -i
  fetch(yA, ynA, -1 - lane);
  fetch(yB, ynB, -lane);
  int yp = 1 - lane;  // the next column to fetch

  // lane 0's FIFO slot, read two steps ahead into the set of the step's parity
  T nf[2][NTF];
  double nd[2][2];
  // lane 63 wrote a slot at least F + 30 steps before this read, and every
  // step issues the NQ loads of this read at least
  constexpr int kFifoWait = fifo_wait_count(F + 30, NQ);
  auto fetch_fifo = [&](int Q, int slot) __attribute__((always_inline)) {
    wait_vmcnt<kFifoWait>();
    const int so = slot * SB;
    {
      const u32x4 t = __builtin_amdgcn_raw_buffer_load_b128(fifo_rs, fin_vo, so, 0);
      nd[Q][0] = __hiloint2double((int)t.y, (int)t.x);
      nd[Q][1] = __hiloint2double((int)t.w, (int)t.z);
    }
#pragma unroll
    for (int qd = 0; qd < NTF / 4; ++qd) {
      const u32x4 t = __builtin_amdgcn_raw_buffer_load_b128(fifo_rs, fin_vo, so + 16 + 16 * qd, 0);
      nf[Q][4 * qd + 0] = __uint_as_float(t.x);
      nf[Q][4 * qd + 1] = __uint_as_float(t.y);
      nf[Q][4 * qd + 2] = __uint_as_float(t.z);
      nf[Q][4 * qd + 3] = __uint_as_float(t.w);
    }
  };
  fetch_fifo(0, 0);
  fetch_fifo(1, 1 % Pc);

  const int total = (nb * Pc + F * (kWave - 1) + 2 * F - 1) / (2 * F) * (2 * F);
  int slot2 = 2 % Pc;                  // (t + 2) mod Pc: lane 0's slot two steps ahead
  int slot63 = Pc - F * (kWave - 1);   // (t - 63F) mod Pc
  int load_band = 2;
  // the last band's lane 63 finishes point column b = Nc - p at its last
  // phase of flipped column p: its column counter
  const int t63 = F * (kWave - 1) + (nb - 1) * Pc + F;  // its step at flipped refined column 0

  auto step_fn = [&](auto phase_tag, auto set_tag, int t, int qq) __attribute__((always_inline)) {
    constexpr int P = decltype(phase_tag)::value;
    constexpr int S = decltype(set_tag)::value;   // column parity: y set and FIFO set
    constexpr int Q = (S * F + P) & 1;            // FIFO prefetch set of this step
    T* ycur = S == 0 ? yA : yB;
    T& ync = S == 0 ? ynA : ynB;
    T* hcur = S == 0 ? hA : hB;   // this column's row, kept for the next column
    T* yprv = S == 0 ? hB : hA;   // the previous column's row
    T imu = shift_down(fl[P][0], nf[Q][0]);
    T ial = shift_down(fl[P][1], nf[Q][1]);
    T ibl = shift_down(fl[P][2], nf[Q][2]);
    double iv = shift_down(fd[P][0], nd[Q][0]);
    double iu = shift_down(fd[P][1], nd[Q][1]);
    int sw_ = qq + F;
    sw_ = sw_ >= Pc ? sw_ - Pc : sw_;
    const T* xb = xslot + xpar;
    if constexpr (P == 0) {  // a new original column
      const bool colok = (unsigned)pcol < (unsigned)Nc;
      T gb = shift_down(flg, nf[Q][3]);  // lower x row of the lowest original row
#pragma unroll
      for (int k = 0; k < NO; ++k) {
        T xv[DMAX];
#pragma unroll
        for (int e = 0; e < DMAX; ++e) xv[e] = xb[k * XS + e];
        const T gt = pde_gval<T, DMAX, kRbf>(xv, kRbf ? xb[k * XS + DMAX] : T(0), ycur, ync);
        T m = (gbr[k] - gb) - (gtr[k] - gt);
        m = colok ? m * rowm[k] : T(0);  // 1/4^lam is folded into the static-kernel values
        gbo[k] = gbr[k];
        gbr[k] = gb;
        gtr[k] = gt;
        // a and b as the forward forms them, so that the rebuild inverts the
        // very coefficients the forward used
        pde_coef(m, ca[k], cb[k]);
        cad[k] = (double)ca[k];
        cbd[k] = (double)cb[k];
        chd[k] = (double)m * (1.0 / 6.0);
        cgd[k] = colok ? (0.5 + chd[k]) * (double)rowm[k] : 0.0;
        // 1 / (1 - b): the float32 reciprocal and two Newton steps
        // (relative error 2^-23 -> 2^-46 -> below float64 rounding)
        const double bb = 1.0 - cbd[k];
        double rinv = (double)__builtin_amdgcn_rcpf((float)bb);
        rinv = fma(rinv, fma(-bb, rinv, 1.0), rinv);
        rinv = fma(rinv, fma(-bb, rinv, 1.0), rinv);
#pragma unroll
        for (int f = 0; f < F; ++f) rbq[k * F + f] = rinv * (double)rmask[k * F + f];
        gb = gt;
      }
      flg = gb;
      ++pcol;
      // keep this column's row, then fetch the row of the column after the
      // next into its set; the lane whose band pass ends with the next
      // column fetches its next band's priming column instead (before a
      // lane starts, such resets fall on its priming column already)
#pragma unroll
      for (int e = 0; e < DMAX; ++e) hcur[e] = ycur[e];
      int sf = qq + 2 * F;
      sf = sf >= Pc ? sf - Pc : sf;
      sf = sf >= Pc ? sf - Pc : sf;
      if (sf < kWave * F) {  // wave-uniform
        if (lane == (sf >> LAM)) yp = -1;
      }
      fetch(ycur, ync, yp);
      ++yp;
    }
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int k = q / F;
      const T mu = imu + ial + alp[q] + blb[q];
      const T lam = lamp[q] + mu;
      const double Ssum = up[q] + iu;
      const double vn = (iv - Ssum * cad[k] - up[q] * cbd[k]) * rbq[q];
      const double un = up[q] - vn;
      gacc[k] += lam * T(Ssum * cgd[k] + un * chd[k]);
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
    fl[P][0] = imu;
    fl[P][1] = ial;
    fl[P][2] = ibl;
    fd[P][0] = iv;
    fd[P][1] = iu;
    if constexpr (P == F - 1) {
      // ---- chain rule of the point column b = Nc - p (p: this column) ----
      const T gmb = shift_down(flm, nf[Q][4]);  // gM of the row below, this column
      T cs[DMAX + 1];
#pragma unroll
      for (int e = 0; e <= DMAX; ++e) cs[e] = shift_down(fcs[e], nf[Q][5 + e]);
      T gmc[NO];
#pragma unroll
      for (int k = 0; k < NO; ++k) gmc[k] = gacc[k];
#pragma unroll
      for (int k = 0; k < NO; ++k) {
        const T below = k == 0 ? gmb : gmc[k - 1];
        const T belowp = k == 0 ? gmbp : gmp[k - 1];
        const T dg = (gmc[k] - gmp[k]) - (below - belowp);
        const T om = kRbf ? dg * gbo[k] : dg * inv4;
        sw[k] += om;
#pragma unroll
        for (int e = 0; e < DMAX; ++e) swy[k][e] = pde_fma(om, yprv[e], swy[k][e]);
        const int xk = k == 0 ? NO : k - 1;  // slot row of point row a = i_k + 1
#pragma unroll
        for (int e = 0; e < DMAX; ++e) cs[e] = pde_fma(om, xb[xk * XS + e], cs[e]);
        cs[DMAX] += om;
      }
#pragma unroll
      for (int k = 0; k < NO; ++k) {
        gmp[k] = gmc[k];
        gacc[k] = T(0);
      }
      gmbp = gmb;
      flm = gmc[NO - 1];
#pragma unroll
      for (int e = 0; e <= DMAX; ++e) fcs[e] = cs[e];
      // the last band's lane 63: column b's sums are complete
      {
        const int p63 = (t - t63) >> LAM;  // its flipped column (phase F-1 step)
        const int bcol = Nc - p63;
        if (t >= t63 && bcol >= 0 && bcol <= Nc) {  // wave-uniform
          if (lane == kWave - 1) {
            // y_b of this column is the previous column's row
            T yn2 = T(0);
#pragma unroll
            for (int e = 0; e < DMAX; ++e)
              if (e < d) yn2 = pde_fma(yprv[e], yprv[e], yn2);
            if (gys) {
#pragma unroll
              for (int e = 0; e < DMAX; ++e)
                if (e < d)
                  unsafeAtomicAdd(gys + (int64_t)bcol * d + e,
                                  T(a.c_self) * yprv[e] * cs[DMAX] +
                                      T(a.c_cross) * cs[e] * rxmul);
            }
            if constexpr (kRbf) gsig += (double)yn2 * (double)cs[DMAX];
          }
        }
      }
      // band change of the lane whose band pass this column ends
      if (t + 1 >= Pc && sw_ < kWave * F) {  // wave-uniform
        const int lw = sw_ >> LAM;
        if (lane == lw) {
          // its point rows are finished: gradients and the parameter part
          const int o0 = (band * kWave + lane) * NO;
#pragma unroll
          for (int k = 0; k < NO; ++k) {
            const int pa = Nr - (o0 + k);  // point row a = i_k + 1
            const int xk = k == 0 ? NO : k - 1;
            if (pa >= 0) {
              T xr[DMAX];
              double xx = 0.0, xy = 0.0;
#pragma unroll
              for (int e = 0; e < DMAX; ++e) {
                xr[e] = xb[xk * XS + e] * rxmul;
                xx += (double)xr[e] * (double)xr[e];
                xy += (double)xr[e] * (double)swy[k][e];
              }
              if (gxs) {
#pragma unroll
                for (int e = 0; e < DMAX; ++e)
                  if (e < d)
                    unsafeAtomicAdd(gxs + (int64_t)pa * d + e,
                                    T(a.c_self) * xr[e] * sw[k] + T(a.c_cross) * swy[k][e]);
              }
              gsig += kRbf ? xx * (double)sw[k] - 2.0 * xy : xy;
            }
          }
          // the next band: right edge, zero adjoint, its rows and masks
#pragma unroll
          for (int k = 0; k < NO; ++k) {
            sw[k] = gmp[k] = T(0);
#pragma unroll
            for (int e = 0; e < DMAX; ++e) swy[k][e] = T(0);
          }
          gmbp = T(0);
          ++band;
#pragma unroll
          for (int q = 0; q < R; ++q) {
            blb[q] = lamp[q] = alp[q] = T(0);
            up[q] = eslot[(band & 1) * R + q];
          }
          xpar = xpar == 0 ? NX * XS : 0;
          pcol = -1;
          set_masks();
        }
        if (lw == kWave - 1) {  // every lane has left its previous band
          load_rows(load_band);
          ++load_band;
        }
      }
    }
    // lane 63 hands its outputs to the next band (any band)
    if (t >= F * (kWave - 1)) {  // wave-uniform
      const int so = slot63 * SB;
      __builtin_amdgcn_raw_buffer_store_b128(
          u32x4{(unsigned)__double2loint(iv), (unsigned)__double2hiint(iv),
                (unsigned)__double2loint(iu), (unsigned)__double2hiint(iu)},
          fifo_rs, fout_vo, so, 0);
      T v[NTF];
      v[0] = imu; v[1] = ial; v[2] = ibl; v[3] = flg; v[4] = flm;
#pragma unroll
      for (int e = 0; e <= DMAX; ++e) v[5 + e] = fcs[e];
#pragma unroll
      for (int f = 6 + DMAX; f < NTF; ++f) v[f] = T(0);
#pragma unroll
      for (int qd = 0; qd < NTF / 4; ++qd)
        __builtin_amdgcn_raw_buffer_store_b128(
            u32x4{__float_as_uint(v[4 * qd]), __float_as_uint(v[4 * qd + 1]),
                  __float_as_uint(v[4 * qd + 2]), __float_as_uint(v[4 * qd + 3])},
            fifo_rs, fout_vo, so + 16 + 16 * qd, 0);
    }
    // lane 0's slot two steps ahead, into this step's input set: issued
    // last, once every input of the step (the chain rule's included) is used
    fetch_fifo(Q, slot2);
    slot2 = slot2 + 1 == Pc ? 0 : slot2 + 1;
    slot63 = slot63 + 1 == Pc ? 0 : slot63 + 1;
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
  // the first two columns outside the loop (the loop is entered with the
  // same loads in flight as at its back edge)
  two_columns(0);
  for (int t = 2 * F; t < total; t += 2 * F) two_columns(t);
  gsig = wave_sum(gsig);
  if (lane == 0 && a.gp) a.gp[pr] = gsig * a.gp_scale;
}

template <typename T, int DMAX, int KIND>
hipError_t launch_fused(const PdeBwdFusedArgs& a, size_t lds, hipStream_t s) {
  const bool vec = a.d == DMAX && ((uintptr_t)a.x.data & 15) == 0 &&
                   ((uintptr_t)a.y.data & 15) == 0;
  auto pick = [&](auto lam_tag) {
    constexpr int L = decltype(lam_tag)::value;
    return vec ? pde_bwd_fused_kernel<T, DMAX, KIND, L, true>
               : pde_bwd_fused_kernel<T, DMAX, KIND, L, false>;
  };
  auto k = a.dyadic == 0 ? pick(std::integral_constant<int, 0>{})
                         : pick(std::integral_constant<int, 1>{});
  hipLaunchKernelGGL(k, dim3((unsigned)a.n_pairs), dim3(kWave), lds, s, a);
  return hipGetLastError();
}

template <typename T>
hipError_t dispatch_fused(const PdeBwdFusedArgs& a, size_t lds, hipStream_t s) {
  if (a.kind == kRBF)
    return a.d <= 4 ? launch_fused<T, 4, kRBF>(a, lds, s) : launch_fused<T, 8, kRBF>(a, lds, s);
  return a.d <= 4 ? launch_fused<T, 4, kLinear>(a, lds, s) : launch_fused<T, 8, kLinear>(a, lds, s);
}

}  // namespace

int pde_fused_dmax(int d) { return d <= 4 ? 4 : (d <= 8 ? 8 : 0); }

size_t pde_fused_lds(int dtype, int d, int dyadic) {
  (void)dtype;  // float32 only
  return fused_lds_bytes(pde_fused_dmax(d), dyadic);
}

size_t pde_fused_slot_bytes(int d) {
  const int dmax = pde_fused_dmax(d);
  return 16 + 4 * (size_t)(dmax <= 4 ? fused_nf<4>() : fused_nf<8>());
}

hipError_t pde_bwd_fused_launch(const PdeBwdFusedArgs& a, int dtype, size_t lds, hipStream_t s) {
  if (dtype == KSIG_F32) return dispatch_fused<float>(a, lds, s);
  return hipErrorInvalidValue;  // float32 only (the FIFO slot holds float flows)
}

}  // namespace ksig
