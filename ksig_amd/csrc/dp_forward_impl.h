// Truncated signature kernel by dynamic programming, forward pass (gfx950).
// Template implementation; instantiated by dp_fwd_*.hip.
//
// Computes, for every requested sequence pair (x, y), the level values
//   K_0 = 1,  K_m = sum_{i,j} R_m[i,j]  (m = 1..NLEV)
// of the reference recursion (ksig/algorithms.py:47-85 first order,
// :88-140 higher order) on M = ΔΔ k(x_i, y_j) (ksig/algorithms.py:64) with
// k the RBF (ksig/static/kernels.py:248-260) or linear
// (ksig/static/kernels.py:145-155) static kernel, or on a caller-supplied
// Gram/M tensor (the functional ksig.algorithms.signature_kern entry).
//
// Dataflow (one 64-lane wave per pair, no L×L tensor ever touches HBM):
//   * lane i owns grid row r = 64*band + i, step t walks column c = t - i - 1
//     (c = -1 is a priming column that evaluates G(r+1, 0));
//   * each lane evaluates ONE static-kernel value per cell, G(r+1, c+1); the
//     other three corners of the ΔΔ stencil are its own previous value and
//     the values lane i-1 produced one step earlier (DPP wave_shr:1);
//   * the exclusive 2-D cumulative sum E(R) is split into a column part
//     C (sum over rows above, flows down lane to lane) and a row part
//     P (running sum of C along the row, lane-local), so every level of the
//     recursion advances in the same step;
//   * lane 63's outgoing column values go to an LDS FIFO slot per column and
//     come back into lane 0 when it starts the next band; after the last band
//     the FIFO holds the full column sums, whose FP64 total is K_m.
// Arithmetic is T (float or double); level totals are reduced in double.
#pragma once

#include "ksig_common.h"
#include "ksig_hip.h"
#include "dp_forward.h"

namespace ksig {

// Side sums of the higher-order recursion: for level m < NLEV the next level
// has dn = min(m+1, ORDER) and needs dn-1 column sums of sum_a R_m[a][s]
// (E_i terms) and dn-1 row sums of sum_b R_m[r][b] (E_j terms).
__host__ __device__ constexpr int dp_num_side(int nlev, int order) {
  int n = 0;
  for (int m = 1; m < nlev; ++m) n += (m + 1 < order ? m + 1 : order) - 1;
  return n;
}
__host__ __device__ constexpr int dp_side_off(int m, int order) {
  int n = 0;
  for (int k = 1; k < m; ++k) n += (k + 1 < order ? k + 1 : order) - 1;
  return n;
}
__host__ __device__ constexpr int cmin(int a, int b) { return a < b ? a : b; }
// Column-flowing values per lane: the static-kernel value, the level column
// sums and the column side sums.
__host__ __device__ constexpr int dp_flow_count_hd(int nlev, int order) {
  return 1 + nlev + dp_num_side(nlev, order);
}

// RBF static-kernel value from a stored scaled dot product v = 2h<x, y> and
// the norm terms hx = h|x|^2, hy = h|y|^2 (h = log2(e) / max(2 sigma^2, eps)).
template <typename T>
__device__ __forceinline__ T gram_rbf(T v, T hx, T hy) {
  return fast_exp2(v - hx - hy);
}

__device__ __forceinline__ float dp_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double dp_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }

// One cell of the level recursion in factored form.  `in` holds the
// column-flowing inputs (index 1..NLEV: column sums of S_m above this cell;
// 1+NLEV+..: column sums of the E_i side terms), PS / HB the lane-local row
// sums (prefix of S_m along the row; the E_j side terms).  Writes the
// outgoing column values into out[1..] and updates PS / HB through
// upd(old, increment) (the row-start reset is the caller's).
//
// Every term of level m' = m + 1 is R[m'][r][s] = M / ((r+1)(s+1)) * src with
//   src = PS_m (r = s = 0), column side sum E_i (r = 0), row side sum E_j
//   (s = 0), or R[m][r-1][s-1] (ksig/algorithms.py:121-129),
// so a cell forms the few scaled copies of M once, materialises only the
// terms the next level reads (r, s < min(m'+1, ORDER) - 1), and folds every
// other term into the sums that consume it as one fused multiply-add:
// 9 operations per level at order 2 instead of a product per term plus a
// product per constant plus the additions.
template <typename T, int NLEV, int ORDER, typename Upd>
__device__ __forceinline__ void dp_cell_core(const T M, const T* in, T* out, T* PS, T* HB,
                                             Upd upd) {
  constexpr int NPS = NLEV > 1 ? NLEV - 1 : 1;
  constexpr int NSIDE = dp_num_side(NLEV, ORDER);
  constexpr int NHB = NSIDE > 0 ? NSIDE : 1;
  T ps[NPS], hb[NHB];  // row sums before this cell
#pragma unroll
  for (int k = 0; k < NPS; ++k) ps[k] = NLEV > 1 ? PS[k] : T(0);
#pragma unroll
  for (int k = 0; k < NHB; ++k) hb[k] = NSIDE > 0 ? HB[k] : T(0);
  const int kNoMat = ORDER > 1 ? ORDER - 1 : 1;
  T mat[NLEV + 1][kNoMat][kNoMat];  // materialised R[m][r][s] read by level m + 1
#pragma unroll
  for (int m = 1; m <= NLEV; ++m) {
    const int dm = cmin(m, ORDER);
    const int nnext = m < NLEV ? cmin(m + 1, ORDER) - 1 : 0;  // side sums / materialised extent
    auto coef = [&](int r, int s) -> T {
      return (r == 0 && s == 0) ? M : M * T(1.0 / ((r + 1) * (s + 1)));
    };
    auto src = [&](int r, int s) -> T {  // m >= 2
      if (r == 0 && s == 0) return ps[m - 2];
      if (r == 0) return in[1 + NLEV + dp_side_off(m - 1, ORDER) + s - 1];
      if (s == 0) return hb[dp_side_off(m - 1, ORDER) + r - 1];
      return mat[m - 1][r - 1][s - 1];
    };
    auto is_mat = [&](int r, int s) { return r < nnext && s < nnext; };
#pragma unroll
    for (int r = 0; r < nnext; ++r)
#pragma unroll
      for (int s = 0; s < nnext; ++s) mat[m][r][s] = m == 1 ? M : coef(r, s) * src(r, s);
    auto first = [&](int r, int s) -> T {
      if (is_mat(r, s)) return mat[m][r][s];
      return m == 1 ? M : coef(r, s) * src(r, s);
    };
    auto add = [&](T acc, int r, int s) -> T {
      if (is_mat(r, s)) return acc + mat[m][r][s];
      return m == 1 ? acc + M : dp_fma(coef(r, s), src(r, s), acc);
    };
    // side sums of this level: A_s = sum_q R[m][q][s], B_r = sum_q R[m][r][q]
    T A[ORDER], B[ORDER];
#pragma unroll
    for (int s = 0; s < nnext; ++s) {
      A[s] = first(0, s);
#pragma unroll
      for (int q = 1; q < dm; ++q) A[s] = add(A[s], q, s);
    }
#pragma unroll
    for (int r = 0; r < nnext; ++r) {
      B[r] = first(r, 0);
#pragma unroll
      for (int q = 1; q < dm; ++q) B[r] = add(B[r], r, q);
    }
    // level total: the column side sums plus the remaining columns
    T o = in[m];
#pragma unroll
    for (int s = 0; s < nnext; ++s) o = o + A[s];
#pragma unroll
    for (int s = nnext; s < dm; ++s)
#pragma unroll
      for (int q = 0; q < dm; ++q) o = add(o, q, s);
    out[m] = o;
    if (m < NLEV) {
      PS[m - 1] = upd(ps[m - 1], in[m]);
#pragma unroll
      for (int s = 0; s < nnext; ++s) {
        const int k = 1 + NLEV + dp_side_off(m, ORDER) + s;
        out[k] = in[k] + A[s];
      }
#pragma unroll
      for (int r = 0; r < nnext; ++r) {
        const int k = dp_side_off(m, ORDER) + r;
        HB[k] = upd(hb[k], B[r]);
      }
    }
  }
}

// One cell with the row sums reset to zero at the start of a row.
template <typename T, int NLEV, int ORDER>
__device__ __forceinline__ void dp_cell(const T M, const T* in, T* out, T* PS, T* HB,
                                        const bool rowstart) {
  dp_cell_core<T, NLEV, ORDER>(M, in, out, PS, HB, [&](T old, T inc) {
    const T v = old + inc;
    return rowstart ? T(0) : v;
  });
}

template <typename T, int NLEV, int ORDER, int DMAX, bool GRAM>
__global__ void __launch_bounds__(kWave)
dp_forward_kernel(const DpFwdArgs a) {
  constexpr int NSIDE = dp_num_side(NLEV, ORDER);
  constexpr int NF = 1 + NLEV + NSIDE;  // [G, CS_1..CS_NLEV, CA...]
  constexpr int NPS = NLEV > 1 ? NLEV - 1 : 1;
  constexpr int NHB = NSIDE > 0 ? NSIDE : 1;

  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* smem = reinterpret_cast<T*>(smem_raw);

  const int lane = threadIdx.x;
  const int64_t pr = blockIdx.x;
  double* out = a.out + pr * (NLEV + 1);
  const int d = a.d;
  const int kind = a.kind;

  const T* xs = nullptr;
  const T* ys = nullptr;
  const T* gm = nullptr;  // caller-supplied Gram / M for this pair
  const T* ghx = nullptr;  // RBF-transformed values: row / column norm terms
  const T* ghy = nullptr;
  int64_t gld = 0;
  int Lx, Ly;
  if (GRAM) {
    Lx = a.gram_len ? a.gram_len[2 * pr] : a.gram_lx;
    Ly = a.gram_len ? a.gram_len[2 * pr + 1] : a.gram_ly;
    gm = reinterpret_cast<const T*>(a.gram) +
         (a.gram_off ? a.gram_off[pr] : pr * (int64_t)a.gram_lx * a.gram_ly);
    gld = a.gram_ld > 0 ? a.gram_ld : (int64_t)Ly;
    if (a.gram_tf == kGramRbfDots) {
      ghx = reinterpret_cast<const T*>(a.gram_hx) + a.gram_hx_off[pr];
      ghy = reinterpret_cast<const T*>(a.gram_hy) + a.gram_hy_off[pr];
    }
  } else {
    const int32_t ia = a.pairs[2 * pr];
    const int32_t ib = a.pairs[2 * pr + 1];
    xs = reinterpret_cast<const T*>(a.x.data) + a.x.off[ia] * d;
    Lx = a.x.len[ia];
    ys = reinterpret_cast<const T*>(a.y.data) + a.y.off[ib] * d;
    Ly = a.y.len[ib];
    if (Ly > Lx) {  // keep the column (FIFO/LDS) side the shorter one
      const T* tp = xs; xs = ys; ys = tp;
      int tl = Lx; Lx = Ly; Ly = tl;
    }
  }
  const int diff = a.difference;
  const int Nr = diff ? Lx - 1 : Lx;
  const int Nc = diff ? Ly - 1 : Ly;
  if (Nr <= 0 || Nc <= 0) {  // empty grid: K_0 = 1, K_m = 0
    if (lane <= NLEV) out[lane] = lane == 0 ? 1.0 : 0.0;
    return;
  }
  const int Pc = Nc + 1 > kWave ? Nc + 1 : kWave;  // columns per band incl. priming
  const int nb = (Nr + kWave - 1) / kWave;
  const int rshift = diff ? 1 : 0;  // grid row r reads x row r+1 (ΔΔ) or r

  // Static-kernel scaling: RBF G = exp2(-(|x'|^2 + |y'|^2 - 2<x',y'>)) with
  // x' = x*sqrt(log2(e)/max(2 sigma^2, eps)); linear G = <scale*x, y>.
  const T xscale = T(a.xscale);
  const T yscale = T(a.yscale);

  // LDS carve (fused kinds): x rows [2][64][DMAX], x norms [2][64],
  // y rows [Ly][DMAX], y norms [Ly]; then the FIFO [Pc][NF].
  T* xl = smem;
  T* xnl = xl + (GRAM ? 0 : 2 * kWave * DMAX);
  T* yl = xnl + (GRAM ? 0 : 2 * kWave);
  T* ynl = yl + (GRAM ? 0 : Ly * DMAX);
  T* fifo = ynl + (GRAM ? 0 : Ly);

  // Stage grid row `r` of x (x row r + rshift) for this lane into buffer b.
  auto stage_x = [&](int b, int r) {
    const int xr = r + rshift;
    const bool ok = xr < Lx;
    T n2 = T(0);
#pragma unroll
    for (int k = 0; k < DMAX; ++k) {
      const T v = (ok && k < d) ? xs[(int64_t)xr * d + k] * xscale : T(0);
      xl[(b * kWave + lane) * DMAX + k] = v;
      n2 += v * v;
    }
    xnl[b * kWave + lane] = n2;
  };
  auto gfused = [&](int b, int yc) -> T {
    const T* xr = xl + (b * kWave + lane) * DMAX;
    const T* yr = yl + yc * DMAX;
    T dot = T(0);
#pragma unroll
    for (int k = 0; k < DMAX; ++k) dot += xr[k] * yr[k];
    return kind == kRBF ? fast_exp2(-(xnl[b * kWave + lane] + ynl[yc] - T(2) * dot)) : dot;
  };

  if (!GRAM) {
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
    stage_x(0, lane);
    stage_x(1, kWave + lane);
  }
  __syncthreads();

  // FIFO initial contents = the row above the grid: G(0, c+1) and zero sums.
  // Slot s holds column c = s - 1.
  {
    T x0[DMAX];
    T x0n = T(0);
    if (!GRAM) {
#pragma unroll
      for (int k = 0; k < DMAX; ++k) {
        x0[k] = k < d ? xs[k] * xscale : T(0);
        x0n += x0[k] * x0[k];
      }
    }
    for (int s = lane; s < Pc; s += kWave) {
      T g = T(0);
      if (diff && s < Ly) {
        if (GRAM) {
          g = gm[s];
          if (ghx) g = gram_rbf(g, ghx[0], ghy[s]);
        } else {
          T dot = T(0);
#pragma unroll
          for (int k = 0; k < DMAX; ++k) dot += x0[k] * yl[s * DMAX + k];
          g = kind == kRBF ? fast_exp2(-(x0n + ynl[s] - T(2) * dot)) : dot;
        }
      }
      fifo[s * NF + 0] = g;
#pragma unroll
      for (int k = 1; k < NF; ++k) fifo[s * NF + k] = T(0);
    }
  }
  __syncthreads();

  T fl[NF];  // this lane's outputs from the previous step
#pragma unroll
  for (int k = 0; k < NF; ++k) fl[k] = T(0);
  T PS[NPS];
  T HB[NHB];
#pragma unroll
  for (int k = 0; k < NPS; ++k) PS[k] = T(0);
#pragma unroll
  for (int k = 0; k < NHB; ++k) HB[k] = T(0);
  T gleft = T(0), gdiag = T(0);

  int c = -1 - lane;  // this lane's column at step 0 (band 0)
  int band = 0;
  int row = lane;     // grid row = 64*band + lane

  // Caller-supplied Gram: next step's value is fetched one step ahead.
  auto gram_at = [&](int rr, int cc) -> T {
    const int xr = rr + rshift, yc = cc + rshift;
    if (xr < 0 || xr >= Lx || yc < 0 || yc >= Ly) return T(0);
    const T v = gm[(int64_t)xr * gld + yc];
    return ghx ? gram_rbf(v, ghx[xr], ghy[yc]) : v;
  };
  T gpre = T(0);
  if (GRAM) gpre = gram_at(row, c);
  T* save = (!GRAM && a.save != nullptr)
                ? reinterpret_cast<T*>(a.save) + a.save_off[pr] : nullptr;

  const int total = nb * Pc + kWave - 1;
  int next_stage = Pc + kWave;  // step at which band b+2 rows replace band b's
  int stage_band = 2;
  for (int step = 0; step < total; ++step) {
    if (!GRAM && step == next_stage) {
      // wave-uniform point: every lane has left band stage_band-2, whose
      // buffer now receives the rows of band stage_band.
      stage_x(stage_band & 1, stage_band * kWave + lane);
      next_stage += Pc;
      ++stage_band;
    }
    const bool active = c >= -1 && band < nb;

    // Inputs from the lane above (lane 0: the FIFO slot of this column).
    T fin[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) fin[k] = T(0);
    if (lane == 0 && active) {
#pragma unroll
      for (int k = 0; k < NF; ++k) fin[k] = fifo[(c + 1) * NF + k];
    }
    T in[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) in[k] = shift_down(fl[k], fin[k]);

    // Static-kernel value G(r+1, c+1) (or G(r, c) without differencing).
    T gnew;
    if (GRAM) {
      gnew = gpre;
      int cn = c + 1, rn = row;
      if (cn == Pc - 1) { cn = -1; rn = row + kWave; }
      gpre = gram_at(rn, cn);
    } else {
      int yc = c + rshift;
      yc = yc < 0 ? 0 : (yc >= Ly ? Ly - 1 : yc);
      gnew = gfused(band & 1, yc);
    }

    const bool cell = active && c >= 0 && c < Nc && row < Nr;
    T M = diff ? (gnew - gleft) - (in[0] - gdiag) : gnew;
    M = cell ? M : T(0);
    gleft = gnew;
    gdiag = in[0];

    T of[NF];
    of[0] = gnew;
    dp_cell<T, NLEV, ORDER>(M, in, of, PS, HB, c == -1);
#pragma unroll
    for (int k = 0; k < NF; ++k) fl[k] = of[k];
    if (ORDER == 1 && NLEV > 1 && save != nullptr && active && c == Nc - 1 && row < Nr) {
      // PS now holds the row-end prefixes P_a(row, Nc) (a = 1..NLEV-1)
#pragma unroll
      for (int m = 0; m < NPS; ++m) save[m * Nr + row] = PS[m];
    }

    if (lane == kWave - 1 && active) {
#pragma unroll
      for (int k = 0; k < NF; ++k) fifo[(c + 1) * NF + k] = of[k];
    }

    // advance this lane to its next column
    ++c;
    if (c == Pc - 1) {
      c = -1;
      ++band;
      row += kWave;
    }
  }
  __syncthreads();

  // Level totals: K_m = sum over columns of the final column sums.
  double acc[NLEV];
#pragma unroll
  for (int m = 0; m < NLEV; ++m) acc[m] = 0.0;
  for (int s = lane + 1; s <= Nc; s += kWave) {
#pragma unroll
    for (int m = 0; m < NLEV; ++m) acc[m] += (double)fifo[s * NF + 1 + m];
  }
#pragma unroll
  for (int m = 0; m < NLEV; ++m) acc[m] = wave_sum(acc[m]);
  if (ORDER == 1 && NLEV > 1 && save != nullptr) {
    // final column sums C_a(Nr, j), a = 1..NLEV-1
    T* cs = save + (NLEV - 1) * Nr;
    for (int s = lane + 1; s <= Nc; s += kWave) {
#pragma unroll
      for (int m = 0; m < NLEV - 1; ++m) cs[m * Nc + s - 1] = fifo[s * NF + 1 + m];
    }
  }
  if (lane == 0) out[0] = 1.0;
#pragma unroll
  for (int m = 0; m < NLEV; ++m)
    if (lane == m + 1) out[m + 1] = acc[m];
}

template <typename T, int NLEV, int ORDER, int DMAX, bool GRAM>
hipError_t launch_dp(const DpFwdArgs& a, size_t lds, hipStream_t s) {
  auto k = dp_forward_kernel<T, NLEV, ORDER, DMAX, GRAM>;
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)k,
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k, dim3((unsigned)a.n_pairs), dim3(kWave), lds, s, a);
  return hipGetLastError();
}

}  // namespace ksig
