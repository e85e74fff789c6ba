// Truncated signature kernel by dynamic programming for ANY (n_levels, order):
// the forward recursion with the level count and the embedding order as
// runtime values, on caller-supplied static-kernel matrices (plain, or dot
// products with the RBF transform applied on read; see DpFwdArgs::gram_*).
//
// Recursion: ksig/algorithms.py:47-85 (first order) and :88-140 (higher
// order) on M = ΔΔ k(x_i, y_j) (ksig/algorithms.py:64).  This is the kernel
// for (n_levels, order) combinations outside the compiled table of the
// register-resident kernels (dp_forward_impl.h, dp_stream.h), which carry
// every level of the state in registers and so need both values at compile
// time.  Here the per-lane state lives in LDS instead, laid out [k][lane]
// (lane-contiguous: conflict-free), and each cell walks the levels in a
// runtime loop, keeping only the run-length R-matrices of the current and the
// next level (2 * order^2 values per lane) — the level-by-level form of
// ksig/algorithms.py:115-132.
//
// Wave structure as in dp_forward_kernel: one 64-lane wave per matrix, lane i
// owns grid row 64*band + i and walks the columns one step behind lane i-1;
// lane i takes its inputs from the outputs lane i-1 left in LDS the step
// before (lane 0: the FIFO slot of the column), the column sums of every
// level flow down the rows and through an LDS FIFO from lane 63 to lane 0 of
// the next band; after the last band the FIFO holds the final column sums,
// whose FP64 totals are the levels.
#include "ksig_common.h"
#include "ksig_hip.h"
#include "dp_forward.h"
#include "dp_forward_impl.h"

namespace ksig {

// Per-lane LDS words of the generic kernel's state (flows and inputs, row
// prefixes, row side sums, two R-matrix buffers) for (n_levels, order).
__host__ __device__ inline int dp_generic_lane_words(int nlev, int order) {
  const int nf = dp_flow_count_hd(nlev, order);
  const int nps = nlev > 1 ? nlev - 1 : 1;
  const int nhb = dp_num_side(nlev, order) > 0 ? dp_num_side(nlev, order) : 1;
  return 2 * nf + nps + nhb + 2 * order * order;
}

template <typename T>
__global__ void __launch_bounds__(kWave)
dp_generic_kernel(const DpFwdArgs a) {
  const int NLEV = a.n_levels;
  const int D = a.order;
  const int NSIDE = dp_num_side(NLEV, D);
  const int NF = 1 + NLEV + NSIDE;
  const int NPS = NLEV > 1 ? NLEV - 1 : 1;
  const int NHB = NSIDE > 0 ? NSIDE : 1;

  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* smem = reinterpret_cast<T*>(smem_raw);
  const int lane = threadIdx.x;
  // [k][lane] arrays
  T* flow = smem;                    // [NF][64] this lane's outputs of the last step
  T* inb = flow + NF * kWave;        // [NF][64] this step's inputs
  T* PS = inb + NF * kWave;          // [NPS][64] row prefixes of levels 1..NLEV-1
  T* HB = PS + NPS * kWave;          // [NHB][64] row side sums (order >= 2)
  T* Ra = HB + NHB * kWave;          // [D*D][64] R-matrix of the current level
  T* Rb = Ra + D * D * kWave;        // [D*D][64] R-matrix of the next level
  T* fifo = Rb + D * D * kWave;      // [Pc][NF]
#define LN(arr, k) arr[(k) * kWave + lane]

  const int64_t pr = blockIdx.x;
  double* out = a.out + pr * (NLEV + 1);
  const int Lx = a.gram_len ? a.gram_len[2 * pr] : a.gram_lx;
  const int Ly = a.gram_len ? a.gram_len[2 * pr + 1] : a.gram_ly;
  const T* gm = reinterpret_cast<const T*>(a.gram) +
                (a.gram_off ? a.gram_off[pr] : pr * (int64_t)a.gram_lx * a.gram_ly);
  const int64_t gld = a.gram_ld > 0 ? a.gram_ld : (int64_t)Ly;
  const T* ghx = nullptr;
  const T* ghy = nullptr;
  if (a.gram_tf == kGramRbfDots) {
    ghx = reinterpret_cast<const T*>(a.gram_hx) + a.gram_hx_off[pr];
    ghy = reinterpret_cast<const T*>(a.gram_hy) + a.gram_hy_off[pr];
  }
  const int diff = a.difference;
  const int Nr = diff ? Lx - 1 : Lx;
  const int Nc = diff ? Ly - 1 : Ly;
  if (Nr <= 0 || Nc <= 0) {  // empty grid: K_0 = 1, K_m = 0
    for (int m = lane; m <= NLEV; m += kWave) out[m] = m == 0 ? 1.0 : 0.0;
    return;
  }
  const int Pc = Nc + 1 > kWave ? Nc + 1 : kWave;
  const int nb = (Nr + kWave - 1) / kWave;
  const int rshift = diff ? 1 : 0;

  auto gram_at = [&](int rr, int cc) -> T {
    const int xr = rr + rshift, yc = cc + rshift;
    if (xr < 0 || xr >= Lx || yc < 0 || yc >= Ly) return T(0);
    const T v = gm[(int64_t)xr * gld + yc];
    return ghx ? gram_rbf(v, ghx[xr], ghy[yc]) : v;
  };

  // FIFO initial contents: the row above the grid, G(0, c+1), and zero sums
  // (slot s holds column s - 1); lane state starts at zero.
  for (int s = lane; s < Pc; s += kWave) {
    T g = T(0);
    if (diff && s < Ly) {
      g = gm[s];
      if (ghx) g = gram_rbf(g, ghx[0], ghy[s]);
    }
    fifo[s * NF] = g;
    for (int k = 1; k < NF; ++k) fifo[s * NF + k] = T(0);
  }
  for (int k = 0; k < NF; ++k) LN(flow, k) = T(0);
  for (int k = 0; k < NPS; ++k) LN(PS, k) = T(0);
  for (int k = 0; k < NHB; ++k) LN(HB, k) = T(0);
  __syncthreads();

  T gleft = T(0), gdiag = T(0);
  int c = -1 - lane;
  int band = 0;
  int row = lane;
  const int total = nb * Pc + kWave - 1;
  for (int step = 0; step < total; ++step) {
    const bool active = c >= -1 && band < nb;
    // inputs: the outputs of lane-1 from the last step, or the FIFO slot
    for (int k = 0; k < NF; ++k) {
      T v;
      if (lane == 0)
        v = active ? fifo[(c + 1) * NF + k] : T(0);
      else
        v = flow[k * kWave + lane - 1];
      LN(inb, k) = v;
    }
    __syncthreads();

    const T gnew = active ? gram_at(row, c) : T(0);
    const bool cell = active && c >= 0 && c < Nc && row < Nr;
    const T in0 = LN(inb, 0);
    T M = diff ? (gnew - gleft) - (in0 - gdiag) : gnew;
    M = cell ? M : T(0);
    gleft = gnew;
    gdiag = in0;
    const bool rowstart = c == -1;
    LN(flow, 0) = gnew;

    T* Rc = Ra;
    T* Rn = Rb;
    LN(Rc, 0) = M;  // R_1 = M (one run-length state)
    for (int m = 1; m <= NLEV; ++m) {
      const int dm = m < D ? m : D;
      T S = T(0);
      for (int r = 0; r < dm; ++r)
        for (int s = 0; s < dm; ++s) S += LN(Rc, r * D + s);
      LN(flow, m) = LN(inb, m) + S;
      if (m < NLEV) {
        const int dn = m + 1 < D ? m + 1 : D;
        const int so = dp_side_off(m, D);
        // R_{m+1} from the old prefixes / side sums and R_m
        // (ksig/algorithms.py:120-131)
        LN(Rn, 0) = M * LN(PS, m - 1);
        for (int s = 0; s < dn - 1; ++s)
          LN(Rn, s + 1) = M * LN(inb, 1 + NLEV + so + s) / T(s + 2);
        for (int r = 0; r < dn - 1; ++r)
          LN(Rn, (r + 1) * D) = M * LN(HB, so + r) / T(r + 2);
        for (int r = 1; r < dn; ++r)
          for (int s = 1; s < dn; ++s)
            LN(Rn, r * D + s) = M * LN(Rc, (r - 1) * D + s - 1) / T((r + 1) * (s + 1));
        // level-m updates: row prefix, column side flows, row side sums
        const T ps = LN(PS, m - 1) + LN(inb, m);
        LN(PS, m - 1) = rowstart ? T(0) : ps;
        for (int s = 0; s < dn - 1; ++s) {
          T A = T(0);
          for (int q = 0; q < dm; ++q) A += LN(Rc, q * D + s);
          LN(flow, 1 + NLEV + so + s) = LN(inb, 1 + NLEV + so + s) + A;
        }
        for (int r = 0; r < dn - 1; ++r) {
          T B = T(0);
          for (int q = 0; q < dm; ++q) B += LN(Rc, r * D + q);
          const T hb = LN(HB, so + r) + B;
          LN(HB, so + r) = rowstart ? T(0) : hb;
        }
        T* t = Rc;
        Rc = Rn;
        Rn = t;
      }
    }
    if (lane == kWave - 1 && active)
      for (int k = 0; k < NF; ++k) fifo[(c + 1) * NF + k] = LN(flow, k);
    __syncthreads();

    ++c;
    if (c == Pc - 1) {
      c = -1;
      ++band;
      row += kWave;
    }
  }

  // K_m = sum over columns of the final column sums (slots 1..Nc)
  if (lane == 0) out[0] = 1.0;
  for (int m = 0; m < NLEV; ++m) {
    double acc = 0.0;
    for (int s = lane + 1; s <= Nc; s += kWave) acc += (double)fifo[s * NF + 1 + m];
    acc = wave_sum(acc);
    if (lane == 0) out[m + 1] = acc;
  }
#undef LN
}

size_t dp_generic_lds(int dtype, int n_levels, int order, int max_cols) {
  const size_t esz = dtype == KSIG_F32 ? 4 : 8;
  const int nf = dp_flow_count_hd(n_levels, order);
  const int pc = max_cols + 1 > kWave ? max_cols + 1 : kWave;
  return esz * ((size_t)kWave * dp_generic_lane_words(n_levels, order) + (size_t)pc * nf);
}

hipError_t dp_generic_launch(const DpFwdArgs& a, int dtype, size_t lds, hipStream_t s) {
  const void* k = dtype == KSIG_F32 ? (const void*)dp_generic_kernel<float>
                                    : (const void*)dp_generic_kernel<double>;
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  if (dtype == KSIG_F32)
    hipLaunchKernelGGL(dp_generic_kernel<float>, dim3((unsigned)a.n_pairs), dim3(kWave), lds, s,
                       a);
  else
    hipLaunchKernelGGL(dp_generic_kernel<double>, dim3((unsigned)a.n_pairs), dim3(kWave), lds,
                       s, a);
  return hipGetLastError();
}

}  // namespace ksig
