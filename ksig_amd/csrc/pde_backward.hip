// Adjoint of the signature-PDE Goursat sweep (gfx950).  No reference
// counterpart: ksig/algorithms.py:344-446 is forward-only.
//
// Forward (ksig/algorithms.py:380-385), refined cell (r, c), M' = M / 4^λ:
//   U(r,c) = (U(r-1,c) + U(r,c-1)) A(r,c) - U(r-1,c-1) B(r,c),
//   A = 1 + a, a = M'/2 + M'^2/12,  B = 1 - b, b = M'^2/12,  K = U(last cell).
// Adjoint λ = dK/dU:
//   λ(r,c) = A(r+1,c) λ(r+1,c) + A(r,c+1) λ(r,c+1) - B(r+1,c+1) λ(r+1,c+1),
// seeded with dL/dK at the last cell, evaluated (like the forward) in
// increment form, μ(r,c) = λ(r,c) - λ(r,c+1):
//   μ(r,c) = μ(r+1,c) + a(r+1,c) λ(r+1,c) + a(r,c+1) λ(r,c+1) + b(r+1,c+1) λ(r+1,c+1),
//   λ(r,c) = λ(r,c+1) + μ(r,c).
// dL/dM'(r,c) = λ(r,c) [(U(r-1,c) + U(r,c-1)) (1/2 + M'/6) + U(r-1,c-1) M'/6],
// summed over the 2^λ x 2^λ refined cells of an original cell and divided by
// 4^λ gives dL/dM of the original cell.
//
// The sweep runs over the flipped grid (bottom-up, right-to-left) with the
// forward's systolic structure: λ, μ and the products a·λ, b·λ flow from the
// lane of row r+1 to the lane of row r through DPP wave_shr:1 and an LDS FIFO
// between 64-row bands.  The forward's U values and cell values are read from
// the dump written by the forward kernel for the same pairs.

#include "ksig_common.h"
#include "ksig_hip.h"
#include "dp_forward.h"

namespace ksig {

template <typename T>
__global__ void __launch_bounds__(kWave)
pde_backward_kernel(const PdeBwdArgs a) {
  constexpr int NF = 4;  // [λ, μ, a λ, b λ]
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* fifo = reinterpret_cast<T*>(smem_raw);

  const int lane = threadIdx.x;
  const int64_t pr = blockIdx.x;
  const int32_t ia = a.pairs[2 * pr];
  const int32_t ib = a.pairs[2 * pr + 1];
  int Lx = a.x_len[ia], Ly = a.y_len[ib];
  if (Ly > Lx) {  // orientation of the forward (rows = longer sequence)
    int t = Lx; Lx = Ly; Ly = t;
  }
  const int Nr = a.difference ? Lx - 1 : Lx;
  const int Nc = a.difference ? Ly - 1 : Ly;
  if (Nr <= 0 || Nc <= 0) return;
  const int lam = a.dyadic;
  const int F = 1 << lam;
  const int Nrr = Nr << lam, Ncr = Nc << lam;
  const int Pc = Ncr > kWave ? Ncr : kWave;
  const int nb = (Nrr + kWave - 1) / kWave;
  const T inv4 = T(1) / T(F * F);
  const T* U = reinterpret_cast<const T*>(a.u_all) + a.u_off[pr];
  const T* Mo = reinterpret_cast<const T*>(a.m_all) + a.m_off[pr];
  T* G = reinterpret_cast<T*>(a.gm) + a.m_off[pr];
  const T w = T(a.gout[pr]);

  for (int s = lane; s < Pc * NF; s += kWave) fifo[s] = T(0);
  __syncthreads();

  T fl[NF] = {T(0), T(0), T(0), T(0)};
  T lam_r = T(0), a_r = T(0), beta_prev = T(0), gacc = T(0);
  int s = -lane;
  int band = 0;
  int fr = lane;
  const int total = nb * Pc + kWave - 1;
  for (int step = 0; step < total; ++step) {
    const bool active = s >= 0 && band < nb;
    T fin[NF] = {T(0), T(0), T(0), T(0)};
    if (lane == 0 && active) {
#pragma unroll
      for (int k = 0; k < NF; ++k) fin[k] = fifo[s * NF + k];
    }
    T in[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) in[k] = shift_down(fl[k], fin[k]);

    const int r = Nrr - 1 - fr;
    const int c = Ncr - 1 - s;
    const bool cell = active && r >= 0 && c >= 0;
    const int64_t mi = (int64_t)(r >> lam) * Nc + (c >> lam);
    const T m = cell ? Mo[mi] : T(0);
    const T bb = m * m * T(1.0 / 12.0);
    const T aa = T(0.5) * m + bb;
    const T eps = in[2] + a_r * lam_r + beta_prev;
    const T seed = (cell && r == Nrr - 1 && c == Ncr - 1) ? w : T(0);
    T mu = in[1] + eps + seed;
    T lv = lam_r + mu;
    mu = cell ? mu : in[1];
    lv = cell ? lv : T(0);
    if (cell) {
      const T uup = r > 0 ? U[(int64_t)(r - 1) * Ncr + c] : T(1);
      const T ulf = c > 0 ? U[(int64_t)r * Ncr + c - 1] : T(1);
      const T udg = (r > 0 && c > 0) ? U[(int64_t)(r - 1) * Ncr + c - 1] : T(1);
      gacc += lv * ((uup + ulf) * (T(0.5) + m * T(1.0 / 6.0)) + udg * m * T(1.0 / 6.0));
      if ((c & (F - 1)) == 0) {  // leaving this original column (reverse order)
        atomicAdd(&G[mi], gacc * inv4);
        gacc = T(0);
      }
    }
    T of[NF] = {lv, mu, cell ? aa * lv : T(0), cell ? bb * lv : T(0)};
#pragma unroll
    for (int k = 0; k < NF; ++k) fl[k] = of[k];
    if (lane == kWave - 1 && active) {
#pragma unroll
      for (int k = 0; k < NF; ++k) fifo[s * NF + k] = of[k];
    }
    lam_r = lv;
    a_r = cell ? aa : T(0);
    beta_prev = in[3];

    ++s;
    if (s == Pc) {
      s = 0;
      ++band;
      fr += kWave;
      lam_r = a_r = beta_prev = gacc = T(0);
    }
  }
}

template <typename T>
static hipError_t launch_pde_bwd(const PdeBwdArgs& a, size_t lds, hipStream_t s) {
  auto k = pde_backward_kernel<T>;
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)k,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k, dim3((unsigned)a.n_pairs), dim3(kWave), lds, s, a);
  return hipGetLastError();
}

hipError_t pde_backward_launch(const PdeBwdArgs& a, int dtype, size_t lds, hipStream_t s) {
  if (dtype == KSIG_F32) return launch_pde_bwd<float>(a, lds, s);
  return launch_pde_bwd<double>(a, lds, s);
}

}  // namespace ksig
