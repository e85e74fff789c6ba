// Truncated signature kernel by dynamic programming, forward pass with the
// static Gram on the matrix cores (gfx950).  Template implementation;
// instantiated by dp_mfma_*.hip.
//
// Same recursion and systolic wave as dp_forward_impl.h (ksig/algorithms.py:
// 47-140 on M = ΔΔ k(x_i, y_j), ksig/algorithms.py:64,
// ksig/static/kernels.py:145-155,248-260), but the static-kernel values are
// not evaluated lane by lane.  Every 16 steps the wave computes a 64-row x
// 16-column tile of <x'_r, y'_c> with MFMA, turns it into G (exp2 of the
// scaled squared distance, or the scaled dot product) and writes it into an
// LDS ring.  The ring is skewed: G(row i, ring column g) lives in slot
// (g + i) mod W, so at step T every lane reads slot T mod W at its own lane
// offset and the per-step sweep needs no address arithmetic.
//
// Ring columns g = band * Pc + c' (c' = column of G, 0..Pc-1, Pc a multiple of
// 16 and >= 64); the producer of chunk k (columns 16k..16k+15) runs at the
// start of step 16k, and W = 80 keeps chunk k-5 until lane 63 has read it.
//
// Beyond the sequence ends the inputs are replicated (rows clamped to Lx-1,
// columns to Ly-1), which makes ΔΔ exactly zero there, so the sweep needs no
// per-cell bounds masks; only the priming column c' = 0 of each row is masked.
//
// Gram precision (template MODE):
//   0: v_mfma_f32_16x16x4f32, float32 products and accumulation;
//   1: v_mfma_f32_16x16x32_bf16 with each operand split into bf16 hi + lo and
//      the three products hi*hi + hi*lo + lo*hi concatenated along K
//      (A = [x_hi, x_hi, x_lo], B = [y_hi, y_lo, y_hi]), float32 accumulation.
#pragma once

#include <type_traits>

#include "ksig_common.h"
#include "ksig_hip.h"
#include "dp_forward.h"
#include "dp_forward_impl.h"

namespace ksig {

typedef float mf4 __attribute__((ext_vector_type(4)));
typedef short mbf8 __attribute__((ext_vector_type(8)));

// float32 -> bfloat16 bits, round to nearest even (finite inputs).
__device__ __forceinline__ unsigned short bf16_rne(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f(unsigned short h) {
  return __uint_as_float((unsigned)h << 16);
}

template <int MODE, int DPAD>
struct MfmaShape {
  // number of K blocks of the MFMA instruction
  static constexpr int KB = MODE == 0 ? DPAD / 4 : (3 * DPAD + 31) / 32;
};

template <int NLEV, int ORDER, int DPAD, int MODE>
__global__ void __launch_bounds__(kWave)
dp_mfma_kernel(const DpFwdArgs a) {
  using T = float;
  constexpr int NSIDE = dp_num_side(NLEV, ORDER);
  constexpr int NF = 1 + NLEV + NSIDE;
  constexpr int NPS = NLEV > 1 ? NLEV - 1 : 1;
  constexpr int NHB = NSIDE > 0 ? NSIDE : 1;
  constexpr int W = 80;   // ring columns
  constexpr int S = 65;   // ring column stride (floats)
  constexpr int KB = MfmaShape<MODE, DPAD>::KB;
  constexpr int XS = DPAD + 1;  // staged x row: DPAD features + norm

  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* ring = reinterpret_cast<T*>(smem_raw);
  T* xst = ring + W * S;       // [64][XS]
  T* fifo = xst + kWave * XS;  // [Pc][NF]

  const int lane = threadIdx.x;
  const int col16 = lane & 15;
  const int grp = lane >> 4;
  const int64_t pr = blockIdx.x;
  double* out = a.out + pr * (NLEV + 1);
  const int d = a.d;
  const int kind = a.kind;

  const int32_t ia = a.pairs[2 * pr];
  const int32_t ib = a.pairs[2 * pr + 1];
  const T* xs = reinterpret_cast<const T*>(a.x.data) + a.x.off[ia] * d;
  int Lx = a.x.len[ia];
  const T* ys = reinterpret_cast<const T*>(a.y.data) + a.y.off[ib] * d;
  int Ly = a.y.len[ib];
  if (Ly > Lx) {
    const T* tp = xs; xs = ys; ys = tp;
    int tl = Lx; Lx = Ly; Ly = tl;
  }
  const int Nr = Lx - 1, Nc = Ly - 1;
  if (Nr <= 0 || Nc <= 0) {
    if (lane <= NLEV) out[lane] = lane == 0 ? 1.0 : 0.0;
    return;
  }
  // Pc >= 80 keeps lane 0's next FIFO slot distinct from the slot lane 63
  // writes in the same step (the sweep reads one step ahead).
  int Pc = Nc + 1 > 80 ? Nc + 1 : 80;
  Pc = (Pc + 15) & ~15;
  const int nb = (Nr + kWave - 1) / kWave;
  const int ncols = nb * Pc;        // ring columns in total
  const int nchunks = ncols >> 4;
  const T xscale = T(a.xscale), yscale = T(a.yscale);
  const bool rbf = kind == kRBF;

  // ---- operand staging ---------------------------------------------------
  // x rows of band b: grid row r = 64b + i reads x row min(r+1, Lx-1).
  T xraw[DPAD];
  auto fetch_x = [&](int b) {
    int xr = b * kWave + lane + 1;
    xr = xr < Lx - 1 ? xr : Lx - 1;
#pragma unroll
    for (int f = 0; f < DPAD; ++f) xraw[f] = f < d ? xs[(int64_t)xr * d + f] : T(0);
  };
  auto stage_x = [&]() {
    T n2 = T(0);
#pragma unroll
    for (int f = 0; f < DPAD; ++f) {
      const T v = xraw[f] * xscale;
      xst[lane * XS + f] = v;
      n2 += v * v;
    }
    xst[lane * XS + DPAD] = n2;
  };
  // A operand (4 row tiles x KB blocks) from the staged rows.
  typedef typename std::conditional<MODE == 0, float, mbf8>::type OpT;
  OpT A[4][KB];
  auto build_A = [&]() {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const T* xr = xst + (16 * t + col16) * XS;
#pragma unroll
      for (int kb = 0; kb < KB; ++kb) {
        if constexpr (MODE == 0) {
          A[t][kb] = xr[4 * kb + grp];
        } else {
          const int idx0 = 32 * kb + 8 * grp;
          const int seg = idx0 / DPAD, f0 = idx0 % DPAD;
          mbf8 v;
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const T x = seg < 3 ? xr[f0 + e] : T(0);
            const unsigned short hi = bf16_rne(x);
            const unsigned short lo = bf16_rne(x - bf16_to_f(hi));
            v[e] = (short)(seg == 2 ? lo : hi);
          }
          A[t][kb] = v;
        }
      }
    }
  };
  // y rows of chunk k: column c' = 16k mod Pc + col16 reads y row min(c', Ly-1).
  constexpr int YR = MODE == 0 ? KB : 8 * KB;  // raw floats per lane
  T yraw[YR];
  auto fetch_y = [&](int k) {
    int yc = ((k << 4) % Pc) + col16;
    yc = yc < Ly - 1 ? yc : Ly - 1;
    const T* yr = ys + (int64_t)yc * d;
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
      if constexpr (MODE == 0) {
        const int f = 4 * kb + grp;
        yraw[kb] = f < d ? yr[f] : T(0);
      } else {
        const int idx0 = 32 * kb + 8 * grp;
        const int seg = idx0 / DPAD, f0 = idx0 % DPAD;
#pragma unroll
        for (int e = 0; e < 8; ++e)
          yraw[8 * kb + e] = (seg < 3 && f0 + e < d) ? yr[f0 + e] : T(0);
      }
    }
  };
  OpT B[KB];
  T ynrm = T(0);
  auto build_B = [&]() {
    T part = T(0);
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
      if constexpr (MODE == 0) {
        const T v = yraw[kb] * yscale;
        B[kb] = v;
        part += v * v;
      } else {
        const int idx0 = 32 * kb + 8 * grp;
        const int seg = idx0 / DPAD;
        mbf8 v;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const T y = yraw[8 * kb + e] * yscale;
          const unsigned short hi = bf16_rne(y);
          const unsigned short lo = bf16_rne(y - bf16_to_f(hi));
          v[e] = (short)(seg == 1 ? lo : hi);
          if (seg == 0) part += y * y;
        }
        B[kb] = v;
      }
    }
    part += __shfl_xor(part, 16, kWave);
    part += __shfl_xor(part, 32, kWave);
    ynrm = part;
  };
  auto produce = [&](int kbase) {  // kbase = (16 k) mod W, wave-uniform
    mf4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = mf4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        if constexpr (MODE == 0)
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[t][kb], B[kb], acc[t], 0, 0, 0);
        else
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[t][kb], B[kb], acc[t], 0, 0, 0);
      }
    const int base = kbase + col16 + 4 * grp;  // < 2W
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = 16 * t + 4 * grp + q;
        const T dot = acc[t][q];
        const T e = fast_exp2(-(xst[row * XS + DPAD] + ynrm - T(2) * dot));
        const T G = rbf ? e : dot;
        int slot = base + 16 * t + q;   // (16k + col16 + row) mod W
        slot = slot >= W ? slot - W : slot;
        slot = slot >= W ? slot - W : slot;
        ring[slot * S + row] = G;
      }
  };

  // ---- prologue ------------------------------------------------------------
  fetch_x(0);
  fetch_y(0);
  // FIFO initial contents: G(0, c') for the row above the grid, zero sums.
  {
    T x0[DPAD];
    T x0n = T(0);
#pragma unroll
    for (int f = 0; f < DPAD; ++f) {
      x0[f] = f < d ? xs[f] * xscale : T(0);
      x0n += x0[f] * x0[f];
    }
    for (int s = lane; s < Pc; s += kWave) {
      const int yc = s < Ly - 1 ? s : Ly - 1;
      T dot = T(0), yn = T(0);
#pragma unroll
      for (int f = 0; f < DPAD; ++f) {
        const T yv = f < d ? ys[(int64_t)yc * d + f] * yscale : T(0);
        dot += x0[f] * yv;
        yn += yv * yv;
      }
      fifo[s * NF] = rbf ? fast_exp2(-(x0n + yn - T(2) * dot)) : dot;
#pragma unroll
      for (int k = 1; k < NF; ++k) fifo[s * NF + k] = T(0);
    }
  }
  stage_x();
  __syncthreads();
  build_A();
  if (nb > 1) fetch_x(1);

  T fl[NF];
#pragma unroll
  for (int k = 0; k < NF; ++k) fl[k] = T(0);
  T PS[NPS];
  T HB[NHB];
#pragma unroll
  for (int k = 0; k < NPS; ++k) PS[k] = T(0);
#pragma unroll
  for (int k = 0; k < NHB; ++k) HB[k] = T(0);
  T gleft = T(0), gdiag = T(0);
  int cp = (Pc - lane) % Pc;   // this lane's column c' (lane i starts i steps late)
  const int lane_off = lane;

  const int total = ncols + kWave - 1;
  int slotT = 0;               // T mod W
  int s0 = 0;                  // lane 0's FIFO slot (T mod Pc)
  int s63 = Pc - (kWave - 1);  // lane 63's FIFO slot ((T - 63) mod Pc)
  int kbase = 0;   // (16 k) mod W
  int kcol = 0;    // (16 k) mod Pc
  int kband = 0;   // 16 k / Pc
  int kchunk = 0;
  auto produce_next = [&]() {
    if (kchunk > 0 && kcol == 0) {  // producer enters a new band
      stage_x();
      __syncthreads();
      build_A();
      if (kband + 1 < nb) fetch_x(kband + 1);
    }
    build_B();
    produce(kbase);
    if (kchunk + 1 < nchunks) fetch_y(kchunk + 1);
    kbase += 16;
    kbase = kbase >= W ? kbase - W : kbase;
    kcol += 16;
    if (kcol == Pc) {
      kcol = 0;
      ++kband;
    }
    ++kchunk;
  };
  // The sweep reads its LDS inputs one step ahead: chunk k is produced at the
  // end of step 16k-1 (chunk 0 before the loop), then the ring slot and the
  // FIFO slot of step T+1 are loaded while step T computes.
  produce_next();
  T pre_g = ring[slotT * S + lane_off];
  T pre_fin[NF];
#pragma unroll
  for (int k = 0; k < NF; ++k) pre_fin[k] = fifo[s0 * NF + k];
  for (int T_ = 0; T_ < total; ++T_) {
    const T gnew = pre_g;
    T fin[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) fin[k] = pre_fin[k];
    if ((T_ & 15) == 15 && kchunk < nchunks) produce_next();
    {
      int sl = slotT + 1;
      sl = sl == W ? 0 : sl;
      int f0 = s0 + 1;
      f0 = f0 == Pc ? 0 : f0;
      pre_g = ring[sl * S + lane_off];
#pragma unroll
      for (int k = 0; k < NF; ++k) pre_fin[k] = fifo[f0 * NF + k];
    }
    // column-flowing inputs: lane 0 takes the FIFO slot of its column (all
    // lanes read the same address; only lane 0's value is kept by the shift)
    T in[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) in[k] = shift_down(fl[k], fin[k]);

    const bool keep = cp != 0;
    T M = (gnew - gleft) - (in[0] - gdiag);
    M = keep ? M : T(0);
    gleft = gnew;
    gdiag = in[0];

    T of[NF];
    of[0] = gnew;
    dp_cell<T, NLEV, ORDER>(M, in, of, PS, HB, !keep);
#pragma unroll
    for (int k = 0; k < NF; ++k) fl[k] = of[k];
    if (lane == kWave - 1 && T_ >= kWave - 1) {
#pragma unroll
      for (int k = 0; k < NF; ++k) fifo[s63 * NF + k] = of[k];
    }

    ++cp;
    cp = cp == Pc ? 0 : cp;
    ++slotT;
    slotT = slotT == W ? 0 : slotT;
    ++s0;
    s0 = s0 == Pc ? 0 : s0;
    ++s63;
    s63 = s63 == Pc ? 0 : s63;
  }
  __syncthreads();

  double acc[NLEV];
#pragma unroll
  for (int m = 0; m < NLEV; ++m) acc[m] = 0.0;
  for (int s = lane + 1; s <= Nc; s += kWave) {
#pragma unroll
    for (int m = 0; m < NLEV; ++m) acc[m] += (double)fifo[s * NF + 1 + m];
  }
#pragma unroll
  for (int m = 0; m < NLEV; ++m) acc[m] = wave_sum(acc[m]);
  if (lane == 0) out[0] = 1.0;
#pragma unroll
  for (int m = 0; m < NLEV; ++m)
    if (lane == m + 1) out[m + 1] = acc[m];
}

template <int NLEV, int ORDER, int DPAD, int MODE>
hipError_t launch_dp_mfma(const DpFwdArgs& a, size_t lds, hipStream_t s) {
  auto k = dp_mfma_kernel<NLEV, ORDER, DPAD, MODE>;
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)k,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k, dim3((unsigned)a.n_pairs), dim3(kWave), lds, s, a);
  return hipGetLastError();
}

}  // namespace ksig
