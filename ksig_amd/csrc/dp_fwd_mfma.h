// First-order DP signature kernel, forward, float32, one pair per wave with
// the static Gram on the matrix cores (gfx950).  Template implementation;
// instantiated by dp_fwd_mfma_*.hip.
//
// The recursion is ksig/algorithms.py:47-85 on M = ΔΔ k(x_i, y_j)
// (ksig/algorithms.py:64, ksig/static/kernels.py:145-155, 248-260), one cell
// per lane and step (dp_cell of dp_forward_impl.h): lane L of band b owns grid
// row r = 64b + L and walks the columns c' = 0 (priming) .. Pc-1 left to
// right, one step behind lane L-1; the column sums flow down the wave by DPP
// and lane 63 hands them to lane 0 of the next band through a FIFO in a
// global-memory workspace (one region per pair).
//
// Static-kernel values: every 16 steps the wave computes a 64-row x 16-column
// tile of <x', y'> on v_mfma_f32_16x16x4f32 (lane (c16, grp) holds features
// grp*KB .. grp*KB+KB-1 of x row 16t + c16 and of y column c16, KB = DMAX/4)
// and writes G into a skewed LDS ring: ring column g = b * Pc + c' of row L
// sits in slot (g + L) mod 80, so at step t every lane reads slot t mod 80.
// Pc is a multiple of 16 (a chunk never straddles a band) and at least 80.
// Chunk k is produced at step 16k - 2 and every ring / FIFO input is read two
// steps before its use into one of two alternating register sets; lane 0's
// FIFO slot is read by every lane (a broadcast) and was written by lane 63
// Pc - 65 >= 15 steps earlier.  LDS holds only the ring (20 KB: 8 waves per
// CU); the x norms of the producing band stay in registers.
//
// Beyond the sequence ends the inputs are replicated (rows clamped to Lx-1,
// columns to Ly-1), which makes ΔΔ exactly zero there, so the sweep needs no
// per-cell bounds masks; only the priming column c' = 0 is masked.
//
// SAVE: also write the adjoint's boundary state (ksig_signature_levels_fwd_save):
// the row-end prefixes P_a(r, Nc) when a lane finishes column Nc, and the
// final column sums C_a(Nr, j) from the FIFO after the sweep.
#pragma once

#include "ksig_common.h"
#include "ksig_hip.h"
#include "dp_forward.h"
#include "dp_forward_impl.h"
#include "dp_backward_mfma.h"  // bwd_mfma_feats, ring constants

namespace ksig {

__host__ __device__ constexpr int fwd_mfma_pc(int nc) {
  // priming column + Nc columns, whole chunks, at least 80
  return (((nc + 1 > 80 ? nc + 1 : 80) + 15) / 16) * 16;
}

template <int NLEV, int DMAX, int KIND, bool VEC, bool SAVE>
__global__ void __launch_bounds__(kWave)
dp_fwd_mfma_kernel(const DpFwdArgs a) {
  using T = float;
  constexpr int NF = 1 + NLEV;             // [G, CS_1 .. CS_NLEV]
  constexpr int NFG = (NF + 3) & ~3;       // FIFO slot stride (16-byte quads)
  constexpr int NPS = NLEV > 1 ? NLEV - 1 : 1;
  constexpr int KB = DMAX / 4;
  constexpr int W = kBwdRingW;
  constexpr bool rbf = KIND == kRBF;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* ring = reinterpret_cast<T*>(smem_raw);  // [W][64] static-kernel values (ring_pos)

  const int lane = threadIdx.x;
  const int c16 = lane & 15, grp = lane >> 4;
  const int64_t pr = blockIdx.x;
  double* out = a.out + pr * (NLEV + 1);
  const int d = a.d;
  const int32_t ia = a.pairs[2 * pr];
  const int32_t ib = a.pairs[2 * pr + 1];
  const T* xs = reinterpret_cast<const T*>(a.x.data) + a.x.off[ia] * d;
  int Lx = a.x.len[ia];
  const T* ys = reinterpret_cast<const T*>(a.y.data) + a.y.off[ib] * d;
  int Ly = a.y.len[ib];
  if (Ly > Lx) {  // rows = the longer sequence
    const T* tp = xs; xs = ys; ys = tp;
    int tl = Lx; Lx = Ly; Ly = tl;
  }
  const int Nr = Lx - 1, Nc = Ly - 1;
  if (Nr <= 0 || Nc <= 0) {
    if (lane <= NLEV) out[lane] = lane == 0 ? 1.0 : 0.0;
    if constexpr (SAVE) {
      // no cells: the boundary state is empty or all zero
      T* sv = reinterpret_cast<T*>(a.save) + a.save_off[pr];
      const int n = (NLEV - 1) * ((Nr > 0 ? Nr : 0) + (Nc > 0 ? Nc : 0));
      for (int e = lane; e < n; e += kWave) sv[e] = T(0);
    }
    return;
  }
  const int Pc = fwd_mfma_pc(Nc);
  const int nb = (Nr + kWave - 1) / kWave;
  const int ncols = nb * Pc;
  const T xscale = T(a.xscale), yscale = T(a.yscale);
  const T xyscale = xscale * yscale, xs2 = xscale * xscale, ys2 = yscale * yscale;
  float4* fifo = reinterpret_cast<float4*>(a.fifo) + (size_t)pr * a.fifo_cols * (NFG / 4);

  // ---- Gram producer -------------------------------------------------------
  T A[4][KB];  // raw x rows of the producing band
  T B[KB];     // raw y rows of the next chunk's 16 columns
  auto load_A = [&](int b) __attribute__((always_inline)) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      int xr = b * kWave + 16 * t + c16 + 1;  // grid row r reads x row r + 1
      xr = xr < Lx - 1 ? xr : Lx - 1;
      bwd_mfma_feats<DMAX, VEC>(A[t], xs + (int64_t)xr * d, d, grp);
    }
  };
  auto load_B = [&](int k) __attribute__((always_inline)) {
    const int g0 = 16 * k;
    int yc = g0 - (g0 / Pc) * Pc + c16;  // column c' reads y row c'
    yc = yc < Ly - 1 ? yc : Ly - 1;
    bwd_mfma_feats<DMAX, VEC>(B, ys + (int64_t)yc * d, d, grp);
  };
  T xnq[4][4];  // scaled squared norms of rows 16t + 4grp + q of the band
  auto x_norms = [&]() __attribute__((always_inline)) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      T s2 = T(0);
#pragma unroll
      for (int e = 0; e < KB; ++e) s2 += A[t][e] * A[t][e];
      s2 += __shfl_xor(s2, 16, kWave);
      s2 += __shfl_xor(s2, 32, kWave);  // every lane: the norm of row 16t + c16
#pragma unroll
      for (int q = 0; q < 4; ++q) xnq[t][q] = __shfl(s2, 4 * grp + q, kWave) * xs2;
    }
  };
  auto produce = [&](int k) __attribute__((always_inline)) {
    bmf4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = bmf4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
      for (int t = 0; t < 4; ++t)
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[t][kb], B[kb], acc[t], 0, 0, 0);
    T yn = T(0);
#pragma unroll
    for (int e = 0; e < KB; ++e) yn += B[e] * B[e];
    yn += __shfl_xor(yn, 16, kWave);
    yn += __shfl_xor(yn, 32, kWave);
    yn *= ys2;
    const int base = (16 * k) % W + c16 + 4 * grp;
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = 16 * t + 4 * grp + q;
        const T dot = acc[t][q] * xyscale;
        const T G = rbf ? fast_exp2(-(xnq[t][q] + yn - T(2) * dot)) : dot;
        int slot = base + 16 * t + q;  // (16k + c16 + row) mod W
        slot = slot >= W ? slot - W : slot;
        slot = slot >= W ? slot - W : slot;
        ring[ring_pos(slot, row)] = G;
      }
  };
  bool norms_due = false;
  // the ring starts at zero: lanes read it before their first column
  for (int e = lane; e < W * kRingS; e += kWave) ring[e] = T(0);
  auto produce_chunk = [&](int k) __attribute__((always_inline)) {  // wave-uniform
    if (16 * k >= ncols) return;
    if (norms_due) {  // first chunk of a band: A was loaded 16 steps ago
      x_norms();
      norms_due = false;
    }
    produce(k);
    const int gn = 16 * k + 16;
    if (gn < ncols) {
      load_B(k + 1);
      if (gn % Pc == 0) {
        load_A(gn / Pc);
        norms_due = true;
      }
    }
  };

  // FIFO: the row above the grid, G(x row 0, y column c') and zero sums.
  {
    T x0[DMAX];
    T x0n = T(0);
#pragma unroll
    for (int f = 0; f < DMAX; ++f) {
      x0[f] = f < d ? xs[f] : T(0);
      x0n += x0[f] * x0[f];
    }
    x0n *= xs2;
    for (int s = lane; s < Pc; s += kWave) {
      const int yc = s < Ly - 1 ? s : Ly - 1;
      T dot = T(0), yn = T(0);
#pragma unroll
      for (int f = 0; f < DMAX; ++f) {
        const T yv = f < d ? ys[(int64_t)yc * d + f] : T(0);
        dot += x0[f] * yv;
        yn += yv * yv;
      }
      dot *= xyscale;
      yn *= ys2;
      float4 q[NFG / 4];
#pragma unroll
      for (int k = 0; k < NFG / 4; ++k) q[k] = float4{0.f, 0.f, 0.f, 0.f};
      q[0].x = rbf ? fast_exp2(-(x0n + yn - T(2) * dot)) : dot;
#pragma unroll
      for (int k = 0; k < NFG / 4; ++k) fifo[s * (NFG / 4) + k] = q[k];
    }
  }
  load_A(0);
  load_B(0);
  x_norms();
  __syncthreads();  // FIFO stores retired

  T fl[NF];
#pragma unroll
  for (int k = 0; k < NF; ++k) fl[k] = T(0);
  T PS[NPS];
  T HB[1];
#pragma unroll
  for (int k = 0; k < NPS; ++k) PS[k] = T(0);
  HB[0] = T(0);
  T gleft = T(0), gdiag = T(0);
  int cp = (Pc - lane) % Pc;  // this lane's column c' (lane i starts i steps late)
  int row = lane;             // grid row of this lane's current band
  int brow = 0;               // band of this lane (counts the wraps of cp)
  // SAVE: a lane's row-end prefixes, held from the end of its band pass
  // until every lane of the band has finished it (63 steps later), then
  // stored by the whole wave at once: one coalesced store per level and
  // band instead of a one-lane store per level and step
  T psh[NPS];
  int hrow = -1;                    // row of the held prefixes (-1: none)
  int next_flush = Pc + kWave - 1;  // band b is flushed at step (b+1)*Pc + 63
  T* sv = SAVE ? reinterpret_cast<T*>(a.save) + a.save_off[pr] : nullptr;
  // row-end prefixes [NPS][Nr] through a buffer descriptor: one VGPR offset
  // per lane, the level as a scalar offset
  const __amdgpu_buffer_rsrc_t sv_rs =
      raw_rsrc(sv, SAVE ? unsigned(NPS * Nr) * 4u : 0u);

  const int total = ncols + kWave - 1;
  int rslot = 0;   // ring slot of the next prefetched step
  // The FIFO through a buffer descriptor: slots are scalar offsets; only lane
  // 0's load and lane 63's store are in range (the other lanes read zero,
  // which the lane shift adds, and their stores are dropped).
  const __amdgpu_buffer_rsrc_t fifo_rs = raw_rsrc(fifo, unsigned(Pc * NFG) * 4u);
  constexpr int kSlotB = NFG * 4;  // FIFO slot bytes
  int fbyte = 0;  // lane 0's FIFO slot of the next prefetched step (bytes)
  const int fin_vo = lane == 0 ? 0 : int(kBufOob);
  int s63b = (Pc - (kWave - 1)) * kSlotB;  // lane 63's FIFO slot ((t - 63) mod Pc, bytes)
  // lane 63 stored lane 0's slot Pc - 65 >= 15 steps earlier; every step
  // issues NFG/4 FIFO loads and NFG/4 FIFO stores
  constexpr int kFifoWait = fifo_wait_count(fwd_mfma_pc(0) - (kWave + 1), NFG / 2);
  auto prefetch = [&](T& g_o, T* fin_o) __attribute__((always_inline)) {  // uniform
    g_o = ring[ring_pos(rslot, lane)];
    rslot = rslot + 1 == W ? 0 : rslot + 1;
    wait_vmcnt<kFifoWait>();
#pragma unroll
    for (int k = 0; k < NFG / 4; ++k) {
      const u32x4 q = __builtin_amdgcn_raw_buffer_load_b128(fifo_rs, fin_vo, fbyte + 16 * k, 0);
      fin_o[4 * k] = __uint_as_float(q.x);
      if (4 * k + 1 < NF) fin_o[4 * k + 1] = __uint_as_float(q.y);
      if (4 * k + 2 < NF) fin_o[4 * k + 2] = __uint_as_float(q.z);
      if (4 * k + 3 < NF) fin_o[4 * k + 3] = __uint_as_float(q.w);
    }
    fbyte = fbyte + kSlotB == Pc * kSlotB ? 0 : fbyte + kSlotB;
  };
  auto flush = [&]() __attribute__((always_inline)) {  // wave-uniform
    if constexpr (SAVE && NLEV > 1) {
      const int vo = hrow >= 0 ? hrow * 4 : int(kBufOob);
#pragma unroll
      for (int m = 0; m < NPS; ++m)
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(psh[m]), sv_rs, vo, m * Nr * 4, 0);
    }
  };
  auto step_fn = [&](int step, T& g_io, T* fin_io) __attribute__((always_inline)) {
    if (((step + 2) & 15) == 0) produce_chunk((step + 2) >> 4);
    if (SAVE && NLEV > 1 && step == next_flush) {  // wave-uniform
      flush();
      next_flush += Pc;
    }
    T in[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) in[k] = shift_add(fl[k], fin_io[k]);
    const T gnew = g_io;

    const bool keep = cp != 0;
    T M = (gnew - gleft) - (in[0] - gdiag);
    M = keep ? M : T(0);
    gleft = gnew;
    gdiag = in[0];
    T of[NF];
    of[0] = gnew;
    // The column sums are zero in the priming column (c' = 0), so that step
    // leaves the row sums as they are: they are cleared when c' wraps.
    dp_cell<T, NLEV, 1>(M, in, of, PS, HB, false);
#pragma unroll
    for (int k = 0; k < NF; ++k) fl[k] = of[k];
    {
      const int vo = lane == kWave - 1 && step >= kWave - 1 ? 0 : int(kBufOob);
#pragma unroll
      for (int k = 0; k < NFG / 4; ++k) {
        u32x4 q;
        q.x = __float_as_uint(of[4 * k]);
        q.y = 4 * k + 1 < NF ? __float_as_uint(of[4 * k + 1]) : 0u;
        q.z = 4 * k + 2 < NF ? __float_as_uint(of[4 * k + 2]) : 0u;
        q.w = 4 * k + 3 < NF ? __float_as_uint(of[4 * k + 3]) : 0u;
        __builtin_amdgcn_raw_buffer_store_b128(q, fifo_rs, vo + 16 * k, s63b, 0);
      }
    }
    ++cp;
    if (cp == Pc) {
      cp = 0;
      if (step >= lane) {
        if constexpr (SAVE && NLEV > 1) {
          // Row-end prefixes P_a(row, Nc): the padding columns past Nc have
          // M = 0 and zero column sums, so PS still holds them here.
#pragma unroll
          for (int m = 0; m < NPS; ++m) psh[m] = PS[m];
          hrow = row < Nr && brow < nb ? row : -1;
        }
        row += kWave;
        ++brow;
      }
#pragma unroll
      for (int k = 0; k < NPS; ++k) PS[k] = T(0);
    }
    s63b = s63b + kSlotB == Pc * kSlotB ? 0 : s63b + kSlotB;
    // This step's inputs are consumed: refill the set for the step after
    // next.  Issued last, so that the loads reuse the registers of this
    // step's inputs instead of copies of them.
    __builtin_amdgcn_sched_barrier(0);
    prefetch(g_io, fin_io);
  };

  T g0, g1, f0[NFG], f1[NFG];
  produce_chunk(0);
  prefetch(g0, f0);
  prefetch(g1, f1);
  // The first two steps run outside the loop, so that the loop is entered
  // with the memory operations of a whole step issued after each prefetch,
  // as on every later iteration; the compiler's wait counts then only wait
  // for the prefetched loads, not for the stores issued after them.
  step_fn(0, g0, f0);
  step_fn(1, g1, f1);
  int step = 2;
  for (; step + 2 <= total; step += 2) {
    step_fn(step, g0, f0);
    step_fn(step + 1, g1, f1);
  }
  if (step < total) step_fn(step, g0, f0);
  if (SAVE && NLEV > 1) flush();  // the last band (its flush step is past the sweep)
  __syncthreads();  // lane 63's last FIFO stores retired

  double acc[NLEV];
#pragma unroll
  for (int m = 0; m < NLEV; ++m) acc[m] = 0.0;
  for (int s = lane + 1; s <= Nc; s += kWave) {
    T v[NFG];
#pragma unroll
    for (int k = 0; k < NFG / 4; ++k) {
      const float4 q = fifo[s * (NFG / 4) + k];
      v[4 * k] = q.x; v[4 * k + 1] = q.y; v[4 * k + 2] = q.z; v[4 * k + 3] = q.w;
    }
#pragma unroll
    for (int m = 0; m < NLEV; ++m) acc[m] += (double)v[1 + m];
    if constexpr (SAVE && NLEV > 1) {
      T* cs = sv + (NLEV - 1) * Nr;  // final column sums C_a(Nr, j), j = s - 1
#pragma unroll
      for (int m = 0; m < NLEV - 1; ++m) cs[m * Nc + s - 1] = v[1 + m];
    }
  }
#pragma unroll
  for (int m = 0; m < NLEV; ++m) acc[m] = wave_sum(acc[m]);
  if (lane == 0) out[0] = 1.0;
#pragma unroll
  for (int m = 0; m < NLEV; ++m)
    if (lane == m + 1) out[m + 1] = acc[m];
}

template <int NLEV, int DMAX, int KIND, bool VEC, bool SAVE>
hipError_t launch_fwd_mfma(const DpFwdArgs& a, size_t lds, hipStream_t s) {
  auto k = dp_fwd_mfma_kernel<NLEV, DMAX, KIND, VEC, SAVE>;
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)k,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k, dim3((unsigned)a.n_pairs), dim3(kWave), lds, s, a);
  return hipGetLastError();
}

#define KSIG_FWD_MFMA_LEVELS(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(12)

template <int DMAX, bool SAVE>
hipError_t dispatch_fwd_mfma(const DpFwdArgs& a, size_t lds, hipStream_t s) {
  const bool vec = a.d == DMAX && ((uintptr_t)a.x.data & 15) == 0 &&
                   ((uintptr_t)a.y.data & 15) == 0;
#define KSIG_CASE(L)                                                              \
  if (a.n_levels == L) {                                                          \
    if (a.kind == kRBF)                                                           \
      return vec ? launch_fwd_mfma<L, DMAX, kRBF, true, SAVE>(a, lds, s)          \
                 : launch_fwd_mfma<L, DMAX, kRBF, false, SAVE>(a, lds, s);        \
    return vec ? launch_fwd_mfma<L, DMAX, kLinear, true, SAVE>(a, lds, s)         \
               : launch_fwd_mfma<L, DMAX, kLinear, false, SAVE>(a, lds, s);       \
  }
  KSIG_FWD_MFMA_LEVELS(KSIG_CASE)
#undef KSIG_CASE
  return hipErrorInvalidValue;
}

}  // namespace ksig
