// Adjoint of the first-order DP signature kernel, float32, with the static
// Gram on the matrix cores (gfx950).  Template implementation; instantiated
// by dp_backward_mfma_d*.hip.
//
// Same reverse sweep as dp_backward.hip (the adjoint of the recursion of
// ksig/algorithms.py:47-85 on M = ΔΔ k(x_i, y_j), ksig/algorithms.py:64,
// ksig/static/kernels.py:145-155, 248-260): lane L of band b owns flipped row
// 64b + L (M row i = Nr-1-(64b+L)) and walks the columns right to left, one
// step behind lane L-1; the column flows move by DPP and lane 63 hands over
// to the next band through a FIFO in a global-memory workspace (one region
// per pair), which keeps LDS to the ring below.
//
// What changes is where the static-kernel values come from.  The lane no
// longer evaluates <x_i, y_j> itself (which needs the whole y row of every
// cell in registers: 4d bytes per cell through the vector-memory path).
// Every 16 steps the wave computes a 64-row x 16-column tile of dot products
// on v_mfma_f32_16x16x4f32 and writes the argument of the static kernel
// (scaled squared distance for RBF, scaled dot product for linear) into an
// LDS ring.  The ring is skewed: ring column g = band * Pc + s of row r lives
// in slot (g + r) mod W, so at step T every lane reads slot T mod W at its
// own row, with no address arithmetic.  Pc is a multiple of 16, so a 16-column
// chunk never straddles a band, and W = 80 keeps chunk k until lane 63 has read
// it.  Chunk k is produced at step 16k - 2 and every ring read is issued two
// steps before its use: chunk k + 5 overwrites chunk k at step 16k + 78, after
// lane 63's read of its last column (issued at step 16k + 76).
//
// MFMA operand layout: lane (c16 = lane & 15, grp = lane >> 4) of instruction
// kb holds feature grp * KB + kb (KB = DMAX / 4) of x row 16t + c16 (A, four
// row tiles t) and of y column c16 (B).  Every feature is covered once, and
// each lane loads KB consecutive features of a row (vector loads when d = DMAX).
#pragma once

#include "ksig_common.h"
#include "ksig_hip.h"
#include "dp_forward.h"

namespace ksig {

typedef float bmf4 __attribute__((ext_vector_type(4)));

constexpr int kBwdRingW = 80;  // ring columns: 64-lane skew + one 16-column chunk
constexpr int kRingS = 64;     // ring slot stride (floats)

// Position of row r within ring slot `slot`: rows are XOR-swizzled by the slot
// so that the producer's writes (16 slots x 4 rows per instruction) and the
// sweep's reads (one slot, 64 rows) both spread over the LDS banks.
__device__ __forceinline__ int ring_pos(int slot, int r) { return slot * kRingS + (r ^ (slot & 31)); }

// Features grp*KB .. grp*KB+KB-1 of one row into v (KB = DMAX / 4): vector
// loads when VEC (d == DMAX, 16-byte aligned rows), else clamped scalar loads
// with the features past d zeroed.  Straight-line either way.
template <int DMAX, bool VEC>
__device__ __forceinline__ void bwd_mfma_feats(float* v, const float* row, int d, int grp) {
  constexpr int KB = DMAX / 4;
  const int f0 = grp * KB;
  if constexpr (VEC && KB % 4 == 0) {
#pragma unroll
    for (int e = 0; e < KB; e += 4) {
      const float4 t = *reinterpret_cast<const float4*>(row + f0 + e);
      v[e] = t.x; v[e + 1] = t.y; v[e + 2] = t.z; v[e + 3] = t.w;
    }
  } else if constexpr (VEC && KB == 2) {
    const float2 t = *reinterpret_cast<const float2*>(row + f0);
    v[0] = t.x; v[1] = t.y;
  } else {
#pragma unroll
    for (int e = 0; e < KB; ++e) {
      const int f = f0 + e;
      const float t = row[f < d ? f : d - 1];
      v[e] = f < d ? t : 0.f;
    }
  }
}

__host__ __device__ constexpr int bwd_mfma_pc(int nc) {
  // priming column, Nc columns, column -1; whole chunks; at least 80 so that
  // lane 63 wrote a FIFO slot 16 steps before lane 0 prefetches it
  return (((nc + 2 > 80 ? nc + 2 : 80) + 15) / 16) * 16;
}

// GP: the parameter gradient sum omega * D' per pair (grad_param); without
// it the static-kernel argument D' does not flow and the FIFO slot is a
// quad shorter at 6 levels.
template <int NLEV, int DMAX, int KIND, bool VEC, bool GP>
__global__ void __launch_bounds__(kWave)
dp_backward_mfma_kernel(const DpBwdArgs a) {
  using T = float;
  constexpr int NL1 = NLEV > 1 ? NLEV - 1 : 1;
  constexpr int NC = NLEV - 1;              // prefix levels 1..NLEV-1
  // flows [G, (D',) gM, C_1.., Dq_1..]
  constexpr int kDv = 1, kGM = GP ? 2 : 1, kC = kGM + 1, kD = kC + NC;
  constexpr int NF = kD + NC;
  constexpr int NFG = (NF + 3) & ~3;        // FIFO slot stride (16-byte quads)
  constexpr int KB = DMAX / 4;
  constexpr int W = kBwdRingW;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* ring = reinterpret_cast<T*>(smem_raw);  // [W][64] static-kernel arguments (20 KB)

  const int lane = threadIdx.x;
  const int c16 = lane & 15, grp = lane >> 4;
  const int64_t pr = blockIdx.x;
  const int d = a.d;
  constexpr bool rbf = KIND == kRBF;
  const int32_t ia = a.pairs[2 * pr];
  const int32_t ib = a.pairs[2 * pr + 1];
  const T* xs = reinterpret_cast<const T*>(a.x.data) + a.x.off[ia] * d;
  const T* ys = reinterpret_cast<const T*>(a.y.data) + a.y.off[ib] * d;
  int Lx = a.x.len[ia], Ly = a.y.len[ib];
  const bool swapped = Ly > Lx;  // same orientation rule as the forward
  if (swapped) {
    const T* tp = xs; xs = ys; ys = tp;
    int tl = Lx; Lx = Ly; Ly = tl;
  }
  const int Nr = Lx - 1, Nc = Ly - 1;
  if (Nr <= 0 || Nc <= 0) {
    if (GP && lane == 0) a.gparam[pr] = 0.0;
    return;
  }
  const int nrows = Nr + 1;  // + virtual row above the grid
  const int Pc = bwd_mfma_pc(Nc);
  const int nb = (nrows + kWave - 1) / kWave;
  const int ncols = nb * Pc;  // ring columns in total
  // band hand-over FIFO in this pair's [fifo_cols][NFG] region of the workspace
  T* fifo = reinterpret_cast<T*>(a.fifo) + (size_t)pr * a.fifo_cols * NFG;

  T w[NLEV + 1];
#pragma unroll
  for (int m = 1; m <= NLEV; ++m) w[m] = T(a.gout[pr * (NLEV + 1) + m]);
  const T* rowend = reinterpret_cast<const T*>(a.save) + a.save_off[pr];  // [NC][Nr]
  const T* colsum = rowend + NC * Nr;                                       // [NC][Nc]
  // this pair's omega block [Lx][Ly] (or [Ly][Lx]); below 2^31 bytes
  const __amdgpu_buffer_rsrc_t omega_rs =
      raw_rsrc(reinterpret_cast<T*>(a.omega) + a.omega_off[pr], unsigned(Lx * Ly) * 4u);

  const T xscale = T(a.xscale), yscale = T(a.yscale);
  const T xyscale = xscale * yscale, xs2 = xscale * xscale, ys2 = yscale * yscale;

  // ---- Gram producer -------------------------------------------------------
  T A[4][KB];  // raw x rows of the producing band (row tile t, feature block grp)
  T B[KB];     // raw y rows of the next chunk's 16 columns
  auto load_A = [&](int b) __attribute__((always_inline)) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      int i = Nr - 1 - (b * kWave + 16 * t + c16);
      i = i < 0 ? 0 : i;  // virtual row and rows past the grid: masked in the sweep
      bwd_mfma_feats<DMAX, VEC>(A[t], xs + (int64_t)i * d, d, grp);
    }
  };
  auto load_B = [&](int k) __attribute__((always_inline)) {  // chunk k: ring columns 16k .. 16k+15, one band
    const int g0 = 16 * k;
    const int s = g0 - (g0 / Pc) * Pc + c16;
    int j = Nc - s;
    j = j < 0 ? 0 : j;  // columns past the grid: masked in the sweep
    bwd_mfma_feats<DMAX, VEC>(B, ys + (int64_t)j * d, d, grp);
  };
  T xnq[4][4];  // scaled squared norms of the band's rows 16t + 4grp + q
  auto x_norms = [&]() __attribute__((always_inline)) {  // after load_A
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
    const int base = (16 * k) % W + c16 + 4 * grp;  // < W + 16 + 12
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = 16 * t + 4 * grp + q;
        const T dot = acc[t][q] * xyscale;
        const T arg = rbf ? xnq[t][q] + yn - T(2) * dot : dot;
        int slot = base + 16 * t + q;  // (16k + c16 + row) mod W
        slot = slot >= W ? slot - W : slot;
        slot = slot >= W ? slot - W : slot;
        ring[ring_pos(slot, row)] = arg;
      }
  };

  // Row-end prefixes of flipped row fr, loaded a band ahead into relnext.
  T relnext[NL1];
  auto stage = [&](int fr) __attribute__((always_inline)) {
    const int i = Nr - 1 - fr;
    const bool ok = i >= 0 && fr < nrows;
    const int ic = ok ? i : 0;
#pragma unroll
    for (int m = 0; m < NL1; ++m) {
      const T v = m < NC ? rowend[m * Nr + ic] : T(0);
      relnext[m] = ok ? v : T(0);
    }
  };

  // FIFO: the row below the grid (G row Nr = Lx-1) with the forward's final
  // column sums; slot s covers column j = Nc - s.
  {
    T xb[DMAX];
    T xbn = T(0);
#pragma unroll
    for (int k = 0; k < DMAX; ++k) {
      const T v = k < d ? xs[(int64_t)Nr * d + k] : T(0);
      xb[k] = v * xyscale;
      xbn += v * v;
    }
    xbn *= xs2;
    for (int s = lane; s < Pc; s += kWave) {
      const int j = Nc - s;
      T g = T(0), dv = T(0);
      if (j >= 0) {
        T dot = T(0), yn = T(0);
#pragma unroll
        for (int k = 0; k < DMAX; ++k) {
          const T v = k < d ? ys[(int64_t)j * d + k] : T(0);
          dot += xb[k] * v;
          yn += v * v;
        }
        yn *= ys2;
        if (rbf) {
          dv = xbn + yn - T(2) * dot;
          g = fast_exp2(-dv);
        } else {
          dv = g = dot;
        }
      }
      T v[NFG];
      v[0] = g;
      if constexpr (GP) v[kDv] = dv;
      v[kGM] = T(0);
#pragma unroll
      for (int m = 0; m < NC; ++m) {
        v[kC + m] = (j >= 0 && j < Nc) ? colsum[m * Nc + j] : T(0);
        v[kD + m] = T(0);
      }
#pragma unroll
      for (int k = NF; k < NFG; ++k) v[k] = T(0);
#pragma unroll
      for (int k = 0; k < NFG; k += 4)
        *reinterpret_cast<float4*>(fifo + s * NFG + k) = float4{v[k], v[k + 1], v[k + 2], v[k + 3]};
    }
  }
  stage(lane);
  load_A(0);
  load_B(0);
  __syncthreads();
  x_norms();

  T fl[NF];
#pragma unroll
  for (int k = 0; k < NF; ++k) fl[k] = T(0);
  // Row state.  The column flows C and D are zero in the priming column (and
  // before it), so the priming step leaves P and S as they are: a lane enters
  // a band with P = its row-end prefixes and S = 0, set at the band change.
  T P[NL1], S[NL1];
#pragma unroll
  for (int m = 0; m < NL1; ++m) {
    P[m] = relnext[m];
    S[m] = T(0);
  }
  T g_right = T(0), gup_right = T(0), dup_right = T(0), gm_right = T(0), gmup_right = T(0);
  T gsig_row = T(0);
  double gsig = 0.0;

  int s = -lane;  // step within the band: column j = Nc - s
  int band = 0;
  int fr = lane;  // flipped row
  const int total = ncols + kWave - 1;
  // band b+1 is staged at step b*Pc + 63: every lane has taken band b by then
  int next_stage = kWave - 1;
  int stage_band = 1;
  int rslot = 0;  // ring slot of the next prefetched step: (step + 2) mod W
  bool norms_due = false;
  // omega index of G point (p, q) = (i+1, j+1) in the caller's orientation:
  // p * Ly + q, or q * Lx + p when swapped; q falls by one per step.
  const int cst = swapped ? Lx : 1;
  const int rst = swapped ? 1 : Ly;
  int oi_row = (Nr - fr) * rst;        // p = i + 1 = Nr - fr
  int oi_col = (Nc + 1 - s) * cst;     // q = j + 1 = Nc + 1 - s

  // One step ahead: the ring value and lane 0's FIFO entry of step t+1 are
  // read at the top of step t (chunk k is produced at step 16k - 1 for that).
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
  // Inputs of step t (the ring value of this lane, lane 0's FIFO entry) are
  // read during step t-2 into one of two register sets: a step consumes its
  // set, then refills it for the step after next.  Every lane reads lane 0's
  // FIFO slot (a broadcast; only lane 0 keeps it), so the read needs no lane
  // mask.  Lane 63 wrote that slot Pc - 65 >= 15 steps earlier.
  // The FIFO through a buffer descriptor: lane 0's slot is a scalar offset;
  // only lane 0's load and lane 63's store are in range (the other lanes
  // read zero, which the lane shift adds, and their stores are dropped).
  const __amdgpu_buffer_rsrc_t fifo_rs = raw_rsrc(fifo, unsigned(Pc * NFG) * 4u);
  constexpr int kSlotB = NFG * 4;  // FIFO slot bytes
  int fbyte = 0;  // lane 0's FIFO slot of the next prefetched step: ((step + 2) mod Pc) * kSlotB
  const int fin_vo = lane == 0 ? 0 : int(kBufOob);
  // every step issues NFG/4 FIFO loads and NFG/4 FIFO stores
  constexpr int kFifoWait = fifo_wait_count(bwd_mfma_pc(0) - (kWave + 1), NFG / 2);
  auto prefetch = [&](T& arg_o, T* fin_o) __attribute__((always_inline)) {  // uniform
    arg_o = ring[ring_pos(rslot, lane)];
    rslot = rslot + 1 == W ? 0 : rslot + 1;
    wait_vmcnt<kFifoWait>();
    const int f0 = fbyte;
    fbyte = fbyte + kSlotB == Pc * kSlotB ? 0 : fbyte + kSlotB;
#pragma unroll
    for (int k = 0; k < NFG; k += 4) {
      const u32x4 t = __builtin_amdgcn_raw_buffer_load_b128(fifo_rs, fin_vo, f0 + 4 * k, 0);
      fin_o[k] = __uint_as_float(t.x);
      if (k + 1 < NF) fin_o[k + 1] = __uint_as_float(t.y);
      if (k + 2 < NF) fin_o[k + 2] = __uint_as_float(t.z);
      if (k + 3 < NF) fin_o[k + 3] = __uint_as_float(t.w);
    }
  };
  auto step_fn = [&](int step, T& arg_io, T* fin_io) __attribute__((always_inline)) {
    if (((step + 2) & 15) == 0) produce_chunk((step + 2) >> 4);
    if (step == next_stage) {  // wave-uniform
      if (stage_band < nb) stage(stage_band * kWave + lane);
      ++stage_band;
      next_stage += Pc;
    }
    const bool active = s >= 0 && band < nb;
    T in[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) in[k] = shift_add(fl[k], fin_io[k]);
    const T arg = arg_io;
    T g = rbf ? fast_exp2(-arg) : arg;

    const int j = Nc - s;
    const int i = Nr - 1 - fr;
    const bool gvalid = active && i >= 0 && j >= 0;
    g = gvalid ? g : T(0);
    const T dv = gvalid ? arg : T(0);
    const bool prim = s == 0;
    const bool cell = active && i >= 0 && j >= 0 && j < Nc;
    const T M = cell ? (gup_right - in[0]) - (g_right - g) : T(0);

    // forward reconstruction: C_a(i,j), P_a(i,j)
    T Pn[NL1], Cout[NL1];
    {
      T R = M;
#pragma unroll
      for (int m = 0; m < NC; ++m) {
        const T C = in[kC + m] - R;
        Pn[m] = P[m] - C;
        Cout[m] = C;
        R = M * Pn[m];
      }
    }
    // adjoint
    T lam[NLEV + 1];
    lam[NLEV] = w[NLEV];
#pragma unroll
    for (int m = NLEV - 1; m >= 1; --m) lam[m] = w[m] + S[m - 1];
    T gM = lam[1];
#pragma unroll
    for (int m = 2; m <= NLEV; ++m) gM += lam[m] * Pn[m - 2];
    gM = cell ? gM : T(0);
    T Dout[NL1];
#pragma unroll
    for (int m = 0; m < NC; ++m) Dout[m] = in[kD + m] + M * lam[m + 2];

    // finalise dL/dG at G point (i+1, j+1)
    const bool fin_ok = active && !prim && i >= -1 && j >= -1;
    const T gG = (gM - gm_right) - (in[kGM] - gmup_right);
    const T om = rbf ? gG * gup_right : gG;
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(om), omega_rs,
                                          fin_ok ? (oi_row + oi_col) * 4 : int(kBufOob), 0, 0);
    if constexpr (GP) gsig_row += (fin_ok ? om : T(0)) * dup_right;

    // state
    g_right = g;
    gup_right = in[0];
    if constexpr (GP) dup_right = in[kDv];
    gm_right = gM;
    gmup_right = in[kGM];
#pragma unroll
    for (int m = 0; m < NC; ++m) {
      P[m] = Pn[m];
      S[m] += in[kD + m];
    }
    fl[0] = g;
    if constexpr (GP) fl[kDv] = dv;
    fl[kGM] = gM;
#pragma unroll
    for (int m = 0; m < NC; ++m) {
      fl[kC + m] = Cout[m];
      fl[kD + m] = Dout[m];
    }
    {
      const int fo = lane == kWave - 1 && active ? s * kSlotB : int(kBufOob);
#pragma unroll
      for (int k = 0; k < NFG; k += 4) {
        u32x4 t;
        t.x = __float_as_uint(fl[k]);
        t.y = k + 1 < NF ? __float_as_uint(fl[k + 1]) : 0u;
        t.z = k + 2 < NF ? __float_as_uint(fl[k + 2]) : 0u;
        t.w = k + 3 < NF ? __float_as_uint(fl[k + 3]) : 0u;
        __builtin_amdgcn_raw_buffer_store_b128(t, fifo_rs, fo + 4 * k, 0, 0);
      }
    }

    ++s;
    oi_col -= cst;
    if (s == Pc) {
      s = 0;
      ++band;
      fr += kWave;
      oi_row = (Nr - fr) * rst;
      oi_col = (Nc + 1) * cst;
      gsig += (double)gsig_row;
      gsig_row = T(0);
#pragma unroll
      for (int m = 0; m < NC; ++m) {
        P[m] = relnext[m];
        S[m] = T(0);
      }
    }
    // This step's inputs are consumed: refill the set for the step after
    // next.  Issued last, so that the loads reuse the registers of this
    // step's inputs instead of copies of them.
    __builtin_amdgcn_sched_barrier(0);
    prefetch(arg_io, fin_io);
  };

  T arg0, arg1, fin0[NFG], fin1[NFG];
  produce_chunk(0);
  prefetch(arg0, fin0);
  prefetch(arg1, fin1);
  // The first two steps run outside the loop, so that the loop is entered
  // with the memory operations of a whole step issued after each prefetch,
  // as on every later iteration; the compiler's wait counts then only wait
  // for the prefetched loads, not for the stores issued after them.
  step_fn(0, arg0, fin0);
  step_fn(1, arg1, fin1);
  int step = 2;
  for (; step + 2 <= total; step += 2) {
    step_fn(step, arg0, fin0);
    step_fn(step + 1, arg1, fin1);
  }
  if (step < total) step_fn(step, arg0, fin0);
  if constexpr (GP) {
    gsig += (double)gsig_row;
    gsig = wave_sum(gsig);
    if (lane == 0) a.gparam[pr] = gsig;
  }
}

template <int NLEV, int DMAX, int KIND, bool VEC, bool GP>
hipError_t launch_bwd_mfma_gp(const DpBwdArgs& a, size_t lds, hipStream_t s) {
  auto k = dp_backward_mfma_kernel<NLEV, DMAX, KIND, VEC, GP>;
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)k,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k, dim3((unsigned)a.n_pairs), dim3(kWave), lds, s, a);
  return hipGetLastError();
}

template <int NLEV, int DMAX, int KIND, bool VEC>
hipError_t launch_bwd_mfma(const DpBwdArgs& a, size_t lds, hipStream_t s) {
  if (a.gparam) return launch_bwd_mfma_gp<NLEV, DMAX, KIND, VEC, true>(a, lds, s);
  return launch_bwd_mfma_gp<NLEV, DMAX, KIND, VEC, false>(a, lds, s);
}

#define KSIG_BWD_MFMA_LEVELS(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(12)

template <int DMAX>
hipError_t dispatch_bwd_mfma(const DpBwdArgs& a, size_t lds, hipStream_t s) {
  const bool vec = a.d == DMAX && ((uintptr_t)a.x.data & 15) == 0 &&
                   ((uintptr_t)a.y.data & 15) == 0;
#define KSIG_CASE(L)                                                             \
  if (a.n_levels == L) {                                                         \
    if (a.kind == kRBF)                                                          \
      return vec ? launch_bwd_mfma<L, DMAX, kRBF, true>(a, lds, s)               \
                 : launch_bwd_mfma<L, DMAX, kRBF, false>(a, lds, s);             \
    return vec ? launch_bwd_mfma<L, DMAX, kLinear, true>(a, lds, s)              \
               : launch_bwd_mfma<L, DMAX, kLinear, false>(a, lds, s);            \
  }
  KSIG_BWD_MFMA_LEVELS(KSIG_CASE)
#undef KSIG_CASE
  return hipErrorInvalidValue;
}

}  // namespace ksig
