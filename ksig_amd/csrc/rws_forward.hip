// Random Warping Series: DTW of sequences against short random warping
// series, forward (gfx950).
//
// Computes P[a, b] of ksig/algorithms.py:565-659: with D the squared
// Euclidean distances between the points of sequence x_a and of warping
// series w_b (ksig/kernels.py:466), the dynamic time warping recursion
//   P(i,j) = D(i,j) + min(P(i-1,j), P(i,j-1), P(i-1,j-1)),
// P(-1,-1) = 0, other out-of-grid predecessors +inf, P[a,b] = P(last, last)
// (:598-606).  The reference runs one block per (a, b) pair, sweeping
// antidiagonals with two barriers per step, on a materialised
// D [n_X, l_X, sum of warp lengths].
//
// Here the warping series are short (at most max_warp points), so one wave
// takes one sequence and a group of consecutive warping series stacked end
// to end along the rows (the row-stream layout of the DP kernel): lane i owns
// R consecutive rows of each 64R-row band of the stack, the sequence's points
// are the columns (staged in LDS with their norms, shared by the R rows of a
// step), and the lane walks them one step behind lane i-1, receiving P of the
// row above its first row through DPP wave_shr:1 (lane 0 from an LDS FIFO
// written by lane 63 in the previous band).  A lane whose row is the
// first of its warping series ignores what flows in (the previous series'
// values) and starts from the boundary; the lane of a series' last row writes
// P at the last column.  D is computed in registers: no distance tensor.
// Lanes are busy on real cells except in the last band of each group.

#include <math.h>

#include "ksig_common.h"
#include "ksig_hip.h"
#include "rws.h"

namespace ksig {

template <typename T, int DMAX, bool GRAMIN, int R>
__global__ void __launch_bounds__(kWave)
rws_forward_kernel(const RwsArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const T kInf = T(INFINITY);
  constexpr int BR = kWave * R;  // rows per band

  const int lane = threadIdx.x;
  const int wave = blockIdx.x;
  const int ia = wave / a.n_groups;
  const int grp = wave % a.n_groups;
  const int d = a.d;
  const int w_beg = a.groups[grp];
  const int w_end = a.groups[grp + 1];
  const int row_beg = a.warp_off[w_beg];
  const int row_end = a.warp_off[w_end];
  if (row_end <= row_beg) return;

  const T* xs = nullptr;
  const T* gm = nullptr;
  int Lx;
  if constexpr (GRAMIN) {
    Lx = a.gram_lx;
    gm = reinterpret_cast<const T*>(a.gram) + (int64_t)ia * Lx * a.warp_rows;
  } else {
    xs = reinterpret_cast<const T*>(a.x.data) + a.x.off[ia] * d;
    Lx = a.x.len[ia];
  }
  double* out = a.out + (int64_t)ia * a.n_warps;
  if (Lx <= 0) {  // no alignment of an empty sequence
    for (int w = w_beg + lane; w < w_end; w += kWave) out[w] = INFINITY;
    return;
  }
  // Columns per band pass; more than 64 so that lane 0's FIFO read, issued
  // one step ahead, comes at least a step after lane 63's write of the slot.
  const int Pc = Lx <= kWave ? kWave + 1 : Lx;
  const int nrows = row_end - row_beg;
  const int nb = (nrows + BR - 1) / BR;

  // LDS: FIFO [Pc], then the sequence's points [Lx][DMAX] and norms [Lx].
  T* fifo = reinterpret_cast<T*>(smem_raw);
  T* xl = fifo + ((Pc + 3) & ~3);
  T* xn = xl + (GRAMIN ? 0 : Lx * DMAX);
  if constexpr (!GRAMIN) {
    for (int j = lane; j < Lx; j += kWave) {
      T n2 = T(0);
#pragma unroll
      for (int k = 0; k < DMAX; ++k) {
        const T v = k < d ? xs[(int64_t)j * d + k] : T(0);
        xl[j * DMAX + k] = v;
        n2 += v * v;
      }
      xn[j] = n2;
    }
  }
  __syncthreads();

  const T* wdata = reinterpret_cast<const T*>(a.warp);
  // this lane's R rows of the current band: warp points, norms, stack row
  // (row_end past the stack), series and first / last flags; and the same of
  // its rows in the next band, staged by the whole wave at one step (a
  // band's worth of steps ahead of the lane's band change, which then only
  // copies registers: loading there would run the loads and the series
  // search, under one lane's mask, at each of the 64 steps where some lane
  // changes band)
  T wv[R][DMAX], wn[R];
  int wid[R];
  bool valid[R], first[R], last[R];
  T pwv[R][DMAX], pwn[R];
  int pwid[R];
  bool pvalid[R], pfirst[R], plast[R];
  auto load_rows = [&](int b) __attribute__((always_inline)) {  // into the pending set
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int row = row_beg + b * BR + lane * R + q;
      const bool ok = row < row_end;  // rows past the stack are padding
      pvalid[q] = ok;
      // series of the row: the last w in [w_beg, w_end) with warp_off[w] <= row
      int lo = w_beg, hi = w_end - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.warp_off[mid] <= row) lo = mid; else hi = mid - 1;
      }
      pwid[q] = lo;
      pfirst[q] = ok && row == a.warp_off[lo];
      plast[q] = ok && row == a.warp_off[lo + 1] - 1;
      T n2 = T(0);
#pragma unroll
      for (int k = 0; k < DMAX; ++k) {
        const T v = (!GRAMIN && ok && k < d) ? wdata[(int64_t)row * d + k] : T(0);
        pwv[q][k] = v;
        n2 += v * v;
      }
      pwn[q] = n2;
    }
  };
  auto take_rows = [&]() __attribute__((always_inline)) {  // the pending set becomes current
#pragma unroll
    for (int q = 0; q < R; ++q) {
#pragma unroll
      for (int k = 0; k < DMAX; ++k) wv[q][k] = pwv[q][k];
      wn[q] = pwn[q];
      valid[q] = pvalid[q];
      wid[q] = pwid[q];
      first[q] = pfirst[q];
      last[q] = plast[q];
    }
  };

  // Inputs of one step, read a step ahead: the column's point and norm (or
  // the caller's distances of the lane's rows), and lane 0's FIFO entry.
  struct StepIn {
    T x[GRAMIN ? R : DMAX];
    T xn;
    T f;
  };
  auto fetch = [&](StepIn& in, int c, int b) __attribute__((always_inline)) {
    const int cx = c < 0 ? 0 : (c >= Lx ? Lx - 1 : c);
    if constexpr (GRAMIN) {
#pragma unroll
      for (int q = 0; q < R; ++q) {
        int rr = row_beg + b * BR + lane * R + q;
        rr = rr < row_end ? rr : row_end - 1;
        in.x[q] = gm[(int64_t)cx * a.warp_rows + rr];
      }
    } else {
#pragma unroll
      for (int k = 0; k < DMAX; ++k) in.x[k] = xl[cx * DMAX + k];
      in.xn = xn[cx];
    }
    in.f = kInf;
    if (lane == 0 && b > 0 && b < nb && c >= 0) in.f = fifo[c];
  };

  int band = 0;
  load_rows(0);
  take_rows();
  // band b+1 is staged at step b*Pc + 64, when every lane has taken band b
  // (lane l takes band b at the end of step b*Pc - 1 + l) and before lane 0
  // takes band b+1 (end of step (b+1)*Pc - 1; Pc >= 65)
  int next_stage = kWave;
  int stage_band = 1;
  T left[R];
#pragma unroll
  for (int q = 0; q < R; ++q) left[q] = kInf;
  T up = kInf, pin = kInf;
  int cc = -lane;
  StepIn cur;
  fetch(cur, cc, band);
  const int total = nb * Pc + kWave - 1;
  for (int step = 0; step < total; ++step) {
    if (step == next_stage) {  // wave-uniform
      if (stage_band < nb) load_rows(stage_band);
      ++stage_band;
      next_stage += Pc;
    }
    const bool active = cc >= 0 && band < nb;
    const bool colok = active && cc < Lx;
    const bool c0 = cc == 0;
    const T in = shift_down(up, cur.f);
    T dist[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      if constexpr (GRAMIN) {
        dist[q] = cur.x[q];
      } else {
        T dot = T(0);
#pragma unroll
        for (int k = 0; k < DMAX; ++k) dot += wv[q][k] * cur.x[k];
        dist[q] = wn[q] + cur.xn - T(2) * dot;
      }
    }
    {
      const bool wrap = cc + 1 == Pc;
      fetch(cur, wrap ? 0 : cc + 1, band + (wrap ? 1 : 0));
    }
    T u = in, dg = pin;  // P(row-1, c), P(row-1, c-1)
#pragma unroll
    for (int q = 0; q < R; ++q) {
      // the first row of a series starts from the boundary: P(-1, c) = inf,
      // P(-1, -1) = 0; column -1 is +inf
      // (left values are reset to +inf at each band start, so column -1
      // needs no select for them, nor for the diagonal of rows q >= 1)
      const T uq = first[q] ? kInf : u;
      T dq = first[q] ? (c0 ? T(0) : kInf) : dg;
      if (q == 0 && !first[q]) dq = c0 ? kInf : dg;
      T val = dist[q] + fminf_t(fminf_t(uq, left[q]), dq);
      const bool cell = colok && valid[q];
      val = cell ? val : kInf;
      if (cell && last[q] && cc == Lx - 1) out[wid[q]] = (double)val;
      dg = left[q];
      left[q] = val;
      u = val;
    }
    if (lane == kWave - 1 && active) fifo[cc] = u;
    up = u;
    pin = in;

    ++cc;
    if (cc == Pc) {  // wrap to the next band: column -1 of the new rows
      cc = 0;
      ++band;
      if (band < nb) take_rows();
#pragma unroll
      for (int q = 0; q < R; ++q) left[q] = kInf;
    }
  }
}

template <typename T, int DMAX, bool GRAMIN, int R>
static hipError_t launch_rws(const RwsArgs& a, size_t lds, hipStream_t s) {
  auto k = rws_forward_kernel<T, DMAX, GRAMIN, R>;
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)k,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  const int64_t waves = a.n_seq * (int64_t)a.n_groups;
  hipLaunchKernelGGL(k, dim3((unsigned)waves), dim3(kWave), lds, s, a);
  return hipGetLastError();
}

// Rows per lane from register pressure: 4 (float32, d <= 16; float64, d <= 8),
// 2 otherwise.
template <typename T>
static hipError_t dispatch_rws(const RwsArgs& a, size_t lds, hipStream_t s) {
  constexpr bool f32 = sizeof(T) == 4;
  if (a.gram) return launch_rws<T, 8, true, 4>(a, lds, s);
  if (a.d <= 8) return launch_rws<T, 8, false, 4>(a, lds, s);
  if (a.d <= 16) return launch_rws<T, 16, false, f32 ? 4 : 2>(a, lds, s);
  return launch_rws<T, 32, false, 2>(a, lds, s);
}

size_t rws_lds_bytes(int dtype, bool gram, int d, int lx) {
  const size_t esz = dtype == KSIG_F32 ? 4 : 8;
  const int dmax = d <= 8 ? 8 : (d <= 16 ? 16 : 32);
  const size_t pc = lx <= kWave ? kWave + 1 : lx;
  const size_t cols = gram ? 0 : (size_t)lx;
  return (((pc + 3) & ~(size_t)3) + cols * dmax + cols) * esz;
}

hipError_t rws_forward_launch(const RwsArgs& a, int dtype, size_t lds, hipStream_t s) {
  if (dtype == KSIG_F32) return dispatch_rws<float>(a, lds, s);
  return dispatch_rws<double>(a, lds, s);
}

}  // namespace ksig
