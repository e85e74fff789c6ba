// Truncated signature kernel by dynamic programming, forward pass, float32,
// fused RBF / linear static kernel on the vector ALUs, difference=True:
// the row-stream sweep (gfx950).  Template implementation; instantiated by
// dp_stream_*.hip.  This is the production path for float32 inputs.
//
// Recursion: ksig/algorithms.py:47-85 (first order) and :88-140 (higher
// order) on M = ΔΔ k(x_i, y_j) (ksig/algorithms.py:64), k from
// ksig/static/kernels.py:248-260 (RBF) or :145-155 (linear).
//
// One wavefront owns a list of pairs and lays their rows end to end:
//   pair k occupies stream rows [start_k, start_k + rows_k), where row 0 of a
//   pair is its priming row (the first sequence point, the grid row above the
//   ΔΔ grid) and rows 1..Nr are the ΔΔ rows; rows_k is padded up to a multiple
//   of R with copies of the last row (their ΔΔ is exactly zero).
// Lane i owns R consecutive stream rows of every 64R-row band and walks the
// band's columns one step behind lane i-1 (the skewed systolic wave of
// dp_sweep.h); the column prefix sums of the recursion flow from lane to
// lane by DPP and from lane 63 to lane 0 of the next band through an LDS FIFO.
// A lane whose rows start a pair ignores what flows in (the previous pair's
// column sums); a lane whose rows end a pair adds the column sums leaving it
// into per-lane FP64 accumulators and writes the pair's level sums when its
// band pass ends.  Consequences:
//   * no lane idles between pairs: a wave pays the 63-step pipeline fill once
//     per pair list, not once per pair, and rows are padded to R, not to 64;
//   * a band pass has one column count Pc for the whole wave (the longest
//     column side in the list; shorter pairs replicate their last column,
//     whose ΔΔ is zero), so pair lists are best made of similar lengths;
//   * the column sequence is read from global memory (L1/L2 resident: the 64
//     lanes of a wave touch the same 64-row window), two steps ahead in
//     registers, so LDS holds only the pair table, each lane's next-band x rows
//     and the FIFO, and occupancy does not depend on the sequence length
//     beyond the FIFO.
//
// Static kernel with all scale factors folded into the row side:
//   RBF:    G = exp2(<gdot x, y> - gn |x|^2 - gn |y|^2),  gdot = 2 gn,
//           gn = log2(e) / max(2 sigma^2, 1e-12)
//   linear: G = <gdot x, y>,                              gdot = scale
#pragma once

#include "ksig_common.h"
#include "ksig_hip.h"
#include "dp_forward.h"
#include "dp_forward_impl.h"
#include "dp_sweep.h"

namespace ksig {

// Row-stream pair table entry (LDS): rows = the longer sequence.
struct StreamPair {
  int32_t q;      // index into the caller's pair list (output row)
  int32_t lr1;    // row sequence length - 1 (= Nr)
  int32_t lc1;    // column sequence length - 1 (= Nc)
  int32_t swap;   // 1: rows come from the y set
  int64_t roff;   // element offset of the row sequence in its set
  int64_t coff;   // element offset of the column sequence in its set
};

template <int DMAX>
struct StreamLayout {
  // Staged x row (lane-private): DMAX features then the norm term, padded to
  // an odd number of 16-byte quads.
  static constexpr int XS = ((DMAX / 4) % 2 == 0) ? DMAX + 4 : DMAX + 8;
};

constexpr int kStreamMaxPairs = 128;       // pair-table entries per wave
constexpr int kStreamMinCols = 68;         // smallest band width Pc (FIFO hazard + staging)
constexpr int kStreamMinColsGlobal = 96;   // same with the FIFO in global memory

// Two-wave slices (global FIFO only): the waves of a block split one slice's
// bands, wave 0 the even and wave 1 the odd ones, and hand the band FIFO to
// each other through global memory.  Each publishes its step count in LDS
// every kStreamCoopSync steps (release) and reads the other's (acquire):
// wave 1 stays at least kStreamCoopSync + 65 steps behind wave 0 (it reads
// what wave 0's lane 63 stored 63 steps after its own lane 0, two steps
// ahead), wave 0 at most Pc - kStreamCoopSync - 65 steps ahead of wave 1 (the
// same for the next band).  Both can wait at once only if
// Pc < 2 kStreamCoopSync + 130, so a two-wave band pass has at least
// kStreamCoopMinCols slots.  A wait is bounded; past the bound the wave stops
// waiting and writes NaN for its pairs instead of values.
constexpr int kStreamCoopSync = 32;
constexpr int kStreamCoopMinCols = 224;
constexpr int kStreamCoopSpinLimit = 1 << 22;
static_assert(kStreamCoopMinCols >= 2 * kStreamCoopSync + 130 + 2, "two-wave deadlock bound");

// Rows per lane of an instantiation: 4 when the per-row registers (x row,
// norm, row prefixes, side sums, left value) stay small, else 2.
__host__ __device__ constexpr int stream_rows(int nlev, int order, int dmax) {
  return (dmax + 3 + (nlev - 1) + dp_num_side(nlev, order)) * 4 <= 100 ? 4 : 2;
}

// Dynamic LDS of the row stream: lane-private x staging, FIFO of `pc` slots,
// row starts and pair table of `cap` entries (cap % 4 == 0).
__host__ __device__ inline size_t stream_lds_bytes(int dmax, int r, int pc, int nf, int cap,
                                                   bool global_fifo) {
  const int xs = ((dmax / 4) % 2 == 0) ? dmax + 4 : dmax + 8;
  return (size_t)4 * kWave * r * xs + (global_fifo ? 0 : (size_t)4 * ((pc * nf + 1) & ~1)) +
         4 * ((size_t)cap + 4) + sizeof(StreamPair) * cap;
}

// Dynamic LDS of a block of `wpb` waves: one region of `per_wave` bytes
// (rounded to 16) per wave and the two-wave slice's published step counts.
__host__ __device__ inline size_t stream_region_bytes(size_t per_wave) {
  return (per_wave + 15) & ~(size_t)15;
}
__host__ __device__ inline size_t stream_block_lds_bytes(size_t per_wave, int wpb) {
  return wpb * stream_region_bytes(per_wave) + 16;
}

// Global FIFO workspace bytes: every wave owns [pc][nf rounded up to 4] floats.
__host__ __device__ inline size_t stream_gfifo_bytes(int pc, int nf, int64_t n_waves) {
  return (size_t)4 * pc * ((nf + 3) & ~3) * (size_t)n_waves;
}

// Load one sequence point (d features) into v[0..DMAX).  VEC: 16-byte vector
// loads (d == DMAX, 16-byte aligned data).  Otherwise one dword load per
// feature slot with the address clamped into the row and the slots past d
// zeroed, so both forms are straight-line code (the compiler then tracks the
// outstanding loads exactly and the two-step prefetch stays in flight).
template <int DMAX, bool VEC>
__device__ __forceinline__ void stream_load_row(float* v, const float* p, int d) {
  if (VEC) {
#pragma unroll
    for (int f = 0; f < DMAX; f += 4) {
      const float4 t = *reinterpret_cast<const float4*>(p + f);
      v[f] = t.x; v[f + 1] = t.y; v[f + 2] = t.z; v[f + 3] = t.w;
    }
  } else {
#pragma unroll
    for (int f = 0; f < DMAX; ++f) {
      const float t = p[f < d ? f : d - 1];
      v[f] = f < d ? t : 0.0f;
    }
  }
}

template <int NLEV, int ORDER, int DMAX, int R, int KIND, bool VEC, bool GF, bool SAVE>
__global__ void __launch_bounds__(2 * kWave)
dp_stream_kernel(const DpStreamArgs a) {
  using T = float;
  constexpr int NSIDE = dp_num_side(NLEV, ORDER);
  constexpr int NF = 1 + NLEV + NSIDE;
  constexpr int NPS = NLEV > 1 ? NLEV - 1 : 1;
  constexpr int NHB = NSIDE > 0 ? NSIDE : 1;
  constexpr int XS = StreamLayout<DMAX>::XS;
  constexpr int NFP = NF;                // LDS FIFO slot stride (unpadded: LDS sets occupancy)
  constexpr int NFG = (NF + 3) & ~3;     // global FIFO slot stride (16-byte quads)
  constexpr int RB = kWave * R;
  constexpr bool RBF = KIND == kRBF;
  constexpr int NC = NLEV - 1;  // saved levels (SAVE: order 1 only)
  static_assert(!SAVE || ORDER == 1, "the adjoint boundary state exists for order 1");

  extern __shared__ __align__(16) unsigned char smem_raw[];
  // one or two waves per block (two with the global FIFO), each with its own
  // LDS region; the published step counts of a two-wave slice follow them
  const int wpb = blockDim.x >> 6;
  const int v = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  const int lane = threadIdx.x & (kWave - 1);
  int* prog = reinterpret_cast<int*>(smem_raw + wpb * a.lds_region);
  if (threadIdx.x < 2) prog[threadIdx.x] = 0;
  __syncthreads();  // the only block-wide barrier: before any wave can leave
  unsigned char* wsm = smem_raw + ((v * a.lds_region) & ~15);  // regions are 16-byte multiples
  const int cap = a.tab_cap;  // pair-table entries (a multiple of 4)
  T* xst = reinterpret_cast<T*>(wsm);                 // [RB][XS], 16-byte rows
  T* fifo = xst + RB * XS;                            // [Pc][NFP] (LDS FIFO)
  // FIFO rounded to an even word count and cap % 4 == 0: the table is 8-byte aligned
  const int lds_fifo_words = GF ? 0 : ((a.fifo_cols * NFP + 1) & ~1);
  int32_t* tstart = reinterpret_cast<int32_t*>(fifo + lds_fifo_words);
  StreamPair* tab = reinterpret_cast<StreamPair*>(tstart + cap + 4);

  const int w0 = wpb * blockIdx.x;
  const int w = w0 + v;  // this wave's slot (its FIFO region)
  const int d = a.d;
  // a two-wave slice: slot w0 holds the slice and slot w0 + 1 is empty
  bool coop = GF && !SAVE && wpb == 2 && a.wave_beg && w0 + 1 < a.n_waves &&
              a.wave_beg[w0 + 1] == a.wave_beg[w0 + 2] && a.wave_beg[w0] < a.wave_beg[w0 + 1];
  const int sw = coop ? w0 : w;  // the slice this wave sweeps
  if (sw >= a.n_waves) return;
  int kb, ke;
  if (a.wave_beg) {
    kb = a.wave_beg[sw];
    ke = a.wave_beg[sw + 1];
  } else {
    kb = sw * a.per_wave;
    ke = kb + a.per_wave;
    ke = ke < a.n_pairs ? ke : (int)a.n_pairs;
  }
  int nk = ke - kb;
  if (nk <= 0) return;
  nk = nk < cap ? nk : cap;  // the host keeps slices within the table

  // ---- pair table: orientation, padded row counts, exclusive row starts ----
  int pcl = 0;
  int carry = 0;
  for (int k0 = 0; k0 < nk; k0 += kWave) {
    const int k = k0 + lane;
    int rows = 0;
    if (k < nk) {
      const int q = a.order ? a.order[kb + k] : kb + k;
      const int ia = a.pairs[2 * q], ib = a.pairs[2 * q + 1];
      const int lx = a.x.len[ia], ly = a.y.len[ib];
      const bool sw = ly > lx;
      const int lr = sw ? ly : lx, lc = sw ? lx : ly;
      const int64_t xo = a.x.off[ia] * d, yo = a.y.off[ib] * d;
      StreamPair e;
      e.q = q;
      e.lr1 = lr - 1;
      e.lc1 = lc - 1;
      e.swap = sw ? 1 : 0;
      e.roff = sw ? yo : xo;
      e.coff = sw ? xo : yo;
      tab[k] = e;
      if (lr >= 2 && lc >= 2) {
        rows = (lr + R - 1) / R * R;
        pcl = pcl > lc ? pcl : lc;
      } else {  // empty ΔΔ grid: K = (1, 0, ..., 0)
        double* o = a.out + (int64_t)q * (NLEV + 1);
        o[0] = 1.0;
        for (int m = 1; m <= NLEV; ++m) o[m] = 0.0;
      }
    }
    int incl = rows;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const int t = __shfl_up(incl, o, kWave);
      if (lane >= o) incl += t;
    }
    if (k < nk) tstart[k] = carry + incl - rows;
    carry += __shfl(incl, kWave - 1, kWave);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const int t = __shfl_xor(pcl, o, kWave);
    pcl = pcl > t ? pcl : t;
  }
  // wave-uniform values live in scalar registers
  const int total_rows = __builtin_amdgcn_readfirstlane(carry);
  if (lane == 0) tstart[nk] = total_rows;
  // the table is wave-private: a wave-level barrier orders its LDS writes
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (total_rows == 0) return;
  constexpr int kMinCols = GF ? kStreamMinColsGlobal : kStreamMinCols;
  int Pc = __builtin_amdgcn_readfirstlane(pcl > kMinCols ? pcl : kMinCols);
  const int nb = (total_rows + RB - 1) / RB;
  if (coop) {  // both waves take the same decision from the same table
    const int pcc = Pc > kStreamCoopMinCols ? Pc : kStreamCoopMinCols;
    if (pcc <= a.fifo_cols && nb >= 2) {
      Pc = pcc;
    } else {  // the first wave sweeps the slice alone
      coop = false;
      if (v == 1) return;
    }
  }
  // bands of this wave: all of them, or every other one of a two-wave slice
  const int nbw = coop ? (nb - v + 1) / 2 : nb;
  // global FIFO: this wave's [fifo_cols][NFG] region of the workspace (lane
  // 63 stores), and the region lane 0 loads (the other wave's when two share
  // the slice)
  T* gfifo = GF ? reinterpret_cast<T*>(a.gfifo) + (size_t)w * a.fifo_cols * NFG : nullptr;
  T* gfifo_in =
      GF ? reinterpret_cast<T*>(a.gfifo) + (size_t)(coop ? w0 + 1 - v : w) * a.fifo_cols * NFG
         : nullptr;

  const T gdot = a.gdot, gn = a.gn;
  const T* xdata = reinterpret_cast<const T*>(a.x.data);
  const T* ydata = reinterpret_cast<const T*>(a.y.data);

  // ---- per-lane band state ----
  // current band: column sequence, its last column, output row, flags
  const T* yptr;
  int ccl, q;
  bool prim, last;
  // next band (written by stage(), taken over at the lane's band change)
  const T* pyptr = nullptr;
  int pccl = 0, pq = 0;
  bool pprim = true, plast = false;
  // SAVE: the pair's boundary-state block, its local first row, Nr
  T* sv = nullptr;
  int sr0 = 0, snr = 0;
  T* psv = nullptr;
  int psr0 = 0, psnr = 0;

  // Stage band `band` for this lane: find the pair of its rows, write its x
  // rows (scaled, with the norm term) to the lane's private LDS slot and its
  // metadata to the pending registers.
  auto stage = [&](int band) {
    const int g0 = band * RB + R * lane;
    const bool live = g0 < total_rows;
    const int gs = live ? g0 : total_rows - 1;
    int lo = 0, hi = nk - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tstart[mid] <= gs) lo = mid; else hi = mid - 1;
    }
    const StreamPair e = tab[lo];
    const int r0 = g0 - tstart[lo];
    const int rows_k = tstart[lo + 1] - tstart[lo];
    pprim = !live || r0 == 0;
    plast = live && r0 + R >= rows_k;
    pq = e.q;
    pccl = e.lc1;
    pyptr = (e.swap ? xdata : ydata) + e.coff;
    if (SAVE) {
      psv = reinterpret_cast<T*>(a.save) + a.save_off[e.q];
      psr0 = r0;
      psnr = e.lr1;
    }
    const T* rb = (e.swap ? ydata : xdata) + e.roff;
#pragma unroll
    for (int j = 0; j < R; ++j) {
      int r = r0 + j;
      r = r < e.lr1 ? r : e.lr1;
      T v[DMAX];
      stream_load_row<DMAX, VEC>(v, rb + (int64_t)r * d, d);
      T n2 = T(0);
      T* dst = xst + (R * lane + j) * XS;
#pragma unroll
      for (int f = 0; f < DMAX; ++f) {
        n2 += v[f] * v[f];
        dst[f] = gdot * v[f];
      }
      dst[DMAX] = RBF ? gn * n2 : T(0);
    }
  };

  T xr[R][DMAX];
  T xn[R];
  // After its last band a lane keeps stepping (at most 62 steps, fewer than a
  // band) on a stale copy of that band's metadata: it must not count as the
  // end of a pair there.
  int lband = -1;
  auto enter_band = [&]() {
    ++lband;
    yptr = pyptr;
    ccl = pccl;
    q = pq;
    prim = pprim;
    last = plast && lband < nbw;
    if (SAVE) {
      sv = psv;
      sr0 = psr0;
      snr = psnr;
    }
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const T* src = xst + (R * lane + j) * XS;
#pragma unroll
      for (int f = 0; f < DMAX; ++f) xr[j][f] = src[f];
      xn[j] = src[DMAX];
    }
  };

  stage(coop ? v : 0);
  enter_band();

  T fl[NF];
#pragma unroll
  for (int k = 0; k < NF; ++k) fl[k] = T(0);
  T PS[R][NPS];
  T HB[R][NHB];
  T gleft[R];
#pragma unroll
  for (int j = 0; j < R; ++j) {
#pragma unroll
    for (int m = 0; m < NPS; ++m) PS[j][m] = T(0);
#pragma unroll
    for (int m = 0; m < NHB; ++m) HB[j][m] = T(0);
    gleft[j] = T(0);
  }
  T gdiag = T(0);
  T accf[NLEV];
  double acc[NLEV];
#pragma unroll
  for (int m = 0; m < NLEV; ++m) {
    accf[m] = T(0);
    acc[m] = 0.0;
  }

  int cp = -lane;             // this lane's column in its band (< 0: not started)
  const int total = nbw * Pc + kWave - 1;
  int s0 = 0;                 // lane 0's FIFO slot (= its column)
  int s63 = Pc - (kWave - 1); // lane 63's FIFO slot
  int next_stage = kWave - 1; // band b+1 is staged at step b*Pc + 63
  int stage_band = 1;
  // Issue priority from the wave's remaining steps, renewed once per band:
  // on a SIMD shared with short slices, the waves that hold the longest
  // pairs (the critical path of an uneven batch) issue first, and short
  // slices fill the gaps.  Results do not depend on it.
  auto set_prio = [&](int remaining) {  // wave-uniform
    if (remaining >= 4096) __builtin_amdgcn_s_setprio(3);
    else if (remaining >= 2048) __builtin_amdgcn_s_setprio(2);
    else if (remaining >= 1024) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
  };
  set_prio(total);

  // Prefetch cursor, two columns ahead of cp: column a_c of the sequence at
  // a_ptr (last column a_cl); it moves to the next band's sequence when a_c
  // wraps (the pending metadata is staged by then).
  int a_c = cp;
  const T* a_ptr = yptr;
  int a_cl = ccl;
  auto yload = [&](T* dst) {
    int cc = a_c > 0 ? a_c : 0;
    cc = cc < a_cl ? cc : a_cl;
    stream_load_row<DMAX, VEC>(dst, a_ptr + (uint32_t)cc * (uint32_t)(VEC ? DMAX : d), d);
    ++a_c;
    if (a_c == Pc) {
      a_c = 0;
      a_ptr = pyptr;
      a_cl = pccl;
    }
  };

  // FIFO slot of lane 0 two steps ahead (prefetched like the y rows)
  int s2 = 2;
  // lane 63 stored the slot Pc - 65 >= kStreamMinColsGlobal - 65 steps
  // earlier; each step issues a y-row load and a FIFO load at least
  constexpr int kFifoWait =
      fifo_wait_count(kStreamMinColsGlobal - (kWave + 1), (VEC ? DMAX / 4 : 1) + NFG / 4);
  auto fload = [&](T* dst) {
    if (GF) {
      wait_vmcnt<kFifoWait>();
#pragma unroll
      for (int k = 0; k < NFG; k += 4) {
        const float4 t = *reinterpret_cast<const float4*>(gfifo_in + s2 * NFG + k);
        dst[k] = t.x;
        if (k + 1 < NF) dst[k + 1] = t.y;
        if (k + 2 < NF) dst[k + 2] = t.z;
        if (k + 3 < NF) dst[k + 3] = t.w;
      }
    } else {
#pragma unroll
      for (int k = 0; k < NF; ++k) dst[k] = fifo[s2 * NFP + k];
    }
    ++s2;
    s2 = s2 == Pc ? 0 : s2;
  };

  // Two-wave slice hand-over: publish this wave's step count t (its FIFO
  // stores of steps < t complete first: release), then wait until the other
  // wave's published count covers the FIFO slots this wave loads in its next
  // kStreamCoopSync steps (two of them prefetched) and no longer reads a slot
  // this wave is about to overwrite.
  bool bail = false;
  auto coop_sync = [&](int t) {
    __hip_atomic_store(&prog[v], t, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    const int need = v == 0 ? t + kStreamCoopSync + 65 - Pc : t + kStreamCoopSync + 65;
    int spins = 0;
    while (!bail) {
      const int got = __builtin_amdgcn_readfirstlane(
          __hip_atomic_load(&prog[1 - v], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
      if (got >= need) break;
      __builtin_amdgcn_s_sleep(2);
      if (++spins >= kStreamCoopSpinLimit) bail = true;
    }
  };

  // Two register sets per prefetched stream: a step consumes its set, then
  // refills it with the data of the step after next.
  T y0[DMAX], y1[DMAX], f0[NFG], f1[NFG];
  auto step = [&](int T_, T* yv, T* fv) {
    if (coop && T_ > 0 && (T_ & (kStreamCoopSync - 1)) == 0) coop_sync(T_);  // wave-uniform
    if (T_ == next_stage) {  // wave-uniform
      if (stage_band < nbw) stage(coop ? v + 2 * stage_band : stage_band);
      ++stage_band;
      next_stage += Pc;
      set_prio(total - T_);
    }
    T in[NF];
    in[0] = shift_down(fl[0], fv[0]);
#pragma unroll
    for (int k = 1; k < NF; ++k) {
      const T t = shift_down(fl[k], fv[k]);
      in[k] = prim ? T(0) : t;
    }
    fload(fv);
    T yn = T(0);
    if (RBF) {
#pragma unroll
      for (int f = 0; f < DMAX; ++f) yn += yv[f] * yv[f];
      yn *= gn;
    }
    T g[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
      if (RBF) {
        T s = -(xn[j] + yn);
#pragma unroll
        for (int f = 0; f < DMAX; ++f) s += xr[j][f] * yv[f];
        g[j] = fast_exp2(s);
      } else {
        T s = T(0);
#pragma unroll
        for (int f = 0; f < DMAX; ++f) s += xr[j][f] * yv[f];
        g[j] = s;
      }
    }
    yload(yv);
    const bool keep = cp > 0;
    const T keepf = keep ? T(1) : T(0);
    T Mk[R];
    {
      T dprev = in[0] - gdiag;
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const T dj = g[j] - gleft[j];
        const T m = dj - dprev;
        const bool kj = j == 0 ? (keep && !prim) : keep;
        Mk[j] = kj ? m : T(0);
        dprev = dj;
      }
    }
#pragma unroll
    for (int j = 0; j < R; ++j) gleft[j] = g[j];
    gdiag = in[0];

    T cur[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) cur[k] = in[k];
#pragma unroll
    for (int j = 0; j < R; ++j) {
      T of[NF];
      of[0] = g[j];
      dp_cell_keep<T, NLEV, ORDER>(Mk[j], cur, of, PS[j], HB[j], keepf);
#pragma unroll
      for (int k = 0; k < NF; ++k) cur[k] = of[k];
    }
#pragma unroll
    for (int k = 0; k < NF; ++k) fl[k] = cur[k];
#pragma unroll
    for (int m = 0; m < NLEV; ++m) accf[m] += cur[1 + m];
    if (SAVE && NC > 0 && last && cp >= 1 && cp <= ccl) {
      // final column sums C_a(Nr, j), a = 1..NLEV-1, j = cp - 1
      T* cs = sv + NC * snr;
#pragma unroll
      for (int m = 0; m < NC; ++m) cs[m * ccl + cp - 1] = cur[1 + m];
    }
    if (lane == kWave - 1 && T_ >= kWave - 1) {
      if (GF) {
        // read back by lane 0 at least Pc - 65 steps later, after a vmcnt
        // wait that completes this store (fload)
#pragma unroll
        for (int k = 0; k < NFG; k += 4) {
          float4 t;
          t.x = cur[k];
          t.y = k + 1 < NF ? cur[k + 1] : T(0);
          t.z = k + 2 < NF ? cur[k + 2] : T(0);
          t.w = k + 3 < NF ? cur[k + 3] : T(0);
          *reinterpret_cast<float4*>(gfifo + s63 * NFG + k) = t;
        }
      } else {
#pragma unroll
        for (int k = 0; k < NF; ++k) fifo[s63 * NFP + k] = cur[k];
      }
    }
    if ((T_ & 15) == 15) {  // wave-uniform flush of the FP32 partial sums
#pragma unroll
      for (int m = 0; m < NLEV; ++m) {
        acc[m] += (double)accf[m];
        accf[m] = T(0);
      }
    }
    ++cp;
    if (cp == Pc) {  // this lane's band pass ends
      if (SAVE && NC > 0) {
        // row-end prefixes P_a(i, Nc) of the lane's rows (ΔΔ row i = local row - 1)
#pragma unroll
        for (int j = 0; j < R; ++j) {
          const int r = sr0 + j;
          if (r >= 1 && r <= snr) {
#pragma unroll
            for (int m = 0; m < NC; ++m) sv[m * snr + r - 1] = PS[j][m];
          }
        }
      }
      if (last) {
        double* o = a.out + (int64_t)q * (NLEV + 1);
        o[0] = bail ? __builtin_nan("") : 1.0;
#pragma unroll
        for (int m = 0; m < NLEV; ++m) o[m + 1] = acc[m] + (double)accf[m];
      }
      cp = 0;
      enter_band();
      if (last) {  // only the accumulators of a pair-ending pass are read
#pragma unroll
        for (int m = 0; m < NLEV; ++m) {
          acc[m] = 0.0;
          accf[m] = T(0);
        }
      }
    }
    ++s0;
    s0 = s0 == Pc ? 0 : s0;
    ++s63;
    s63 = s63 == Pc ? 0 : s63;
  };

  yload(y0);
  yload(y1);
  // lane 0 starts band 0 with a priming row: slots 0 and 1 are read but unused
  // (a two-wave slice's second wave waits for its first FIFO slots here)
  if (coop) coop_sync(0);
  s2 = 0;
  fload(f0);
  fload(f1);
  int T_ = 0;
  for (; T_ + 2 <= total; T_ += 2) {
    step(T_, y0, f0);
    step(T_ + 1, y1, f1);
  }
  if (T_ < total) step(T_, y0, f0);
  if (coop)  // done: the other wave no longer waits for this one
    __hip_atomic_store(&prog[v], 0x7fffffff, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <int NLEV, int ORDER, int DMAX, int KIND, bool VEC, bool GF, bool SAVE = false>
StreamEntry stream_entry() {
  constexpr int R = stream_rows(NLEV, ORDER, DMAX);
  return StreamEntry{(const void*)dp_stream_kernel<NLEV, ORDER, DMAX, R, KIND, VEC, GF, SAVE>, R};
}

}  // namespace ksig
