// Shapes of the low-rank level scans (lowrank_scan.hip), shared by the
// kernels and the host entry points that decide what is built.
#pragma once

namespace ksig {

// Register state of the diagonal scan (lowrank_diag_scan_kernel): level i
// holds q^(i+1) entries per component, the running sums of every level
// diag_total(q, nl) doubles, and at order p > 1 the terms of two adjacent
// levels are live at once (diag_state).  Shapes above kDiagStateMax doubles
// are not built (they would spill); the scan keeps the loads of up to 4 time
// steps in flight in a register ring, as deep as the registers left over
// allow (one thread per sequence and component: about one wave per SIMD at
// 1000 sequences, so the walk is bound by load latency unless several steps
// are outstanding; 0.428 -> 0.323 ms at 1000 x 199 steps, 50 components, 5
// levels, q = 2).
constexpr int kDiagStateMax = 96;
// (loops, not recursion: the kernel calls these with unrolled loop indices,
// which the compiler folds only through inlinable code)
__host__ __device__ constexpr int ipow(int b, int e) {
  int r = 1;
  for (int i = 0; i < e; ++i) r *= b;
  return r;
}
// entries of levels 0..nl-1: the offset of level nl's block
__host__ __device__ constexpr int diag_total(int q, int nl) {
  int t = 0, p = 1;
  for (int i = 0; i < nl; ++i) {
    p *= q;
    t += p;
  }
  return t;
}
// entries of level i's terms at order p: min(i+1, p) orders of q^(i+1)
__host__ __device__ constexpr int diag_terms(int q, int i, int p) {
  return (i + 1 < p ? i + 1 : p) * ipow(q, i + 1);
}
// the largest non-top level's terms (at least level 0's q entries)
__host__ __device__ constexpr int diag_rbuf(int q, int nl, int p) {
  if (p <= 1) return 1;
  int r = q;
  for (int i = 0; i + 1 < nl; ++i) r = r > diag_terms(q, i, p) ? r : diag_terms(q, i, p);
  return r;
}
// running sums plus, at order p > 1, the terms of two adjacent levels
__host__ __device__ constexpr int diag_state(int q, int nl, int p) {
  int pair = q;
  if (p > 1)
    for (int i = 1; i < nl; ++i) {
      const int v = diag_terms(q, i - 1, p) + (i < nl - 1 ? diag_terms(q, i, p) : 0);
      pair = pair > v ? pair : v;
    }
  return diag_total(q, nl) + (p <= 1 ? 0 : pair);
}
__host__ __device__ constexpr bool diag_fits(int q, int nl, int p) {
  return q >= 1 && q <= 4 && nl >= 1 && nl <= 5 && p >= 1 && p <= nl &&
         diag_state(q, nl, p) <= kDiagStateMax;
}
// ring depth: 2 VGPRs per double of state plus the ring within 224 VGPRs
__host__ __device__ constexpr int diag_ahead(int q, int nl, int p, int tsz) {
  return 2 * diag_state(q, nl, p) + 4 * q * nl * (tsz / 4) <= 224   ? 4
         : 2 * diag_state(q, nl, p) + 2 * q * nl * (tsz / 4) <= 224 ? 2
                                                                     : 1;
}

// LDS rows (2m doubles each) of the TensorSketch scan for nl levels at the
// given order: e_0, the running sums of levels 0..nl-2 and, at order p > 1,
// this step's terms of levels 0..nl-2 (min(i+1, p) rows for level i)
__host__ __device__ inline int sketch_rows(int nl, int order) {
  int rows = nl;  // e_0 and C_0..C_{nl-2}
  if (order > 1)
    for (int i = 0; i + 1 < nl; ++i) rows += i + 1 < order ? i + 1 : order;
  return rows;
}

}  // namespace ksig
