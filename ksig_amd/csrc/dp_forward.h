// Internal launch interface of the DP and PDE sweeps (not part of the C-ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ksig_common.h"

// (n_levels, order) combinations compiled into the library.  order is the
// normalised order of ksig/algorithms.py:37 (1 <= order <= n_levels).
#define KSIG_DP_LEVEL_ORDER_TABLE(X)                                              \
  X(1, 1)                                                                         \
  X(2, 1) X(2, 2)                                                                 \
  X(3, 1) X(3, 2) X(3, 3)                                                         \
  X(4, 1) X(4, 2) X(4, 3) X(4, 4)                                                 \
  X(5, 1) X(5, 2) X(5, 3) X(5, 4) X(5, 5)                                         \
  X(6, 1) X(6, 2) X(6, 3) X(6, 6)                                                 \
  X(7, 1) X(7, 2) X(7, 3)                                                         \
  X(8, 1) X(8, 2) X(8, 3)                                                         \
  X(9, 1) X(9, 2)                                                                 \
  X(10, 1) X(10, 2)                                                               \
  X(12, 1) X(12, 2)

namespace ksig {

struct DpFwdArgs {
  SeqSet x, y;
  int d;
  const int32_t* pairs;  // [n_pairs, 2] (index into x, index into y)
  int64_t n_pairs;
  int kind;              // StaticKind
  double xscale, yscale; // per-side input scaling folded into the static kernel
  int n_levels, order, difference;
  const void* gram;      // kind == kGram: [n_pairs, gram_lx, gram_ly]
  int gram_lx, gram_ly;
  // kind == kGram, general layout: matrix p at element offset gram_off[p]
  // (null: p * gram_lx * gram_ly), rows gram_ld elements apart (0: gram_ly),
  // per-pair sizes gram_len[2p], gram_len[2p+1] (null: gram_lx, gram_ly), and
  // the static-kernel value formed from the stored value v as
  //   gram_tf == kGramRbfDots: exp2(v - hx[hx_off[p] + i] - hy[hy_off[p] + j])
  //   gram_tf == kGramPlain:    v
  const int64_t* gram_off;
  int64_t gram_ld;
  const int32_t* gram_len;
  int gram_tf;
  const void* gram_hx;
  const int64_t* gram_hx_off;
  const void* gram_hy;
  const int64_t* gram_hy_off;
  double* out;           // [n_pairs, n_levels + 1]
  // Optional boundary state for the adjoint (order 1, difference only):
  // per pair at save_off[p], [n_levels-1][Nr] row-end prefixes P_a(i, Nc)
  // then [n_levels-1][Nc] final column sums C_a(Nr, j), in the kernel's
  // row/column orientation (rows = the longer sequence).
  void* save;
  const int64_t* save_off;
  // MFMA-Gram forward (dp_fwd_mfma.h): band FIFO workspace, one region of
  // fifo_cols 16-byte-quad slots per pair
  void* fifo;
  int fifo_cols;
};

// Row-stream forward (dp_stream.h): each wave takes a contiguous slice of the
// processing order, [wave_beg[w], wave_beg[w+1]) or [w*per_wave, ...) when
// wave_beg is null, at most kStreamMaxPairs pairs.
struct DpStreamArgs {
  SeqSet x, y;
  int d;
  int tab_cap;             // pair-table entries in LDS (>= every slice, multiple of 4)
  int fifo_cols;           // FIFO slots (>= every wave's column count)
  void* gfifo;             // global-memory FIFO, [n_waves][fifo_cols][4-rounded flows] floats
  const int32_t* pairs;    // [n_pairs, 2]
  int64_t n_pairs;
  const int32_t* order;    // processing order of the pair list (null: identity)
  const int32_t* wave_beg; // [n_waves + 1] or null
  int n_waves;             // slices; block b runs slices 2b and 2b + 1
  int per_wave;
  int lds_region;          // dynamic LDS bytes of one wave (16-byte multiple)
  float gdot, gn;          // static-kernel factors folded into the row side
  double* out;             // [n_pairs, n_levels + 1]
  // training forward (order 1): the adjoint's boundary state, as in DpFwdArgs
  void* save;
  const int64_t* save_off;
};

// A compiled row-stream kernel: entry point and rows per lane.
struct StreamEntry {
  const void* fn;
  int r;
};

struct DpBwdArgs {
  SeqSet x, y;
  int d;
  const int32_t* pairs;
  int64_t n_pairs;
  int kind;
  double xscale, yscale;
  int n_levels;
  const double* gout;      // [n_pairs, n_levels + 1] d loss / d K_m
  const void* save;        // boundary state written by the forward
  const int64_t* save_off;
  void* omega;             // per pair [len_x, len_y] at omega_off[p]: dL/dG * dG/d(arg)
  const int64_t* omega_off;
  double* gparam;          // [n_pairs]: sum omega * D' (RBF) or omega * G (linear)
  void* fifo;              // MFMA variant: band FIFO workspace, [n_pairs][fifo_cols][4-rounded flows]
  int fifo_cols;
};

struct PdeFwdArgs {
  SeqSet x, y;
  int d;
  const int32_t* pairs;
  int64_t n_pairs;
  int kind;
  double xscale, yscale;
  int dyadic, difference;
  const void* gram;
  int gram_lx, gram_ly;
  double* out;           // [n_pairs]
  // Optional edges for the rebuilding adjoint (kernel orientation): U of the
  // last refined row [Ncr] then of the last refined column [Nrr] at e_off[p].
  void* edges;
  const int64_t* e_off;
};

// Contraction of a per-pair Gram adjoint omega with the inputs (the chain
// rule through the static kernel); gx / gy may be null.
struct ContractArgs {
  SeqSet x, y;
  int d;
  const int32_t* pairs;
  int64_t n_pairs;
  const void* omega;          // per pair [len_x, len_y] at omega_off[p]
  const int64_t* omega_off;
  double c_cross, c_self;
  void* gx;
  void* gy;
  int group;                  // consecutive pairs per workgroup (set by the launcher)
  // signature-PDE chain rule (mode 1: with difference, 2: without): omega is
  // dL/dM per pair in the adjoint's orientation and omega is built from it
  int mode;
  int kind;
  double gsc;                 // RBF: log2(e) / max(2 sigma^2, 1e-12)
  double gp_scale;            // factor of the accumulated dL/dparam
  double* gparam;             // may be null
};

// Adjoint of the PDE sweep that rebuilds U from the saved edges.
struct PdeBwdEdgeArgs {
  SeqSet x, y;
  int d;
  const int32_t* pairs;
  int64_t n_pairs;
  int kind;
  double xscale, yscale;
  int dyadic, difference;
  const double* gout;    // [n_pairs] dL/dK
  const void* edges;     // as PdeFwdArgs::edges
  const int64_t* e_off;
  void* gm;              // [Nr][Nc] at m_off[p]: dL/dM of every original cell
  const int64_t* m_off;
  void* fifo;            // band FIFO workspace: [n_pairs][fifo_cols] 16-byte-quad slots
  int fifo_cols;
};

// Signature-PDE adjoint fused with the chain rule (pde_bwd_fused.hip):
// gradients of the caller's x / y rows (either may be null), the parameter
// part per pair, the band FIFO workspace.
struct PdeBwdFusedArgs {
  SeqSet x, y;
  int d;
  const int32_t* pairs;
  int64_t n_pairs;
  int kind;
  double xscale, yscale;
  int dyadic;
  const double* gout;    // [n_pairs] dL/dK
  const void* edges;     // as PdeFwdArgs::edges
  const int64_t* e_off;
  void* gx;
  void* gy;
  double* gp;            // [n_pairs] (may be null)
  double c_cross, c_self, gp_scale;
  void* fifo;            // [n_pairs][fifo_cols] slots of pde_fused_slot_bytes(d)
  int fifo_cols;
};

int dp_flow_count(int nlev, int order);
// fused VALU sweep for difference=1 (dp_sweep.h): LDS bytes for a column side of `cols`
size_t dp_sweep_lds(int dtype, int d, int cols, int n_levels, int order);
hipError_t dp_sweep_launch(const DpFwdArgs& a, int dtype, size_t lds, hipStream_t s);
bool dp_supported(int n_levels, int order);
// any (n_levels, order) on caller matrices (dp_generic.hip): LDS bytes for
// matrices of at most max_cols columns, and the launch
size_t dp_generic_lds(int dtype, int n_levels, int order, int max_cols);
hipError_t dp_generic_launch(const DpFwdArgs& a, int dtype, size_t lds, hipStream_t s);
hipError_t dp_forward_launch(const DpFwdArgs& a, int dtype, size_t lds, hipStream_t s);
size_t dp_backward_lds(int dtype, int d, int cols, int n_levels);
hipError_t dp_backward_launch(const DpBwdArgs& a, int dtype, size_t lds, hipStream_t s);
bool dp_backward_supported(int n_levels);
// float32 adjoint with the static Gram on MFMA (dp_backward_mfma.h)
size_t dp_backward_mfma_lds(int d, int cols, int n_levels);
// float32 first-order forward with the static Gram on MFMA (dp_fwd_mfma.h)
size_t dp_fwd_mfma_lds();
int dp_fwd_mfma_fifo_cols(int cols);
size_t dp_fwd_mfma_fifo_bytes(int cols, int n_levels);
bool dp_fwd_mfma_supported(int d, int n_levels);
hipError_t dp_fwd_mfma_launch(const DpFwdArgs& a, bool save, size_t lds, hipStream_t s);
// FIFO slots per pair and workspace bytes per pair of the MFMA variant
int dp_backward_mfma_fifo_cols(int cols);
size_t dp_backward_mfma_fifo_bytes(int cols, int n_levels);
hipError_t dp_backward_mfma_launch(const DpBwdArgs& a, size_t lds, hipStream_t s);
bool dp_mfma_supported(int n_levels, int order);
// mode 1: float32 MFMA Gram, mode 2: bf16 hi/lo split MFMA Gram
hipError_t dp_mfma_launch(const DpFwdArgs& a, int mode, size_t lds, hipStream_t s);
// vec: d equals the bucket width and both data pointers are 16-byte aligned;
// global_fifo: band hand-over through a global-memory workspace (vec only)
StreamEntry dp_stream_entry(int kind, int dmax, bool vec, bool global_fifo, int n_levels,
                            int order);
// training forward (order 1, vector loads, LDS FIFO) that also writes the
// adjoint's boundary state
StreamEntry dp_stream_save_entry(int kind, int dmax, int n_levels);
// wpb: waves per block (1, or 2 with the global FIFO: two-wave slices); lds is
// the block's dynamic LDS; the occupancy is in waves per CU
hipError_t dp_stream_launch(const StreamEntry& e, const DpStreamArgs& a, size_t lds, int wpb,
                            hipStream_t s);
int dp_stream_occupancy(const StreamEntry& e, size_t lds, int wpb);
hipError_t pde_forward_launch(const PdeFwdArgs& a, int dtype, size_t lds, hipStream_t s);
// FIFO slots (band pass length) of the phase-aligned PDE forward (dyadic
// order 0 and 1) for nc original columns: a multiple of 2^lam, at least
// 64 * 2^lam + 2.
__host__ __device__ inline int pde_aligned_pc(int nc, int lam) {
  const int f = 1 << lam;
  int pc = (nc + 1) * f;
  if (pc < 64 * f + 2) pc = 64 * f + 2;
  return (pc + f - 1) / f * f;
}
// Band pass of the phase-aligned PDE adjoint: the forward's, and at least
// 64 * 2^lam + 32 steps, so that lane 63 writes a global-memory FIFO slot
// 2^lam + 30 steps before lane 0 reads it.
__host__ __device__ inline int pde_bwd_aligned_pc(int nc, int lam) {
  const int f = 1 << lam;
  const int pc = pde_aligned_pc(nc, lam);
  const int lo = 64 * f + 32;
  return pc >= lo ? pc : (lo + f - 1) / f * f;
}
// Band pass of the fused adjoint: priming column, Nc columns and one more
// (the static-kernel column 0), at least 64 * 2^lam + 32 steps (lane 63
// writes a global-memory FIFO slot 2^lam + 30 steps before lane 0 reads it).
__host__ __device__ inline int pde_fused_pc(int nc, int lam) {
  const int f = 1 << lam;
  int pc = (nc + 2) * f;
  if (pc < 64 * f + 32) pc = 64 * f + 32;
  return (pc + f - 1) / f * f;
}
int pde_fused_dmax(int d);  // 4 or 8, 0 when d > 8
size_t pde_fused_lds(int dtype, int d, int dyadic);
size_t pde_fused_slot_bytes(int d);
hipError_t pde_bwd_fused_launch(const PdeBwdFusedArgs& a, int dtype, size_t lds, hipStream_t s);
size_t pde_bwd_edges_lds(int dtype, int d, int ly, int ncr, int nrr, int dyadic);
int pde_bwd_fifo_cols(int ncr, int dyadic);
size_t pde_bwd_fifo_bytes(int dtype, int ncr, int dyadic);
hipError_t pde_bwd_edges_launch(const PdeBwdEdgeArgs& a, int dtype, size_t lds, hipStream_t s);
hipError_t grad_contract_launch(const ContractArgs& a, bool f64, int max_len_x, int max_len_y,
                                int n_cus, hipStream_t s);

}  // namespace ksig
