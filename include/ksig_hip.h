/*
 * ksig_hip.h — C ABI of the MI355X (gfx950) signature-kernel library
 * (libksig_hip.so).  Plain pointers and sizes only; every pointer argument
 * named x, y, pairs, out, grad_* or M is a DEVICE pointer, and `stream` is a
 * hipStream_t (NULL = the default stream).  All calls are asynchronous on
 * `stream`; none allocates device memory or synchronises.
 *
 * Sequence sets are packed rows: sequence a occupies rows
 * [off[a], off[a] + len[a]) of a row-major [total_rows, d] array.  A dense
 * [n, L, d] array is the case off[a] = a*L, len[a] = L.
 *
 * Return value: 0 on success, otherwise a KSIG_E* code or a hipError_t;
 * ksig_last_error() describes the last failure of the calling thread.
 */
#ifndef KSIG_HIP_H_
#define KSIG_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* dtype codes */
#define KSIG_F32 0
#define KSIG_F64 1

/* static kernels */
#define KSIG_STATIC_RBF 0    /* ksig/static/kernels.py:245-260, param = bandwidth */
#define KSIG_STATIC_LINEAR 1 /* ksig/static/kernels.py:134-166, param = scale */

/* static-kernel Gram evaluation inside the DP sweep (float32 inputs) */
#define KSIG_GRAM_AUTO -1         /* the library's choice: the VALU sweep (faster than both MFMA forms at d <= 32) */
#define KSIG_GRAM_VALU 0          /* one float32 (or float64) VALU dot product per cell */
#define KSIG_GRAM_MFMA_F32 1      /* v_mfma_f32_16x16x4f32 tiles */
#define KSIG_GRAM_MFMA_BF16X3 2   /* bf16 hi/lo split, 3 products, v_mfma_f32_16x16x32_bf16 */

/* error codes (besides hipError_t values) */
#define KSIG_EINVAL 1001      /* bad argument (shape, dtype, kind, null pointer) */
#define KSIG_EUNSUPPORTED 1002 /* (n_levels, order) or d outside the built set */
#define KSIG_ELDS 1003        /* sequence too long for the per-pair LDS state */

/* Library version string. */
const char* ksig_version(void);

/* Description of the last error raised on the calling thread. */
const char* ksig_last_error(void);

/* 1 if (n_levels, order) is compiled into the DP kernels (order already
 * normalised as in ksig/algorithms.py:37), else 0. */
int ksig_dp_supported(int n_levels, int order);

/*
 * Truncated signature kernel, level values, for a list of sequence pairs.
 *
 * Replaces, per pair, the composition used by SignatureKernel._compute_kernel
 * (ksig/kernels.py:147-164): the static Gram of ksig/static/kernels.py:248-260
 * (RBF) or :145-155 (linear) followed by signature_kern(M, n_levels, order,
 * difference, return_levels=True) of ksig/algorithms.py:22-140.
 *
 *   out[p, m] = K_m(x_{pairs[p,0]}, y_{pairs[p,1]}),  m = 0..n_levels,
 *   with K_0 = 1, as float64.
 *
 * max_len_x / max_len_y bound the sequence lengths (they size LDS).
 * dtype selects float32 or float64 inputs and arithmetic; gram_mode selects
 * how the static-kernel values are produced (KSIG_GRAM_*; float64 and
 * difference=0 always use KSIG_GRAM_VALU).
 */
int ksig_signature_levels(int dtype, const void* x, const int64_t* x_off,
                          const int32_t* x_len, int max_len_x, const void* y,
                          const int64_t* y_off, const int32_t* y_len,
                          int max_len_y, int d, const int32_t* pairs,
                          int64_t n_pairs, int static_kind, double static_param,
                          int n_levels, int order, int difference, int gram_mode,
                          double* out, void* stream);

/*
 * Resident-wave capacity of the row-stream kernel behind
 * ksig_signature_levels_batched for a problem shape on the current device,
 * with slices of at most max_slice pairs (<= 0: the largest accepted) and the
 * band hand-over FIFO in LDS (global_fifo = 0) or in a global-memory
 * workspace (global_fifo = 1; long sequences, where an LDS FIFO would cut
 * occupancy): *slots = waves that run concurrently (0 if no row-stream kernel
 * or LDS layout fits), *max_pairs_per_wave = the largest slice one wave
 * accepts, *workspace_per_wave = workspace bytes per launched wave (0 for the
 * LDS FIFO).  A scheduling query with no reference counterpart (the
 * reference launches one CuPy kernel sequence per Gram,
 * ksig/kernels.py:147-164).
 */
int ksig_stream_slots(int d, int static_kind, int n_levels, int order, int max_len_x,
                      int max_len_y, int max_slice, int global_fifo, int* slots,
                      int* max_pairs_per_wave, size_t* workspace_per_wave);

/*
 * ksig_signature_levels for float32 inputs with difference=1 and a caller
 * schedule, in one persistent launch: slice w is the pairs
 * sched_order[wave_beg[w] .. wave_beg[w+1]) (sched_order null: the identity),
 * n_waves slices of at most max_slice pairs each (1 <= max_slice <= the
 * *max_pairs_per_wave of ksig_stream_slots); the wave that sweeps a slice
 * writes out[q] for each pair index q it takes.  A wave stacks its pairs'
 * rows and sweeps them with one column count (the longest shorter side among
 * them), so slices of similar lengths waste least.  wave_beg null: slices of
 * max_slice consecutive pairs of the order.  workspace non-null selects the
 * global-memory FIFO and must hold n_waves * *workspace_per_wave bytes
 * (16-byte aligned; contents are scratch); null keeps the FIFO in LDS.  With
 * the global FIFO a launch runs two waves per block (slices 2b and 2b + 1),
 * and an empty slice 2b + 1 after a non-empty slice 2b makes both waves
 * sweep slice 2b, every other band each (slices of at least 224 columns;
 * shorter ones are swept by the first wave alone).  Replaces the same reference
 * composition as ksig_signature_levels (ksig/kernels.py:147-164 over
 * ksig/algorithms.py:22-140) for pair lists the reference cannot take
 * (uneven lengths; the reference needs one dense [n, L, d] array,
 * ksig/kernels.py:87).
 */
int ksig_signature_levels_batched(const void* x, const int64_t* x_off, const int32_t* x_len,
                                  int max_len_x, const void* y, const int64_t* y_off,
                                  const int32_t* y_len, int max_len_y, int d,
                                  const int32_t* pairs, int64_t n_pairs, int static_kind,
                                  double static_param, int n_levels, int order,
                                  const int32_t* sched_order, const int32_t* wave_beg,
                                  int n_waves, int max_slice, void* workspace,
                                  size_t workspace_bytes, double* out, void* stream);

/* Size of the device scratch array behind ksig_signature_levels_batched_noalloc. */
#define KSIG_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

/*
 * Allocation-free variant of ksig_signature_levels_batched with the band FIFO
 * in global memory (the long-sequence choice), for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation).  No counterpart in the reference.
 * Same arguments and results as ksig_signature_levels_batched with a
 * workspace, except that the FIFO lives in a device array of the library
 * (KSIG_NOALLOC_SCRATCH_BYTES = 64 MiB per GPU, held until the process ends,
 * its device address looked up on the first call on each GPU).  Fails with
 * KSIG_EINVAL, launching nothing, when the FIFO of n_waves slices needs more
 * than the scratch (n_waves * the *workspace_per_wave of ksig_stream_slots
 * with global_fifo = 1).  Needs d in {8, 16, 32} and 16-byte aligned inputs.
 * Every call on one GPU shares that GPU's scratch: calls that overlap on one
 * GPU (concurrent streams or host threads) race and must be serialised by
 * the caller.  The FIFO contents need no clearing between calls.  Calls no
 * hipMalloc, hipFree or allocating library routine.
 */
int ksig_signature_levels_batched_noalloc(const void* x, const int64_t* x_off,
                                          const int32_t* x_len, int max_len_x, const void* y,
                                          const int64_t* y_off, const int32_t* y_len,
                                          int max_len_y, int d, const int32_t* pairs,
                                          int64_t n_pairs, int static_kind, double static_param,
                                          int n_levels, int order, const int32_t* sched_order,
                                          const int32_t* wave_beg, int n_waves, int max_slice,
                                          double* out, void* stream);

/* KSIG_NOALLOC_SCRATCH_BYTES, as compiled into the library. */
size_t ksig_noalloc_scratch_bytes(void);

/*
 * Training forward: ksig_signature_levels for order 1 with difference, which
 * also writes the boundary state the adjoint needs.  For pair p, save holds
 * (n_levels-1) * (len_x + len_y - 2) values of dtype at element offset
 * save_off[p] (row-end prefixes of the recursion and its final column sums).
 * save may be NULL: levels only.  Same reference interface as
 * ksig_signature_levels; the state itself has no reference counterpart (the
 * reference has no gradients).
 */
int ksig_signature_levels_fwd_save(int dtype, const void* x, const int64_t* x_off,
                                   const int32_t* x_len, int max_len_x, const void* y,
                                   const int64_t* y_off, const int32_t* y_len,
                                   int max_len_y, int d, const int32_t* pairs,
                                   int64_t n_pairs, int static_kind, double static_param,
                                   int n_levels, double* out, void* save,
                                   const int64_t* save_off, void* workspace,
                                   size_t workspace_bytes, void* stream);

/*
 * Workspace bytes with which ksig_signature_levels_fwd_save takes its fastest
 * kernel for the shape (the static Gram on MFMA: float32, 8 < d <= 32); 0
 * when no workspace is used.  The workspace is optional: NULL or smaller
 * selects a kernel without one.  Contents are scratch.  No reference
 * counterpart.
 */
size_t ksig_signature_levels_fwd_save_workspace(int dtype, int d, int n_levels,
                                                int max_len_x, int max_len_y,
                                                int64_t n_pairs);

/*
 * Adjoint of ksig_signature_levels_fwd_save (no reference counterpart: the
 * reference, ksig/algorithms.py:47-85, is forward-only).  Given
 * grad_levels[p, m] = dL/dK_m (m = 0..n_levels), writes for every pair
 *   omega[omega_off[p] + a*len_y + b] = dL/dG(a, b) * dG/dz
 * over the static-kernel Gram G(a, b) = k(x_a, y_b) (times G for RBF, plain
 * dL/dG for linear), and grad_param[p] = sum omega * D'(a,b) (RBF,
 * D' = log2(e) |x_a - y_b|^2 / max(2 sigma^2, 1e-12)) or sum omega * G (linear);
 * grad_param may be NULL when no parameter gradient is wanted (the float32
 * adjoint then carries one flow fewer per step).
 * The caller contracts omega with the inputs to obtain dL/dx, dL/dy and the
 * bandwidth / scale gradient (ksig_static_grad_contract).  workspace: device
 * scratch of at least ksig_signature_levels_bwd_workspace(...) bytes, 16-byte
 * aligned (may be NULL when that is 0).
 */
int ksig_signature_levels_bwd(int dtype, const void* x, const int64_t* x_off,
                              const int32_t* x_len, int max_len_x, const void* y,
                              const int64_t* y_off, const int32_t* y_len,
                              int max_len_y, int d, const int32_t* pairs,
                              int64_t n_pairs, int static_kind, double static_param,
                              int n_levels, const double* grad_levels, const void* save,
                              const int64_t* save_off, void* omega,
                              const int64_t* omega_off, double* grad_param,
                              void* workspace, size_t workspace_bytes, void* stream);

/*
 * Workspace bytes ksig_signature_levels_bwd needs for n_pairs pairs of at most
 * max_len_x x max_len_y points (0 when the chosen adjoint keeps all of its
 * state on chip: float64, and float32 with d <= 16).  The workspace holds the
 * adjoint's band hand-over state, one region per pair; contents are scratch.
 * No reference counterpart (a sizing query of this library).
 */
size_t ksig_signature_levels_bwd_workspace(int dtype, int d, int n_levels, int max_len_x,
                                           int max_len_y, int64_t n_pairs);

/*
 * Chain rule of the signature-PDE adjoint through ΔΔ and the static kernel
 * (no reference counterpart: ksig/algorithms.py:402-446 and the static kernels
 * of ksig/static/kernels.py:134-166, 245-260 are forward-only).  For every
 * pair p = (i, j), grad_m + m_off[p] holds dL/dM of
 * ksig_signature_pde_bwd_edges (dtype, [Nr][Nc] in the kernel's orientation:
 * rows = the longer sequence, Nr/Nc = lengths - difference).  With
 * dL/dG = ΔΔ^T(dL/dM) (difference=1) or dL/dM (difference=0) over the static
 * Gram G(a, b) = k(x_a, y_b), this ADDS
 *   grad_x[x_off[i] + a] += sum_b dL/dG(a,b) dG(a,b)/dx_a,
 *   grad_y[y_off[j] + b] += sum_a dL/dG(a,b) dG(a,b)/dy_b,
 *   *grad_param += sum over all pairs of dL/dG dG/dparam
 * (RBF: param = sigma; linear: param = scale).  Any of grad_x, grad_y,
 * grad_param may be NULL; grad_x and grad_y may alias.  Device atomics: the
 * result is order-dependent in its last bits.  Any d >= 1 (float32: MFMA tiles up
 * to d = 64, a VALU reduction beyond).
 */
int ksig_signature_pde_grad_contract(int dtype, const void* x, const int64_t* x_off,
                                     const int32_t* x_len, int max_len_x, const void* y,
                                     const int64_t* y_off, const int32_t* y_len,
                                     int max_len_y, int d, const int32_t* pairs,
                                     int64_t n_pairs, int static_kind, double static_param,
                                     int difference, const void* grad_m,
                                     const int64_t* m_off, void* grad_x, void* grad_y,
                                     double* grad_param, void* stream);

/*
 * Chain rule through the static kernel (no reference counterpart: the
 * reference's static kernels, ksig/static/kernels.py:134-166 and 245-260, have
 * no gradients).  For every pair p = (i, j) with omega_p = omega + omega_off[p]
 * of shape [len_x[i], len_y[j]] (row-major), adds
 *   grad_x[x_off[i] + a] += c_cross * sum_b omega_p[a, b] y_b
 *                         + c_self * (sum_b omega_p[a, b]) x_a
 *   grad_y[y_off[j] + b] += c_cross * sum_a omega_p[a, b] x_a
 *                         + c_self * (sum_a omega_p[a, b]) y_b
 * (d-vectors, dtype of x).  RBF: c_cross = -c_self = 1 / sigma^2,
 * omega = dL/dG * G; linear: c_cross = scale, c_self = 0, omega = dL/dG.
 * Either gradient may be NULL; both may alias (x and y the same set).
 * Accumulates with device atomics: the result is order-dependent in its last
 * bits.  Any d >= 1 (float32: MFMA tiles up to d = 64, a VALU reduction beyond).
 */
int ksig_static_grad_contract(int dtype, const void* x, const int64_t* x_off,
                              const int32_t* x_len, int max_len_x, const void* y,
                              const int64_t* y_off, const int32_t* y_len, int max_len_y,
                              int d, const int32_t* pairs, int64_t n_pairs,
                              const void* omega, const int64_t* omega_off, double c_cross,
                              double c_self, void* grad_x, void* grad_y, void* stream);

/*
 * Truncated signature kernel on caller-supplied static-kernel matrices.
 *
 * Replaces signature_kern(M, n_levels, order, difference, return_levels=True)
 * of ksig/algorithms.py:22-44 for M of shape [n_mats, lx, ly] (contiguous,
 * any leading shape flattened): out[p, m] = K_m of matrix p, float64.
 */
int ksig_signature_levels_gram(int dtype, const void* M, int64_t n_mats,
                               int lx, int ly, int n_levels, int order,
                               int difference, double* out, void* stream);

/*
 * Truncated signature kernel from static-kernel dot products computed by a
 * library GEMM (matrix cores, K = d): the route for inputs the fused kernels
 * do not take (d > 32, or an (n_levels, order) outside the compiled table,
 * see ksig_dp_supported).  Replaces, per matrix, the composition of
 * SignatureKernel._compute_kernel (ksig/kernels.py:147-164): the static Gram
 * of ksig/static/kernels.py:248-260 (RBF) or :145-155 (linear), whose inner
 * products ksig/utils.py:105-157 forms with a GEMM, followed by
 * signature_kern(M, n_levels, order, difference, return_levels=True) of
 * ksig/algorithms.py:22-140.
 *
 * Matrix p holds the values v(i, j) = dots[mat_off[p] + i * ld + j] for
 * i < lx_p, j < ly_p, with (lx_p, ly_p) = (lens[2p], lens[2p+1]) or, lens
 * NULL, (lx, ly); max_ly bounds every ly_p (it sizes LDS).  The static-kernel
 * value is exp2(v - hx[hx_off[p] + i] - hy[hy_off[p] + j]) for
 * KSIG_STATIC_RBF (v = 2h<x_i, y_j>, hx = h|x_i|^2, hy = h|y_j|^2,
 * h = log2(e) / max(2 sigma^2, 1e-12)) and v itself for KSIG_STATIC_LINEAR
 * (v = scale <x_i, y_j>).  out[p, m] = K_m, float64.  Orders outside the
 * compiled table run a kernel whose level count and order are runtime values
 * (KSIG_ELDS if its per-lane state does not fit).
 */
int ksig_signature_levels_dots(int dtype, const void* dots, const int64_t* mat_off,
                               int64_t ld, const int32_t* lens, int lx, int ly, int max_ly,
                               int static_kind, const void* hx, const int64_t* hx_off,
                               const void* hy, const int64_t* hy_off, int64_t n_mats,
                               int n_levels, int order, int difference, double* out,
                               void* stream);

/*
 * Signature-PDE kernel for a list of sequence pairs.
 *
 * Replaces, per pair, SignaturePDEKernel._compute_kernel
 * (ksig/kernels.py:242-256): the static Gram followed by
 * signature_kern_pde(M, difference) of ksig/algorithms.py:402-446
 * (device kernel :344-399).  dyadic_order refines each ΔΔ cell into
 * 2^λ × 2^λ cells carrying M/4^λ (λ = 0 is the reference scheme).
 *   out[p] = K(x_{pairs[p,0]}, y_{pairs[p,1]}) as float64.
 */
int ksig_signature_pde(int dtype, const void* x, const int64_t* x_off,
                       const int32_t* x_len, int max_len_x, const void* y,
                       const int64_t* y_off, const int32_t* y_len,
                       int max_len_y, int d, const int32_t* pairs,
                       int64_t n_pairs, int static_kind, double static_param,
                       int dyadic_order, int difference, double* out,
                       void* stream);


/*
 * Training forward of the signature-PDE kernel: ksig_signature_pde, which also
 * writes, per pair at element offset e_off[p] of `edges` (float64 whatever
 * dtype; kernel orientation: rows = the longer sequence), U of the last
 * refined row (Ncr values) then U of the last refined column (Nrr values),
 * Nrr = (len_rows - difference) << dyadic_order, Ncr likewise.  At dyadic
 * order 0 and 1 U is accumulated in float64 here, since the adjoint rebuilds
 * it backwards from these edges.  Same reference interface as
 * ksig_signature_pde; the edges have no reference counterpart (the
 * reference has no gradients).
 */
int ksig_signature_pde_fwd_edges(int dtype, const void* x, const int64_t* x_off,
                                 const int32_t* x_len, int max_len_x, const void* y,
                                 const int64_t* y_off, const int32_t* y_len,
                                 int max_len_y, int d, const int32_t* pairs,
                                 int64_t n_pairs, int static_kind, double static_param,
                                 int dyadic_order, int difference, double* out,
                                 void* edges, const int64_t* e_off, void* stream);

/*
 * Adjoint of ksig_signature_pde_fwd_edges (no reference counterpart: the
 * scheme of ksig/algorithms.py:380-385 is forward-only).  Given grad_out[p] =
 * dL/dK, writes dL/dM for every original ΔΔ cell, [Nr][Nc] in kernel
 * orientation at element offset m_off[p] of grad_m (dtype).  The solution U is
 * rebuilt in float64 in the reverse sweep from the (float64) edges; the
 * static kernel is recomputed.  workspace: device scratch of at least
 * ksig_signature_pde_bwd_workspace(...) bytes, 16-byte aligned (none at
 * dyadic order 0 and 1).
 */
int ksig_signature_pde_bwd_edges(int dtype, const void* x, const int64_t* x_off,
                                 const int32_t* x_len, int max_len_x, const void* y,
                                 const int64_t* y_off, const int32_t* y_len,
                                 int max_len_y, int d, const int32_t* pairs,
                                 int64_t n_pairs, int static_kind, double static_param,
                                 int dyadic_order, int difference, const double* grad_out,
                                 const void* edges, const int64_t* e_off, void* grad_m,
                                 const int64_t* m_off, void* workspace,
                                 size_t workspace_bytes, void* stream);

/*
 * Workspace bytes of ksig_signature_pde_bwd_edges for n_pairs pairs of at most
 * max_len_x x max_len_y points: one region per pair for the adjoint's band
 * hand-over state (contents are scratch).  No reference counterpart.
 */
size_t ksig_signature_pde_bwd_workspace(int dtype, int max_len_x, int max_len_y,
                                        int dyadic_order, int difference, int64_t n_pairs);

/*
 * Adjoint of ksig_signature_pde_fwd_edges fused with the chain rule through
 * ΔΔ and the static kernel: the work of ksig_signature_pde_bwd_edges followed
 * by ksig_signature_pde_grad_contract in one sweep, with no dL/dM array (no
 * reference counterpart: ksig/algorithms.py:344-446 and the static kernels
 * of ksig/static/kernels.py:134-166, 245-260 are forward-only).  Given
 * grad_out[p] = dL/dK of pair p, ADDS to grad_x / grad_y (dtype, either may
 * be NULL, both may alias) the gradients of the pairs' sequences and writes
 * grad_param_pairs[p] (float64, may be NULL) = this pair's part of dL/dparam
 * (RBF: sigma; linear: scale); the caller sums it over pairs.  Device atomics:
 * the gradients are order-dependent in their last bits.  Built for float32,
 * difference = 1, dyadic_order 0 or 1, 1 <= d <= 8, RBF or linear: other
 * shapes return KSIG_EUNSUPPORTED, launching nothing (the two-call path
 * takes them).  workspace: at least
 * ksig_signature_pde_bwd_fused_workspace(...) bytes, 16-byte aligned.
 */
int ksig_signature_pde_bwd_fused(int dtype, const void* x, const int64_t* x_off,
                                 const int32_t* x_len, int max_len_x, const void* y,
                                 const int64_t* y_off, const int32_t* y_len,
                                 int max_len_y, int d, const int32_t* pairs,
                                 int64_t n_pairs, int static_kind, double static_param,
                                 int dyadic_order, int difference, const double* grad_out,
                                 const void* edges, const int64_t* e_off, void* grad_x,
                                 void* grad_y, double* grad_param_pairs, void* workspace,
                                 size_t workspace_bytes, void* stream);

/*
 * Workspace bytes of ksig_signature_pde_bwd_fused for n_pairs pairs of at
 * most max_len_x x max_len_y points (0 for shapes it does not take): one band
 * hand-over region per pair, contents scratch.  No reference counterpart.
 */
size_t ksig_signature_pde_bwd_fused_workspace(int dtype, int d, int max_len_x, int max_len_y,
                                              int dyadic_order, int difference,
                                              int64_t n_pairs);

/*
 * Signature-PDE kernel on caller-supplied matrices M [n_mats, lx, ly].
 * Replaces signature_kern_pde(M, difference) of ksig/algorithms.py:402-446.
 */
int ksig_signature_pde_gram(int dtype, const void* M, int64_t n_mats, int lx,
                            int ly, int dyadic_order, int difference,
                            double* out, void* stream);

/*
 * Global alignment kernel in log space for a list of sequence pairs, with the
 * static kernel fused (RBF or linear): out[p] = log GA(x_a, y_b) of
 * pairs[p] = (a, b), float64.  Replaces the device recursion
 * _global_align_kern_log of ksig/algorithms.py:453-511 behind
 * global_align_kern_log(M) (:514-559), as called by
 * GlobalAlignmentKernel._compute_kernel (ksig/kernels.py:395-406), including
 * the transform M / (2 - M) and the clamp at 1e-12 (:532-534).  One wave per
 * pair; the shorter sequence of a pair (at most min(max_len_x, max_len_y)
 * points) is staged in LDS: KSIG_ELDS when it does not fit.
 */
int ksig_global_align_log(int dtype, const void* x, const int64_t* x_off,
                          const int32_t* x_len, int max_len_x, const void* y,
                          const int64_t* y_off, const int32_t* y_len, int max_len_y,
                          int d, const int32_t* pairs, int64_t n_pairs, int static_kind,
                          double static_param, double* out, void* stream);

/*
 * The same on caller-supplied static-kernel matrices M [n_mats, lx, ly]
 * (untransformed values, dtype): global_align_kern_log(M) of
 * ksig/algorithms.py:514-559 for M of shape [n, lx, ly].
 */
int ksig_global_align_log_gram(int dtype, const void* M, int64_t n_mats, int lx, int ly,
                               double* out, void* stream);

/*
 * Random Warping Series: the DTW value of every sequence a against every
 * warping series b, out[a * n_warps + b] (float64), with the squared
 * Euclidean distances computed in the kernel.  Replaces the device kernel
 * _random_warping_series_dtw behind random_warping_series(D, warp_lens)
 * (ksig/algorithms.py:565-659) as called by
 * RandomWarpingSeries._compute_features (ksig/kernels.py:459-477), including
 * the distance tensor D of ksig/kernels.py:466.  warp: the series' points
 * [warp_rows, d] (dtype), warp_off: [n_warps + 1] row offsets (non-decreasing,
 * warp_off[0] = 0, warp_off[n_warps] = warp_rows), groups: [n_groups + 1]
 * non-decreasing split points of the series from 0 to n_warps; one wave runs
 * one sequence against one group of consecutive series.  All device arrays.
 */
int ksig_random_warping_series(int dtype, const void* x, const int64_t* x_off,
                               const int32_t* x_len, int max_len_x, int64_t n_seq, int d,
                               const void* warp, const int32_t* warp_off, int n_warps,
                               int warp_rows, const int32_t* groups, int n_groups,
                               double* out, void* stream);

/*
 * The same on caller-supplied distances D [n_seq, lx, warp_rows] (dtype):
 * random_warping_series(D, warp_lens) of ksig/algorithms.py:630-659.
 */
int ksig_random_warping_series_gram(int dtype, const void* D, int64_t n_seq, int lx,
                                    int warp_rows, const int32_t* warp_off, int n_warps,
                                    const int32_t* groups, int n_groups, double* out,
                                    void* stream);

/*
 * Level recursion of the low-rank signature features with a tensorized
 * random projection: the level loops of signature_kern_first_order_low_rank
 * (ksig/algorithms.py:218-246) and signature_kern_higher_order_low_rank
 * (:296-337), dispatched by signature_kern_low_rank (:146-172), with the
 * TensorizedRandomProjection outer product (ksig/projections.py:566-584), as
 * called by SignatureFeatures._compute_features (ksig/kernels.py:330-361).
 * z: the projected increments [n_levels, n_seq, steps, n_comp] (dtype; level 0
 * projected alone with its scaling, level m >= 1 the projection of U_m);
 * order: the embedding order (<= 0 or >= n_levels: the full recursion, as
 * :165); out: the level features before normalisation,
 * [n_seq, n_levels, n_comp] float64.  1 <= n_levels <= 8.
 */
int ksig_lowrank_levels(int dtype, const void* z, int n_levels, int order, int64_t n_seq,
                        int steps, int n_comp, double* out, void* stream);

/*
 * The same level recursion with a DiagonalProjection of internal size 2
 * (ksig/projections.py:587-674) on random Fourier features (RFSF-DP):
 * du: the differenced features [n_levels, n_seq, steps, 2 * half] (dtype; sin
 * half then cos half); out: the level features before normalisation,
 * [n_seq, sum_i 2^(i+1) * half] float64, level i's block laid out as
 * [2^(i+1), half].  1 <= n_levels <= 5.
 */
int ksig_lowrank_levels_diag(int dtype, const void* du, int n_levels, int64_t n_seq,
                             int steps, int half, double* out, void* stream);

/*
 * The level recursion with a DiagonalProjection of any internal size q
 * (ksig/projections.py:587-674) at any embedding order, as the reference runs
 * it (first order ksig/algorithms.py:218-246; higher order :296-337 with its
 * [d, n, L, q^(i+1), H] terms for this projection, :302-306): du: the
 * differenced lifted paths [n_levels, n_seq, steps, q * half] (dtype; the
 * internal index b of feature b * half + h); order as ksig_lowrank_levels;
 * out: the level features before normalisation, [n_seq, sum_i q^(i+1) *
 * half] float64, level i's block laid out as [q^(i+1), half] (the internal
 * index of level 0 most significant).  Built for 1 <= q <= 4 and up to 5
 * levels where every level's running sums, with two adjacent levels' terms
 * at order > 1, fit 96 float64 registers per thread (q = 2: 5 levels at
 * order 1, 4 levels at any order; q = 3: 3 levels; q = 4: 3 levels at order
 * 1, 2 at any order); other shapes return KSIG_EUNSUPPORTED and launch
 * nothing.  ksig_lowrank_levels_diag is this entry at q = 2, order 1.
 */
int ksig_lowrank_levels_diagonal(int dtype, const void* du, int n_levels, int order,
                                 int internal_size, int64_t n_seq, int steps, int half,
                                 double* out, void* stream);

/*
 * The level recursion with TensorSketch projections (ksig/projections.py:
 * 379-458: the count sketch of ksig/utils.py:279-298 and the FFT convolution
 * of :301-313) at any embedding order (ksig/algorithms.py:175-337), as
 * SignatureFeatures._compute_features calls it with projection=TensorSketch
 * (ksig/kernels.py:330-361).  du: the (differenced) lifted paths [n_levels,
 * n_seq, steps, width] (dtype); hash_idx [n_levels, width] int32 and hash_bit
 * [n_levels, width] (dtype): level i's projection's hash_idx_ and hash_bit_
 * (a value of hash_idx outside [0, n_comp) selects no component, as the
 * reference's 0/1 hash mask); order as
 * ksig_lowrank_levels; out: the level features before normalisation,
 * [n_seq, n_levels, n_comp] float64.  The convolution of a sketch is
 * evaluated directly from the sparse count sketch (width products per
 * component; with fewer components than features, each step's count sketch
 * first and then n_comp products per component), equal to the FFT route up
 * to rounding.  1 <= n_levels <= 8,
 * n_comp <= 2048 and the workgroup's LDS ((n_levels + the order terms of
 * the levels below the top) * 2 * n_comp doubles, plus 20 bytes per level
 * and input feature for the staged hashes, signs and increments, and with
 * fewer components than features 12 bytes per level and component and 4 per
 * level and feature for the sketches and their buckets) within the device's
 * limit, else KSIG_EUNSUPPORTED and nothing is launched.
 */
int ksig_lowrank_levels_sketch(int dtype, const void* du, int n_levels, int order, int64_t n_seq,
                               int steps, int width, const int32_t* hash_idx,
                               const void* hash_bit, int n_comp, double* out, void* stream);

/*
 * Random Fourier features of projected points and their time increments in
 * one pass: proj = X W [n_seq, len, half] (dtype; rows ld >= half elements
 * apart, so one GEMM may hold several levels' weights), out [n_seq, len - 1, 2 * half]
 * (or [n_seq, len, 2 * half] without difference) = the sin half then the cos
 * half divided by sqrt(half), differenced along time.  Replaces
 * RandomFourierFeatures._compute_features after its GEMM
 * (ksig/static/features.py:322-342) followed by the increments of
 * ksig/algorithms.py:196, as SignatureFeatures._compute_features calls them.
 */
int ksig_rff_increments(int dtype, const void* proj, int64_t n_seq, int len, int half, int ld,
                        int difference, void* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* KSIG_HIP_H_ */
