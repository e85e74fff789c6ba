"""FP64 CPU oracle for the signature-kernel hot path.

TEST INFRASTRUCTURE ONLY.  This module is the checker used by ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``.
The product path (``ksig_amd``) never imports it; it runs the HIP library.

It restates, in NumPy float64 (or any dtype given), the reference algorithms
of tgcsaba/KSig v1.0.0 with the same dataflow (materialise the static Gram,
difference it, per-level exclusive 2-D cumulative sums):

* static kernels ............ ksig/static/kernels.py:145-155 (Linear),
                              :248-260 (RBF), ksig/utils.py:125-157
* exclusive multi-cumsum .... ksig/utils.py:63-90
* DP signature kernel ....... ksig/algorithms.py:22-44 (dispatch),
                              :47-85 (first order), :88-140 (higher order)
* signature-PDE kernel ...... ksig/algorithms.py:344-399 (Goursat update),
                              :402-446 (launcher), written as the plain 2-D
                              recurrence the antidiagonal sweep evaluates
* normalisation / level mean  ksig/kernels.py:166-214, ksig/utils.py:16,93-102,173-182
* class-level embedding ..... ksig/kernels.py:124-164,242-256
* global alignment kernel ... ksig/algorithms.py:453-559 (log-space
                              recursion), ksig/kernels.py:366-409 (class,
                              log-space normalisation :191-209)
* random warping series ..... ksig/algorithms.py:565-659 (DTW recursion),
                              ksig/kernels.py:459-477 (features)
* low-rank signature features ksig/algorithms.py:148-246 (first order),
                              :249-337 (higher order), Nystroem features
                              ksig/static/features.py:205-285, random
                              Fourier features with phases :344-390,
                              ksig/kernels.py:263-361, random Fourier features
                              ksig/static/features.py:288-342, tensorized
                              random projection ksig/projections.py:465-584,
                              diagonal projection :587-674, TensorSketch
                              :379-458 with ksig/utils.py:279-313

Additions that have no reference counterpart (build extensions):
``dyadic_order`` for the PDE (each ΔΔ cell refined into 2^λ×2^λ sub-cells
carrying M/4^λ), and the closed forms used to pin the restatement
(Chen signatures, iterated sums, chain enumeration, Bessel I0).

Parity pins (all in tests/test_oracle.py): DP vs Chen-signature inner
products (Linear kernel, full order), DP order 1 vs truncated iterated-sum
inner products, DP (any order, RBF) vs brute-force chain enumeration,
PDE one-cell closed form, PDE dyadic refinement vs I0(2*sqrt(ab)),
stuttering / translation invariance, symmetry and the normalised diagonal;
GA vs explicit enumeration of alignment paths in exp space and vs Delannoy
numbers (M' = 1); DTW vs explicit enumeration of warping paths and the
one-point closed forms; low-rank features without projection vs the
first-order DP with a linear static kernel on the lifted path, with a
tensorized projection vs brute-force iterated-sum tensors contracted with
the rank-1 projection tensors; higher-order low-rank features without
projection vs the order-p DP with a linear static kernel, with a tensorized
projection vs the unprojected ones contracted with the projection tensors;
diagonal projection (any internal size and order) vs the unprojected
features of each component's q-dimensional path scaled by sqrt(H)^(m-1);
TensorSketch (any order) vs the count sketch of the unprojected features
under the summed hash and the multiplied signs (tests/test_oracle_lowrank.py).
"""

from __future__ import annotations

import itertools
import math
from typing import Optional

import numpy as np

EPS = 1e-12  # ksig/utils.py:16 (_EPS)


# -----------------------------------------------------------------------------
# Static kernels (ksig/static/kernels.py, ksig/utils.py)
# -----------------------------------------------------------------------------

def squared_norm(X: np.ndarray, axis: int = -1) -> np.ndarray:
  """ksig/utils.py:125-135."""
  return np.sum(np.square(X), axis=axis)


def squared_euclid_dist(X: np.ndarray, Y: Optional[np.ndarray] = None
                        ) -> np.ndarray:
  """ksig/utils.py:138-157: ||x||^2 + ||y||^2 - 2<x, y>, no clamp."""
  Xn2 = squared_norm(X)
  if Y is None:
    return Xn2[..., :, None] + Xn2[..., None, :] - 2 * X @ np.swapaxes(X, -1, -2)
  Yn2 = squared_norm(Y)
  return Xn2[..., :, None] + Yn2[..., None, :] - 2 * X @ np.swapaxes(Y, -1, -2)


def rbf_gram(X: np.ndarray, Y: Optional[np.ndarray] = None,
             bandwidth: float = 1.0) -> np.ndarray:
  """ksig/static/kernels.py:248-260: exp(-D2 / max(2 sigma^2, EPS))."""
  return np.exp(-squared_euclid_dist(X, Y) / max(2 * bandwidth**2, EPS))


def linear_gram(X: np.ndarray, Y: Optional[np.ndarray] = None,
                scale: float = 1.0) -> np.ndarray:
  """ksig/static/kernels.py:145-155: scale * X Y^T."""
  Y = X if Y is None else Y
  return scale * (X @ np.swapaxes(Y, -1, -2))


def static_gram(X, Y, kind: str, param: float):
  if kind == 'rbf':
    return rbf_gram(X, Y, param)
  if kind == 'linear':
    return linear_gram(X, Y, param)
  raise ValueError(kind)


def pairwise_gram(X: np.ndarray, Y: Optional[np.ndarray], kind: str,
                  param: float, diag: bool = False) -> np.ndarray:
  """Embedding of ksig/kernels.py:124-145.

  Returns M of shape [nX, lX, lX] if ``diag`` else [nX, nY, lX, lY]."""
  if diag:
    return static_gram(X, X, kind, param)
  Y = X if Y is None else Y
  return static_gram(X[:, None], Y[None, :], kind, param)


# -----------------------------------------------------------------------------
# Exclusive cumulative sums (ksig/utils.py:63-90)
# -----------------------------------------------------------------------------

def multi_cumsum(M: np.ndarray, exclusive: bool = False, axis=-1) -> np.ndarray:
  """Cumulative sum over one or several axes; ``exclusive`` drops the last
  entry along each axis and pre-pads with a zero (ksig/utils.py:63-90)."""
  ndim = M.ndim
  axes = [axis] if np.isscalar(axis) else list(axis)
  axes = [ndim + a if a < 0 else a for a in axes]
  if exclusive:
    M = M[tuple(slice(-1) if a in axes else slice(None) for a in range(ndim))]
  for a in axes:
    M = np.cumsum(M, axis=a)
  if exclusive:
    M = np.pad(M, [(1, 0) if a in axes else (0, 0) for a in range(ndim)])
  return M


def double_diff(M: np.ndarray) -> np.ndarray:
  """ΔΔ of ksig/algorithms.py:64 (diff over axis -2, then -1)."""
  return np.diff(np.diff(M, axis=-2), axis=-1)


# -----------------------------------------------------------------------------
# DP signature kernel (ksig/algorithms.py:22-140)
# -----------------------------------------------------------------------------

def normalize_order(n_levels: int, order: int) -> int:
  """ksig/algorithms.py:37 and ksig/kernels.py:49-50."""
  return n_levels if order <= 0 or order >= n_levels else order


def signature_kern(M: np.ndarray, n_levels: int, order: int = -1,
                   difference: bool = True, return_levels: bool = False
                   ) -> np.ndarray:
  """ksig/algorithms.py:22-44."""
  order = normalize_order(n_levels, order)
  if order == 1:
    return signature_kern_first_order(M, n_levels, difference, return_levels)
  return signature_kern_higher_order(M, n_levels, order, difference,
                                     return_levels)


def signature_kern_first_order(M, n_levels, difference=True,
                               return_levels=False):
  """ksig/algorithms.py:47-85.

  K_0 = 1, R_1 = M, R_m = M * E(R_{m-1}), K_m = sum(R_m), with E the
  exclusive 2-D cumulative sum over the two time axes."""
  if difference:
    M = double_diff(M)
  lead = M.shape[:-2]
  levels = [np.ones(lead, dtype=M.dtype), np.sum(M, axis=(-2, -1))]
  R = M.copy()
  for _ in range(1, n_levels):
    R = M * multi_cumsum(R, exclusive=True, axis=(-2, -1))
    levels.append(np.sum(R, axis=(-2, -1)))
  if return_levels:
    return np.stack(levels, axis=0)
  return sum(levels[1:], levels[0].copy())


def signature_kern_higher_order(M, n_levels, order, difference=True,
                                return_levels=False):
  """ksig/algorithms.py:88-140.

  State R[r, s] with r (s) the current repeat count of the X (Y) time index;
  repeats carry 1/(r+1) weights, i.e. 1/l! per run of length l."""
  if difference:
    M = double_diff(M)
  lead = M.shape[:-2]
  levels = [np.ones(lead, dtype=M.dtype), np.sum(M, axis=(-2, -1))]
  R = M.copy()[None, None]
  for i in range(1, n_levels):
    d = min(i + 1, order)
    Rn = np.empty((d, d) + M.shape, dtype=M.dtype)
    Rn[0, 0] = M * multi_cumsum(np.sum(R, axis=(0, 1)), exclusive=True,
                                axis=(-2, -1))
    for r in range(1, d):
      Rn[0, r] = 1. / (r + 1) * M * multi_cumsum(
        np.sum(R[:, r - 1], axis=0), exclusive=True, axis=-2)
      Rn[r, 0] = 1. / (r + 1) * M * multi_cumsum(
        np.sum(R[r - 1, :], axis=0), exclusive=True, axis=-1)
      for s in range(1, d):
        Rn[r, s] = 1. / ((r + 1) * (s + 1)) * M * R[r - 1, s - 1]
    R = Rn
    levels.append(np.sum(R, axis=(0, 1, -2, -1)))
  if return_levels:
    return np.stack(levels, axis=0)
  return sum(levels[1:], levels[0].copy())


# -----------------------------------------------------------------------------
# Signature-PDE kernel (ksig/algorithms.py:344-446)
# -----------------------------------------------------------------------------

def refine_dyadic(M: np.ndarray, dyadic_order: int) -> np.ndarray:
  """Build extension (no reference counterpart): each cell of the ΔΔ grid
  becomes a 2^λ × 2^λ block of cells carrying M / 4^λ."""
  if dyadic_order == 0:
    return M
  f = 1 << dyadic_order
  Mr = np.repeat(np.repeat(M, f, axis=-2), f, axis=-1)
  return Mr / float(f * f)


def goursat_solve(M: np.ndarray) -> np.ndarray:
  """Explicit Goursat scheme of ksig/algorithms.py:380-385 as a 2-D recurrence:
  U(i,j) = (U(i-1,j) + U(i,j-1)) (1 + M/2 + M^2/12) - U(i-1,j-1) (1 - M^2/12)
  with U = 1 on the boundary (i = -1 or j = -1); returns U(lX-1, lY-1).

  Vectorised over the leading axes; loops over rows, sequential over columns
  inside each row (the antidiagonal order of the reference evaluates the same
  expressions on the same operands)."""
  lead = M.shape[:-2]
  lx, ly = M.shape[-2:]
  A = 1. + 0.5 * M + M * M / 12.
  B = 1. - M * M / 12.
  prev = np.ones(lead + (ly + 1,), dtype=M.dtype)  # row i-1, index 0 = j=-1
  for i in range(lx):
    cur = np.ones(lead + (ly + 1,), dtype=M.dtype)
    for j in range(ly):
      cur[..., j + 1] = ((prev[..., j + 1] + cur[..., j]) * A[..., i, j]
                         - prev[..., j] * B[..., i, j])
    prev = cur
  return prev[..., ly]


def signature_kern_pde(M: np.ndarray, difference: bool = True,
                       dyadic_order: int = 0) -> np.ndarray:
  """ksig/algorithms.py:402-446 (plus the dyadic extension)."""
  if M.ndim not in (3, 4):
    raise ValueError('The `M` matrix must have `.ndim==3` or `.ndim==4`.')
  if difference:
    M = double_diff(M)
  if M.shape[-1] == 0 or M.shape[-2] == 0:
    return np.ones(M.shape[:-2], dtype=M.dtype)
  return goursat_solve(refine_dyadic(M, dyadic_order))


# -----------------------------------------------------------------------------
# Class-level semantics (ksig/kernels.py:124-256)
# -----------------------------------------------------------------------------

def _normalise(K, Kxd, Kyd):
  """ksig/kernels.py:191-207 with ksig/utils.py:173-182."""
  sx = np.maximum(np.sqrt(np.maximum(Kxd, EPS)), EPS)
  sy = np.maximum(np.sqrt(np.maximum(Kyd, EPS)), EPS)
  return K / (sx[..., :, None] * sy[..., None, :])


def signature_kernel(X, Y=None, diag=False, n_levels=5, order=1,
                     difference=True, normalize=True, kind='rbf', param=1.0):
  """SignatureKernel(...)(X, Y, diag) of ksig/kernels.py:97-214."""
  order = normalize_order(n_levels, order)
  if diag:
    if normalize:
      return np.full((X.shape[0],), 1.)  # ksig/kernels.py:175-176
    M = pairwise_gram(X, None, kind, param, diag=True)
    return signature_kern(M, n_levels, order, difference, False)
  M = pairwise_gram(X, Y, kind, param)
  K = signature_kern(M, n_levels, order, difference, return_levels=normalize)
  if normalize:
    if Y is None:
      Kd = np.einsum('...ii->...i', K)
      K = _normalise(K, Kd, Kd)
    else:
      Kxd = signature_kern(pairwise_gram(X, None, kind, param, diag=True),
                           n_levels, order, difference, True)
      Kyd = signature_kern(pairwise_gram(Y, None, kind, param, diag=True),
                           n_levels, order, difference, True)
      K = _normalise(K, Kxd, Kyd)
    K = np.mean(K, axis=0)
  return K


def signature_pde_kernel(X, Y=None, diag=False, difference=True,
                         normalize=True, kind='rbf', param=1.0,
                         dyadic_order=0):
  """SignaturePDEKernel(...)(X, Y, diag) of ksig/kernels.py:218-256."""
  if diag:
    if normalize:
      return np.full((X.shape[0],), 1.)
    M = pairwise_gram(X, None, kind, param, diag=True)
    return signature_kern_pde(M, difference, dyadic_order)
  M = pairwise_gram(X, Y, kind, param)
  K = signature_kern_pde(M, difference, dyadic_order)
  if normalize:
    if Y is None:
      Kd = np.einsum('...ii->...i', K)
      K = _normalise(K, Kd, Kd)
    else:
      Kxd = signature_kern_pde(pairwise_gram(X, None, kind, param, True),
                               difference, dyadic_order)
      Kyd = signature_kern_pde(pairwise_gram(Y, None, kind, param, True),
                               difference, dyadic_order)
      K = _normalise(K, Kxd, Kyd)
  return K


# -----------------------------------------------------------------------------
# Per-pair helpers (ragged batches: each pair at its native lengths)
# -----------------------------------------------------------------------------

def pair_levels(x, y, n_levels, order=1, difference=True, kind='rbf',
                param=1.0):
  """Level values K_0..K_n of one pair (x [lx, d], y [ly, d])."""
  M = static_gram(x, y, kind, param)
  return signature_kern(M, n_levels, order, difference, return_levels=True)


def pair_pde(x, y, difference=True, kind='rbf', param=1.0, dyadic_order=0):
  M = static_gram(x, y, kind, param)
  return signature_kern_pde(M[None], difference, dyadic_order)[0]


# -----------------------------------------------------------------------------
# Global alignment kernel (ksig/algorithms.py:453-559, ksig/kernels.py:366-409)
# -----------------------------------------------------------------------------

def global_align_kern_log(M: np.ndarray) -> np.ndarray:
  """log GA of ksig/algorithms.py:514-559 for M [..., lX, lY]: M' = M/(2-M)
  (:532), l = log(max(M', EPS)) (:534), then the log-space alignment
  recursion of the device kernel (:453-511), written as the plain 2-D
  recurrence the antidiagonal sweep evaluates (each cell reads only its
  three predecessors, so the traversal order does not change any value):
    logK(i,j) = l(i,j) + LSE(logK(i-1,j), logK(i,j-1), logK(i-1,j-1)),
  logK(-1,-1) = 0, other out-of-grid predecessors -inf, LSE with the
  max-shift of :494-497.  Returns logK(lX-1, lY-1) per leading index."""
  M = np.asarray(M)
  Mp = M / (2. - M)
  with np.errstate(divide='ignore'):
    lg = np.log(np.maximum(Mp, EPS))
  lX, lY = lg.shape[-2:]
  lead = lg.shape[:-2]
  ninf = np.full(lead, -np.inf, dtype=lg.dtype)
  prev = [ninf] * lY  # row i-1
  for i in range(lX):
    cur = []
    for j in range(lY):
      k01 = prev[j]
      k10 = cur[j - 1] if j > 0 else ninf
      if i == 0 and j == 0:
        k00 = np.zeros(lead, dtype=lg.dtype)
      elif i == 0 or j == 0:
        k00 = ninf
      else:
        k00 = prev[j - 1]
      m = np.maximum(np.maximum(k01, k10), k00)
      with np.errstate(invalid='ignore'):
        s = np.exp(k01 - m) + np.exp(k10 - m) + np.exp(k00 - m)
      cur.append(lg[..., i, j] + m + np.log(s))
    prev = cur
  return prev[lY - 1]


def global_alignment_kernel(X, Y=None, diag=False, kind='rbf', param=1.0):
  """GlobalAlignmentKernel(...)(X, Y, diag) of ksig/kernels.py:366-409 with
  the log-space normalisation of ksig/kernels.py:191-209."""
  if diag:
    return np.full((X.shape[0],), 1.)
  logK = global_align_kern_log(pairwise_gram(X, Y, kind, param))
  if Y is None:
    d = np.diagonal(logK)
    logK = logK - .5 * (d[:, None] + d[None, :])
  else:
    dx = global_align_kern_log(pairwise_gram(X, None, kind, param, diag=True))
    dy = global_align_kern_log(pairwise_gram(Y, None, kind, param, diag=True))
    logK = logK - .5 * (dx[:, None] + dy[None, :])
  return np.exp(logK)


def pair_gak_log(x, y, kind='rbf', param=1.0):
  """log GA of one pair (x [lx, d], y [ly, d])."""
  return global_align_kern_log(static_gram(x, y, kind, param)[None])[0]


def delannoy(m: int, n: int) -> int:
  """Number of monotone lattice paths (0,0) -> (m,n) with steps (1,0), (0,1),
  (1,1): sum_k C(m,k) C(n,k) 2^k."""
  return sum(math.comb(m, k) * math.comb(n, k) * 2 ** k for k in range(min(m, n) + 1))


def gak_path_sum(Mp: np.ndarray) -> float:
  """GA by explicit enumeration of alignments (tiny grids): the sum over
  monotone paths (0,0) -> (lX-1,lY-1) with steps (1,0), (0,1), (1,1) of the
  product of the transformed values M' on the path, in exp space."""
  lX, lY = Mp.shape
  total = 0.0

  def walk(i, j, prod):
    nonlocal total
    prod = prod * Mp[i, j]
    if i == lX - 1 and j == lY - 1:
      total += prod
      return
    if i + 1 < lX:
      walk(i + 1, j, prod)
    if j + 1 < lY:
      walk(i, j + 1, prod)
    if i + 1 < lX and j + 1 < lY:
      walk(i + 1, j + 1, prod)

  walk(0, 0, 1.0)
  return total


# -----------------------------------------------------------------------------
# Random Warping Series (ksig/algorithms.py:565-659, ksig/kernels.py:413-477)
# -----------------------------------------------------------------------------

def random_warping_series(D: np.ndarray, warp_lens) -> np.ndarray:
  """P [n_X, n_Y] of ksig/algorithms.py:630-659 from distances D [n_X, l_X,
  sum of warp lengths]: for each warping series b (columns warp_off[b] ..
  warp_off[b+1]-1 of D) the dynamic time warping recursion of the device
  kernel (:565-627),
    P(i,j) = D(i,j) + min(P(i-1,j), P(i,j-1), P(i-1,j-1)),
  P(-1,-1) = 0, other out-of-grid predecessors +inf, written as the plain 2-D
  recurrence the antidiagonal sweep evaluates."""
  D = np.asarray(D)
  lens = np.asarray(warp_lens, dtype=np.int64)
  n_X, l_X = D.shape[:2]
  off = np.concatenate([[0], np.cumsum(lens)])
  P = np.empty((n_X, len(lens)), dtype=D.dtype)
  inf = np.full((n_X,), np.inf, dtype=D.dtype)
  for b in range(len(lens)):
    Db = D[:, :, off[b]:off[b + 1]]
    l_Y = Db.shape[2]
    prev = [inf] * l_Y
    for i in range(l_X):
      cur = []
      for j in range(l_Y):
        p01 = prev[j]
        p10 = cur[j - 1] if j > 0 else inf
        if i == 0 and j == 0:
          p00 = np.zeros_like(inf)
        elif i == 0 or j == 0:
          p00 = inf
        else:
          p00 = prev[j - 1]
        cur.append(Db[:, i, j] + np.minimum(np.minimum(p01, p10), p00))
      prev = cur
    P[:, b] = prev[l_Y - 1]
  return P


def rws_features(X, warp_series, warp_lens, normalize=True):
  """RandomWarpingSeries._compute_features of ksig/kernels.py:459-477 for
  given warping series: squared distances, DTW, then unit-norm rows
  (ksig/utils.py:125-135,173-182)."""
  D = squared_euclid_dist(X, warp_series[None])
  P = random_warping_series(D, warp_lens)
  if normalize:
    nrm = np.sqrt(np.maximum(np.sum(P * P, -1), EPS))
    P = P / np.maximum(nrm, EPS)[..., None]
  return P


def dtw_path_min(D: np.ndarray) -> float:
  """DTW by explicit enumeration (tiny grids): the minimum over monotone
  paths (0,0) -> (lX-1,lY-1) with steps (1,0), (0,1), (1,1) of the sum of D
  on the path."""
  lX, lY = D.shape
  best = np.inf

  def walk(i, j, acc):
    nonlocal best
    acc = acc + D[i, j]
    if i == lX - 1 and j == lY - 1:
      best = min(best, acc)
      return
    if i + 1 < lX:
      walk(i + 1, j, acc)
    if j + 1 < lY:
      walk(i, j + 1, acc)
    if i + 1 < lX and j + 1 < lY:
      walk(i + 1, j + 1, acc)

  walk(0, 0, 0.0)
  return best


# -----------------------------------------------------------------------------
# Low-rank signature features (ksig/algorithms.py:148-246, ksig/kernels.py:263-361,
# ksig/static/features.py:288-342, ksig/projections.py:465-584)
# -----------------------------------------------------------------------------

def rff_features(X: np.ndarray, weights: np.ndarray) -> np.ndarray:
  """RandomFourierFeatures._compute_features (ksig/static/features.py:322-342)
  for given weights [d, D]: [sin(XW), cos(XW)] / sqrt(D)."""
  proj = X @ weights
  return np.concatenate([np.sin(proj), np.cos(proj)], -1) / np.sqrt(weights.shape[1])


def rff1d_features(X: np.ndarray, weights: np.ndarray, shift: np.ndarray) -> np.ndarray:
  """RandomFourierFeatures1D._compute_features (ksig/static/features.py:381-390)
  for given weights [d, D] and phases [1, D]: cos(XW + b) sqrt(2/D)."""
  return np.cos(X @ weights + shift) * np.sqrt(2 / weights.shape[1])


def nystroem_features(X: np.ndarray, basis_idx, kind='rbf', param=1.0) -> np.ndarray:
  """NystroemFeatures (ksig/static/features.py:227-285) for given basis
  indices into the flattened points of X [..., d]: K(X, B) U_+ diag(s_+)^-1/2
  with K(B, B) = U diag(s) U^T and |s| > 1e-12 kept."""
  Xf = X.reshape(-1, X.shape[-1])
  B = Xf[np.asarray(basis_idx)]
  s, U = np.linalg.eigh(static_gram(B, B, kind, param))
  keep = np.abs(s) > EPS
  norm = U[:, keep] / np.sqrt(np.maximum(s[keep], EPS))[None, :]
  return (static_gram(Xf, B, kind, param) @ norm).reshape(X.shape[:-1] + (-1,))


def low_rank_features_normalised(P, n_levels):
  """Level features normalised to unit norm, concatenated, divided by
  sqrt(n_levels + 1) (ksig/kernels.py:354-361)."""
  P = [p / np.maximum(np.sqrt(np.maximum(np.sum(p * p, -1), EPS)), EPS)[..., None] for p in P]
  return np.concatenate(P, -1) / np.sqrt(n_levels + 1)


def signature_kern_low_rank_fo(U, n_levels, difference=True, projections=None):
  """First-order low-rank signature features of ksig/algorithms.py:175-246
  with ``return_levels=True``: U is a list of n_levels arrays [n, L, D] (one
  per level, independent static features) or one array; ``projections`` is
  None or a list of TRP (components [D, P], scaling) pairs, level 0 projected
  alone (:199-201) and level i >= 1 as the elementwise product of the
  exclusive time cumsum with the projection of U_i (:224-237, TRP outer
  product ksig/projections.py:566-584).  Returns the list of level
  features [n, 1], [n, D or P], ..."""
  Us = U if isinstance(U, list) else [U] * n_levels
  if difference:
    Us = [np.diff(u, axis=1) for u in Us]
  n = Us[0].shape[0]
  P = [np.ones((n, 1), dtype=Us[0].dtype)]
  if projections == 'diag':
    # DiagonalProjection of internal size 2 (ksig/projections.py:649-674):
    # level 0 split into [2, H]; level i the component-wise product with an
    # outer product over the internal axis, scaled by sqrt(H)
    H = Us[0].shape[-1] // 2
    R = Us[0].reshape(Us[0].shape[:-1] + (2, H))
    P.append(R.sum(1).reshape(n, -1))
    for i in range(1, n_levels):
      R = multi_cumsum(R, exclusive=True, axis=1)
      Y = Us[i].reshape(Us[i].shape[:-1] + (2, H))
      R = np.sqrt(H) * (R[..., :, None, :] * Y[..., None, :, :]).reshape(
          R.shape[:2] + (-1, H))
      P.append(R.sum(1).reshape(n, -1))
    return P
  if projections is None:
    R = Us[0].copy()
  else:
    C0, s0 = projections[0]
    R = s0 * (Us[0] @ C0)
  P.append(R.sum(1))
  for i in range(1, n_levels):
    R = multi_cumsum(R, exclusive=True, axis=1)
    if projections is None:
      R = (R[..., :, None] * Us[i][..., None, :]).reshape(R.shape[:2] + (-1,))
    else:
      Ci, si = projections[i]
      R = si * R * (Us[i] @ Ci)
    P.append(R.sum(1))
  return P


def signature_kern_low_rank(U, n_levels, order=1, difference=True, projections=None):
  """Low-rank signature features of any embedding order, ksig/algorithms.py:146-172
  dispatching to :175-246 (order 1) or :249-337 (order p > 1), with
  ``return_levels=True``.  At level i the order-(r+1) terms R[r], r < min(i+1, p):
  R[0] is the exclusive time cumsum of the previous level's sum over r times
  the (projected) U_i (:312-316), R[r] = R_prev[r-1] times U_i / (r+1)
  (:317-324); the level feature sums over r and time (:325-334).
  ``projections``: None or TRP (components [D, P], scaling) pairs, as in
  signature_kern_low_rank_fo."""
  order = normalize_order(n_levels, order)
  Us = U if isinstance(U, list) else [U] * n_levels
  if difference:
    Us = [np.diff(u, axis=1) for u in Us]
  n, steps = Us[0].shape[:2]
  P = [np.ones((n, 1), dtype=Us[0].dtype)]
  if projections is None:
    Rs = [Us[0].copy()]
  else:
    C0, s0 = projections[0]
    Rs = [s0 * (Us[0] @ C0)]
  P.append(Rs[0].sum(1))
  for i in range(1, n_levels):
    Q = multi_cumsum(sum(Rs), exclusive=True, axis=1)
    if projections is None:
      outer = lambda A: (A[..., :, None] * Us[i][..., None, :]).reshape(n, steps, -1)
    else:
      Ci, si = projections[i]
      Z = si * (Us[i] @ Ci)
      outer = lambda A: A * Z
    nxt = [outer(Q)]
    for r in range(1, min(i + 1, order)):
      nxt.append(outer(Rs[r - 1]) / (r + 1))
    Rs = nxt
    P.append(sum(R.sum(1) for R in Rs))
  return P


def signature_kern_low_rank_gauss(U, n_levels, comps, order=1, difference=True):
  """Low-rank signature features (ksig/algorithms.py:146-337, any order) with
  GaussianRandomProjection components (ksig/projections.py:189-242): comps[i]
  = (C [P, F], s); level 0 projects U_0 as s C U, level i >= 1 projects the
  outer product of a running sum (or order term) with U_i, X index major."""
  order = normalize_order(n_levels, order)
  Us = U if isinstance(U, list) else [U] * n_levels
  if difference:
    Us = [np.diff(u, axis=1) for u in Us]
  n, steps = Us[0].shape[:2]

  def proj(i, A, B=None):
    C, sc = comps[i]
    if B is not None:
      A = (A[..., :, None] * B[..., None, :]).reshape(n, steps, -1)
    return sc * (A @ C.T)

  Rs = [proj(0, Us[0])]
  P = [np.ones((n, 1)), Rs[0].sum(1)]
  for i in range(1, n_levels):
    Q = multi_cumsum(sum(Rs), exclusive=True, axis=1)
    nxt = [proj(i, Q, Us[i])]
    for r in range(1, min(i + 1, order)):
      nxt.append(proj(i, Rs[r - 1], Us[i]) / (r + 1))
    Rs = nxt
    P.append(sum(R.sum(1) for R in Rs))
  return P


def signature_kern_low_rank_diag(U, n_levels, internal_size=2, order=1, difference=True):
  """Low-rank signature features with a DiagonalProjection of any internal
  size q (ksig/projections.py:587-674) at any embedding order (first order
  ksig/algorithms.py:175-246; higher order :249-337, whose R_next holds
  [d, n, L, q^(i+1), H] entries for this projection, :302-306): level 0 is
  U_0 split into [q, H] (:649-659, H = D / q); the outer-product step is
  sqrt(H) A[..., :, None, :] * Y[..., None, :, :] flattened over the two
  internal axes with Y = U_i split into [q, H] (:661-674; H is the fitted
  step's n_features_, the last axis of the projected level 0).  Level m's
  features are [n, q^m * H], the internal index of level 0 most significant."""
  order = normalize_order(n_levels, order)
  Us = U if isinstance(U, list) else [U] * n_levels
  if difference:
    Us = [np.diff(u, axis=1) for u in Us]
  n, steps, D = Us[0].shape
  q = internal_size
  H = D // q
  split = lambda A: A.reshape(A.shape[:-1] + (q, -1))

  def outer(A, i):
    Y = split(Us[i])
    return np.sqrt(H) * (A[..., :, None, :] * Y[..., None, :, :]).reshape(A.shape[:2] + (-1, H))

  Rs = [split(Us[0])]
  P = [np.ones((n, 1), dtype=Us[0].dtype), Rs[0].sum(1).reshape(n, -1)]
  for i in range(1, n_levels):
    Q = multi_cumsum(sum(Rs), exclusive=True, axis=1)
    nxt = [outer(Q, i)]
    for r in range(1, min(i + 1, order)):
      nxt.append(outer(Rs[r - 1], i) / (r + 1))
    Rs = nxt
    P.append(sum(R.sum(1) for R in Rs).reshape(n, -1))
  return P


def count_sketch(X, hash_idx, hash_bit, n_components):
  """compute_count_sketch (ksig/utils.py:279-298): the einsum of X with the
  0/1 mask [hash_idx[i] == j] and the signs, out[..., j] = sum over i with
  hash_idx[i] = j of hash_bit[i] X[..., i]."""
  mask = (np.asarray(hash_idx)[:, None] == np.arange(n_components)[None, :]).astype(X.dtype)
  return np.einsum('...i,ij,i->...j', X, mask, np.asarray(hash_bit, dtype=X.dtype))


def convolve_fft(X, Y):
  """convolve_fft (ksig/utils.py:301-313): the real part of the inverse FFT
  of the product of the FFTs along the last axis (circular convolution)."""
  return np.real(np.fft.ifft(np.fft.fft(X, axis=-1) * np.fft.fft(Y, axis=-1), axis=-1))


def signature_kern_low_rank_sketch(U, n_levels, sketches, order=1, difference=True):
  """Low-rank signature features with TensorSketch projections
  (ksig/projections.py:379-458) at any embedding order (ksig/algorithms.py:
  175-246, :249-337): sketches[i] = (hash_idx [D], hash_bit [D], m) of level
  i's fitted projection; level 0 is the count sketch of U_0 (_project,
  :438-448); level i >= 1 projects (A, U_i) to convolve_fft(A, count sketch
  of U_i) (_project_outer_prod, :450-465).  Level features [n, m]."""
  order = normalize_order(n_levels, order)
  Us = U if isinstance(U, list) else [U] * n_levels
  if difference:
    Us = [np.diff(u, axis=1) for u in Us]
  n = Us[0].shape[0]
  cs = lambda i, A: count_sketch(A, *sketches[i])
  Rs = [cs(0, Us[0])]
  P = [np.ones((n, 1), dtype=Us[0].dtype), Rs[0].sum(1)]
  for i in range(1, n_levels):
    Z = cs(i, Us[i])
    Q = multi_cumsum(sum(Rs), exclusive=True, axis=1)
    nxt = [convolve_fft(Q, Z)]
    for r in range(1, min(i + 1, order)):
      nxt.append(convolve_fft(Rs[r - 1], Z) / (r + 1))
    Rs = nxt
    P.append(sum(R.sum(1) for R in Rs))
  return P


def signature_features(X, n_levels, rff_weights, trp=None, difference=True, normalize=True,
                       order=1):
  """SignatureFeatures._compute_features (ksig/kernels.py:330-361) with
  RandomFourierFeatures static features (one weight matrix per level,
  ``indep_feat``) and optionally TensorizedRandomProjection components:
  level features, each normalised to unit norm, concatenated and divided by
  sqrt(n_levels + 1)."""
  U = [rff_features(X, w) for w in rff_weights]
  if normalize_order(n_levels, order) == 1:
    P = signature_kern_low_rank_fo(U, n_levels, difference, trp)  # trp: list, 'diag' or None
  else:
    P = signature_kern_low_rank(U, n_levels, order, difference, trp)
  if not normalize:
    return np.concatenate(P, -1)
  P = [p / np.maximum(np.sqrt(np.maximum(np.sum(p * p, -1), EPS)), EPS)[..., None] for p in P]
  return np.concatenate(P, -1) / np.sqrt(n_levels + 1)


# -----------------------------------------------------------------------------
# MMD^2 (build extension; unbiased estimator)
# -----------------------------------------------------------------------------

def mmd2_unbiased(Kxx, Kyy, Kxy):
  """mean_{i!=j} Kxx + mean_{i!=j} Kyy - 2 mean Kxy."""
  n, m = Kxx.shape[0], Kyy.shape[0]
  sxx = (np.sum(Kxx) - np.trace(Kxx)) / (n * (n - 1))
  syy = (np.sum(Kyy) - np.trace(Kyy)) / (m * (m - 1))
  return sxx + syy - 2. * np.mean(Kxy)


# -----------------------------------------------------------------------------
# Closed forms and brute force used to pin the restatement
# -----------------------------------------------------------------------------

def _tensor_exp(v: np.ndarray, depth: int):
  """Truncated tensor exponential of a vector: [1, v, v⊗v/2!, ...]."""
  out = [np.ones(())]
  for m in range(1, depth + 1):
    out.append(np.multiply.outer(out[-1], v) / m)
  return out


def _tensor_mul(a, b, depth):
  out = []
  for m in range(depth + 1):
    acc = np.zeros(a[m].shape if m else ())
    for k in range(m + 1):
      acc = acc + np.multiply.outer(a[k], b[m - k])
    out.append(acc)
  return out


def chen_signature(x: np.ndarray, depth: int):
  """Signature of the piecewise-linear path through the rows of x, truncated
  at ``depth``, via Chen's identity (product of segment exponentials)."""
  d = x.shape[1]
  sig = [np.ones(())] + [np.zeros((d,) * m) for m in range(1, depth + 1)]
  for inc in np.diff(x, axis=0):
    sig = _tensor_mul(sig, _tensor_exp(inc, depth), depth)
  return sig


def iterated_sums(x: np.ndarray, depth: int):
  """Discrete iterated sums S_m = sum_{i1<...<im} dx_i1 ⊗ ... ⊗ dx_im."""
  dx = np.diff(x, axis=0)
  d = x.shape[1]
  S = [np.ones(())] + [np.zeros((d,) * m) for m in range(1, depth + 1)]
  for inc in dx:
    for m in range(depth, 0, -1):
      S[m] = S[m] + np.multiply.outer(S[m - 1], inc)
  return S


def brute_force_levels(M: np.ndarray, n_levels: int, order: int):
  """K_m by explicit enumeration of index chains (tiny grids only).

  K_m = sum over non-decreasing i_1<=...<=i_m and j_1<=...<=j_m whose runs of
  equal indices have length <= order, of prod_k M[i_k, j_k] times
  prod over runs of 1/(run length)! on each side.  Independent of the DP."""
  lx, ly = M.shape
  order = normalize_order(n_levels, order)

  def sequences(n, m):
    out = []
    for seq in itertools.combinations_with_replacement(range(n), m):
      runs = [len(list(g)) for _, g in itertools.groupby(seq)]
      if max(runs) <= order:
        w = 1.
        for r in runs:
          w /= math.factorial(r)
        out.append((seq, w))
    return out

  K = [1.]
  for m in range(1, n_levels + 1):
    tot = 0.
    sx, sy = sequences(lx, m), sequences(ly, m)
    for (ix, wx) in sx:
      for (jy, wy) in sy:
        tot += wx * wy * np.prod(M[list(ix), list(jy)])
    K.append(tot)
  return np.array(K)


def bessel_i0(z: float, terms: int = 60) -> float:
  """I_0(z) = sum_k (z/2)^{2k} / (k!)^2."""
  return float(sum((z / 2.) ** (2 * k) / math.factorial(k) ** 2
                   for k in range(terms)))


# -----------------------------------------------------------------------------
# Synthetic inputs shared by tests and the benchmark
# -----------------------------------------------------------------------------

def brownian(rng: np.random.Generator, n: int, L: int, d: int,
             dtype=np.float64) -> np.ndarray:
  """n paths: cumsum over L steps of i.i.d. N(0, 1/L) increments per channel
  (the first point equals the first increment)."""
  inc = rng.standard_normal((n, L, d)) / math.sqrt(L)
  return np.cumsum(inc, axis=1).astype(dtype)
