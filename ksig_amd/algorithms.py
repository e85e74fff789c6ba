"""Functional signature-kernel operators on a precomputed static-kernel
tensor M (ksig/algorithms.py counterparts), evaluated by libksig_hip.so."""

from __future__ import annotations

import torch

from . import _ops, utils


def _normalize_order(n_levels: int, order: int) -> int:
  """ksig/algorithms.py:37."""
  return n_levels if order <= 0 or order >= n_levels else order


def signature_kern(M, n_levels: int, order: int = -1, difference: bool = True,
                   return_levels: bool = False) -> torch.Tensor:
  """Full-rank signature kernel by dynamic programming (ksig/algorithms.py:22-140).

  Args:
    M: static-kernel evaluations of shape [n_X, n_Y, l_X, l_Y] or [n, l_X, l_Y].
    n_levels: number of signature levels.
    order: embedding order (<= 0 or >= n_levels means full order).
    difference: take increments of the lifted sequences (ΔΔ of M).
    return_levels: return the per-level values stacked on a new first axis.

  Returns:
    [n_X, n_Y] or [n] (with a leading n_levels+1 axis if return_levels),
    dtype of M, on the GPU.
  """
  M = utils.as_device(M)
  order = _normalize_order(n_levels, order)
  lead = tuple(M.shape[:-2])
  lev = _ops.signature_levels_gram(M, n_levels, order, difference)
  lev = lev.t().reshape((n_levels + 1,) + lead)
  if return_levels:
    return lev.to(M.dtype)
  return lev.sum(0).to(M.dtype)


def signature_kern_pde(M, difference: bool = True, dyadic_order: int = 0
                       ) -> torch.Tensor:
  """Signature-PDE kernel (ksig/algorithms.py:402-446).

  ``dyadic_order`` (no reference counterpart; 0 reproduces the reference)
  refines every ΔΔ cell into 2^λ x 2^λ cells carrying M / 4^λ."""
  M = utils.as_device(M)
  is_diag = M.ndim == 3
  if is_diag:
    M = M[:, None, :, :]
  if M.ndim != 4:
    raise ValueError('The `M` matrix must have `.ndim==3` or `.ndim==4`.')
  n_X, n_Y = M.shape[:2]
  K = _ops.signature_pde_gram(M, difference, dyadic_order).reshape(n_X, n_Y).to(M.dtype)
  if is_diag:
    K = K[:, 0]
  return K


def global_align_kern_log(M) -> torch.Tensor:
  """Logarithm of the global alignment kernel (ksig/algorithms.py:514-559):
  M' = M / (2 - M), log-space alignment recursion over max(M', 1e-12).

  Args:
    M: static-kernel evaluations of shape [n_X, n_Y, l_X, l_Y] or [n, l_X, l_Y].

  Returns:
    log GA of shape [n_X, n_Y] or [n], dtype of M, on the GPU.
  """
  M = utils.as_device(M)
  is_diag = M.ndim == 3
  if is_diag:
    M = M[:, None, :, :]
  if M.ndim != 4:
    raise ValueError('The `M` matrix must have `.ndim==3` or `.ndim==4`.')
  n_X, n_Y = M.shape[:2]
  K = _ops.global_align_log_gram(M).reshape(n_X, n_Y).to(M.dtype)
  if is_diag:
    K = K[:, 0]
  return K


def random_warping_series(D, warp_lens) -> torch.Tensor:
  """Random Warping Series DTW values (ksig/algorithms.py:630-659).

  Args:
    D: squared distances of shape [n_X, l_X, sum of warp lengths].
    warp_lens: lengths of the warping series, shape [n_Y].

  Returns:
    [n_X, n_Y], dtype of D, on the GPU.
  """
  D = utils.as_device(D)
  if D.ndim != 3:
    raise ValueError('`D` distances array must have `.ndim==3`.')
  lens = warp_lens.detach().cpu().numpy() if isinstance(warp_lens, torch.Tensor) else warp_lens
  return _ops.random_warping_series_gram(D, lens).to(D.dtype)
