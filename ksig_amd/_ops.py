"""Host-side launchers: torch device tensors -> C-ABI calls -> torch results.

Every function here hands device pointers of contiguous torch tensors to
libksig_hip.so and returns freshly allocated float64 device tensors; nothing
is computed on the host.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Optional, Sequence

import torch

from . import _lib


@dataclass
class SeqSet:
  """Packed sequence set: rows [total, d], per-sequence offsets and lengths."""
  data: torch.Tensor     # [total_rows, d], contiguous
  off: torch.Tensor      # int64 [n]
  len: torch.Tensor      # int32 [n]
  max_len: int
  n: int

  @property
  def d(self) -> int:
    return self.data.shape[-1]

  @staticmethod
  def dense(X: torch.Tensor) -> "SeqSet":
    """From a dense [n, L, d] device tensor."""
    n, L, d = X.shape
    data = X.reshape(n * L, d).contiguous()
    off = torch.arange(n, device=X.device, dtype=torch.int64) * L
    lens = torch.full((n,), L, device=X.device, dtype=torch.int32)
    return SeqSet(data, off, lens, L, n)

  @staticmethod
  def ragged(seqs: Sequence[torch.Tensor], device=None, dtype=None) -> "SeqSet":
    """From a list of [L_a, d] tensors of differing lengths."""
    device = device or seqs[0].device
    dtype = dtype or seqs[0].dtype
    lens_list = [int(s.shape[0]) for s in seqs]
    data = torch.cat([s.to(device=device, dtype=dtype) for s in seqs], 0).contiguous()
    offs = [0]
    for L in lens_list[:-1]:
      offs.append(offs[-1] + L)
    off = torch.tensor(offs, device=device, dtype=torch.int64)
    lens = torch.tensor(lens_list, device=device, dtype=torch.int32)
    return SeqSet(data, off, lens, max(lens_list) if lens_list else 0, len(seqs))


def _scalar(p) -> float:
  return float(p.detach()) if isinstance(p, torch.Tensor) else float(p)


def static_code(static_kernel, raw: bool = False) -> tuple:
  """(kind, param) of a fused static kernel, or None for a generic one.
  With ``raw`` the parameter is returned as stored (it may be a tensor that
  requires grad)."""
  from .static.kernels import LinearKernel, RBFKernel
  if type(static_kernel) is RBFKernel:
    p = static_kernel.bandwidth
    return _lib.STATIC_RBF, (p if raw else _scalar(p))
  if type(static_kernel) is LinearKernel:
    p = static_kernel.scale
    return _lib.STATIC_LINEAR, (p if raw else _scalar(p))
  return None


def dense_pairs(nx: int, ny: int, device) -> torch.Tensor:
  ix = torch.arange(nx, device=device, dtype=torch.int32)
  iy = torch.arange(ny, device=device, dtype=torch.int32)
  return torch.stack(torch.meshgrid(ix, iy, indexing="ij"), -1).reshape(-1, 2).contiguous()


def triu_pairs(n: int, device) -> torch.Tensor:
  return torch.triu_indices(n, n, device=device).t().to(torch.int32).contiguous()


def diag_pairs(n: int, device) -> torch.Tensor:
  i = torch.arange(n, device=device, dtype=torch.int32)
  return torch.stack([i, i], -1).contiguous()


def _check_sets(X: SeqSet, Y: SeqSet):
  if X.data.dtype != Y.data.dtype:
    raise TypeError("X and Y must share a dtype")
  if X.d != Y.d:
    raise ValueError("X and Y must have the same number of features")
  if X.data.device != Y.data.device:
    raise ValueError("X and Y must live on the same device")


def signature_levels(X: SeqSet, Y: SeqSet, pairs: torch.Tensor, n_levels: int,
                     order: int, difference: bool, kind: int, param: float,
                     gram_mode: int = -1) -> torch.Tensor:
  """Level values [n_pairs, n_levels+1] (float64) of the fused DP kernel."""
  lib = _lib.load()
  _check_sets(X, Y)
  n_pairs = pairs.shape[0]
  out = torch.empty((n_pairs, n_levels + 1), device=X.data.device, dtype=torch.float64)
  if n_pairs == 0:
    return out
  _lib.check(lib.ksig_signature_levels(
      _lib.dtype_code(X.data.dtype), X.data.data_ptr(), X.off.data_ptr(),
      X.len.data_ptr(), X.max_len, Y.data.data_ptr(), Y.off.data_ptr(),
      Y.len.data_ptr(), Y.max_len, X.d, pairs.data_ptr(), n_pairs, kind,
      float(param), n_levels, order, int(bool(difference)), int(gram_mode),
      out.data_ptr(), _lib.stream_handle(X.data.device)))
  return out


def signature_levels_gram(M: torch.Tensor, n_levels: int, order: int,
                          difference: bool) -> torch.Tensor:
  """Level values [prod(lead), n_levels+1] for matrices M [..., lx, ly]."""
  lib = _lib.load()
  lx, ly = M.shape[-2:]
  Mc = M.reshape(-1, lx, ly).contiguous()
  n = Mc.shape[0]
  out = torch.empty((n, n_levels + 1), device=M.device, dtype=torch.float64)
  if n == 0:
    return out
  _lib.check(lib.ksig_signature_levels_gram(
      _lib.dtype_code(Mc.dtype), Mc.data_ptr(), n, lx, ly, n_levels, order,
      int(bool(difference)), out.data_ptr(), _lib.stream_handle(M.device)))
  return out


def signature_pde(X: SeqSet, Y: SeqSet, pairs: torch.Tensor, difference: bool,
                  kind: int, param: float, dyadic_order: int) -> torch.Tensor:
  """Signature-PDE values [n_pairs] (float64)."""
  lib = _lib.load()
  _check_sets(X, Y)
  n_pairs = pairs.shape[0]
  out = torch.empty((n_pairs,), device=X.data.device, dtype=torch.float64)
  if n_pairs == 0:
    return out
  _lib.check(lib.ksig_signature_pde(
      _lib.dtype_code(X.data.dtype), X.data.data_ptr(), X.off.data_ptr(),
      X.len.data_ptr(), X.max_len, Y.data.data_ptr(), Y.off.data_ptr(),
      Y.len.data_ptr(), Y.max_len, X.d, pairs.data_ptr(), n_pairs, kind,
      float(param), int(dyadic_order), int(bool(difference)), out.data_ptr(),
      _lib.stream_handle(X.data.device)))
  return out


def signature_pde_gram(M: torch.Tensor, difference: bool, dyadic_order: int
                       ) -> torch.Tensor:
  lib = _lib.load()
  lx, ly = M.shape[-2:]
  Mc = M.reshape(-1, lx, ly).contiguous()
  n = Mc.shape[0]
  out = torch.empty((n,), device=M.device, dtype=torch.float64)
  if n == 0:
    return out
  _lib.check(lib.ksig_signature_pde_gram(
      _lib.dtype_code(Mc.dtype), Mc.data_ptr(), n, lx, ly, int(dyadic_order),
      int(bool(difference)), out.data_ptr(), _lib.stream_handle(M.device)))
  return out
