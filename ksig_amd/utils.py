"""Validation and small linear-algebra helpers (ksig/utils.py counterparts).

Arrays are torch tensors on the GPU; host numpy arrays are moved there by
``as_device`` (the counterpart of ``cp.asarray``).
"""

from __future__ import annotations

from numbers import Integral, Number
from typing import Optional, Sequence, Union

import numpy as np
import torch

_EPS = 1e-12  # ksig/utils.py:16

ArrayOnCPU = np.ndarray
ArrayOnGPU = torch.Tensor
ArrayOnCPUOrGPU = Union[np.ndarray, torch.Tensor]
RandomState = np.random.RandomState
RandomStateOrSeed = Union[Integral, np.random.RandomState]


def check_positive_value(scalar: Number, name: str) -> Number:
  """ksig/utils.py:23-35."""
  if scalar <= 0:
    raise ValueError(f'The parameter \'{name}\' should have a positive value.')
  return scalar


def device() -> torch.device:
  if not torch.cuda.is_available():
    raise RuntimeError('ksig_amd needs a visible AMD GPU (torch.cuda.is_available() is False)')
  return torch.device('cuda', torch.cuda.current_device())


def as_device(X) -> torch.Tensor:
  """Counterpart of cp.asarray: host arrays are copied to the GPU, device
  tensors are used as they are.  Non-floating inputs become float64."""
  if isinstance(X, torch.Tensor):
    T = X if X.is_cuda else X.to(device())
  else:
    T = torch.as_tensor(np.asarray(X), device=device())
  if not T.is_floating_point():
    T = T.to(torch.float64)
  elif T.dtype not in (torch.float32, torch.float64):
    T = T.to(torch.float32)
  return T


def to_host(K: torch.Tensor) -> np.ndarray:
  """Counterpart of cp.asnumpy (device -> host copy, implicit sync)."""
  return K.detach().cpu().numpy()


def matrix_diag(A: torch.Tensor) -> torch.Tensor:
  """ksig/utils.py:93-102: diagonals of a batch of matrices."""
  return torch.diagonal(A, dim1=-2, dim2=-1)


def matrix_mult(X: torch.Tensor, Y: Optional[torch.Tensor] = None,
                transpose_X: bool = False, transpose_Y: bool = False
                ) -> torch.Tensor:
  """ksig/utils.py:105-122."""
  Y = X if Y is None else Y
  if transpose_X:
    X = X.transpose(-1, -2)
  if transpose_Y:
    Y = Y.transpose(-1, -2)
  return torch.matmul(X, Y.contiguous())


def squared_norm(X: torch.Tensor, axis: int = -1) -> torch.Tensor:
  """ksig/utils.py:125-135."""
  return torch.sum(X * X, dim=axis)


def squared_euclid_dist(X: torch.Tensor, Y: Optional[torch.Tensor] = None
                        ) -> torch.Tensor:
  """ksig/utils.py:138-157 (no clamp of negative round-off)."""
  Xn2 = squared_norm(X)
  if Y is None:
    return Xn2[..., :, None] + Xn2[..., None, :] - 2 * matrix_mult(X, X, transpose_Y=True)
  Yn2 = squared_norm(Y)
  return Xn2[..., :, None] + Yn2[..., None, :] - 2 * matrix_mult(X, Y, transpose_Y=True)


def robust_nonzero(X: torch.Tensor) -> torch.Tensor:
  """ksig/utils.py:199-208: |X| > eps elementwise."""
  return torch.abs(X) > _EPS


def robust_sqrt(X: torch.Tensor) -> torch.Tensor:
  """ksig/utils.py:173-182."""
  return torch.sqrt(torch.clamp(X, min=_EPS))


def euclid_dist(X: torch.Tensor, Y: Optional[torch.Tensor] = None) -> torch.Tensor:
  """Pairwise Euclidean distances, the robust square root of
  ``squared_euclid_dist`` (ksig/utils.py:185-196, without its stray ``self``
  parameter)."""
  return robust_sqrt(squared_euclid_dist(X, Y))


def outer_prod(X: torch.Tensor, Y: torch.Tensor) -> torch.Tensor:
  """ksig/utils.py:160-170: batched outer product flattened to [..., d1*d2]."""
  return (X[..., :, None] * Y[..., None, :]).reshape(X.shape[:-1] + (-1,))


def multi_cumsum(M: torch.Tensor, exclusive: bool = False, axis=-1) -> torch.Tensor:
  """ksig/utils.py:63-90: cumulative sum over one or several axes;
  ``exclusive`` drops the last entry along each axis and pre-pads a zero."""
  ndim = M.dim()
  axes = [axis] if isinstance(axis, int) else list(axis)
  axes = [a % ndim for a in axes]
  if exclusive:
    M = M[tuple(slice(0, -1) if a in axes else slice(None) for a in range(ndim))]
  for a in axes:
    M = torch.cumsum(M, a)
  if exclusive:
    pad = []
    for a in reversed(range(ndim)):
      pad += [1, 0] if a in axes else [0, 0]
    M = torch.nn.functional.pad(M, pad)
  return M


def check_random_state(random_state=None) -> np.random.RandomState:
  """ksig/utils.py:38-60 with a NumPy RandomState in place of CuPy's: None
  or an integer seeds a new one, a RandomState is returned as it is."""
  if random_state is None or isinstance(random_state, Integral):
    return np.random.RandomState(random_state)
  if isinstance(random_state, np.random.RandomState):
    return random_state
  raise ValueError(f'{random_state} cannot seed a `numpy.random.RandomState` instance')
