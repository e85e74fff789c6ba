"""Static (state-space) kernels: the ksig.static.kernels counterparts
(ksig/static/kernels.py:14-343).

``RBFKernel`` and ``LinearKernel`` are recognised by the sequence kernels and
fused into the HIP sweep (their Gram is never materialised there); called on
their own they evaluate the Gram with GPU tensor ops.  The polynomial, Matern
and rational quadratic kernels are evaluated with GPU tensor ops and reach the
signature kernels as a caller-supplied static-kernel tensor.
"""

from __future__ import annotations

from abc import ABCMeta, abstractmethod
from typing import Optional

import torch
from sklearn.base import BaseEstimator

from .. import utils
from ..utils import _EPS


class Kernel(BaseEstimator, metaclass=ABCMeta):
  """Base class for kernels (ksig/static/kernels.py:18-97)."""

  def fit(self, X, y=None) -> 'Kernel':
    return self

  @abstractmethod
  def _K(self, X: torch.Tensor, Y: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Kernel matrix between X and Y (or X and X)."""

  @abstractmethod
  def _Kdiag(self, X: torch.Tensor) -> torch.Tensor:
    """Diagonal kernel entries of X."""

  @abstractmethod
  def _validate_data(self, X, reset: bool = False):
    """Validate / reshape an input array."""

  def __call__(self, X, Y=None, diag: bool = False, return_on_gpu: bool = False):
    """ksig/static/kernels.py:71-97: validate, move to the GPU, compute
    K(X, Y) (or its diagonal), and return it on the host unless
    ``return_on_gpu`` (then a torch tensor on the GPU)."""
    X = utils.as_device(self._validate_data(X))
    if diag:
      K = self._Kdiag(X)
    else:
      Y = utils.as_device(self._validate_data(Y)) if Y is not None else None
      K = self._K(X, Y)
    if not return_on_gpu:
      K = utils.to_host(K)
    return K


class StaticKernel(Kernel, metaclass=ABCMeta):
  """Static kernels merge the last two axes of inputs with ndim > 2
  (ksig/static/kernels.py:100-127)."""

  def _validate_data(self, X, reset: bool = False):
    if X.ndim > 2:
      X = X.reshape(tuple(X.shape[:-2]) + (-1,))
    return X


class LinearKernel(StaticKernel):
  """k(x, y) = scale * <x, y> (ksig/static/kernels.py:134-166)."""

  def __init__(self, scale: float = 1.):
    self.scale = utils.check_positive_value(scale, 'scale')

  def _K(self, X, Y=None):
    return self.scale * utils.matrix_mult(X, Y, transpose_Y=True)

  def _Kdiag(self, X):
    return self.scale * utils.squared_norm(X, axis=-1)


class StationaryKernel(StaticKernel):
  """Stationary kernels with unit diagonal (ksig/static/kernels.py:214-242)."""

  def __init__(self, bandwidth: float = 1.) -> None:
    self.bandwidth = utils.check_positive_value(bandwidth, 'bandwidth')

  def _Kdiag(self, X):
    # integer ones, as cp.full((n,), 1) of ksig/static/kernels.py:242
    return torch.full((X.shape[0],), 1, device=X.device)


class RBFKernel(StationaryKernel):
  """k(x, y) = exp(-|x - y|^2 / max(2 sigma^2, eps))
  (ksig/static/kernels.py:245-260)."""

  def _K(self, X, Y=None):
    D2_scaled = utils.squared_euclid_dist(X, Y) / max(2 * self.bandwidth**2, _EPS)
    return torch.exp(-D2_scaled)


class PolynomialKernel(StaticKernel):
  """k(x, y) = scale * (<x, y> + gamma)^degree (ksig/static/kernels.py:169-211)."""

  def __init__(self, degree: float = 3., gamma: float = 1., scale: float = 1.):
    self.degree = utils.check_positive_value(degree, 'degree')
    self.gamma = gamma
    self.scale = utils.check_positive_value(scale, 'scale')

  def _K(self, X, Y=None):
    return self.scale * torch.pow(utils.matrix_mult(X, Y, transpose_Y=True) + self.gamma,
                                  self.degree)

  def _Kdiag(self, X):
    return self.scale * torch.pow(utils.squared_norm(X, axis=-1) + self.gamma, self.degree)


class Matern12Kernel(StationaryKernel):
  """k(x, y) = exp(-|x - y| / max(sigma, eps)) (ksig/static/kernels.py:263-277).

  The reference computes |x - y| through ``utils.euclid_dist``, which is
  declared with a stray ``self`` parameter (ksig/utils.py:185), so its call
  ``euclid_dist(X, Y)`` evaluates the distances of Y with itself (and fails
  for Y = None); this class evaluates the intended distances of X and Y."""

  def _K(self, X, Y=None):
    return torch.exp(-utils.euclid_dist(X, Y) / max(self.bandwidth, _EPS))


class Matern32Kernel(StationaryKernel):
  """k = (1 + r) exp(-r), r = sqrt(3) |x - y| / max(sigma, eps)
  (ksig/static/kernels.py:280-296; distances of X and Y as in
  ``Matern12Kernel``)."""

  def _K(self, X, Y=None):
    D = 3.0 ** 0.5 * utils.euclid_dist(X, Y) / max(self.bandwidth, _EPS)
    return (1. + D) * torch.exp(-D)


class Matern52Kernel(StationaryKernel):
  """k = (1 + r + r^2 / 3) exp(-r), r^2 = 5 |x - y|^2 / max(sigma^2, eps)
  with r the robust square root (ksig/static/kernels.py:299-315)."""

  def _K(self, X, Y=None):
    D2 = 5 * utils.squared_euclid_dist(X, Y) / max(self.bandwidth**2, _EPS)
    D = utils.robust_sqrt(D2)
    return (1. + D + D2 / 3.) * torch.exp(-D)


class RationalQuadraticKernel(StationaryKernel):
  """k = (1 + |x - y|^2 / max(2 alpha sigma^2, eps))^(-alpha)
  (ksig/static/kernels.py:318-343)."""

  def __init__(self, bandwidth: float = 1., alpha: float = 1.):
    super().__init__(bandwidth=bandwidth)
    self.alpha = utils.check_positive_value(alpha, 'alpha')

  def _K(self, X, Y=None):
    D2 = utils.squared_euclid_dist(X, Y) / max(2 * self.alpha * self.bandwidth**2, _EPS)
    return torch.pow(1 + D2, -self.alpha)
