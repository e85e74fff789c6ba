"""Featurized kernels: the ``KernelFeatures`` base of
ksig/static/features.py:21-177 (fit / transform / kernel through the
Euclidean inner product of the features), used by RandomWarpingSeries and
SignatureFeatures, and the static feature maps RandomFourierFeatures,
RandomFourierFeatures1D and NystroemFeatures."""

from __future__ import annotations

import warnings
from abc import ABCMeta, abstractmethod
from typing import Optional

import numpy as np
import torch
from sklearn.base import TransformerMixin
from sklearn.utils.validation import check_is_fitted

from .. import utils
from .kernels import Kernel


def check_random_state(random_state=None) -> np.random.RandomState:
  """ksig/utils.py:38-56 with a NumPy random state in place of CuPy's (the
  host draws the random components; the streams differ from CuPy's)."""
  if random_state is None or isinstance(random_state, (int, np.integer)):
    return np.random.RandomState(random_state)
  if isinstance(random_state, np.random.RandomState):
    return random_state
  raise ValueError(f'{random_state} cannot seed a `numpy.random.RandomState` instance')


class KernelFeatures(Kernel, TransformerMixin, metaclass=ABCMeta):
  """Base class for featurized kernels (ksig/static/features.py:21-177)."""

  def __init__(self, n_components: int = 100, random_state=None):
    self.n_components = utils.check_positive_value(n_components, 'n_components')
    self.random_state = check_random_state(random_state)

  def _validate_format(self, X, reset: bool = False):
    """ksig/static/features.py:43-59."""
    n_features = X.shape[-1]
    if reset or not hasattr(self, 'n_features_') or self.n_features_ is None:
      self.n_features_ = n_features
    elif n_features != self.n_features_:
      raise ValueError(
          'Data `X` has a different number of features than during fit.',
          f'({n_features} != {self.n_features_})')

  def _validate_data(self, X, reset: bool = False):
    self._validate_format(X, reset=reset)
    return X

  @abstractmethod
  def _make_feature_components(self, X):
    """Initializes internal variables, called by ``fit``."""

  def fit(self, X, y=None) -> 'KernelFeatures':
    X = self._validate_data(utils.as_device(X), reset=True)
    self._make_feature_components(X)
    return self

  @abstractmethod
  def _compute_features(self, X: torch.Tensor) -> torch.Tensor:
    """The feature map, called by ``transform``."""

  def transform(self, X, return_on_gpu: bool = False):
    """ksig/static/features.py:108-129."""
    check_is_fitted(self)
    X = utils.as_device(self._validate_data(X))
    F = self._compute_features(X)
    return F if return_on_gpu else utils.to_host(F)

  def _K(self, X, Y=None):
    Fx = self._compute_features(X)
    Fy = self._compute_features(Y) if Y is not None else None
    return utils.matrix_mult(Fx, Fy, transpose_Y=True)

  def _Kdiag(self, X):
    return utils.squared_norm(self._compute_features(X), axis=-1)

  def __call__(self, X, Y: Optional[object] = None, diag: bool = False,
               return_on_gpu: bool = False):
    """ksig/static/features.py:157-177."""
    check_is_fitted(self)
    return super().__call__(X, Y=Y, diag=diag, return_on_gpu=return_on_gpu)


class StaticFeatures(KernelFeatures):
  """Static features flatten the time axis (ksig/static/features.py:180-202)."""

  def _validate_data(self, X, reset: bool = False):
    if X.ndim > 2:
      X = X.reshape(tuple(X.shape[:-2]) + (-1,))
    return super()._validate_data(X, reset=reset)


class RandomFourierFeatures(StaticFeatures):
  """Random Fourier features of the RBF kernel (ksig/static/features.py:288-342):
  [sin(X W), cos(X W)] / sqrt(n_components) with W ~ N(0, 1/bandwidth^2)
  of shape [n_features, n_components].  The product X W is a library GEMM;
  the weights are drawn on the host with NumPy (the reference draws them with
  CuPy, so equal seeds give different weights)."""

  def __init__(self, bandwidth: float = 1., n_components: int = 100, random_state=None):
    super().__init__(n_components=n_components, random_state=random_state)
    self.bandwidth = utils.check_positive_value(bandwidth, 'bandwidth')

  def _make_feature_components(self, X):
    self.n_components_ = self.n_components
    w = self.random_state.normal(size=[self.n_features_, self.n_components_])
    self.random_weights_ = torch.as_tensor(w / self.bandwidth, device=X.device, dtype=X.dtype)

  def _compute_features(self, X):
    W = self.random_weights_.to(device=X.device, dtype=X.dtype)
    proj = X.reshape(-1, self.n_features_) @ W
    F = torch.cat([torch.sin(proj), torch.cos(proj)], -1) / np.sqrt(self.n_components)
    return F.reshape(tuple(X.shape[:-1]) + (-1,))


class NystroemFeatures(StaticFeatures):
  """Nystroem feature map (ksig/static/features.py:205-285): ``n_components``
  basis points drawn without replacement from the fitted points, the basis
  Gram K_bb = U diag(s) U^T, and features K(X, basis) U_+ diag(s_+)^(-1/2)
  over the eigenvalues with |s| > 1e-12 (:249-255).  Both Grams run through
  the base kernel on the device; the small symmetric eigenproblem is solved
  in float64 (the reference solves it in the data dtype), the factor is kept
  in float64 and the small product K(X, basis) @ factor runs in float64
  before the features are cast to the data dtype, so a float32 result
  carries only the rounding of K(X, basis) and of the final cast.  The basis
  indices are drawn with NumPy (the
  reference draws them with CuPy)."""

  def __init__(self, base_kernel: Kernel, n_components: int = 100, random_state=None):
    super().__init__(n_components=n_components, random_state=random_state)
    self.base_kernel = base_kernel

  def _make_feature_components(self, X):
    X = X.reshape(-1, self.n_features_)
    n_samples = X.shape[0]
    if self.n_components >= n_samples:
      warnings.warn('`n_samples <= n_components`, hence `n_components` was set to '
                    '`n_samples`, which results in inefficient evaluation of the full kernel.')
      self.n_components_ = n_samples
    else:
      self.n_components_ = self.n_components
    idx = self.random_state.choice(n_samples, size=self.n_components_, replace=False)
    basis = X[torch.as_tensor(idx, device=X.device)]
    Kb = self.base_kernel(basis, return_on_gpu=True).to(torch.float64)
    S, U = torch.linalg.eigh(Kb)
    keep = utils.robust_nonzero(S)
    self.normalization_ = U[:, keep] / utils.robust_sqrt(S[keep])[None, :]
    self.components_ = basis
    self.component_idx_ = idx

  def _compute_features(self, X):
    emb = self.base_kernel(X.reshape(-1, self.n_features_), self.components_,
                           return_on_gpu=True)
    F = (emb.to(torch.float64) @ self.normalization_.to(emb.device)).to(X.dtype)
    return F.reshape(tuple(X.shape[:-1]) + (-1,))


class RandomFourierFeatures1D(StaticFeatures):
  """Random Fourier features of the RBF kernel with random phases
  (ksig/static/features.py:344-390): cos(X W + b) sqrt(2 / n_components),
  W ~ N(0, 1/bandwidth^2) of shape [n_features, n_components], b ~ U(0, 2 pi).
  The product is a library GEMM; weights and phases are drawn on the host
  with NumPy (weights first, then phases, as the reference draws them)."""

  def __init__(self, bandwidth: float = 1., n_components: int = 100, random_state=None):
    super().__init__(n_components=n_components, random_state=random_state)
    self.bandwidth = utils.check_positive_value(bandwidth, 'bandwidth')

  def _make_feature_components(self, X):
    self.n_components_ = self.n_components
    w = self.random_state.normal(size=[self.n_features_, self.n_components_]) / self.bandwidth
    b = 2 * np.pi * self.random_state.uniform(size=[1, self.n_components_])
    self.random_weights_ = torch.as_tensor(w, device=X.device, dtype=X.dtype)
    self.random_shift_ = torch.as_tensor(b, device=X.device, dtype=X.dtype)

  def _compute_features(self, X):
    W = self.random_weights_.to(device=X.device, dtype=X.dtype)
    b = self.random_shift_.to(device=X.device, dtype=X.dtype)
    F = torch.cos(X.reshape(-1, self.n_features_) @ W + b) * np.sqrt(2 / self.n_components)
    return F.reshape(tuple(X.shape[:-1]) + (-1,))
