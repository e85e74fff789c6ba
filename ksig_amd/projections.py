"""Random projections for the low-rank signature features: the
``GaussianRandomProjection`` of ksig/projections.py:189-242, the
``SubsamplingProjection`` of :245-284, the ``VerySparseRandomProjection`` of
:287-376, the ``TensorSketch`` of :379-458, the
``TensorizedRandomProjection`` of :465-584 and the ``DiagonalProjection`` of
:587-674 over the ``RandomProjection`` base (:20-187).  Components are drawn
on the host with NumPy (the reference draws them with CuPy)."""

from __future__ import annotations

from abc import ABCMeta, abstractmethod
from typing import Optional

import numpy as np
import torch
from sklearn.base import BaseEstimator, TransformerMixin
from sklearn.utils.validation import check_is_fitted

from . import utils
from .static.features import check_random_state


class RandomProjection(BaseEstimator, TransformerMixin, metaclass=ABCMeta):
  """Base class of random projections (ksig/projections.py:20-187)."""

  def __init__(self, n_components: int = 100, random_state=None):
    self.n_components = utils.check_positive_value(n_components, 'n_components')
    self.random_state = check_random_state(random_state)

  @abstractmethod
  def _make_projection_components(self, X, Y=None):
    """Initializes internal variables, called by ``fit``."""

  def _validate_format(self, X, Y=None, reset: bool = False):
    n_features = X.shape[-1] * (Y.shape[-1] if Y is not None else 1)
    if reset or not hasattr(self, 'n_features_') or self.n_features_ is None:
      self.n_features_ = n_features
    elif n_features != self.n_features_:
      raise ValueError('Data has a different number of features than during fit.',
                       f'({n_features} != {self.n_features_})')

  def fit(self, X, Y=None) -> 'RandomProjection':
    X = utils.as_device(X)
    Y = utils.as_device(Y) if Y is not None else None
    self._validate_format(X, Y=Y, reset=True)
    self._make_projection_components(X, Y=Y)
    return self

  @abstractmethod
  def _project(self, X):
    """Projects X."""

  @abstractmethod
  def _project_outer_prod(self, X, Y):
    """Projects the outer product of X and Y along the last axis."""

  def transform(self, X, Y=None, return_on_gpu: bool = False):
    check_is_fitted(self)
    X = utils.as_device(X)
    Y = utils.as_device(Y) if Y is not None else None
    self._validate_format(X, Y=Y)
    P = self._project_outer_prod(X, Y) if Y is not None else self._project(X)
    return P if return_on_gpu else utils.to_host(P)

  def __call__(self, X, Y=None, return_on_gpu: bool = False):
    return self.transform(X, Y=Y, return_on_gpu=return_on_gpu)


class GaussianRandomProjection(RandomProjection):
  """Gaussian random projection (ksig/projections.py:189-242): components
  N(0, 1) of shape [n_components, n_features], X -> X C^T / sqrt(n_components)
  as one library GEMM; (X, Y) projects the outer product X (x) Y flattened
  with X's index major (ksig/utils.py:160-172), formed explicitly as the
  reference does."""

  def _make_projection_components(self, X, Y=None):
    c = self.random_state.normal(size=[self.n_components, self.n_features_])
    self.components_ = torch.as_tensor(c, device=X.device, dtype=X.dtype)
    self.scaling_ = 1. / np.sqrt(self.n_components)

  def _project(self, X):
    C = self.components_.to(device=X.device, dtype=X.dtype)
    P = self.scaling_ * (X.reshape(-1, self.n_features_) @ C.T)
    return P.reshape(tuple(X.shape[:-1]) + (self.n_components,))

  def _project_outer_prod(self, X, Y):
    return self._project((X[..., :, None] * Y[..., None, :]).reshape(tuple(X.shape[:-1]) + (-1,)))


class SubsamplingProjection(RandomProjection):
  """Subsampling projection (ksig/projections.py:245-284): n_components
  coordinates drawn without replacement, scaled by
  sqrt(n_features / n_components); on an outer product only the drawn
  entries are formed (index k = i * d_Y + j takes X_i Y_j, ksig/utils.py
  subsample_outer_prod)."""

  def _make_projection_components(self, X, Y=None):
    idx = self.random_state.choice(self.n_features_, size=self.n_components, replace=False)
    self.sampled_idx_ = torch.as_tensor(idx, device=X.device)
    self.scaling_ = float(np.sqrt(self.n_features_ / self.n_components))

  def _project(self, X):
    return self.scaling_ * X[..., self.sampled_idx_.to(X.device)]

  def _project_outer_prod(self, X, Y):
    idx = self.sampled_idx_.to(X.device)
    dy = Y.shape[-1]
    return self.scaling_ * X[..., idx // dy] * Y[..., idx % dy]


class VerySparseRandomProjection(RandomProjection):
  """Very sparse random projection (ksig/projections.py:287-376) as the
  reference computes it: the Bernoulli mask of density p (log(F)/F or
  1/sqrt(F)) is drawn and then replaced by the Rademacher draw (:340-345
  assign rather than multiply), so the components are dense +-1 over every
  column, scaled by sqrt(1 / (p n_components)); kept as is so that the
  features equal the reference's."""

  def __init__(self, n_components: int = 100, sparsity: str = 'log', random_state=None):
    if sparsity not in ('sqrt', 'log'):
      raise ValueError(f'Unknown sparsity mode ({sparsity}).',
                       "Possible values are ['sqrt', 'log'].")
    self.sparsity = sparsity
    super().__init__(n_components=n_components, random_state=random_state)

  def _make_projection_components(self, X, Y=None):
    F = self.n_features_
    prob = np.log(F) / F if self.sparsity == 'log' else 1. / np.sqrt(F)
    mask = self.random_state.uniform(size=[self.n_components, F]) < prob
    mask[0, 0] = True
    full = np.where(self.random_state.uniform(size=[self.n_components, F]) < 0.5, 1., -1.)
    idx = np.where(np.any(np.abs(full) > 1e-12, axis=0))[0]
    self.sampled_idx_ = torch.as_tensor(idx, device=X.device)
    self.n_sampled_ = len(idx)
    self.components_ = torch.as_tensor(full[:, idx], device=X.device, dtype=X.dtype)
    self.scaling_ = float(np.sqrt(1. / (prob * self.n_components)))

  def _project(self, X):
    C = self.components_.to(device=X.device, dtype=X.dtype)
    Xs = X[..., self.sampled_idx_.to(X.device)].reshape(-1, self.n_sampled_)
    return (self.scaling_ * (Xs @ C.T)).reshape(tuple(X.shape[:-1]) + (-1,))

  def _project_outer_prod(self, X, Y):
    idx = self.sampled_idx_.to(X.device)
    dy = Y.shape[-1]
    XY = X[..., idx // dy] * Y[..., idx % dy]
    C = self.components_.to(device=X.device, dtype=X.dtype)
    return (self.scaling_ * (XY.reshape(-1, self.n_sampled_) @ C.T)).reshape(
        tuple(XY.shape[:-1]) + (-1,))


class TensorSketch(RandomProjection):
  """Tensor sketch by convolution of count sketches (ksig/projections.py:
  379-458): each input feature j gets a hash h[j] in [0, n_components) and a
  Rademacher sign s[j]; X -> the count sketch CS(X)[k] = sum_{j: h[j] = k}
  s[j] X[..., j] (ksig/utils.py:279-298), and (X, Y) -> the circular
  convolution of an already-sketched X with CS(Y) along the last axis
  (ksig/utils.py:301-313; an FFT of length n_components here as there).
  Inside SignatureFeatures the whole level recursion runs in one HIP scan
  (ksig_lowrank_levels_sketch)."""

  def _validate_format(self, X, Y=None, reset: bool = False):
    """ksig/projections.py:393-418."""
    n_features = Y.shape[-1] if Y is not None else X.shape[-1]
    if reset or not hasattr(self, 'n_features_') or self.n_features_ is None:
      self.n_features_ = n_features
    elif n_features != self.n_features_:
      raise ValueError('Received data with a different number of features than at fit time.',
                       f'({n_features} != {self.n_features_})')

  def _make_projection_components(self, X, Y=None):
    """ksig/projections.py:420-430: the hashes, then the signs
    (draw_rademacher_matrix, ksig/utils.py:215-230: +1 where a uniform draw
    is below 1/2)."""
    idx = self.random_state.randint(self.n_components, size=[self.n_features_])
    bit = np.where(self.random_state.uniform(size=[self.n_features_]) < 0.5, 1., -1.)
    self.hash_idx_ = torch.as_tensor(idx, device=X.device, dtype=torch.int64)
    self.hash_bit_ = torch.as_tensor(bit, device=X.device, dtype=X.dtype)

  def _count_sketch(self, X):
    idx = self.hash_idx_.to(X.device)
    bit = self.hash_bit_.to(device=X.device, dtype=X.dtype)
    out = torch.zeros(tuple(X.shape[:-1]) + (self.n_components,), device=X.device,
                      dtype=X.dtype)
    return out.index_add_(-1, idx, X * bit)

  def _project(self, X):
    return self._count_sketch(X)

  def _project_outer_prod(self, X, Y):
    n = self.n_components
    return torch.fft.irfft(torch.fft.rfft(X, n=n, dim=-1) *
                           torch.fft.rfft(self._count_sketch(Y), n=n, dim=-1), n=n, dim=-1)


CountSketch = TensorSketch
CountSketchRandomProjection = TensorSketch


class TensorizedRandomProjection(RandomProjection):
  """Tensorized random projection (ksig/projections.py:465-584): X -> X C /
  sqrt(n_components) alone, and (X, Y) -> X * (Y C) for an already-projected
  X (the recursive step of a tensor's projection).  ``rank`` behaves as in
  the reference: the components hold n_components columns whatever the rank
  (:538-548), the outer-product step expects X with n_components * rank
  features (:527-531), so a rank above 1 projects standalone data and is
  refused with the reference's ValueError when fitted as a recursive step
  (the low-rank signature features with more than one level)."""

  def __init__(self, n_components: int = 100, rank: int = 1, random_state=None):
    super().__init__(n_components=n_components, random_state=random_state)
    self.rank = utils.check_positive_value(rank, 'rank')

  def _validate_format(self, X, Y=None, reset: bool = False):
    """ksig/projections.py:500-536."""
    n_features = Y.shape[-1] if Y is not None else X.shape[-1]
    if reset or not hasattr(self, 'n_features_') or self.n_features_ is None:
      self.n_features_ = n_features
    if Y is not None:
      if Y.shape[-1] != self.n_features_:
        raise ValueError('Data `Y` has a different number of features than during fit.',
                         f'({Y.shape[-1]} != {self.n_features_})')
      if X.shape[-1] != self.n_components * self.rank:
        raise ValueError('Data `X` has incorrect number of features.',
                         f'({X.shape[-1]} != {self.n_components * self.rank})')
    elif X.shape[-1] != self.n_features_:
      raise ValueError('Data `X` has a different number of features than during fit.',
                       f'({X.shape[-1]} != {self.n_features_})')

  def _make_projection_components(self, X, Y=None):
    """ksig/projections.py:538-548."""
    c = self.random_state.normal(size=[self.n_features_, self.n_components])
    self.components_ = torch.as_tensor(c, device=X.device, dtype=X.dtype)
    self.scaling_ = 1. / np.sqrt(self.n_components) if Y is None else 1.

  def _project(self, X):
    C = self.components_.to(device=X.device, dtype=X.dtype)
    return self.scaling_ * (X.reshape(-1, self.n_features_) @ C).reshape(
        tuple(X.shape[:-1]) + (-1,))

  def _project_outer_prod(self, X, Y):
    C = self.components_.to(device=Y.device, dtype=Y.dtype)
    return self.scaling_ * X * (Y.reshape(-1, self.n_features_) @ C).reshape(
        tuple(Y.shape[:-1]) + (-1,))


class DiagonalProjection(RandomProjection):
  """Diagonal projection (ksig/projections.py:587-674): only for the low-rank
  signature features on random Fourier features (RFSF-DP).  ``_project``
  splits the last axis into [internal_size, -1]; the outer-product step is
  the Hadamard product over the component axis with an outer product over
  the internal axis, scaled by sqrt(n_features_).  No random components."""

  def __init__(self, internal_size: int = 2):
    self.internal_size = internal_size

  def _validate_format(self, X, Y=None, reset: bool = False):
    """ksig/projections.py:611-637."""
    q = self.internal_size
    if reset or not hasattr(self, 'n_features_') or self.n_features_ is None:
      self.n_features_ = X.shape[-1]
    if X.shape[-1] != self.n_features_:
      raise ValueError('Data `X` has a different number of features than during fit.',
                       f'({X.shape[-1]} != {self.n_features_})')
    if Y is not None and Y.shape[-1] // q != self.n_features_:
      raise ValueError('Data `Y` has a different number of features than during fit.',
                       f'({Y.shape[-1]} != {self.n_features_})')

  def _make_projection_components(self, X, Y=None):
    pass

  def _project(self, X):
    return X.reshape(tuple(X.shape[:-1]) + (self.internal_size, -1))

  def _project_outer_prod(self, X, Y):
    q = self.internal_size
    Y = Y.reshape(tuple(Y.shape[:-1]) + (q, -1))
    return np.sqrt(self.n_features_) * (X[..., :, None, :] * Y[..., None, :, :]).reshape(
        tuple(X.shape[:-2]) + (-1, X.shape[-1]))
