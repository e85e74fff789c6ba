"""SVC with a precomputed sequence-kernel Gram: the counterpart of
``ksig.models.PrecomputedKernelSVC`` (ksig/models/pre_base.py:17-282,
ksig/models/pre_svc.py:25-106), the caller that feeds the signature kernels
to ``sklearn.svm.SVC(kernel='precomputed')``.

The reference fills the Gram tile by tile when ``batch_size`` is set
(ksig/models/pre_base.py:88-142): a Python double loop over
[batch_size x batch_size] blocks, one kernel call and one device-to-host copy
per block, symmetric blocks mirrored on the host.  Its blocks exist to bound
the [b, b, L, L] static-kernel tensor the CuPy recursion materialises.  The
kernels of this library keep O(levels * L) state per pair and never
materialise that tensor, so for them the whole (upper-triangle) Gram is one
persistent launch over the pair list and one copy to the host, whatever
``batch_size`` says; the values are the same, since per-block normalisation
by the diagonals equals normalising the whole matrix.  Any other callable
kernel goes through the reference's block loop unchanged.
"""

from __future__ import annotations

from math import ceil
from typing import Dict, Optional

import numpy as np
import torch
from sklearn.base import BaseEstimator, ClassifierMixin
from sklearn.model_selection import GridSearchCV, StratifiedKFold
from sklearn.svm import SVC, LinearSVC
from sklearn.utils.validation import check_is_fitted

from .static.kernels import Kernel

# ksig/models/pre_svc.py:19-22
_DEFAULT_SVC_HPARAMS = {
  'decision_function_shape': 'ovo',
  'max_iter': 10000,
}
# ksig/models/pre_lin_svc.py:13-17
_DEFAULT_LIN_SVC_HPARAMS = {
  'dual': False,
  'fit_intercept': False,
  'tol': 1e-3,
}


def _as_matrix(F) -> torch.Tensor:
  """A feature block as a 2-D tensor (host arrays of foreign kernels too)."""
  F = F if isinstance(F, torch.Tensor) else torch.as_tensor(np.asarray(F))
  return F.reshape(F.shape[0], -1)


def _is_library_kernel(kernel) -> bool:
  """Kernels of this library: one launch for any Gram, no L x L tensor."""
  from .kernels import SignatureBase
  return isinstance(kernel, SignatureBase)


class PrecomputedKernelSVC(BaseEstimator, ClassifierMixin):
  """``SVC`` on a precomputed kernel matrix with optional ``GridSearchCV``
  (ksig/models/pre_svc.py:25-106 over ksig/models/pre_base.py:17-282).

  Same constructor arguments and stored attributes as the reference (sklearn
  ``get_params`` / ``clone`` read them).  ``has_transform=True`` takes a
  feature-map kernel (``SignatureFeatures`` or any object with
  ``transform(X, return_on_gpu=...)``): the Gram is the product of the
  feature matrices, one device GEMM in float64 (the reference multiplies
  blocks of its ``_MAT_MUL_BSIZE`` rows, ksig/models/pre_base.py:172-224, to
  bound device memory; one [n, n] float64 product fits HBM for any n SVC can
  take)."""

  def __init__(self,
               kernel: Kernel,
               svc_hparams: Dict = {},
               svc_grid: Optional[Dict] = None,
               cv: int = 3,
               n_jobs: int = -1,
               need_kernel_fit: bool = False,
               batch_size: Optional[int] = None,
               has_transform: bool = False):
    self.kernel = kernel
    self.svc_hparams = svc_hparams
    self.svc_grid = svc_grid
    self.cv = cv
    self.n_jobs = n_jobs
    self.need_kernel_fit = need_kernel_fit
    self.batch_size = batch_size
    self.has_transform = has_transform
    # ksig/models/pre_svc.py:55-59 (the reference mutates the dict it is given)
    for key, val in _DEFAULT_SVC_HPARAMS.items():
      self.svc_hparams.setdefault(key, val)
    self.svc_hparams['kernel'] = 'precomputed'

  def _get_svc_model(self) -> SVC:
    return SVC(**self.svc_hparams)

  def _precompute_kernel_mat(self, X, X2=None) -> np.ndarray:
    """Kernel matrix on the host, [n, n] (X2 None) or [n, n2]
    (ksig/models/pre_base.py:88-142)."""
    if self.batch_size is None or _is_library_kernel(self.kernel):
      K = self.kernel(X) if X2 is None else self.kernel(X, X2)
      return np.asarray(K, dtype=np.float64)
    b = int(self.batch_size)
    n = X.shape[0]
    if X2 is None:
      K = np.zeros([n, n])
      nb = ceil(n / b)
      for i in range(nb):
        i0, i1 = i * b, min((i + 1) * b, n)
        for j in range(i, nb):
          j0, j1 = j * b, min((j + 1) * b, n)
          if i0 == j0 and i1 == j1:
            K[i0:i1, j0:j1] = self.kernel(X[i0:i1])
          else:
            K[i0:i1, j0:j1] = self.kernel(X[i0:i1], X[j0:j1])
            K[j0:j1, i0:i1] = K[i0:i1, j0:j1].T
      return K
    n2 = X2.shape[0]
    K = np.zeros([n, n2])
    for i in range(ceil(n / b)):
      i0, i1 = i * b, min((i + 1) * b, n)
      for j in range(ceil(n2 / b)):
        j0, j1 = j * b, min((j + 1) * b, n2)
        K[i0:i1, j0:j1] = self.kernel(X[i0:i1], X2[j0:j1])
    return K

  def _precompute_feature_mat(self, X):
    """Feature matrix [n, n_features], on the device when the kernel is this
    library's (ksig/models/pre_base.py:143-170: in blocks of ``batch_size``
    sequences when set)."""
    if not hasattr(self.kernel, 'transform'):
      raise ValueError('has_transform=True needs a kernel with a transform method')
    if self.batch_size is None:
      return _as_matrix(self.kernel.transform(X, return_on_gpu=True))
    b = int(self.batch_size)
    parts = [_as_matrix(self.kernel.transform(X[i:i + b], return_on_gpu=True))
             for i in range(0, X.shape[0], b)]
    return torch.cat(parts, dim=0)

  @staticmethod
  def _precompute_matrix_mult(F, F2=None) -> np.ndarray:
    """F @ F2^T on the host as float64 (ksig/models/pre_base.py:172-224);
    with F2 None the product is mirrored from its upper triangle so that it
    is exactly symmetric, as the reference's block mirroring makes it."""
    A = F.to(torch.float64)
    if F2 is None:
      P = A @ A.T
      P = torch.triu(P) + torch.triu(P, diagonal=1).T
    else:
      P = A @ F2.to(device=A.device, dtype=torch.float64).T
    return P.cpu().numpy()

  def _precompute_model_inputs(self, X=None) -> np.ndarray:
    """ksig/models/pre_svc.py:69-104."""
    if X is None:
      if self.has_transform:
        self.X_feat = self._precompute_feature_mat(self.X)
        return self._precompute_matrix_mult(self.X_feat)
      return self._precompute_kernel_mat(self.X)
    if self.has_transform:
      return self._precompute_matrix_mult(self._precompute_feature_mat(X), self.X_feat)
    return self._precompute_kernel_mat(X, self.X)

  def fit(self, X, y) -> 'PrecomputedKernelSVC':
    """ksig/models/pre_base.py:225-261."""
    self.X, self.y = X, y
    self.model = self._get_svc_model()
    if self.svc_grid is not None:
      self.kfold = StratifiedKFold(n_splits=min(self.cv, np.min(np.bincount(y))))
      self.model = GridSearchCV(estimator=self.model, param_grid=self.svc_grid,
                                cv=self.kfold, n_jobs=self.n_jobs)
    if self.need_kernel_fit:
      if self.batch_size is not None:
        idx_fit = np.random.choice(range(X.shape[0]), min(self.batch_size, X.shape[0]),
                                   replace=False)
        X_fit = X[idx_fit]
      else:
        X_fit = X
      self.kernel.fit(X_fit)
    self.model.fit(self._precompute_model_inputs(), y)
    return self

  def predict(self, X) -> np.ndarray:
    """ksig/models/pre_base.py:263-282."""
    check_is_fitted(self.model)
    y_pred = self.model.predict(self._precompute_model_inputs(X))
    return np.asarray(y_pred)


class PrecomputedFeatureLinSVC(PrecomputedKernelSVC):
  """``LinearSVC`` on a precomputed feature matrix with optional
  ``GridSearchCV`` (ksig/models/pre_lin_svc.py:24-103 over
  ksig/models/pre_base.py:17-282), for feature-map kernels such as
  ``SignatureFeatures`` or ``RandomWarpingSeries``.

  The features are computed on the GPU by this library (in blocks of
  ``batch_size`` sequences when set) and handed to scikit-learn's
  ``LinearSVC`` on the host.  ``on_gpu=True`` selects cuML's GPU LinearSVC in
  the reference (ksig/models/pre_lin_svc.py:62-68), a CUDA library with no
  ROCm build: here it is accepted for drop-in use and the model still runs
  on the host, with the reference's CPU defaults."""

  def __init__(self,
               kernel: Kernel,
               svc_hparams: Dict = {},
               svc_grid: Optional[Dict] = None,
               cv: int = 5,
               n_jobs: int = -1,
               need_kernel_fit: bool = False,
               batch_size: Optional[int] = None,
               on_gpu: bool = False):
    self.kernel = kernel
    self.svc_hparams = svc_hparams
    self.svc_grid = svc_grid
    self.cv = cv
    self.n_jobs = 1 if on_gpu else n_jobs  # ksig/models/pre_lin_svc.py:49
    self.need_kernel_fit = need_kernel_fit
    self.batch_size = batch_size
    self.has_transform = True
    self.on_gpu = on_gpu
    for key, val in _DEFAULT_LIN_SVC_HPARAMS.items():
      self.svc_hparams.setdefault(key, val)

  def _get_svc_model(self) -> LinearSVC:
    return LinearSVC(**self.svc_hparams)

  def _precompute_model_inputs(self, X=None) -> np.ndarray:
    """ksig/models/pre_lin_svc.py:70-103: the feature matrix of the training
    sequences (X None) or of X, on the host."""
    F = self._precompute_feature_mat(self.X if X is None else X)
    return F.detach().cpu().numpy()
