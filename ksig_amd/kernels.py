"""Signature kernels for sequential data: drop-in counterparts of
``ksig.kernels.SignatureKernel`` and ``ksig.kernels.SignaturePDEKernel``
(ksig/kernels.py:21-256) computed by the gfx950 HIP library.

With an ``RBFKernel`` or ``LinearKernel`` static kernel the static Gram, its
double difference and the recursion are fused into one sweep per sequence
pair (nothing of size L x L is materialised).  Any other ``StaticKernel`` is
evaluated as in the reference (``static_kernel(X', Y', return_on_gpu=True)``)
and its tensor is fed to the same sweep.
"""

from __future__ import annotations

import warnings
from abc import ABCMeta
from typing import Optional

import torch

from . import _lib, _ops, algorithms, static, utils
import numpy as np
from .static.kernels import Kernel, StaticKernel
from .utils import _EPS


class SignatureBase(Kernel, metaclass=ABCMeta):
  """Base class for signature kernels (ksig/kernels.py:25-90)."""

  def __init__(self, n_levels: int = 5, order: int = 1, difference: bool = True,
               normalize: bool = True, n_features: Optional[int] = None):
    self.n_levels = utils.check_positive_value(n_levels, 'n_levels')
    self.order = (self.n_levels if order <= 0 or order >= self.n_levels
                  else order)
    self.normalize = normalize
    self.difference = difference
    self.n_features = (utils.check_positive_value(n_features, 'n_features')
                       if n_features is not None else None)

  def _validate_data(self, X, reset: Optional[bool] = False):
    """ksig/kernels.py:56-90.  For ndim == 2 inputs without ``n_features`` the
    reference intends a warning (its malformed ``warnings.warn`` call raises
    TypeError instead); this build emits the warning and proceeds with
    univariate sequences."""
    n_features = (self.n_features_ if hasattr(self, 'n_features_')
                  and self.n_features_ is not None else self.n_features)
    if X.ndim == 2:
      if n_features is None or reset:
        warnings.warn(
          '`X` has ndim==2. Assuming inputs are univariate time series. '
          'Recommend passing an `n_features` parameter during init when using '
          'flattened arrays of ndim==2.')
        n_features = 1
    elif X.ndim == 3:
      if n_features is None or reset:
        n_features = X.shape[-1]
      elif X.shape[-1] != n_features:
        raise ValueError(
          'Number of features in `X` does not match saved `n_features` param.')
    else:
      raise ValueError(
        'Only input sequence arrays with ndim==2 or ndim==3 are supported.')
    X = X.reshape([X.shape[0], -1, n_features])
    if reset:
      self.n_features_ = n_features
    return X


def _normalise(K: torch.Tensor, Kxd: torch.Tensor, Kyd: torch.Tensor) -> torch.Tensor:
  """ksig/kernels.py:194-207 with ksig/utils.py:173-182."""
  sx = torch.clamp(utils.robust_sqrt(Kxd), min=_EPS)
  sy = torch.clamp(utils.robust_sqrt(Kyd), min=_EPS)
  return K / (sx[..., :, None] * sy[..., None, :])


def _mirror(vals: torch.Tensor, pairs: torch.Tensor, n: int) -> torch.Tensor:
  """Symmetric [..., n, n] matrix from upper-triangle pair values [..., P]
  (out-of-place index_add, so autograd sums both mirror images)."""
  i, j = pairs[:, 0].long(), pairs[:, 1].long()
  off = i != j
  lead = vals.shape[:-1]
  K = torch.zeros(lead + (n * n,), device=vals.device, dtype=vals.dtype)
  K = K.index_add(-1, i * n + j, vals)
  K = K.index_add(-1, (j * n + i)[off], vals[..., off])
  return K.reshape(lead + (n, n))


class SignatureKernel(SignatureBase):
  """Full-rank signature kernel (ksig/kernels.py:97-214)."""

  def __init__(self, n_levels: int = 5, order: int = 1, difference: bool = True,
               normalize: bool = True, n_features: Optional[int] = None,
               static_kernel: Optional[StaticKernel] = None,
               gram_precision: str = 'auto'):
    """``gram_precision`` (no reference counterpart) selects how float32
    static-kernel values are produced inside the sweep: 'auto' (the library's
    choice, which is the vector-unit sweep: measured faster than both
    matrix-core forms for the fused kernels), 'valu' (one float32 dot product
    per cell on the vector units), 'fp32' (float32 MFMA tiles), 'bf16x3' (bf16
    hi/lo split, three MFMA products).  float64 inputs always use exact float64
    arithmetic."""
    super().__init__(n_levels=n_levels, order=order, difference=difference,
                     normalize=normalize, n_features=n_features)
    self.static_kernel = static_kernel or static.kernels.RBFKernel()
    if gram_precision not in _lib.GRAM_MODES:
      raise ValueError(f'gram_precision must be one of {sorted(_lib.GRAM_MODES)}')
    self.gram_precision = gram_precision

  # -- embedding (generic static kernels only) --------------------------------
  def _compute_embedding(self, X, Y=None, diag: bool = False) -> torch.Tensor:
    """ksig/kernels.py:124-145: pairwise static-kernel evaluations."""
    if diag:
      return self.static_kernel(X[..., None, :], return_on_gpu=True)
    if Y is None:
      return self.static_kernel(X[:, None, :, None, :], X[None, :, :, None, :],
                                return_on_gpu=True)
    return self.static_kernel(X[:, None, :, None, :], Y[None, :, :, None, :],
                              return_on_gpu=True)

  # -- level values ------------------------------------------------------------
  def _levels(self, X, Y=None, diag: bool = False) -> torch.Tensor:
    """Per-level kernel values [n_levels+1, n_X, n_Y] (or [n_levels+1, n]
    for ``diag``) in float64."""
    code = _ops.static_code(self.static_kernel)
    if code is None:
      M = self._compute_embedding(X, Y, diag=diag)
      return algorithms.signature_kern(
          M, self.n_levels, order=self.order, difference=self.difference,
          return_levels=True).to(torch.float64)
    kind, param = code
    dev = X.device
    Xs = _ops.SeqSet.dense(X)
    Ys = Xs if (Y is None or diag) else _ops.SeqSet.dense(Y)
    grad = self._wants_grad(X, Y, self.static_kernel)

    def levels(pairs):
      if grad:
        from . import autograd
        return autograd.signature_levels(
            Xs, Ys, pairs, self.n_levels, kind, _ops.static_code(self.static_kernel, True)[1],
            xdata=X.reshape(-1, X.shape[-1]),
            ydata=(X if (Y is None or diag) else Y).reshape(-1, X.shape[-1]))
      return _ops.signature_levels(Xs, Ys, pairs, self.n_levels, self.order,
                                   self.difference, kind, param,
                                   _lib.GRAM_MODES[getattr(self, 'gram_precision', 'auto')])

    if diag:
      return levels(_ops.diag_pairs(X.shape[0], dev)).t()
    if Y is None:
      n = X.shape[0]
      pairs = _ops.triu_pairs(n, dev)
      lev = levels(pairs).t()
      return _mirror(lev, pairs, n)
    pairs = _ops.dense_pairs(X.shape[0], Y.shape[0], dev)
    return levels(pairs).t().reshape(self.n_levels + 1, X.shape[0], Y.shape[0])

  def _wants_grad(self, X, Y, static_kernel) -> bool:
    """Use the differentiable path when autograd needs gradients of the
    sequences or of the static-kernel parameter (first order, difference)."""
    if not torch.is_grad_enabled():
      return False
    p = _ops.static_code(static_kernel, True)[1]
    need = (X.requires_grad or (Y is not None and Y.requires_grad)
            or (isinstance(p, torch.Tensor) and p.requires_grad))
    if need and (self.order != 1 or not self.difference):
      raise NotImplementedError(
          'gradients are implemented for order=1, difference=True')
    return need

  def _compute_kernel(self, X, Y=None, diag: bool = False) -> torch.Tensor:
    """ksig/kernels.py:147-164: levels if ``normalize`` else their sum."""
    lev = self._levels(X, Y, diag=diag)
    return lev if self.normalize else lev.sum(0)

  def _Kdiag(self, X) -> torch.Tensor:
    """ksig/kernels.py:166-178."""
    if self.normalize:
      return torch.full((X.shape[0],), 1., device=X.device, dtype=torch.float64)
    return self._compute_kernel(X, diag=True).to(X.dtype)

  def _K(self, X, Y=None) -> torch.Tensor:
    """ksig/kernels.py:180-214: normalise each level by the diagonal and
    average over levels (including the constant level 0)."""
    K = self._compute_kernel(X, Y)
    if self.normalize:
      if Y is None:
        Kd = utils.matrix_diag(K)
        K = _normalise(K, Kd, Kd)
      else:
        K = _normalise(K, self._compute_kernel(X, diag=True),
                       self._compute_kernel(Y, diag=True))
    if K.ndim == 3:
      K = torch.mean(K, dim=0)
    return K.to(X.dtype)


class SignaturePDEKernel(SignatureKernel):
  """Signature-PDE kernel (ksig/kernels.py:218-256).

  ``dyadic_order`` has no reference counterpart: it refines every ΔΔ cell
  into 2^λ x 2^λ cells carrying M / 4^λ; the default 0 is the reference."""

  def __init__(self, difference: bool = True, normalize: bool = True,
               n_features: Optional[int] = None,
               static_kernel: Optional[StaticKernel] = None,
               dyadic_order: int = 0):
    self.difference = difference
    self.normalize = normalize
    self.n_features = (utils.check_positive_value(n_features, 'n_features')
                       if n_features is not None else None)
    self.static_kernel = static_kernel or static.kernels.RBFKernel()
    if dyadic_order < 0:
      raise ValueError('The parameter \'dyadic_order\' should be non-negative.')
    self.dyadic_order = dyadic_order

  def _compute_kernel(self, X, Y=None, diag: bool = False) -> torch.Tensor:
    """ksig/kernels.py:242-256 (float64 values)."""
    code = _ops.static_code(self.static_kernel)
    if code is None:
      M = self._compute_embedding(X, Y, diag=diag)
      return algorithms.signature_kern_pde(
          M, difference=self.difference,
          dyadic_order=self.dyadic_order).to(torch.float64)
    kind, param = code
    dev = X.device
    Xs = _ops.SeqSet.dense(X)
    Ys = Xs if (Y is None or diag) else _ops.SeqSet.dense(Y)
    p_raw = _ops.static_code(self.static_kernel, True)[1]
    grad = torch.is_grad_enabled() and (
        X.requires_grad or (Y is not None and Y.requires_grad)
        or (isinstance(p_raw, torch.Tensor) and p_raw.requires_grad))

    def values(pairs):
      if grad:
        from . import autograd
        return autograd.signature_pde(
            Xs, Ys, pairs, kind, p_raw, self.dyadic_order, self.difference,
            xdata=X.reshape(-1, X.shape[-1]),
            ydata=(X if (Y is None or diag) else Y).reshape(-1, X.shape[-1]))
      return _ops.signature_pde(Xs, Ys, pairs, self.difference, kind, param,
                                self.dyadic_order)

    if diag:
      return values(_ops.diag_pairs(X.shape[0], dev))
    if Y is None:
      n = X.shape[0]
      pairs = _ops.triu_pairs(n, dev)
      return _mirror(values(pairs), pairs, n)
    pairs = _ops.dense_pairs(X.shape[0], Y.shape[0], dev)
    return values(pairs).reshape(X.shape[0], Y.shape[0])


class GlobalAlignmentKernel(SignatureKernel):
  """Normalised global alignment kernel (ksig/kernels.py:366-409).

  As in the reference it takes only ``n_features`` and ``static_kernel``
  (it does not call the signature-kernel initializer), is always normalised
  and works in log space: K(x, y) = exp(log GA(x, y) - (log GA(x, x) +
  log GA(y, y)) / 2) (ksig/kernels.py:191-209)."""

  def __init__(self, n_features: Optional[int] = None,
               static_kernel: Optional[StaticKernel] = None):
    self.n_features = (utils.check_positive_value(n_features, 'n_features')
                       if n_features is not None else None)
    self.static_kernel = static_kernel or static.kernels.RBFKernel()

  @property
  def normalize(self):
    return True

  @property
  def is_log_space(self):
    return True

  def _compute_kernel(self, X, Y=None, diag: bool = False) -> torch.Tensor:
    """ksig/kernels.py:395-406: log GA(X, Y) or its diagonal (float64)."""
    code = _ops.static_code(self.static_kernel)
    if code is None:
      M = self._compute_embedding(X, Y, diag=diag)
      return algorithms.global_align_kern_log(M).to(torch.float64)
    kind, param = code
    dev = X.device
    Xs = _ops.SeqSet.dense(X)
    Ys = Xs if (Y is None or diag) else _ops.SeqSet.dense(Y)
    if diag:
      return _ops.global_align_log(Xs, Xs, _ops.diag_pairs(X.shape[0], dev), kind, param)
    if Y is None:
      n = X.shape[0]
      pairs = _ops.triu_pairs(n, dev)
      return _mirror(_ops.global_align_log(Xs, Xs, pairs, kind, param), pairs, n)
    pairs = _ops.dense_pairs(X.shape[0], Y.shape[0], dev)
    return _ops.global_align_log(Xs, Ys, pairs, kind, param).reshape(X.shape[0], Y.shape[0])

  def _K(self, X, Y=None) -> torch.Tensor:
    """ksig/kernels.py:180-214, log-space branch."""
    K = self._compute_kernel(X, Y)
    if Y is None:
      Kd = utils.matrix_diag(K)
      K = K - 0.5 * (Kd[..., :, None] + Kd[..., None, :])
    else:
      Kxd = self._compute_kernel(X, diag=True)
      Kyd = self._compute_kernel(Y, diag=True)
      K = K - 0.5 * (Kxd[..., :, None] + Kyd[..., None, :])
    return torch.exp(K).to(X.dtype)


class RandomWarpingSeries(static.features.KernelFeatures):
  """Normalised Random Warping Series features (ksig/kernels.py:413-477):
  the DTW value of a sequence against each of ``n_components`` random
  warping series (Gaussian points around the data's per-channel mean and
  spread, lengths uniform in 1..max_warp), normalised to unit norm.  The
  random series are drawn on the host with NumPy (the reference draws them
  with CuPy, so equal seeds give different series); the DTW values for given
  series follow ksig/algorithms.py:565-659 exactly."""

  def __init__(self, n_components: int = 100, stdev: float = 1.,
               max_warp: int = 32, normalize: bool = True,
               n_features: Optional[int] = None, random_state=None):
    self.n_components = utils.check_positive_value(n_components, 'n_components')
    self.stdev = utils.check_positive_value(stdev, 'stdev')
    self.max_warp = utils.check_positive_value(max_warp, 'max_warp')
    self.normalize = normalize
    self.n_features = (utils.check_positive_value(n_features, 'n_features')
                       if n_features is not None else None)
    self.random_state = static.features.check_random_state(random_state)

  def _make_feature_components(self, X):
    """ksig/kernels.py:443-457."""
    Xf = X.reshape(-1, X.shape[-1])
    self.data_means_ = torch.mean(Xf, 0)
    self.data_stdevs_ = torch.std(Xf, 0, unbiased=False)
    self.warp_lens_ = self.random_state.randint(1, self.max_warp + 1, size=[self.n_components])
    z = self.random_state.normal(size=[int(np.sum(self.warp_lens_)), self.n_features_])
    self.warp_series_ = self.data_means_[None, :] + self.stdev * (
        self.data_stdevs_[None, :] * torch.as_tensor(z, device=Xf.device, dtype=Xf.dtype))

  def _compute_features(self, X) -> torch.Tensor:
    """ksig/kernels.py:459-477: DTW against every warping series (distances
    computed inside the sweep), then unit-norm rows."""
    if X.ndim != 3:
      raise ValueError('RandomWarpingSeries takes sequence arrays of ndim==3.')
    W = self.warp_series_.to(device=X.device, dtype=X.dtype)
    P = _ops.random_warping_series(_ops.SeqSet.dense(X.contiguous()), W,
                                   self.warp_lens_).to(X.dtype)
    if self.normalize:
      nrm = utils.robust_sqrt(utils.squared_norm(P, axis=-1))
      P = P / torch.clamp(nrm, min=_EPS)[..., None]
    return P


class SignatureFeatures(SignatureBase, static.features.KernelFeatures):
  """Low-rank signature features (ksig/kernels.py:263-361): static features
  per level (``indep_feat``), optional random projection, first-order
  low-rank recursion (ksig/algorithms.py:175-246).  With a
  ``TensorizedRandomProjection`` (RFSF-TRP) the projected increments come
  from library GEMMs and the level recursion (exclusive time cumsum,
  elementwise product, time sum of every level) is one HIP pass
  (``ksig_lowrank_levels``).  With a ``DiagonalProjection`` (RFSF-DP) the
  differenced lifted features go straight into the diagonal level scan
  (``ksig_lowrank_levels_diagonal``: internal size 1-4, up to 5 levels, any
  order whose state fits the registers); with a ``TensorSketch`` every
  level's count sketch and convolution run in one scan per sequence
  (``ksig_lowrank_levels_sketch``, up to 8 levels and 2048 components).
  Gaussian, subsampling and very sparse projections, and the diagonal and
  sketch shapes those scans are not built for, run the reference's
  projected recursion at any order in float64 torch ops on the GPU (one
  projection call per outer product); TRP also runs at any order (rank > 1
  raises the reference's ValueError).  Without a projection the level-m
  features are the D^m iterated-sum coordinates (torch ops on the GPU, small
  D only)."""

  def __init__(self, n_levels: int = 5, order: int = 1, difference: bool = True,
               normalize: bool = True, indep_feat: bool = True,
               n_features: Optional[int] = None, static_features=None, projection=None):
    super().__init__(n_levels=n_levels, order=order, difference=difference,
                     normalize=normalize, n_features=n_features)
    self.indep_feat = indep_feat
    self.static_features = static_features
    self.projection = projection

  def _validate_data(self, X, reset: bool = False):
    X = SignatureBase._validate_data(self, X, reset=reset)
    self._validate_format(X, reset=reset)
    return X

  def _make_feature_components(self, X):
    """ksig/kernels.py:295-328."""
    from .projections import RandomProjection
    if self.projection is not None and not isinstance(self.projection, RandomProjection):
      raise NotImplementedError('low-rank signature features take a projection of '
                                'ksig_amd.projections, or none')
    if self.static_features is not None:
      seq_shape = tuple(X.shape[:-1]) + (-1,)
      Xf = X.reshape(-1, X.shape[-1])
      if self.indep_feat:
        hps = self.static_features.get_params(deep=False)
        cls = self.static_features.__class__
        self.static_features_ = [cls(**hps).fit(Xf) for _ in range(self.n_levels)]
        U = self.static_features_[0].transform(Xf, return_on_gpu=True)
      else:
        self.static_features_ = self.static_features.fit(Xf)
        U = self.static_features_.transform(Xf, return_on_gpu=True)
      U = U.reshape(seq_shape)
    else:
      U = X
    if self.projection is not None:
      hps = self.projection.get_params(deep=False)
      cls = self.projection.__class__
      self.projections_ = [cls(**hps).fit(U)]
      V = self.projections_[0](U, return_on_gpu=True)
      self.projections_ += [cls(**hps).fit(V, Y=U) for _ in range(1, self.n_levels)]
    else:
      self.projections_ = None

  def _order(self) -> int:
    """Embedding order of the recursion (ksig/algorithms.py:165)."""
    o = self.order
    return self.n_levels if o <= 0 or o >= self.n_levels else o

  def _lifted(self, X):
    """Per-level static features U_m [n, L, D] (or X itself)."""
    if self.static_features is None:
      return [X] * self.n_levels
    seq_shape = tuple(X.shape[:-1]) + (-1,)
    Xf = X.reshape(-1, X.shape[-1])
    if isinstance(self.static_features_, list):
      return [f.transform(Xf, return_on_gpu=True).reshape(seq_shape)
              for f in self.static_features_]
    U = self.static_features_.transform(Xf, return_on_gpu=True).reshape(seq_shape)
    return [U] * self.n_levels

  def _increments(self, X) -> torch.Tensor:
    """Per-level lifted paths [n_levels, n, steps, D], differenced along time
    when ``difference``.  Random Fourier features take one GEMM for every
    level's weights and go from it to their increments in one HIP pass per
    level (``ksig_rff_increments``: sin / cos, scaling and time difference
    without the intermediate feature tensor)."""
    RFF = static.features.RandomFourierFeatures
    fitted = getattr(self, 'static_features_', None) if self.static_features is not None else None
    fs = fitted if isinstance(fitted, list) else ([fitted] if fitted is not None else [])
    n, L = X.shape[0], X.shape[1]
    if fs and all(type(f) is RFF for f in fs):
      D = fs[0].random_weights_.shape[1]
      W = torch.cat([f.random_weights_.to(device=X.device, dtype=X.dtype) for f in fs], 1)
      proj = (X.reshape(-1, X.shape[-1]) @ W).reshape(n, L, -1)
      steps = L - 1 if self.difference else L
      dU = torch.empty((len(fs), n, max(steps, 0), 2 * D), device=X.device, dtype=X.dtype)
      for i in range(len(fs)):
        _ops.rff_increments(proj[:, :, i * D:(i + 1) * D], self.difference, out=dU[i])
      return dU if len(fs) == self.n_levels else dU.expand(self.n_levels, -1, -1, -1)
    U = self._lifted(X)
    if self.difference:
      U = [torch.diff(u, dim=1) for u in U]
    return torch.stack(U, 0)

  def _compute_features(self, X) -> torch.Tensor:
    """ksig/kernels.py:330-361."""
    if X.ndim != 3:
      raise ValueError('SignatureFeatures takes sequence arrays of ndim==3.')
    U = self._increments(X)
    n = X.shape[0]
    ones = torch.ones((n, 1), device=X.device, dtype=torch.float64)
    from .projections import DiagonalProjection, TensorizedRandomProjection, TensorSketch
    proj0 = self.projections_[0] if self.projections_ is not None else None
    P = None
    if isinstance(proj0, DiagonalProjection):
      # RFSF-DP: the lifted increments go straight into the diagonal level
      # scan (any internal size and order whose state fits its registers)
      q = proj0.internal_size
      try:
        lev = _ops.lowrank_levels_diagonal(U.contiguous(), q, self._order())
      except NotImplementedError:
        lev = None
      if lev is not None:
        H = U.shape[-1] // q
        P, off = [ones], 0
        for i in range(self.n_levels):
          w = q ** (i + 1) * H
          P.append(lev[:, off:off + w])
          off += w
    elif isinstance(proj0, TensorSketch):
      # every level's count sketch and convolution in one scan per sequence
      hidx = torch.stack([p.hash_idx_.to(U.device) for p in self.projections_], 0)
      hbit = torch.stack([p.hash_bit_.to(device=U.device, dtype=U.dtype)
                          for p in self.projections_], 0)
      try:
        lev = _ops.lowrank_levels_sketch(U, hidx, hbit, proj0.n_components, self._order())
        P = [ones] + [lev[:, m] for m in range(self.n_levels)]
      except NotImplementedError:
        P = None
    if P is not None:
      pass  # a HIP scan above produced the level features
    elif proj0 is not None and not isinstance(proj0, TensorizedRandomProjection):
      # any other projection, and the diagonal / sketch shapes the scans are
      # not built for: the reference's recursion (ksig/algorithms.py:218-246,
      # :296-337) with the projection called on each (running sum,
      # increments) pair, in float64 torch ops on the GPU
      proj = self.projections_
      order = self._order()
      Ud = U.to(torch.float64)
      Rs = [proj[0](Ud[0], return_on_gpu=True)]
      P = [ones, Rs[0].sum(1).reshape(n, -1)]
      for i in range(1, self.n_levels):
        S = sum(Rs)
        Q = torch.cumsum(S, 1) - S  # exclusive cumsum over time
        nxt = [proj[i](Q, Ud[i], return_on_gpu=True)]
        for r in range(1, min(i + 1, order)):
          nxt.append(proj[i](Rs[r - 1], Ud[i], return_on_gpu=True) / (r + 1))
        Rs = nxt
        P.append(sum(R.sum(1) for R in Rs).reshape(n, -1))
    elif self.projections_ is not None:
      # the reference reshapes each level's projection to [..., n_components,
      # rank] (ksig/algorithms.py:210-214), which fails for rank > 1 since the
      # components hold n_components columns; fitting refuses that rank for
      # the levels after the first, this is the single-level case
      p0 = self.projections_[0]
      if p0.rank != 1:
        raise ValueError(f'cannot reshape array of size {n * U.shape[2] * p0.n_components} '
                         f'into shape {(n, U.shape[2], p0.n_components, p0.rank)}')
      # projected increments, one batched GEMM over the levels: level 0
      # alone (scaled), level m >= 1 the projection of U_m that multiplies
      # the running sums
      C = torch.stack([p.scaling_ * p.components_.to(device=U.device, dtype=U.dtype)
                       for p in self.projections_], 0)
      nl, _, steps, width = U.shape
      Z = torch.bmm(U.reshape(nl, n * steps, width), C.contiguous()).reshape(nl, n, steps, -1)
      lev = _ops.lowrank_levels(Z, self._order())
      P = [ones] + [lev[:, m] for m in range(self.n_levels)]
    else:
      # no projection: the exact (exponentially wide) features of
      # ksig/algorithms.py:218-246 (order 1) and :296-337 (order p > 1),
      # Rs[r] the order-(r+1) terms of the current level
      order = self._order()
      Rs = [U[0].to(torch.float64)]
      P = [ones, Rs[0].sum(1)]
      for i in range(1, self.n_levels):
        Ui = U[i].to(torch.float64)[..., None, :]
        S = sum(Rs)
        Q = torch.cumsum(S, 1) - S  # exclusive cumsum over time
        steps = Q.shape[1]
        nxt = [(Q[..., :, None] * Ui).reshape(n, steps, -1)]
        for r in range(1, min(i + 1, order)):
          nxt.append((Rs[r - 1][..., :, None] * Ui).reshape(n, steps, -1) / (r + 1))
        Rs = nxt
        P.append(sum(R.sum(1) for R in Rs))
    if not self.normalize:
      return torch.cat(P, -1).to(X.dtype)
    P = [p / torch.clamp(utils.robust_sqrt(utils.squared_norm(p, axis=-1)), min=_EPS)[..., None]
         for p in P]
    return (torch.cat(P, -1) / np.sqrt(self.n_levels + 1)).to(X.dtype)

  def _K(self, X, Y=None):
    return static.features.KernelFeatures._K(self, X, Y)

  def _Kdiag(self, X):
    return static.features.KernelFeatures._Kdiag(self, X)
