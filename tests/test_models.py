"""PrecomputedKernelSVC host logic (no GPU): the reference's block loop for
foreign kernels, SVC plumbing, sklearn parameter introspection."""

import numpy as np
import pytest
from sklearn.base import clone

from ksig_amd.models import PrecomputedKernelSVC


class FlatRBF:
  """A callable kernel that is not one of this library's (host NumPy)."""

  def __init__(self, gamma=0.05):
    self.gamma = gamma
    self.calls = 0

  def fit(self, X):
    return self

  def __call__(self, X, Y=None):
    self.calls += 1
    A = X.reshape(X.shape[0], -1)
    B = A if Y is None else Y.reshape(Y.shape[0], -1)
    d2 = (A * A).sum(1)[:, None] + (B * B).sum(1)[None, :] - 2 * A @ B.T
    return np.exp(-self.gamma * d2)


def _toy(rng, n):
  y = np.arange(n) % 2
  X = rng.standard_normal((n, 10, 3)) * np.where(y == 0, 0.3, 1.5)[:, None, None]
  return X, y


@pytest.mark.parametrize("b", [None, 1, 4, 7, 100])
def test_block_loop_matches_direct_gram(b):
  rng = np.random.default_rng(0)
  X, _ = _toy(rng, 23)
  X2, _ = _toy(rng, 9)
  k = FlatRBF()
  m = PrecomputedKernelSVC(k, svc_hparams={}, batch_size=b)
  np.testing.assert_allclose(m._precompute_kernel_mat(X), k(X), rtol=0, atol=1e-14)
  np.testing.assert_allclose(m._precompute_kernel_mat(X2, X), k(X2, X), rtol=0, atol=1e-14)
  if b is not None and b < 23:
    nb = -(-23 // b)
    assert k.calls >= nb * (nb + 1) // 2  # one call per upper-triangle block


def test_fit_predict_and_grid_search():
  rng = np.random.default_rng(1)
  X, y = _toy(rng, 40)
  Xt, yt = _toy(rng, 20)
  m = PrecomputedKernelSVC(FlatRBF(), svc_hparams={}, batch_size=8).fit(X, y)
  assert m.svc_hparams["kernel"] == "precomputed"
  assert m.svc_hparams["decision_function_shape"] == "ovo"
  assert (m.predict(Xt) == yt).mean() >= 0.9
  g = PrecomputedKernelSVC(FlatRBF(), svc_hparams={}, svc_grid={"C": [0.1, 1.0, 10.0]},
                           n_jobs=1).fit(X, y)
  assert (g.predict(Xt) == yt).mean() >= 0.9


class MomentFeatures:
  """A feature-map kernel on the host: per-channel mean and mean square."""

  def fit(self, X):
    return self

  def transform(self, X, return_on_gpu=False):
    return np.concatenate([X.mean(1), (X * X).mean(1)], axis=1)


def test_params_and_transform_without_map():
  m = PrecomputedKernelSVC(FlatRBF(), svc_hparams={}, cv=4, batch_size=5)
  p = m.get_params(deep=False)
  assert p["cv"] == 4 and p["batch_size"] == 5 and p["has_transform"] is False
  c = clone(m)
  assert c.batch_size == 5 and c.cv == 4
  t = PrecomputedKernelSVC(FlatRBF(), svc_hparams={}, has_transform=True)
  rng = np.random.default_rng(2)
  X, y = _toy(rng, 10)
  with pytest.raises(ValueError):
    t.fit(X, y)


@pytest.mark.parametrize("b", [None, 3, 100])
def test_transform_gram_is_feature_product(b):
  """has_transform: Gram = F F^T (train) and F_test F^T (test), exactly
  symmetric, whatever the feature blocks (ksig/models/pre_base.py:143-224)."""
  rng = np.random.default_rng(3)
  X, y = _toy(rng, 20)
  Xt, yt = _toy(rng, 8)
  m = PrecomputedKernelSVC(MomentFeatures(), svc_hparams={}, batch_size=b, has_transform=True)
  m.X = X
  K = m._precompute_model_inputs()
  F = MomentFeatures().transform(X)
  np.testing.assert_allclose(K, F @ F.T, rtol=1e-12, atol=1e-12)
  assert np.array_equal(K, K.T)
  Kt = m._precompute_model_inputs(Xt)
  np.testing.assert_allclose(Kt, MomentFeatures().transform(Xt) @ F.T, rtol=1e-12, atol=1e-12)
  m.fit(X, y)
  assert (m.predict(Xt) == yt).mean() >= 0.9
