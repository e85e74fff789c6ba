"""PrecomputedKernelSVC over the HIP signature kernels: the Gram it hands to
SVC is the oracle's (one launch, whatever batch_size), and the classifier
separates two classes of paths.  Tolerance 1e-5 (float32 inputs)."""

import math

import numpy as np
import pytest

import ksig_amd as ksig
from ksig_amd.models import PrecomputedKernelSVC
from oracle import sigkern as O
from tests.gpu_util import normwise

pytestmark = pytest.mark.gpu


def _paths(rng, n, L, d):
  """Two classes of Brownian paths with opposite drift in channel 0 (the
  normalised kernel divides out overall scale, not direction)."""
  y = np.arange(n) % 2
  drift = np.zeros((n, L, d))
  drift[:, :, 0] = np.where(y == 0, 1.0, -1.0)[:, None] * np.linspace(0, 1, L)[None, :]
  return (0.3 * O.brownian(rng, n, L, d) + drift).astype(np.float32), y


@pytest.mark.parametrize("batch_size", [None, 7])
def test_svc_gram_and_classification(batch_size):
  rng = np.random.default_rng(12)
  X, y = _paths(rng, 30, 24, 3)
  Xt, yt = _paths(rng, 12, 24, 3)
  sig = math.sqrt(3)
  sk = ksig.kernels.SignatureKernel(4, static_kernel=ksig.static.kernels.RBFKernel(sig))
  m = PrecomputedKernelSVC(sk, svc_hparams={}, batch_size=batch_size)
  K = m._precompute_kernel_mat(X)
  ref = O.signature_kernel(X.astype(np.float64), n_levels=4, param=sig)
  assert K.shape == (30, 30) and K.dtype == np.float64
  assert normwise(K, ref) < 1e-5
  Kt = m._precompute_kernel_mat(Xt, X)
  reft = O.signature_kernel(Xt.astype(np.float64), X.astype(np.float64), n_levels=4, param=sig)
  assert Kt.shape == (12, 30) and normwise(Kt, reft) < 1e-5
  m.fit(X, y)
  pred = m.predict(Xt)
  assert (pred == yt).mean() >= 0.9
  # the same SVC on the oracle's Gram decides the same way
  from sklearn.svm import SVC
  svc = SVC(kernel='precomputed', decision_function_shape='ovo', max_iter=10000).fit(ref, y)
  np.testing.assert_array_equal(pred, svc.predict(reft))


@pytest.mark.parametrize("batch_size", [None, 11])
def test_svc_on_rfsf_features(batch_size):
  """has_transform with RFSF-TRP: the Gram handed to SVC is the product of
  the oracle's feature matrices (ksig/models/pre_svc.py:69-104), and the
  classifier separates the two drift classes."""
  rng = np.random.default_rng(13)
  X, y = _paths(rng, 30, 24, 3)
  Xt, yt = _paths(rng, 12, 24, 3)
  sf = ksig.kernels.SignatureFeatures(
      n_levels=3,
      static_features=ksig.static.features.RandomFourierFeatures(n_components=32, random_state=1),
      projection=ksig.projections.TensorizedRandomProjection(n_components=64, random_state=2))
  m = PrecomputedKernelSVC(sf, svc_hparams={}, batch_size=batch_size, has_transform=True,
                           need_kernel_fit=True)
  m.fit(X, y)
  w = [f.random_weights_.double().cpu().numpy() for f in sf.static_features_]
  trp = [(p.components_.double().cpu().numpy(), p.scaling_) for p in sf.projections_]
  F = O.signature_features(X.astype(np.float64), 3, w, trp)
  Ft = O.signature_features(Xt.astype(np.float64), 3, w, trp)
  K = m._precompute_model_inputs()
  assert K.shape == (30, 30) and normwise(K, F @ F.T) < 1e-5
  assert np.array_equal(K, K.T)
  Kt = m._precompute_model_inputs(Xt)
  assert Kt.shape == (12, 30) and normwise(Kt, Ft @ F.T) < 1e-5
  assert (m.predict(Xt) == yt).mean() >= 0.9
