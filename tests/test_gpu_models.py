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
