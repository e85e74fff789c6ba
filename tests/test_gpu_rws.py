"""GPU parity of the Random Warping Series kernel against the FP64 oracle:
fused distances and caller-supplied distances, series stacked across band
and group boundaries, short and long sequences, and the feature class.

Tolerances: DTW values within 1e-11 relative (float64 inputs) and 1e-5
(float32 inputs, normwise per sequence); normalised features 1e-5."""

import numpy as np
import pytest
import torch

import ksig_amd as ksig
from ksig_amd import _ops
from oracle import sigkern as O
from tests.gpu_util import normwise, to_dev

pytestmark = pytest.mark.gpu


def _rel(got, ref):
  got = np.asarray(got, dtype=np.float64)
  ref = np.asarray(ref, dtype=np.float64)
  return float(np.max(np.abs(got - ref).max(-1) / np.maximum(np.abs(ref).max(-1), 1e-300)))


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-11), (np.float32, 1e-5)])
@pytest.mark.parametrize("n,L,d,n_w", [(5, 50, 3, 20), (3, 1, 2, 7), (2, 64, 8, 9),
                                       (2, 65, 16, 40), (1, 130, 32, 300), (7, 200, 5, 100)])
def test_fused_dtw(dtype, tol, n, L, d, n_w):
  rng = np.random.default_rng(L + d + n_w)
  X = O.brownian(rng, n, L, d).astype(dtype)
  lens = rng.integers(1, 33, n_w)
  W = rng.standard_normal((lens.sum(), d)).astype(dtype)
  got = _ops.random_warping_series(_ops.SeqSet.dense(to_dev(X)), to_dev(W), lens).cpu().numpy()
  ref = O.random_warping_series(O.squared_euclid_dist(X.astype(np.float64),
                                                      W.astype(np.float64)[None]), lens)
  assert got.shape == (n, n_w) and _rel(got, ref) < tol


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-11), (np.float32, 1e-5)])
def test_distance_matrices(dtype, tol):
  rng = np.random.default_rng(4)
  lens = rng.integers(1, 20, 30)
  D = rng.uniform(0.0, 3.0, (4, 70, lens.sum())).astype(dtype)
  got = ksig.algorithms.random_warping_series(to_dev(D), lens).cpu().numpy()
  ref = O.random_warping_series(D.astype(np.float64), lens)
  assert got.shape == (4, 30) and _rel(got, ref) < tol
  with pytest.raises(ValueError):
    ksig.algorithms.random_warping_series(to_dev(D[0]), lens)


def test_ragged_sequences():
  rng = np.random.default_rng(5)
  lens_x = [1, 3, 64, 65, 100]
  xs = [O.brownian(rng, 1, L, 4)[0] for L in lens_x]
  X = _ops.SeqSet.ragged([to_dev(s) for s in xs])
  lens = rng.integers(1, 33, 50)
  W = rng.standard_normal((lens.sum(), 4))
  got = _ops.random_warping_series(X, to_dev(W), lens).cpu().numpy()
  for a, x in enumerate(xs):
    ref = O.random_warping_series(O.squared_euclid_dist(x[None], W[None]), lens)[0]
    assert _rel(got[a:a + 1], ref[None]) < 1e-11


def test_feature_class():
  rng = np.random.default_rng(6)
  X = O.brownian(rng, 8, 60, 3).astype(np.float32)
  Y = O.brownian(rng, 5, 45, 3).astype(np.float32)
  rws = ksig.kernels.RandomWarpingSeries(n_components=40, max_warp=16, random_state=3)
  rws.fit(X)
  assert rws.warp_series_.shape == (int(rws.warp_lens_.sum()), 3)
  F = rws.transform(X)
  W = rws.warp_series_.double().cpu().numpy()
  ref = O.rws_features(X.astype(np.float64), W, rws.warp_lens_)
  assert F.shape == (8, 40) and normwise(F, ref) < 1e-5
  K = rws(X, Y)
  refK = ref @ O.rws_features(Y.astype(np.float64), W, rws.warp_lens_).T
  assert K.shape == (8, 5) and normwise(K, refK) < 1e-5
  np.testing.assert_allclose(rws(X, diag=True), 1.0, rtol=1e-5)
  Fg = rws.transform(to_dev(X), return_on_gpu=True)
  assert isinstance(Fg, torch.Tensor) and Fg.is_cuda
  # refitting with the same seed draws the same series
  rws2 = ksig.kernels.RandomWarpingSeries(n_components=40, max_warp=16, random_state=3).fit(X)
  np.testing.assert_array_equal(rws2.warp_lens_, rws.warp_lens_)
