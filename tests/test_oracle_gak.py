"""Pins of the global alignment kernel oracle (no GPU): explicit enumeration
of alignment paths in exp space, Delannoy numbers, symmetry and the
normalised Gram.  The reference ships no tests or fixtures for it."""

import math

import numpy as np
import pytest

from oracle import sigkern as O


@pytest.mark.parametrize("shape", [(1, 1), (1, 4), (3, 1), (2, 2), (3, 4), (5, 5), (4, 6)])
def test_log_recursion_equals_path_enumeration(shape):
  rng = np.random.default_rng(sum(shape))
  M = rng.uniform(0.02, 1.0, shape)
  got = O.global_align_kern_log(M[None])[0]
  ref = math.log(O.gak_path_sum(M / (2. - M)))
  assert abs(got - ref) < 1e-12 * max(1., abs(ref))


def test_clamp_and_linear_kernel_values():
  """Values beyond 2 give a negative M' and are clamped at EPS, as the
  reference's log(max(M', EPS))."""
  rng = np.random.default_rng(3)
  M = rng.uniform(-1.0, 3.0, (3, 4))
  M[M == 2.] = 1.
  Mp = np.maximum(M / (2. - M), O.EPS)
  got = O.global_align_kern_log(M[None])[0]
  assert abs(got - math.log(O.gak_path_sum(Mp))) < 1e-10


@pytest.mark.parametrize("lx,ly", [(1, 1), (7, 3), (50, 20), (33, 33)])
def test_delannoy_closed_form(lx, ly):
  """Constant sequences: RBF values 1, M' = 1, GA = number of alignments."""
  X = np.ones((1, lx, 2))
  Y = np.ones((1, ly, 2))
  logK = O.global_align_kern_log(O.pairwise_gram(X, Y, 'rbf', 1.0))[0, 0]
  assert abs(logK - math.log(O.delannoy(lx - 1, ly - 1))) < 1e-12 * max(1., logK)


def test_symmetry_and_normalised_gram():
  rng = np.random.default_rng(4)
  X = O.brownian(rng, 6, 12, 3)
  Y = O.brownian(rng, 4, 9, 3)
  lxy = O.global_align_kern_log(O.pairwise_gram(X, Y, 'rbf', 1.5))
  lyx = O.global_align_kern_log(O.pairwise_gram(Y, X, 'rbf', 1.5))
  np.testing.assert_allclose(lxy, lyx.T, rtol=1e-13)
  K = O.global_alignment_kernel(X, kind='rbf', param=1.5)
  np.testing.assert_allclose(np.diag(K), 1.0, rtol=1e-13)
  np.testing.assert_allclose(K, K.T, rtol=1e-13)
  assert np.linalg.eigvalsh(K).min() > -1e-12
  assert np.all(K <= 1. + 1e-12)
  Kxy = O.global_alignment_kernel(X, Y, kind='rbf', param=1.5)
  assert Kxy.shape == (6, 4) and np.all(Kxy > 0) and np.all(Kxy <= 1. + 1e-12)
  assert np.all(O.global_alignment_kernel(X, diag=True) == 1.)


def test_stuttering_changes_ga():
  """GA counts alignments, so (unlike the signature kernels) repeating a
  point changes it: a guard against confusing the two recursions."""
  rng = np.random.default_rng(5)
  x = O.brownian(rng, 1, 6, 2)[0]
  y = O.brownian(rng, 1, 5, 2)[0]
  xs = np.concatenate([x[:3], x[2:3], x[3:]])
  assert abs(O.pair_gak_log(x, y) - O.pair_gak_log(xs, y)) > 1e-3


def test_planted_mistakes_are_caught():
  """Dropping the diagonal predecessor, or skipping the M/(2-M) transform,
  breaks the enumeration pin (the pin is not a tautology)."""
  rng = np.random.default_rng(6)
  M = rng.uniform(0.1, 1.0, (4, 5))
  ref = math.log(O.gak_path_sum(M / (2. - M)))

  def no_diag(M):
    lg = np.log(np.maximum(M / (2. - M), O.EPS))
    K = np.full((M.shape[0] + 1, M.shape[1] + 1), -np.inf)
    K[0, 0] = 0.
    for i in range(M.shape[0]):
      for j in range(M.shape[1]):
        a, b = K[i, j + 1], K[i + 1, j]
        if i == 0 and j == 0:
          a = 0.
        m = max(a, b)
        K[i + 1, j + 1] = lg[i, j] + m + math.log(math.exp(a - m) + math.exp(b - m))
    return K[-1, -1]

  assert abs(no_diag(M) - ref) > 1e-3
  untransformed = O.global_align_kern_log((2. * M / (1. + M))[None])[0]  # M' = M
  assert abs(O.global_align_kern_log(M[None])[0] - untransformed) > 1e-3
  assert abs(untransformed - math.log(O.gak_path_sum(M))) < 1e-12
