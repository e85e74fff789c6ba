"""Pins of the Random Warping Series (DTW) oracle (no GPU): explicit
enumeration of warping paths, one-point closed forms, zero distance along
the diagonal, per-series independence.  The reference ships no tests or
fixtures for it."""

import numpy as np
import pytest

from oracle import sigkern as O


@pytest.mark.parametrize("shape", [(1, 1), (1, 4), (3, 1), (2, 2), (3, 4), (5, 5), (4, 6)])
def test_dtw_equals_path_enumeration(shape):
  rng = np.random.default_rng(sum(shape) + 3)
  D = rng.uniform(-0.2, 2.0, shape)
  got = O.random_warping_series(D[None], [shape[1]])[0, 0]
  assert abs(got - O.dtw_path_min(D)) < 1e-12


def test_one_point_closed_forms_and_zero_diagonal():
  rng = np.random.default_rng(1)
  x = rng.standard_normal((7, 3))
  w = rng.standard_normal((1, 3))
  D = O.squared_euclid_dist(x[None], w[None])  # [1, 7, 1]
  assert abs(O.random_warping_series(D, [1])[0, 0] - D[0, :, 0].sum()) < 1e-12
  D2 = O.squared_euclid_dist(x[None, :1], x[None])  # one sequence point, 7 warp points
  assert abs(O.random_warping_series(D2, [7])[0, 0] - D2[0, 0].sum()) < 1e-12
  D3 = O.squared_euclid_dist(x[None], x[None])
  assert abs(O.random_warping_series(D3, [7])[0, 0]) < 1e-12


def test_series_are_independent_columns():
  rng = np.random.default_rng(2)
  X = rng.standard_normal((3, 9, 2))
  lens = np.array([4, 1, 6, 2])
  W = rng.standard_normal((lens.sum(), 2))
  P = O.random_warping_series(O.squared_euclid_dist(X, W[None]), lens)
  off = np.concatenate([[0], np.cumsum(lens)])
  for b in range(len(lens)):
    Pb = O.random_warping_series(O.squared_euclid_dist(X, W[None, off[b]:off[b + 1]]), [lens[b]])
    np.testing.assert_allclose(P[:, b], Pb[:, 0], rtol=1e-14)
  F = O.rws_features(X, W, lens)
  np.testing.assert_allclose(np.linalg.norm(F, axis=-1), 1.0, rtol=1e-13)


def test_planted_mistake_is_caught():
  """Dropping the diagonal predecessor breaks the enumeration pin."""
  rng = np.random.default_rng(3)
  D = rng.uniform(0.0, 1.0, (5, 4))

  def no_diag(D):
    P = np.full((D.shape[0] + 1, D.shape[1] + 1), np.inf)
    for i in range(D.shape[0]):
      for j in range(D.shape[1]):
        best = 0.0 if i == 0 and j == 0 else min(P[i, j + 1], P[i + 1, j])
        P[i + 1, j + 1] = D[i, j] + best
    return P[-1, -1]

  assert abs(no_diag(D) - O.dtw_path_min(D)) > 1e-3
