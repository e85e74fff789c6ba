"""Pins of the FP64 oracle (oracle/sigkern.py) against closed forms, brute
force and invariants.  The reference ships no tests or golden vectors and
cannot run here (CuPy/numba only), so every oracle function is pinned by
mathematics that a transcription mistake (dropped term, wrong axis, wrong
weight) would violate."""

import math

import numpy as np
import pytest

from oracle import sigkern as O


def _sig_inner(Sx, Sy, m):
  return float(np.sum(Sx[m] * Sy[m]))


@pytest.mark.parametrize("order", [3, 4, -1])
def test_dp_full_order_linear_equals_chen_signature_inner_product(order):
  # LinearKernel + difference: M = <dx_i, dy_j>; full order = exact truncated
  # signature kernel of the piecewise-linear paths.
  rng = np.random.default_rng(1)
  x = rng.standard_normal((5, 2)) * 0.7
  y = rng.standard_normal((4, 2)) * 0.7
  n = 3
  lev = O.pair_levels(x, y, n, order=order if order > 0 else -1, kind='linear', param=1.0)
  Sx, Sy = O.chen_signature(x, n), O.chen_signature(y, n)
  for m in range(n + 1):
    assert abs(lev[m] - _sig_inner(Sx, Sy, m)) < 1e-12, m


def test_dp_order1_linear_equals_iterated_sums():
  rng = np.random.default_rng(2)
  x = rng.standard_normal((7, 3))
  y = rng.standard_normal((6, 3))
  n = 4
  lev = O.pair_levels(x, y, n, order=1, kind='linear', param=1.0)
  Sx, Sy = O.iterated_sums(x, n), O.iterated_sums(y, n)
  for m in range(n + 1):
    assert abs(lev[m] - _sig_inner(Sx, Sy, m)) < 1e-11, m


@pytest.mark.parametrize("order", [1, 2, 3])
def test_dp_rbf_matches_chain_enumeration(order):
  rng = np.random.default_rng(3)
  x = rng.standard_normal((5, 2))
  y = rng.standard_normal((4, 2))
  n = 3
  M = O.double_diff(O.rbf_gram(x, y, 0.8))
  lev = O.signature_kern(M, n, order, difference=False, return_levels=True)
  bf = O.brute_force_levels(M, n, order)
  np.testing.assert_allclose(lev, bf, rtol=1e-12, atol=1e-14)


def test_higher_order_with_order_one_equals_first_order():
  rng = np.random.default_rng(4)
  M = rng.standard_normal((3, 2, 6, 5))
  a = O.signature_kern_first_order(M, 4, difference=False, return_levels=True)
  b = O.signature_kern_higher_order(M, 4, order=1, difference=False, return_levels=True)
  np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-15)


def test_multi_cumsum_exclusive_definition():
  rng = np.random.default_rng(5)
  R = rng.standard_normal((4, 5))
  E = O.multi_cumsum(R, exclusive=True, axis=(-2, -1))
  for i in range(4):
    for j in range(5):
      assert abs(E[i, j] - R[:i, :j].sum()) < 1e-13
  Ei = O.multi_cumsum(R, exclusive=True, axis=-2)
  Ej = O.multi_cumsum(R, exclusive=True, axis=-1)
  for i in range(4):
    for j in range(5):
      assert abs(Ei[i, j] - R[:i, j].sum()) < 1e-13
      assert abs(Ej[i, j] - R[i, :j].sum()) < 1e-13


def test_rbf_gram_definition():
  rng = np.random.default_rng(6)
  x, y = rng.standard_normal((3, 4)), rng.standard_normal((5, 4))
  G = O.rbf_gram(x, y, 1.7)
  for i in range(3):
    for j in range(5):
      assert abs(G[i, j] - math.exp(-np.sum((x[i] - y[j]) ** 2) / (2 * 1.7**2))) < 1e-14


def test_pde_one_cell_closed_form():
  for a in [0.3, -0.7, 1.9]:
    K = O.signature_kern_pde(np.array([[[a]]]), difference=False)
    assert abs(K[0] - (1 + a + a * a / 4)) < 1e-15


def test_pde_dyadic_converges_to_bessel():
  # straight 1-D segments x: 0 -> a, y: 0 -> b with LinearKernel: the exact
  # signature kernel is I0(2 sqrt(ab)); refinement error must shrink.
  a, b = 1.3, 0.9
  x = np.array([[0.], [a]])
  y = np.array([[0.], [b]])
  exact = O.bessel_i0(2 * math.sqrt(a * b))
  errs = [abs(O.pair_pde(x, y, kind='linear', dyadic_order=lam) - exact) for lam in range(7)]
  assert errs[0] < 0.1
  assert errs[6] < 2e-5
  # second-order scheme: error falls ~4x per refinement once resolved
  assert all(errs[k + 1] < errs[k] / 3.4 for k in range(3, 6))


def test_pde_matches_truncated_sum_for_small_increments():
  rng = np.random.default_rng(7)
  x = np.cumsum(rng.standard_normal((12, 2)) * 0.05, 0)
  y = np.cumsum(rng.standard_normal((10, 2)) * 0.05, 0)
  pde = O.pair_pde(x, y, kind='linear', dyadic_order=3)
  lev = O.pair_levels(x, y, 6, order=-1, kind='linear')
  assert abs(pde - lev.sum()) < 1e-4 * abs(lev.sum())


def test_goursat_recurrence_is_the_antidiagonal_sweep():
  # the reference evaluates the same update antidiagonal by antidiagonal;
  # an explicit antidiagonal traversal must agree bit for bit.
  rng = np.random.default_rng(8)
  M = rng.standard_normal((6, 9)) * 0.3
  lx, ly = M.shape
  U = np.ones((lx + 1, ly + 1))
  for t in range(lx + ly - 1):
    for i in range(max(0, t - ly + 1), min(t, lx - 1) + 1):
      j = t - i
      m = M[i, j]
      U[i + 1, j + 1] = ((U[i, j + 1] + U[i + 1, j]) * (1. + 0.5 * m + m * m / 12.)
                         - U[i, j] * (1. - m * m / 12.))
  assert O.goursat_solve(M) == U[lx, ly]


@pytest.mark.parametrize("order", [1, 2])
def test_stuttering_invariance(order):
  rng = np.random.default_rng(9)
  x = rng.standard_normal((6, 3))
  y = rng.standard_normal((5, 3))
  xs = np.insert(x, 3, x[3], axis=0)  # duplicate a time step
  a = O.pair_levels(x, y, 4, order=order)
  b = O.pair_levels(xs, y, 4, order=order)
  np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-14)
  assert abs(O.pair_pde(x, y) - O.pair_pde(xs, y)) < 1e-12


def test_translation_invariance():
  rng = np.random.default_rng(10)
  x, y = rng.standard_normal((6, 3)), rng.standard_normal((5, 3))
  s = rng.standard_normal(3)
  np.testing.assert_allclose(O.pair_levels(x + s, y + s, 4), O.pair_levels(x, y, 4),
                             rtol=1e-11, atol=1e-13)
  np.testing.assert_allclose(O.pair_levels(x + s, y, 4, kind='linear'),
                             O.pair_levels(x, y, 4, kind='linear'), rtol=1e-11, atol=1e-13)


def test_class_semantics_symmetry_diag_and_normalisation():
  rng = np.random.default_rng(11)
  X = rng.standard_normal((4, 7, 3))
  Y = rng.standard_normal((3, 5, 3))
  Kxx = O.signature_kernel(X, n_levels=4)
  np.testing.assert_allclose(Kxx, Kxx.T, rtol=1e-13)
  np.testing.assert_allclose(np.diag(Kxx), 1.0, rtol=1e-13)
  assert np.all(np.abs(Kxx) <= 1 + 1e-12)
  assert np.linalg.eigvalsh(Kxx).min() > -1e-10
  Kxy = O.signature_kernel(X, Y, n_levels=4)
  Kyx = O.signature_kernel(Y, X, n_levels=4)
  np.testing.assert_allclose(Kxy, Kyx.T, rtol=1e-12)
  # unnormalised: diag(K(X, X)) == K(X, diag=True)
  Ku = O.signature_kernel(X, n_levels=4, normalize=False)
  Kd = O.signature_kernel(X, diag=True, n_levels=4, normalize=False)
  np.testing.assert_allclose(np.diag(Ku), Kd, rtol=1e-13)
  # normalised = mean over levels of per-level normalised values
  lev = O.signature_kern(O.pairwise_gram(X, Y, 'rbf', 1.0), 4, 1, True, True)
  lx = O.signature_kern(O.pairwise_gram(X, None, 'rbf', 1.0, True), 4, 1, True, True)
  ly = O.signature_kern(O.pairwise_gram(Y, None, 'rbf', 1.0, True), 4, 1, True, True)
  man = np.mean(lev / np.sqrt(lx[:, :, None] * ly[:, None, :]), axis=0)
  np.testing.assert_allclose(Kxy, man, rtol=1e-12)
  np.testing.assert_array_equal(O.signature_kernel(X, diag=True), np.ones(4))


def test_pde_class_semantics():
  rng = np.random.default_rng(12)
  X = rng.standard_normal((3, 6, 2)) * 0.5
  K = O.signature_pde_kernel(X)
  np.testing.assert_allclose(np.diag(K), 1.0, rtol=1e-13)
  Ku = O.signature_pde_kernel(X, normalize=False)
  Kd = O.signature_pde_kernel(X, diag=True, normalize=False)
  np.testing.assert_allclose(np.diag(Ku), Kd, rtol=1e-13)


def test_short_sequences_give_empty_grid():
  x = np.zeros((1, 3))
  y = np.ones((4, 3))
  lev = O.pair_levels(x, y, 3)
  np.testing.assert_array_equal(lev, [1., 0., 0., 0.])
  assert O.pair_pde(x, y) == 1.0


def test_mmd_unbiased_formula():
  rng = np.random.default_rng(13)
  A = rng.standard_normal((5, 5)); Kxx = A @ A.T
  B = rng.standard_normal((4, 4)); Kyy = B @ B.T
  Kxy = rng.standard_normal((5, 4))
  ref = ((Kxx.sum() - np.trace(Kxx)) / 20 + (Kyy.sum() - np.trace(Kyy)) / 12
         - 2 * Kxy.mean())
  assert abs(O.mmd2_unbiased(Kxx, Kyy, Kxy) - ref) < 1e-12


def test_planted_mistakes_are_caught():
  """A dropped 1/(r+1) weight or a non-exclusive cumsum in the oracle must
  break the chain-enumeration pin (the pins are not tautologies)."""
  rng = np.random.default_rng(14)
  M = rng.standard_normal((4, 4)) * 0.5
  bf = O.brute_force_levels(M, 3, 2)

  def bad_higher(M):  # repeats without the 1/(r+1) weights
    K = [1., M.sum()]
    R = M[None, None].copy()
    for i in range(1, 3):
      d = min(i + 1, 2)
      Rn = np.empty((d, d) + M.shape)
      Rn[0, 0] = M * O.multi_cumsum(R.sum((0, 1)), True, (-2, -1))
      for r in range(1, d):
        Rn[0, r] = M * O.multi_cumsum(R[:, r - 1].sum(0), True, -2)
        Rn[r, 0] = M * O.multi_cumsum(R[r - 1, :].sum(0), True, -1)
        for s in range(1, d):
          Rn[r, s] = M * R[r - 1, s - 1]
      R = Rn
      K.append(R.sum())
    return np.array(K)

  assert np.abs(bad_higher(M) - bf).max() > 1e-6
  inclusive = [1., M.sum(), (M * np.cumsum(np.cumsum(M, 0), 1)).sum()]
  assert abs(inclusive[2] - bf[2] if False else
             inclusive[2] - O.brute_force_levels(M, 2, 1)[2]) > 1e-6
