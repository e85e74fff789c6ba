"""Pins of the low-rank signature feature oracle (no GPU): without a
projection the features' inner products are the first-order DP signature
kernel with a linear static kernel on the lifted path; with a tensorized
random projection each level is the iterated-sum tensor contracted with the
rank-1 projection tensors (brute force); the normalised map has unit norm
per level."""

import itertools

import numpy as np
import pytest

from oracle import sigkern as O


def test_unprojected_features_equal_first_order_dp_linear_kernel():
  rng = np.random.default_rng(0)
  X = rng.standard_normal((3, 7, 2))
  Y = rng.standard_normal((2, 6, 2))
  for n_levels in (1, 2, 4):
    Px = O.signature_kern_low_rank_fo(X, n_levels)
    Py = O.signature_kern_low_rank_fo(Y, n_levels)
    M = O.pairwise_gram(X, Y, 'linear', 1.0)
    ref = O.signature_kern(M, n_levels, 1, True, return_levels=True)  # [n_levels+1, 3, 2]
    for m in range(n_levels + 1):
      np.testing.assert_allclose(Px[m] @ Py[m].T, ref[m], rtol=1e-11, atol=1e-12)


def _iterated_sum_tensor(dU, m):
  """S_m = sum_{t1 < ... < tm} dU[t1] (x) ... (x) dU[tm] by enumeration."""
  L, D = dU.shape
  S = np.zeros((D,) * m)
  for ts in itertools.combinations(range(L), m):
    T = dU[ts[0]]
    for t in ts[1:]:
      T = np.multiply.outer(T, dU[t])
    S += T
  return S


@pytest.mark.parametrize("n_levels", [1, 3])
def test_trp_features_equal_contracted_iterated_sums(n_levels):
  rng = np.random.default_rng(1)
  U = rng.standard_normal((2, 6, 3))
  P = 4
  comps = [rng.standard_normal((3, P)) for _ in range(n_levels)]
  proj = [(comps[0], 1.0 / np.sqrt(P))] + [(c, 1.0) for c in comps[1:]]
  feats = O.signature_kern_low_rank_fo(U, n_levels, True, proj)
  for a in range(2):
    dU = np.diff(U[a], axis=0)
    for m in range(1, n_levels + 1):
      S = _iterated_sum_tensor(dU, m)
      for p in range(P):
        T = comps[0][:, p] / np.sqrt(P)
        for k in range(1, m):
          T = np.multiply.outer(T, comps[k][:, p])
        assert abs(feats[m][a, p] - np.sum(S * T)) < 1e-10 * max(1.0, np.abs(S).max())


def test_normalised_feature_map():
  rng = np.random.default_rng(2)
  X = O.brownian(rng, 5, 20, 3)
  w = [rng.standard_normal((3, 8)) for _ in range(4)]
  c = [rng.standard_normal((16, 6)) for _ in range(4)]
  trp = [(c[0], 1 / np.sqrt(6))] + [(ci, 1.0) for ci in c[1:]]
  F = O.signature_features(X, 4, w, trp)
  assert F.shape == (5, 1 + 4 * 6)
  np.testing.assert_allclose(np.sum(F * F, -1), 1.0, rtol=1e-12)
  # the planted mistake: an inclusive cumsum breaks the brute-force pin
  U = rng.standard_normal((1, 5, 2))
  good = O.signature_kern_low_rank_fo(U, 2)[2][0]
  dU = np.diff(U[0], axis=0)
  bad = np.einsum('ti,tj->ij', np.cumsum(dU, 0), dU).reshape(-1)
  assert np.abs(good - _iterated_sum_tensor(dU, 2).reshape(-1)).max() < 1e-12
  assert np.abs(bad - good).max() > 1e-3


@pytest.mark.parametrize("n_levels", [1, 2, 3])
def test_diagonal_projection_picks_matching_component_entries(n_levels):
  """RFSF-DP level m equals the unprojected iterated-sum tensor at the
  entries whose m factors share one component k (any sin / cos choice),
  scaled by sqrt(H)^(m-1)."""
  rng = np.random.default_rng(3)
  H = 3
  U = rng.standard_normal((2, 6, 2 * H))
  dp = O.signature_kern_low_rank_fo(U, n_levels, True, 'diag')
  full = O.signature_kern_low_rank_fo(U, n_levels, True, None)
  for m in range(1, n_levels + 1):
    S = full[m].reshape((2,) + (2 * H,) * m)
    D = dp[m].reshape(2, 2 ** m, H)
    for a in range(2):
      for j, qs in enumerate(itertools.product(range(2), repeat=m)):
        for k in range(H):
          idx = tuple(q * H + k for q in qs)
          assert abs(D[a, j, k] - np.sqrt(H) ** (m - 1) * S[(a,) + idx]) < 1e-10


@pytest.mark.parametrize("n_levels,order", [(2, 2), (3, 2), (3, 3), (4, 2), (4, 3), (4, 4)])
def test_higher_order_unprojected_features_equal_dp_linear_kernel(n_levels, order):
  """Without a projection the order-p features' inner products are the
  order-p DP signature kernel with a linear static kernel on the path; the
  order changes that kernel, so a recursion that ignored it would fail."""
  rng = np.random.default_rng(5)
  X = rng.standard_normal((3, 7, 2))
  Y = rng.standard_normal((2, 6, 2))
  Px = O.signature_kern_low_rank(X, n_levels, order)
  Py = O.signature_kern_low_rank(Y, n_levels, order)
  M = O.pairwise_gram(X, Y, 'linear', 1.0)
  ref = O.signature_kern(M, n_levels, order, True, return_levels=True)
  first = O.signature_kern(M, n_levels, 1, True, return_levels=True)
  assert np.abs(ref[n_levels] - first[n_levels]).max() > 1e-3
  for m in range(n_levels + 1):
    np.testing.assert_allclose(Px[m] @ Py[m].T, ref[m], rtol=1e-11, atol=1e-12)


def test_general_recursion_at_order_one_is_first_order():
  rng = np.random.default_rng(6)
  U = rng.standard_normal((2, 8, 3))
  comps = [(rng.standard_normal((3, 5)), 0.4)] + [(rng.standard_normal((3, 5)), 1.0)] * 3
  for proj in (None, comps):
    a = O.signature_kern_low_rank(U, 4, 1, True, proj)
    b = O.signature_kern_low_rank_fo(U, 4, True, proj)
    for x, y in zip(a, b):
      np.testing.assert_allclose(x, y, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("order", [2, 3])
def test_higher_order_trp_features_contract_unprojected_ones(order):
  """TRP at order p: component p of level m is the unprojected order-p level
  feature (a flattened [D]*m tensor) contracted with the rank-1 tensor
  c_0[:, p] s_0 (x) c_1[:, p] s_1 (x) ... (the projection is linear in each
  tensor factor)."""
  rng = np.random.default_rng(7)
  n_levels, D, P = 3, 3, 4
  U = rng.standard_normal((2, 6, D))
  proj = [(rng.standard_normal((D, P)), 0.5)] + [(rng.standard_normal((D, P)), 1.3)
                                                  for _ in range(n_levels - 1)]
  feats = O.signature_kern_low_rank(U, n_levels, order, True, proj)
  full = O.signature_kern_low_rank(U, n_levels, order, True, None)
  for m in range(1, n_levels + 1):
    for p in range(P):
      T = proj[0][0][:, p] * proj[0][1]
      for k in range(1, m):
        T = np.multiply.outer(T, proj[k][0][:, p] * proj[k][1])
      np.testing.assert_allclose(feats[m][:, p], full[m] @ T.reshape(-1), rtol=1e-11,
                                 atol=1e-12)


def test_nystroem_oracle_reproduces_kernel_on_its_span():
  """With every point in the basis the Nystroem features reproduce the
  Gram exactly (F F^T = K for a full-rank K), and any point's features are
  the projection of its kernel section onto the basis span (so for basis
  points F_b F^T = K(B, X))."""
  rng = np.random.default_rng(9)
  X = rng.standard_normal((2, 6, 3))
  Xf = X.reshape(-1, 3)
  F = O.nystroem_features(X, np.arange(12), 'rbf', 1.3).reshape(12, -1)
  np.testing.assert_allclose(F @ F.T, O.static_gram(Xf, Xf, 'rbf', 1.3), atol=1e-9)
  idx = rng.choice(12, 5, replace=False)
  Fb = O.nystroem_features(X, idx, 'rbf', 1.3).reshape(12, -1)
  np.testing.assert_allclose(Fb[idx] @ Fb.T, O.static_gram(Xf[idx], Xf, 'rbf', 1.3),
                             atol=1e-9)


def test_rff1d_oracle_approximates_rbf_kernel():
  """Bochner: with W ~ N(0, 1/sigma^2) and b ~ U(0, 2 pi), E[cos(w.x + b)
  cos(w.y + b)] * 2 = exp(-|x-y|^2 / (2 sigma^2)); at D = 40000 the Monte
  Carlo error is a few 1e-3 (a dropped sqrt(2), a phase left out or a wrong
  bandwidth scale is off by far more)."""
  rng = np.random.default_rng(10)
  X = rng.standard_normal((8, 3)) * 0.7
  sigma, D = 1.3, 40000
  W = rng.standard_normal((3, D)) / sigma
  b = 2 * np.pi * rng.uniform(size=(1, D))
  F = O.rff1d_features(X, W, b)
  np.testing.assert_allclose(F @ F.T, O.static_gram(X, X, 'rbf', sigma), atol=0.03)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_gaussian_projection_features_are_linear_images(order):
  """Gaussian projections: level m is A_m vec(unprojected level m) with
  A_1 = s_0 C_0 and A_m = s_m C_m (A_{m-1} (x) I_D) (each projection acts on
  the outer product with the previous level's index major), at every order."""
  rng = np.random.default_rng(11)
  n_levels, D, P = 3, 3, 5
  U = rng.standard_normal((2, 7, D))
  comps = [(rng.standard_normal((P, D)), 1 / np.sqrt(P))] + [
      (rng.standard_normal((P, P * D)), 1 / np.sqrt(P)) for _ in range(n_levels - 1)]
  feats = O.signature_kern_low_rank_gauss(U, n_levels, comps, order)
  full = O.signature_kern_low_rank(U, n_levels, order, True, None)
  A = comps[0][1] * comps[0][0]
  np.testing.assert_allclose(feats[1], full[1] @ A.T, rtol=1e-11, atol=1e-12)
  for m in range(2, n_levels + 1):
    C, sc = comps[m - 1]
    A = sc * C @ np.kron(A, np.eye(D))
    np.testing.assert_allclose(feats[m], full[m] @ A.T, rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("q,n_levels,order", [(2, 3, 1), (2, 3, 2), (3, 3, 3), (1, 4, 2),
                                              (4, 2, 1)])
def test_diagonal_features_are_per_component_unprojected_features(q, n_levels, order):
  """DiagonalProjection of internal size q at order p: for each component k
  the level-m block [q^m, H] at column k equals the unprojected order-p
  features of the q-dimensional path (U[..., b H + k])_b, scaled by
  sqrt(H)^(m-1) (the projection is diagonal over k and the outer product
  runs over the internal axis only).  A wrong scale, a swapped internal
  order or an order term dropped fails it."""
  rng = np.random.default_rng(20 + q + 3 * order)
  H = 3
  U = rng.standard_normal((2, 6, q * H))
  dp = O.signature_kern_low_rank_diag(U, n_levels, q, order)
  for k in range(H):
    full = O.signature_kern_low_rank(U[..., k::H], n_levels, order, True, None)
    for m in range(1, n_levels + 1):
      got = dp[m].reshape(2, q ** m, H)[..., k]
      np.testing.assert_allclose(got, np.sqrt(H) ** (m - 1) * full[m], rtol=1e-11, atol=1e-12)
  if q == 2 and order == 1:
    fo = O.signature_kern_low_rank_fo(U, n_levels, True, 'diag')
    for a, b in zip(dp, fo):
      np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-13)


def _sketch_of_tensor(T, sketches):
  """Count sketch of a flattened [D0 x D1 x ... x Dm] tensor (index of
  level 0 most significant) under the summed hash (h_0(i_0) + ... + h_m(i_m))
  mod n and the multiplied signs, by scatter-add."""
  h = np.zeros((), dtype=np.int64)
  s = np.ones(())
  for idx, bit, n in sketches:
    h = np.add.outer(h, np.asarray(idx))
    s = np.multiply.outer(s, np.asarray(bit))
  n = sketches[0][2]
  h = (h % n).reshape(-1)
  s = s.reshape(-1)
  out = np.zeros(T.shape[:-1] + (n,))
  for j in range(h.size):
    out[..., h[j]] += s[j] * T[..., j]
  return out


def test_convolved_count_sketches_are_the_sketch_of_the_outer_product():
  """convolve_fft(CS_1(x), CS_2(y)) = CS_h(x (x) y) with h = h_1 + h_2 mod n
  and signs s_1 s_2; a cross-correlation (conjugated FFT) instead fails."""
  rng = np.random.default_rng(30)
  n = 7
  sk = [(rng.integers(0, n, 5), rng.choice([-1.0, 1.0], 5), n),
        (rng.integers(0, n, 4), rng.choice([-1.0, 1.0], 4), n)]
  x, y = rng.standard_normal(5), rng.standard_normal(4)
  got = O.convolve_fft(O.count_sketch(x, *sk[0]), O.count_sketch(y, *sk[1]))
  want = _sketch_of_tensor(np.multiply.outer(x, y).reshape(-1), sk)
  np.testing.assert_allclose(got, want, atol=1e-12)
  a, b = O.count_sketch(x, *sk[0]), O.count_sketch(y, *sk[1])
  corr = np.real(np.fft.ifft(np.conj(np.fft.fft(a)) * np.fft.fft(b)))
  assert np.abs(corr - want).max() > 1e-3


@pytest.mark.parametrize("n_levels,order", [(1, 1), (3, 1), (3, 2), (4, 3)])
def test_sketch_features_are_sketches_of_unprojected_features(n_levels, order):
  """TensorSketch features of level m are the count sketch of the
  unprojected level-m features (a flattened [D]^m tensor) under the summed
  hash of levels 0..m-1 and their multiplied signs: every step is linear,
  and a convolution of count sketches sketches the outer product."""
  rng = np.random.default_rng(40 + n_levels + order)
  D, n = 3, 5
  U = rng.standard_normal((2, 6, D))
  sk = [(rng.integers(0, n, D), rng.choice([-1.0, 1.0], D), n) for _ in range(n_levels)]
  feats = O.signature_kern_low_rank_sketch(U, n_levels, sk, order)
  full = O.signature_kern_low_rank(U, n_levels, order, True, None)
  for m in range(1, n_levels + 1):
    np.testing.assert_allclose(feats[m], _sketch_of_tensor(full[m], sk[:m]), rtol=1e-11,
                               atol=1e-12)
