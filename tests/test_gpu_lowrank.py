"""GPU parity of the low-rank signature features (RFSF-TRP: random Fourier
features + tensorized random projection, first and higher order; RFSF-DP)
against the FP64 oracle, given the object's random weights and components.

Tolerances: level features within 1e-11 (float64 inputs) / 1e-5 (float32,
normwise per sequence and level); normalised features 1e-5."""

import numpy as np
import pytest
import torch

import ksig_amd as ksig
from ksig_amd import _ops
from oracle import sigkern as O
from tests.gpu_util import normwise, to_dev

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-11), (np.float32, 1e-5)])
@pytest.mark.parametrize("n_levels", [1, 3, 5, 8])
def test_level_scan(dtype, tol, n_levels):
  rng = np.random.default_rng(n_levels)
  n, L, D, P = 6, 40, 5, 33
  U = rng.standard_normal((n, L, D)) * 0.3
  comps = [rng.standard_normal((D, P)) for _ in range(n_levels)]
  proj = [(comps[0], 1 / np.sqrt(P))] + [(c, 1.0) for c in comps[1:]]
  ref = O.signature_kern_low_rank_fo(U, n_levels, True, proj)
  dU = np.diff(U, axis=1)
  Z = np.stack([dU @ comps[0] / np.sqrt(P)] + [dU @ c for c in comps[1:]], 0).astype(dtype)
  got = _ops.lowrank_levels(to_dev(Z)).cpu().numpy()
  assert got.shape == (n, n_levels, P)
  for m in range(n_levels):
    num = np.abs(got[:, m] - ref[m + 1]).max(-1)
    den = np.maximum(np.abs(ref[m + 1]).max(-1), 1e-300)
    assert float(np.max(num / den)) < tol, m


@pytest.mark.parametrize("normalize", [True, False])
def test_rfsf_trp_class(normalize):
  rng = np.random.default_rng(7)
  X = O.brownian(rng, 9, 30, 4).astype(np.float32)
  Y = O.brownian(rng, 5, 25, 4).astype(np.float32)
  sf = ksig.kernels.SignatureFeatures(
      n_levels=4, normalize=normalize,
      static_features=ksig.static.features.RandomFourierFeatures(n_components=16, random_state=1),
      projection=ksig.projections.TensorizedRandomProjection(n_components=24, random_state=2))
  sf.fit(X)
  w = [f.random_weights_.double().cpu().numpy() for f in sf.static_features_]
  trp = [(p.components_.double().cpu().numpy(), p.scaling_) for p in sf.projections_]
  F = sf.transform(X)
  ref = O.signature_features(X.astype(np.float64), 4, w, trp, normalize=normalize)
  assert F.shape == (9, 1 + 4 * 24) and normwise(F, ref) < 1e-5
  K = sf(X, Y)
  refK = ref @ O.signature_features(Y.astype(np.float64), 4, w, trp, normalize=normalize).T
  assert K.shape == (9, 5) and normwise(K, refK) < 1e-5
  if normalize:
    np.testing.assert_allclose(sf(X, diag=True), 1.0, rtol=1e-5)


def test_unprojected_features_match_oracle():
  rng = np.random.default_rng(8)
  X = O.brownian(rng, 4, 12, 2)
  sf = ksig.kernels.SignatureFeatures(n_levels=3, normalize=False).fit(X)
  F = sf.transform(X)
  ref = np.concatenate(O.signature_kern_low_rank_fo(X, 3), -1)
  assert F.shape == (4, 1 + 2 + 4 + 8) and normwise(F, ref) < 1e-11
  for order in (2, 3):
    sf = ksig.kernels.SignatureFeatures(n_levels=3, order=order, normalize=False).fit(X)
    ref = np.concatenate(O.signature_kern_low_rank(X, 3, order), -1)
    assert normwise(sf.transform(X), ref) < 1e-11, order


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-11), (np.float32, 1e-5)])
@pytest.mark.parametrize("n_levels,order", [(2, 2), (3, 2), (5, 2), (5, 3), (5, 0), (8, 4),
                                            (8, 8)])
def test_level_scan_higher_order(dtype, tol, n_levels, order):
  """ksig_lowrank_levels at embedding order p > 1 (ksig/algorithms.py:249-337)."""
  rng = np.random.default_rng(10 * n_levels + order)
  n, L, D, P = 5, 30, 4, 37
  U = rng.standard_normal((n, L, D)) * 0.3
  comps = [rng.standard_normal((D, P)) for _ in range(n_levels)]
  proj = [(comps[0], 1 / np.sqrt(P))] + [(c, 1.0) for c in comps[1:]]
  ref = O.signature_kern_low_rank(U, n_levels, order, True, proj)
  dU = np.diff(U, axis=1)
  Z = np.stack([dU @ comps[0] / np.sqrt(P)] + [dU @ c for c in comps[1:]], 0).astype(dtype)
  got = _ops.lowrank_levels(to_dev(Z), order).cpu().numpy()
  for m in range(n_levels):
    num = np.abs(got[:, m] - ref[m + 1]).max(-1)
    den = np.maximum(np.abs(ref[m + 1]).max(-1), 1e-300)
    assert float(np.max(num / den)) < tol, m


def test_rfsf_trp_class_higher_order():
  rng = np.random.default_rng(11)
  X = O.brownian(rng, 7, 28, 3).astype(np.float32)
  sf = ksig.kernels.SignatureFeatures(
      n_levels=4, order=2,
      static_features=ksig.static.features.RandomFourierFeatures(n_components=16, random_state=3),
      projection=ksig.projections.TensorizedRandomProjection(n_components=20, random_state=4))
  sf.fit(X)
  w = [f.random_weights_.double().cpu().numpy() for f in sf.static_features_]
  trp = [(p.components_.double().cpu().numpy(), p.scaling_) for p in sf.projections_]
  F = sf.transform(X)
  ref = O.signature_features(X.astype(np.float64), 4, w, trp, order=2)
  first = O.signature_features(X.astype(np.float64), 4, w, trp, order=1)
  assert F.shape == (7, 1 + 4 * 20) and normwise(F, ref) < 1e-5
  assert normwise(first, ref) > 1e-3  # the order matters at these sizes


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-11), (np.float32, 1e-5)])
@pytest.mark.parametrize("n_levels", [1, 3, 5])
def test_diagonal_level_scan(dtype, tol, n_levels):
  rng = np.random.default_rng(10 + n_levels)
  n, L, H = 5, 30, 37
  U = rng.standard_normal((n, L, 2 * H)) * 0.3
  ref = np.concatenate(O.signature_kern_low_rank_fo(U, n_levels, True, 'diag')[1:], -1)
  dU = np.stack([np.diff(U, axis=1)] * n_levels, 0).astype(dtype)
  got = _ops.lowrank_levels_diag(to_dev(dU)).cpu().numpy()
  assert got.shape == ref.shape
  num = np.abs(got - ref).max(-1)
  assert float(np.max(num / np.abs(ref).max(-1))) < tol


def test_rfsf_dp_class():
  rng = np.random.default_rng(11)
  X = O.brownian(rng, 7, 25, 3).astype(np.float32)
  sf = ksig.kernels.SignatureFeatures(
      n_levels=4, static_features=ksig.static.features.RandomFourierFeatures(
          n_components=20, random_state=4),
      projection=ksig.projections.DiagonalProjection())
  sf.fit(X)
  w = [f.random_weights_.double().cpu().numpy() for f in sf.static_features_]
  F = sf.transform(X)
  ref = O.signature_features(X.astype(np.float64), 4, w, 'diag')
  assert F.shape == (7, 1 + (2 + 4 + 8 + 16) * 20) and normwise(F, ref) < 1e-5
  np.testing.assert_allclose(sf(X, diag=True), 1.0, rtol=1e-5)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 1e-5)])
@pytest.mark.parametrize("difference", [True, False])
def test_rff_increments(dtype, tol, difference):
  rng = np.random.default_rng(12)
  proj = (rng.standard_normal((3, 17, 10)) * 2).astype(dtype)
  got = _ops.rff_increments(to_dev(proj), difference).cpu().numpy()
  F = np.concatenate([np.sin(proj.astype(np.float64)), np.cos(proj.astype(np.float64))], -1)
  F /= np.sqrt(10)
  ref = np.diff(F, axis=1) if difference else F
  assert got.shape == ref.shape and np.abs(got - ref).max() < tol


def test_signature_features_shared_static_map():
  """indep_feat=False: one Fourier map for every level."""
  rng = np.random.default_rng(13)
  X = O.brownian(rng, 6, 20, 3).astype(np.float32)
  sf = ksig.kernels.SignatureFeatures(
      n_levels=3, indep_feat=False,
      static_features=ksig.static.features.RandomFourierFeatures(n_components=8, random_state=5),
      projection=ksig.projections.TensorizedRandomProjection(n_components=12, random_state=6))
  sf.fit(X)
  w = sf.static_features_.random_weights_.double().cpu().numpy()
  trp = [(p.components_.double().cpu().numpy(), p.scaling_) for p in sf.projections_]
  F = sf.transform(X)
  ref = O.signature_features(X.astype(np.float64), 3, [w] * 3, trp)
  assert normwise(F, ref) < 1e-5


def test_nystroem_features():
  """NystroemFeatures (ksig/static/features.py:205-285) vs the oracle given
  the drawn basis indices; with every point in the basis F F^T = K."""
  rng = np.random.default_rng(14)
  X = rng.standard_normal((40, 3))
  nf = ksig.static.features.NystroemFeatures(ksig.static.kernels.RBFKernel(1.5),
                                             n_components=12, random_state=0).fit(X)
  F = nf.transform(X)
  ref = O.nystroem_features(X, nf.component_idx_, 'rbf', 1.5)
  # eigenvectors are fixed up to sign (and rotation in repeated eigenspaces):
  # the approximate Gram F F^T is what the map defines
  assert F.shape == ref.shape and normwise(F @ F.T, ref @ ref.T) < 1e-9
  B = X[nf.component_idx_]
  own = O.static_gram(X, B, 'rbf', 1.5) @ nf.normalization_.double().cpu().numpy()
  assert normwise(F, own) < 1e-9
  with pytest.warns(UserWarning):
    full = ksig.static.features.NystroemFeatures(ksig.static.kernels.RBFKernel(1.5),
                                                 n_components=64).fit(X[:20])
  Ff = full.transform(X[:20])
  assert normwise(Ff @ Ff.T, O.static_gram(X[:20], X[:20], 'rbf', 1.5)) < 1e-8


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-9), (np.float32, 2e-3)])
def test_signature_features_with_nystroem_and_trp(dtype, tol):
  """RFSF with Nystroem static features.  The factor U_+ s_+^(-1/2) and the
  product K(X, B) @ factor are float64; float32 inputs still carry the
  rounding of K(X, B), scaled by s^(-1/2) (eigenvalues down to 1e-12) before
  the time differences cancel (2.6e-4 at worst over four seeds); float64
  pins the dataflow."""
  rng = np.random.default_rng(15)
  X = O.brownian(rng, 6, 20, 2).astype(dtype)
  sf = ksig.kernels.SignatureFeatures(
      n_levels=3,
      static_features=ksig.static.features.NystroemFeatures(ksig.static.kernels.RBFKernel(1.0),
                                                            n_components=16, random_state=5),
      projection=ksig.projections.TensorizedRandomProjection(n_components=24, random_state=6))
  sf.fit(X)
  Xd = X.astype(np.float64)
  # the static features with the object's basis and factor (eigenvector
  # signs are the solver's)
  U = [(O.static_gram(Xd.reshape(-1, 2), Xd.reshape(-1, 2)[f.component_idx_], 'rbf', 1.0)
        @ f.normalization_.double().cpu().numpy()).reshape(6, 20, -1)
       for f in sf.static_features_]
  trp = [(p.components_.double().cpu().numpy(), p.scaling_) for p in sf.projections_]
  ref = O.low_rank_features_normalised(O.signature_kern_low_rank_fo(U, 3, True, trp), 3)
  F = sf.transform(X)
  assert F.shape == ref.shape and normwise(F, ref) < tol


def test_rff1d_features_and_signature_features():
  rng = np.random.default_rng(16)
  X = O.brownian(rng, 5, 18, 3).astype(np.float32)
  rf = ksig.static.features.RandomFourierFeatures1D(bandwidth=1.2, n_components=32,
                                                    random_state=2).fit(X.reshape(-1, 3))
  F = rf.transform(X.reshape(-1, 3))
  ref = O.rff1d_features(X.reshape(-1, 3).astype(np.float64),
                         rf.random_weights_.double().cpu().numpy(),
                         rf.random_shift_.double().cpu().numpy())
  assert normwise(F, ref) < 1e-5
  sf = ksig.kernels.SignatureFeatures(
      n_levels=3,
      static_features=ksig.static.features.RandomFourierFeatures1D(n_components=16,
                                                                   random_state=3),
      projection=ksig.projections.TensorizedRandomProjection(n_components=20, random_state=4))
  sf.fit(X)
  Xd = X.astype(np.float64)
  U = [O.rff1d_features(Xd, f.random_weights_.double().cpu().numpy(),
                        f.random_shift_.double().cpu().numpy()) for f in sf.static_features_]
  trp = [(p.components_.double().cpu().numpy(), p.scaling_) for p in sf.projections_]
  ref = O.low_rank_features_normalised(O.signature_kern_low_rank_fo(U, 3, True, trp), 3)
  assert normwise(sf.transform(X), ref) < 1e-4


@pytest.mark.parametrize("order", [1, 2])
def test_signature_features_gaussian_projection(order):
  """RFSF with GaussianRandomProjection (ksig/projections.py:189-242): the
  reference's recursion with the projection on each (running sum,
  increments) outer product, vs the oracle given the object's components."""
  rng = np.random.default_rng(17)
  X = O.brownian(rng, 5, 16, 2).astype(np.float32)
  sf = ksig.kernels.SignatureFeatures(
      n_levels=3, order=order,
      static_features=ksig.static.features.RandomFourierFeatures(n_components=8, random_state=1),
      projection=ksig.projections.GaussianRandomProjection(n_components=12, random_state=2))
  sf.fit(X)
  w = [f.random_weights_.double().cpu().numpy() for f in sf.static_features_]
  U = [O.rff_features(X.astype(np.float64), wi) for wi in w]
  comps = [(p.components_.double().cpu().numpy(), p.scaling_) for p in sf.projections_]
  assert comps[0][0].shape == (12, 16) and comps[1][0].shape == (12, 12 * 16)
  ref = O.low_rank_features_normalised(
      O.signature_kern_low_rank_gauss(U, 3, comps, order), 3)
  F = sf.transform(X)
  assert F.shape == (5, 1 + 3 * 12) and normwise(F, ref) < 1e-5


@pytest.mark.parametrize("kind", ["subsampling", "very_sparse"])
def test_signature_features_subsampling_and_very_sparse(kind):
  """SubsamplingProjection (ksig/projections.py:245-284) and
  VerySparseRandomProjection (:287-376) are linear maps: a row selection of
  the identity, or the +-1 components on the kept columns, each scaled; the
  pinned Gaussian-projection oracle takes them as its components."""
  rng = np.random.default_rng(18)
  X = O.brownian(rng, 5, 16, 2).astype(np.float32)
  if kind == "subsampling":
    proj = ksig.projections.SubsamplingProjection(n_components=10, random_state=2)
  else:
    proj = ksig.projections.VerySparseRandomProjection(n_components=10, random_state=2)
  sf = ksig.kernels.SignatureFeatures(
      n_levels=3, order=2,
      static_features=ksig.static.features.RandomFourierFeatures(n_components=6, random_state=1),
      projection=proj)
  sf.fit(X)
  w = [f.random_weights_.double().cpu().numpy() for f in sf.static_features_]
  U = [O.rff_features(X.astype(np.float64), wi) for wi in w]
  comps = []
  for p in sf.projections_:
    idx = p.sampled_idx_.cpu().numpy()
    if kind == "subsampling":
      C = np.eye(p.n_features_)[idx]
    else:
      C = np.zeros((p.n_components, p.n_features_))
      C[:, idx] = p.components_.double().cpu().numpy()
    comps.append((C, p.scaling_))
  ref = O.low_rank_features_normalised(O.signature_kern_low_rank_gauss(U, 3, comps, 2), 3)
  F = sf.transform(X)
  assert F.shape == (5, 1 + 3 * 10) and normwise(F, ref) < 1e-5
