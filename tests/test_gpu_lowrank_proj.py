"""GPU parity of the low-rank signature features with the diagonal
projection at any internal size and order (ksig_lowrank_levels_diagonal),
with TensorSketch (ksig_lowrank_levels_sketch), and of the shapes those
scans are not built for (the projected recursion in torch ops on the GPU),
against the FP64 oracle given the objects' random weights, hashes and
signs; TensorizedRandomProjection with rank > 1 raises the reference's
ValueError.

Tolerances: level features within 1e-11 (float64 inputs) / 1e-5 (float32),
normwise per sequence and level; normalised features 1e-5 (float32 data)."""

import numpy as np
import pytest
import torch

import ksig_amd as ksig
from ksig_amd import _ops
from oracle import sigkern as O
from tests.gpu_util import normwise, to_dev

pytestmark = pytest.mark.gpu


def _levels_err(got, ref_levels):
  """max over levels and sequences of the normwise error of a level block."""
  worst, off = 0.0, 0
  for r in ref_levels:
    w = r.shape[-1]
    g = got[:, off:off + w]
    num = np.abs(g - r).max(-1)
    den = np.maximum(np.abs(r).max(-1), 1e-300)
    worst = max(worst, float(np.max(num / den)))
    off += w
  assert off == got.shape[-1]
  return worst


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-11), (np.float32, 1e-5)])
@pytest.mark.parametrize("q,n_levels,order", [(2, 5, 1), (2, 4, 2), (2, 4, 4), (2, 3, 3),
                                              (1, 5, 3), (3, 3, 3), (3, 2, 1), (4, 3, 1),
                                              (4, 2, 2), (2, 1, 1)])
def test_diagonal_scan_any_internal_size_and_order(dtype, tol, q, n_levels, order):
  rng = np.random.default_rng(100 + 10 * q + n_levels + order)
  n, L, H = 5, 23, 37
  U = rng.standard_normal((n, L, q * H)) * 0.3
  ref = O.signature_kern_low_rank_diag(U, n_levels, q, order)[1:]
  dU = np.stack([np.diff(U, axis=1)] * n_levels, 0).astype(dtype)
  got = _ops.lowrank_levels_diagonal(to_dev(dU), q, order).cpu().numpy()
  assert _levels_err(got, ref) < tol


def test_diagonal_scan_entry_at_q2_order1_matches_the_general_entry():
  rng = np.random.default_rng(120)
  dU = to_dev(rng.standard_normal((4, 6, 30, 2 * 19)) * 0.2)
  a = _ops.lowrank_levels_diag(dU)
  b = _ops.lowrank_levels_diagonal(dU, 2, 1)
  assert torch.equal(a, b)


def test_diagonal_scan_refuses_unbuilt_shapes():
  dU = to_dev(np.zeros((4, 2, 5, 3 * 4)))
  with pytest.raises(NotImplementedError):
    _ops.lowrank_levels_diagonal(dU, 3, 1)  # 3 + 9 + 27 + 81 sums per component
  with pytest.raises(ValueError):
    _ops.lowrank_levels_diagonal(dU, 5, 1)  # 12 features do not split in 5


def _rff_dp(n_levels, order, q, n_rff, seed):
  return ksig.kernels.SignatureFeatures(
      n_levels=n_levels, order=order,
      static_features=ksig.static.features.RandomFourierFeatures(n_components=n_rff,
                                                                 random_state=seed),
      projection=ksig.projections.DiagonalProjection(internal_size=q))


@pytest.mark.parametrize("q,n_levels,order,n_rff", [(2, 4, 2, 20), (2, 6, 1, 6), (3, 3, 2, 12),
                                                    (3, 4, 1, 3), (4, 2, 2, 10)])
def test_rfsf_dp_class_any_shape(q, n_levels, order, n_rff):
  """The class at scan-built shapes and at shapes that take the projected
  recursion (q = 2 with 6 levels, q = 3 with 4 levels), normalised features
  against the oracle on the object's Fourier weights."""
  rng = np.random.default_rng(130 + q + n_levels)
  X = O.brownian(rng, 6, 18, 3).astype(np.float32)
  sf = _rff_dp(n_levels, order, q, n_rff, 7).fit(X)
  w = [f.random_weights_.double().cpu().numpy() for f in sf.static_features_]
  F = sf.transform(X)
  U = [O.rff_features(X.astype(np.float64), wi) for wi in w]
  ref = O.low_rank_features_normalised(O.signature_kern_low_rank_diag(U, n_levels, q, order),
                                       n_levels)
  H = 2 * n_rff // q
  assert F.shape == (6, 1 + sum(q ** (i + 1) for i in range(n_levels)) * H)
  assert normwise(F, ref) < 1e-5
  np.testing.assert_allclose(sf(X, diag=True), 1.0, rtol=1e-5)


def _sketches(rng, n_levels, width, m):
  idx = rng.integers(0, m, (n_levels, width))
  bit = rng.choice([-1.0, 1.0], (n_levels, width))
  return idx, bit, [(idx[i], bit[i], m) for i in range(n_levels)]


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-11), (np.float32, 1e-5)])
@pytest.mark.parametrize("n_levels,order,m,width", [(1, 1, 50, 7), (3, 1, 100, 24),
                                                    (4, 2, 70, 16), (5, 5, 33, 9),
                                                    (3, 1, 300, 40), (2, 2, 1000, 12),
                                                    (8, 3, 64, 5),
                                                    # fewer components than features:
                                                    # the count sketch first
                                                    (3, 1, 20, 50), (4, 2, 33, 100),
                                                    (5, 5, 8, 40), (8, 2, 64, 200),
                                                    (1, 1, 1, 9)])
def test_sketch_scan(dtype, tol, n_levels, order, m, width):
  rng = np.random.default_rng(140 + n_levels + order + m)
  n, L = 4, 21
  U = rng.standard_normal((n, L, width)) * 0.4
  idx, bit, sk = _sketches(rng, n_levels, width, m)
  ref = O.signature_kern_low_rank_sketch(U, n_levels, sk, order)[1:]
  dU = np.stack([np.diff(U, axis=1)] * n_levels, 0).astype(dtype)
  got = _ops.lowrank_levels_sketch(to_dev(dU), to_dev(idx), to_dev(bit), m,
                                   order).cpu().numpy()
  assert got.shape == (n, n_levels, m)
  assert _levels_err(got.reshape(n, -1), ref) < tol


def test_sketch_scan_largest_built_size_and_dropped_hashes():
  """n_comp = 2048 (two components per thread, 1024 threads); hash values
  outside [0, n_comp) select no component, as the reference's mask."""
  rng = np.random.default_rng(150)
  n_levels, width, m = 2, 10, 2048
  U = rng.standard_normal((3, 9, width)) * 0.5
  idx, bit, _ = _sketches(rng, n_levels, width, m)
  idx[0, 3] = m + 5
  idx[1, 7] = -1
  keep = [(np.where((idx[i] >= 0) & (idx[i] < m), idx[i], 0),
           np.where((idx[i] >= 0) & (idx[i] < m), bit[i], 0.0), m) for i in range(n_levels)]
  ref = O.signature_kern_low_rank_sketch(U, n_levels, keep, 1)[1:]
  dU = np.stack([np.diff(U, axis=1)] * n_levels, 0)
  got = _ops.lowrank_levels_sketch(to_dev(dU), to_dev(idx), to_dev(bit), m).cpu().numpy()
  assert _levels_err(got.reshape(3, -1), ref) < 1e-11
  with pytest.raises(NotImplementedError):
    _ops.lowrank_levels_sketch(to_dev(dU), to_dev(idx), to_dev(bit), 4096)


def test_tensorsketch_standalone_transform():
  """TensorSketch alone: the count sketch, and the convolution of a sketch
  with the next input's count sketch (the FFT route), against the oracle."""
  rng = np.random.default_rng(160)
  X = rng.standard_normal((5, 7, 6))
  ts = ksig.projections.TensorSketch(n_components=11, random_state=3).fit(X)
  idx = ts.hash_idx_.cpu().numpy()
  bit = ts.hash_bit_.cpu().numpy()
  assert idx.min() >= 0 and idx.max() < 11 and set(np.unique(bit)) <= {-1.0, 1.0}
  got = ts.transform(X)
  assert normwise(got, O.count_sketch(X, idx, bit, 11)) < 1e-13
  A = rng.standard_normal((5, 7, 11))
  ts2 = ksig.projections.TensorSketch(n_components=11, random_state=4).fit(A, Y=X)
  got2 = ts2.transform(A, Y=X)
  ref2 = O.convolve_fft(A, O.count_sketch(X, ts2.hash_idx_.cpu().numpy(),
                                          ts2.hash_bit_.cpu().numpy(), 11))
  assert normwise(got2, ref2) < 1e-12
  assert ksig.projections.CountSketch is ksig.projections.TensorSketch


@pytest.mark.parametrize("n_levels,order,m", [(4, 1, 40), (3, 2, 64), (3, 1, 2500), (4, 3, 10)])
def test_rfsf_tensorsketch_class(n_levels, order, m):
  """SignatureFeatures with TensorSketch on random Fourier features against
  the oracle on the object's weights, hashes and signs; n_comp = 2500 takes
  the projected recursion (the scan is built up to 2048)."""
  rng = np.random.default_rng(170 + m)
  X = O.brownian(rng, 5, 16, 3).astype(np.float32)
  sf = ksig.kernels.SignatureFeatures(
      n_levels=n_levels, order=order,
      static_features=ksig.static.features.RandomFourierFeatures(n_components=12,
                                                                 random_state=8),
      projection=ksig.projections.TensorSketch(n_components=m, random_state=9)).fit(X)
  w = [f.random_weights_.double().cpu().numpy() for f in sf.static_features_]
  sk = [(p.hash_idx_.cpu().numpy(), p.hash_bit_.double().cpu().numpy(), m)
        for p in sf.projections_]
  F = sf.transform(X)
  U = [O.rff_features(X.astype(np.float64), wi) for wi in w]
  ref = O.low_rank_features_normalised(
      O.signature_kern_low_rank_sketch(U, n_levels, sk, order), n_levels)
  assert F.shape == (5, 1 + n_levels * m) and normwise(F, ref) < 1e-5


def test_tensorized_projection_rank_above_one_follows_the_reference():
  """rank > 1: a standalone projection works (n_components columns); as a
  recursive step the fitted level expects n_components * rank features and
  raises ValueError, and a single level fails the rank reshape."""
  rng = np.random.default_rng(180)
  X = rng.standard_normal((4, 9, 3))
  trp = ksig.projections.TensorizedRandomProjection(n_components=5, rank=2,
                                                    random_state=1).fit(X)
  assert tuple(trp.transform(X).shape) == (4, 9, 5)
  Xs = O.brownian(rng, 4, 10, 3).astype(np.float32)
  for n_levels in (1, 3):
    sf = ksig.kernels.SignatureFeatures(
        n_levels=n_levels,
        static_features=ksig.static.features.RandomFourierFeatures(n_components=6),
        projection=ksig.projections.TensorizedRandomProjection(n_components=5, rank=2))
    with pytest.raises(ValueError):
      sf.fit(Xs)
      sf.transform(Xs)


def test_scans_on_empty_inputs():
  """No time steps (sequences of one point, differenced): every level
  feature is an empty sum; no sequences: an empty result."""
  z = to_dev(np.zeros((3, 4, 0, 12)))
  dz = _ops.lowrank_levels_diagonal(z, 3, 2)
  assert dz.shape == (4, (3 + 9 + 27) * 4) and not bool(dz.abs().max() > 0)
  idx = to_dev(np.zeros((3, 12), dtype=np.int32))
  bit = to_dev(np.ones((3, 12)))
  for m in (5, 30):  # the count sketch first, and from the sparse sketch
    out = _ops.lowrank_levels_sketch(z, idx, bit, m, 2)
    assert out.shape == (4, 3, m) and not bool(out.abs().max() > 0)
  e = to_dev(np.zeros((3, 0, 7, 12)))
  assert _ops.lowrank_levels_sketch(e, idx, bit, 5).shape == (0, 3, 5)
  assert _ops.lowrank_levels_diagonal(e, 3, 1).shape == (0, (3 + 9 + 27) * 4)
