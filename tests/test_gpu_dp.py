"""GPU parity of the DP signature-kernel sweep (libksig_hip.so) against the
FP64 oracle and the committed golden fixtures.

Tolerances: float64 inputs run the float64 instantiation and must agree with
the oracle to 1e-11 (normwise); float32 inputs must agree to 1e-5 normwise
relative error over each pair's level vector."""

import math
import os

import numpy as np
import pytest
import torch

import ksig_amd as ksig
from oracle import sigkern as O
from tests.gpu_util import dev, level_err, normwise, to_dev

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
TOL64 = 1e-11
TOL32 = 1e-5


def test_c0_golden_fixture():
  g = np.load(os.path.join(GOLDEN, "c0_signature_rbf.npz"))
  X, Y = g["X"], g["Y"]
  sk = ksig.kernels.SignatureKernel(5, order=1, static_kernel=ksig.static.kernels.RBFKernel(1.0))
  assert normwise(sk(X), g["K_XX"]) < TOL64
  assert normwise(sk(X, Y), g["K_XY"]) < TOL64
  np.testing.assert_array_equal(sk(X, diag=True), g["K_X_diag"])
  su = ksig.kernels.SignatureKernel(5, order=1, normalize=False)
  assert normwise(su(X), g["K_XX_unnorm"]) < TOL64
  assert normwise(su(X, Y), g["K_XY_unnorm"]) < TOL64
  assert normwise(su(X, diag=True), g["K_X_diag_unnorm"]) < TOL64
  lev = ksig.algorithms.signature_kern(
      ksig.static.kernels.RBFKernel(1.0)(X[:, None, :, None, :], Y[None, :, :, None, :],
                                         return_on_gpu=True), 5, order=1, return_levels=True)
  assert normwise(lev.cpu().numpy(), g["levels_XY"]) < TOL64


def test_c0_output_types_and_shapes():
  g = np.load(os.path.join(GOLDEN, "c0_signature_rbf.npz"))
  sk = ksig.kernels.SignatureKernel(5)
  K = sk(g["X"])
  assert isinstance(K, np.ndarray) and K.shape == (10, 10) and K.dtype == np.float64
  Kg = sk(g["X"], g["Y"], return_on_gpu=True)
  assert isinstance(Kg, torch.Tensor) and Kg.is_cuda and Kg.shape == (10, 8)
  assert sk(g["X"], diag=True).shape == (10,)
  K32 = sk(g["X"].astype(np.float32))
  assert K32.dtype == np.float32


def test_higher_order_golden():
  g = np.load(os.path.join(GOLDEN, "higher_order.npz"))
  X, Y = g["X"], g["Y"]
  for order in (2, 3):
    sk = ksig.kernels.SignatureKernel(4, order=order, normalize=False,
                                      static_kernel=ksig.static.kernels.RBFKernel(0.9))
    lev = sk._levels(to_dev(X), to_dev(Y)).cpu().numpy()
    assert normwise(lev, g[f"levels_rbf_o{order}"]) < TOL64, order
  sk = ksig.kernels.SignatureKernel(4, order=-1, normalize=False,
                                    static_kernel=ksig.static.kernels.LinearKernel(1.0))
  lev = sk._levels(to_dev(X), to_dev(Y)).cpu().numpy()
  assert normwise(lev, g["levels_lin_full"]) < TOL64


@pytest.mark.parametrize("lx,ly", [(2, 2), (3, 70), (65, 66), (66, 65), (130, 40), (200, 200),
                                   (129, 257)])
@pytest.mark.parametrize("nlev,order", [(5, 1), (3, 2), (8, 2), (4, 4), (6, 1)])
def test_pair_shapes_float64(lx, ly, nlev, order):
  """Band boundaries (64-row bands), short column sides (< 64 padded period),
  transposed orientation, and the level/order table, in float64."""
  rng = np.random.default_rng(lx * 1000 + ly + 7 * nlev + order)
  X = O.brownian(rng, 2, lx, 3)
  Y = O.brownian(rng, 2, ly, 3)
  sig = 0.7
  sk = ksig.kernels.SignatureKernel(nlev, order=order, normalize=False,
                                    static_kernel=ksig.static.kernels.RBFKernel(sig))
  got = sk._levels(to_dev(X), to_dev(Y)).cpu().numpy()
  ref = O.signature_kern(O.pairwise_gram(X, Y, "rbf", sig), nlev, order, True, True)
  assert level_err(got, ref) < TOL64


@pytest.mark.parametrize("d", [1, 5, 8, 9, 16, 17, 32])
def test_feature_buckets_float32(d):
  rng = np.random.default_rng(100 + d)
  X = O.brownian(rng, 3, 90, d)
  Y = O.brownian(rng, 2, 70, d)
  sig = math.sqrt(d)
  sk = ksig.kernels.SignatureKernel(5, normalize=False,
                                    static_kernel=ksig.static.kernels.RBFKernel(sig))
  got = sk._levels(to_dev(X, torch.float32), to_dev(Y, torch.float32)).cpu().numpy()
  ref = O.signature_kern(O.pairwise_gram(X.astype(np.float32).astype(np.float64),
                                         Y.astype(np.float32).astype(np.float64), "rbf", sig),
                         5, 1, True, True)
  assert level_err(got, ref) < TOL32


def test_c1_shape_float32_subset():
  """512 x 200 x 8 Brownian, sigma = sqrt(8), 5 levels: a 12-sequence subset
  of the symmetric Gram against the FP64 oracle."""
  rng = np.random.default_rng(1)
  X = O.brownian(rng, 12, 200, 8)
  Xf = X.astype(np.float32)
  sk = ksig.kernels.SignatureKernel(5, static_kernel=ksig.static.kernels.RBFKernel(math.sqrt(8)))
  K = sk(Xf)
  ref = O.signature_kernel(Xf.astype(np.float64), n_levels=5, param=math.sqrt(8))
  assert normwise(K, ref) < TOL32
  su = ksig.kernels.SignatureKernel(5, normalize=False,
                                    static_kernel=ksig.static.kernels.RBFKernel(math.sqrt(8)))
  lev = su._levels(to_dev(Xf)).cpu().numpy()
  ref_lev = O.signature_kern(O.pairwise_gram(Xf.astype(np.float64), None, "rbf", math.sqrt(8)),
                             5, 1, True, True)
  assert level_err(lev, ref_lev) < TOL32


@pytest.mark.parametrize("kind", ["linear", "rbf"])
@pytest.mark.parametrize("difference", [True, False])
def test_static_kinds_and_difference(kind, difference):
  rng = np.random.default_rng(5)
  X = rng.standard_normal((3, 40, 4)) * 0.3
  Y = rng.standard_normal((2, 33, 4)) * 0.3
  param = 0.8
  st = (ksig.static.kernels.LinearKernel(param) if kind == "linear"
        else ksig.static.kernels.RBFKernel(param))
  n = 4
  sk = ksig.kernels.SignatureKernel(n, order=2, difference=difference, normalize=False,
                                    static_kernel=st)
  got = sk._levels(to_dev(X), to_dev(Y)).cpu().numpy()
  ref = O.signature_kern(O.pairwise_gram(X, Y, kind, param), n, 2, difference, True)
  assert level_err(got, ref) < TOL64


def test_gram_input_functional_api():
  """signature_kern(M) on a caller-supplied tensor (generic static kernels)."""
  rng = np.random.default_rng(6)
  M = rng.standard_normal((3, 2, 70, 45)) * 0.05
  for order, difference in [(1, True), (2, False), (-1, True)]:
    got = ksig.algorithms.signature_kern(M, 4, order=order, difference=difference,
                                         return_levels=True).cpu().numpy()
    ref = O.signature_kern(M, 4, order, difference, True)
    assert normwise(got, ref) < TOL64
    got1 = ksig.algorithms.signature_kern(M, 4, order=order, difference=difference).cpu().numpy()
    assert normwise(got1, O.signature_kern(M, 4, order, difference, False)) < TOL64
  M3 = rng.standard_normal((4, 66, 66)) * 0.05
  got = ksig.algorithms.signature_kern(to_dev(M3, torch.float32), 5).cpu().numpy()
  ref = O.signature_kern(M3.astype(np.float32).astype(np.float64), 5, -1, True, False)
  assert normwise(got, ref) < TOL32


class _Poly(ksig.static.kernels.StaticKernel):
  """A static kernel the sweep does not fuse: goes through the M path."""

  def __init__(self, c: float = 1.0):
    self.c = c

  def _K(self, X, Y=None):
    Y = X if Y is None else Y
    return (torch.matmul(X, Y.transpose(-1, -2).contiguous()) + self.c) ** 2

  def _Kdiag(self, X):
    return ((X * X).sum(-1) + self.c) ** 2


def test_generic_static_kernel_path():
  rng = np.random.default_rng(7)
  X = rng.standard_normal((3, 12, 2)) * 0.3
  Y = rng.standard_normal((2, 9, 2)) * 0.3
  sk = ksig.kernels.SignatureKernel(3, static_kernel=_Poly(0.5))
  K = sk(X, Y)
  M = (np.einsum("aik,bjk->abij", X, Y) + 0.5) ** 2
  lev = O.signature_kern(M, 3, 1, True, True)
  Mx = (np.einsum("aik,ajk->aij", X, X) + 0.5) ** 2
  My = (np.einsum("aik,ajk->aij", Y, Y) + 0.5) ** 2
  lx = O.signature_kern(Mx, 3, 1, True, True)
  ly = O.signature_kern(My, 3, 1, True, True)
  ref = np.mean(lev / np.sqrt(lx[:, :, None] * ly[:, None, :]), 0)
  assert normwise(K, ref) < TOL64


def test_empty_grid_and_tiny_inputs():
  X = np.zeros((2, 1, 3))
  Y = np.random.default_rng(8).standard_normal((3, 5, 3))
  su = ksig.kernels.SignatureKernel(3, normalize=False)
  lev = su._levels(to_dev(X), to_dev(Y)).cpu().numpy()
  np.testing.assert_array_equal(lev[0], 1.0)
  np.testing.assert_array_equal(lev[1:], 0.0)
  K = su(X, Y)
  np.testing.assert_array_equal(K, 1.0)


def test_ragged_pairs_native_lengths():
  from ksig_amd import _ops
  rng = np.random.default_rng(9)
  lens_x = [3, 64, 65, 130, 17, 200]
  lens_y = [90, 2, 64, 128, 300, 33]
  xs = [O.brownian(rng, 1, L, 4)[0] for L in lens_x]
  ys = [O.brownian(rng, 1, L, 4)[0] for L in lens_y]
  Xs = _ops.SeqSet.ragged([to_dev(x) for x in xs])
  Ys = _ops.SeqSet.ragged([to_dev(y) for y in ys])
  pairs = torch.tensor([[i, i] for i in range(6)] + [[0, 5], [5, 0]], device=dev(),
                       dtype=torch.int32)
  got = _ops.signature_levels(Xs, Ys, pairs, 8, 2, True, 0, 2.0).cpu().numpy()
  for p, (a, b) in enumerate(pairs.cpu().numpy()):
    ref = O.pair_levels(xs[a], ys[b], 8, 2, True, "rbf", 2.0)
    assert level_err(got[p][:, None], ref[:, None]) < TOL64, (a, b)


def test_symmetric_gram_matches_full_evaluation():
  rng = np.random.default_rng(10)
  X = O.brownian(rng, 7, 50, 3)
  sk = ksig.kernels.SignatureKernel(4, normalize=False)
  Kxx = sk(X)
  Kxy = sk(X, X)
  assert normwise(Kxx, Kxy) < 1e-13
  np.testing.assert_allclose(Kxx, Kxx.T, rtol=0, atol=0)


def test_untabulated_combination_routes_around_the_fused_entry():
  """(40, 1) is outside the compiled table: the fused entry refuses it, the
  class takes the GEMM route (runtime-levels kernel) and matches the oracle."""
  from ksig_amd import _lib, _ops
  rng = np.random.default_rng(40)
  X = O.brownian(rng, 2, 5, 3)
  Xs = _ops.SeqSet.dense(to_dev(X))
  lib = _lib.load()
  out = torch.empty((1, 41), device=dev(), dtype=torch.float64)
  pairs = to_dev(np.array([[0, 1]], dtype=np.int32))
  rc = lib.ksig_signature_levels(_lib.F64, Xs.data.data_ptr(), Xs.off.data_ptr(),
                                 Xs.len.data_ptr(), 5, Xs.data.data_ptr(), Xs.off.data_ptr(),
                                 Xs.len.data_ptr(), 5, 3, pairs.data_ptr(), 1, _lib.STATIC_RBF,
                                 1.0, 40, 1, 1, -1, out.data_ptr(), _lib.stream_handle(dev()))
  assert rc == 1002
  K = ksig.kernels.SignatureKernel(40, order=1, normalize=False)(X)
  assert normwise(K, O.signature_kernel(X, n_levels=40, normalize=False)) < TOL64
