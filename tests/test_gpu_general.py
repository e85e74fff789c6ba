"""GPU parity of the general route: inputs the fused sweeps do not take —
d > 32, and (n_levels, order) outside the compiled table — through library
GEMMs on the matrix cores (K = d) and the recursion kernels of
ksig_signature_levels_dots / ksig_signature_levels_gram (the runtime-levels
kernel for untabulated orders), against the FP64 oracle.

Tolerances: float64 1e-11 normwise, float32 1e-5 normwise relative error over
each pair's level vector (the bars of the fused kernels)."""

import math

import numpy as np
import pytest
import torch

import ksig_amd as ksig
from ksig_amd import _lib, _ops
from oracle import sigkern as O
from tests.gpu_util import level_err, normwise, to_dev

pytestmark = pytest.mark.gpu
TOL64 = 1e-11
TOL32 = 1e-5


def _paths(rng, n, L, d):
  return O.brownian(rng, n, L, d)


@pytest.mark.parametrize("d", [33, 64, 129, 963])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_signature_kernel_beyond_d32(d, dtype):
  rng = np.random.default_rng(d)
  X = _paths(rng, 5, 21, d)
  Y = _paths(rng, 3, 34, d)   # Y longer than X: the route swaps the sides
  sig = math.sqrt(d)
  tol = TOL64 if dtype == np.float64 else TOL32
  Xc, Yc = X.astype(dtype), Y.astype(dtype)
  Xr, Yr = Xc.astype(np.float64), Yc.astype(np.float64)
  sk = ksig.kernels.SignatureKernel(5, order=1, normalize=False,
                                    static_kernel=ksig.static.kernels.RBFKernel(sig))
  lev = sk._levels(to_dev(Xc), to_dev(Yc)).cpu().numpy()
  ref = O.signature_kern(O.pairwise_gram(Xr, Yr, "rbf", sig), 5, 1, True, True)
  assert level_err(lev.reshape(6, -1), ref.reshape(6, -1)) < tol
  levxx = sk._levels(to_dev(Xc)).cpu().numpy()
  refxx = O.signature_kern(O.pairwise_gram(Xr, None, "rbf", sig), 5, 1, True, True)
  assert level_err(levxx.reshape(6, -1), refxx.reshape(6, -1)) < tol
  skn = ksig.kernels.SignatureKernel(5, order=1, static_kernel=ksig.static.kernels.RBFKernel(sig))
  assert normwise(skn(Xc, Yc), O.signature_kernel(Xr, Yr, n_levels=5, param=sig)) < \
      max(tol, 1e-10)
  assert normwise(sk(Xc, diag=True),
                  O.signature_kernel(Xr, diag=True, n_levels=5, normalize=False,
                                     param=sig)) < tol


@pytest.mark.parametrize("nlev,order", [(6, 4), (15, 1), (6, 5), (11, 2), (13, 3), (7, 4)])
def test_untabulated_levels_and_orders(nlev, order):
  assert not _lib.load().ksig_dp_supported(nlev, order)  # outside the compiled table
  rng = np.random.default_rng(100 + nlev * 10 + order)
  L = 150 if (nlev, order) == (7, 4) else 23   # 149 rows: three 64-row bands
  X = _paths(rng, 3, L, 4)
  Y = _paths(rng, 2, 17, 4)
  for dtype, tol in ((np.float64, TOL64), (np.float32, TOL32)):
    Xc, Yc = X.astype(dtype), Y.astype(dtype)
    sk = ksig.kernels.SignatureKernel(nlev, order=order, normalize=False,
                                      static_kernel=ksig.static.kernels.RBFKernel(1.3))
    lev = sk._levels(to_dev(Xc), to_dev(Yc)).cpu().numpy()
    ref = O.signature_kern(O.pairwise_gram(Xc.astype(np.float64), Yc.astype(np.float64),
                                           "rbf", 1.3), nlev, order, True, True)
    assert level_err(lev.reshape(nlev + 1, -1), ref.reshape(nlev + 1, -1)) < tol
  # the functional form on a caller-supplied M (the runtime-levels kernel on
  # plain matrices), with and without differencing
  M = O.pairwise_gram(X, Y, "linear", 0.8)
  for diff in (True, False):
    got = ksig.algorithms.signature_kern(M, nlev, order=order, difference=diff,
                                         return_levels=True).cpu().numpy()
    ref = O.signature_kern(M, nlev, order, diff, True)
    assert level_err(got.reshape(nlev + 1, -1), ref.reshape(nlev + 1, -1)) < TOL64


def test_ragged_pairs_beyond_d32():
  rng = np.random.default_rng(7)
  d = 40
  lx = [2, 9, 30, 71, 1, 15]
  ly = [5, 33, 12, 80, 3]
  xs = [O.brownian(rng, 1, L, d)[0] for L in lx]
  ys = [O.brownian(rng, 1, L, d)[0] for L in ly]
  X = _ops.SeqSet.ragged([to_dev(s) for s in xs])
  Y = _ops.SeqSet.ragged([to_dev(s) for s in ys])
  pr = np.array([[a, b] for a in range(len(lx)) for b in range(len(ly))], dtype=np.int32)
  for kind, code, param in (("rbf", _lib.STATIC_RBF, 5.0), ("linear", _lib.STATIC_LINEAR, 0.4)):
    got = _ops.signature_levels(X, Y, to_dev(pr), 4, 2, True, code, param).cpu().numpy()
    ref = np.stack([O.pair_levels(xs[a], ys[b], 4, 2, True, kind, param) for a, b in pr])
    assert level_err(got.T, ref.T) < TOL64


def test_tabulated_orders_through_the_dot_route():
  """(5, 1) and (8, 2) run the register-resident kernels on the strided
  dot products (the d > 32 route), also without differencing."""
  rng = np.random.default_rng(11)
  d = 48
  X = _paths(rng, 4, 40, d).astype(np.float32)
  Y = _paths(rng, 6, 25, d).astype(np.float32)
  Xr, Yr = X.astype(np.float64), Y.astype(np.float64)
  for nlev, order in ((5, 1), (8, 2)):
    for diff in (True, False):
      sk = ksig.kernels.SignatureKernel(nlev, order=order, difference=diff, normalize=False,
                                        static_kernel=ksig.static.kernels.RBFKernel(7.0))
      lev = sk._levels(to_dev(X), to_dev(Y)).cpu().numpy()
      ref = O.signature_kern(O.pairwise_gram(Xr, Yr, "rbf", 7.0), nlev, order, diff, True)
      assert level_err(lev.reshape(nlev + 1, -1), ref.reshape(nlev + 1, -1)) < TOL32
  lin = ksig.kernels.SignatureKernel(5, normalize=False,
                                     static_kernel=ksig.static.kernels.LinearKernel(0.3))
  lev = lin._levels(to_dev(X)).cpu().numpy()
  ref = O.signature_kern(O.pairwise_gram(Xr, None, "linear", 0.3), 5, 1, True, True)
  assert level_err(lev.reshape(6, -1), ref.reshape(6, -1)) < TOL32


def test_pde_gak_rws_beyond_d32():
  rng = np.random.default_rng(12)
  d = 40
  X = _paths(rng, 4, 30, d)
  Y = _paths(rng, 3, 26, d)
  pk = ksig.kernels.SignaturePDEKernel(normalize=False, dyadic_order=1,
                                       static_kernel=ksig.static.kernels.RBFKernel(6.0))
  ref = O.signature_pde_kernel(X, Y, normalize=False, param=6.0, dyadic_order=1)
  assert normwise(pk(X, Y), ref) < TOL64
  gk = ksig.kernels.GlobalAlignmentKernel(static_kernel=ksig.static.kernels.RBFKernel(6.0))
  ref = O.global_alignment_kernel(X, Y, kind="rbf", param=6.0)
  assert normwise(gk(X, Y), ref) < 1e-10
  W = rng.standard_normal((17, d)) * 0.3
  lens = np.array([5, 1, 11])
  got = _ops.random_warping_series(_ops.SeqSet.dense(to_dev(X)), to_dev(W), lens).cpu().numpy()
  ref = O.random_warping_series(O.squared_euclid_dist(X, W[None]), lens)
  assert normwise(got, ref) < TOL64
