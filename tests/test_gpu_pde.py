"""GPU parity of the signature-PDE sweep against the FP64 oracle."""

import math
import os

import numpy as np
import pytest
import torch

import ksig_amd as ksig
from oracle import sigkern as O
from tests.gpu_util import normwise, to_dev

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_pde_dyadic_golden():
  g = np.load(os.path.join(GOLDEN, "pde_dyadic.npz"))
  for lam in range(3):
    pk = ksig.kernels.SignaturePDEKernel(normalize=False, dyadic_order=lam)
    assert normwise(pk(g["X"], g["Y"]), g[f"pde_lam{lam}"]) < 1e-11, lam


def test_pde_c0_golden():
  g = np.load(os.path.join(GOLDEN, "c0_signature_rbf.npz"))
  pk = ksig.kernels.SignaturePDEKernel(normalize=False)
  assert normwise(pk(g["X"][:, :20] * 0.3, g["Y"] * 0.3), g["pde_XY_unnorm"]) < 1e-11


@pytest.mark.parametrize("lx,ly", [(2, 2), (5, 90), (65, 65), (130, 64), (257, 100)])
@pytest.mark.parametrize("lam", [0, 1, 2])
def test_pde_shapes_float64(lx, ly, lam):
  rng = np.random.default_rng(lx * 31 + ly + lam)
  X = O.brownian(rng, 2, lx, 3)
  Y = O.brownian(rng, 2, ly, 3)
  pk = ksig.kernels.SignaturePDEKernel(normalize=False, dyadic_order=lam)
  got = pk(X, Y)
  ref = O.signature_pde_kernel(X, Y, normalize=False, dyadic_order=lam)
  assert normwise(got, ref) < 1e-11


@pytest.mark.parametrize("d", [1, 4, 5, 8, 16, 20])
@pytest.mark.parametrize("lam", [0, 1])
@pytest.mark.parametrize("kind", ["rbf", "linear"])
def test_pde_feature_buckets(d, lam, kind):
  """Every feature bucket of the phase-aligned sweep, with 16-byte y-row loads
  (d = 4, 8, 16) and per-feature loads (d = 1, 5, 20), in float64 and float32,
  on pairs over one and several bands."""
  rng = np.random.default_rng(100 * d + 10 * lam + (kind == "rbf"))
  X = O.brownian(rng, 3, 150, d)
  Y = O.brownian(rng, 2, 70, d)
  static = (ksig.static.kernels.RBFKernel(1.3) if kind == "rbf"
            else ksig.static.kernels.LinearKernel(0.7))
  ref = O.signature_pde_kernel(X, Y, normalize=False, dyadic_order=lam, kind=kind,
                               param=1.3 if kind == "rbf" else 0.7)
  pk = ksig.kernels.SignaturePDEKernel(normalize=False, dyadic_order=lam, static_kernel=static)
  assert normwise(pk(X, Y), ref) < 1e-11
  got32 = pk(X.astype(np.float32), Y.astype(np.float32))
  ref32 = O.signature_pde_kernel(X.astype(np.float32).astype(np.float64),
                                 Y.astype(np.float32).astype(np.float64), normalize=False,
                                 dyadic_order=lam, kind=kind,
                                 param=1.3 if kind == "rbf" else 0.7)
  assert normwise(got32, ref32) < 2e-5


def test_pde_c2_shape_float32_subset():
  """C2 shapes (L = 256, d = 4, dyadic 1) on 4 x 3 pairs in float32."""
  rng = np.random.default_rng(2)
  X = O.brownian(rng, 4, 256, 4).astype(np.float32)
  Y = O.brownian(rng, 3, 256, 4).astype(np.float32)
  pk = ksig.kernels.SignaturePDEKernel(normalize=False, dyadic_order=1)
  got = pk(X, Y)
  ref = O.signature_pde_kernel(X.astype(np.float64), Y.astype(np.float64), normalize=False,
                               dyadic_order=1)
  assert normwise(got, ref) < 1e-5
  Kn = ksig.kernels.SignaturePDEKernel(dyadic_order=1)(X)
  refn = O.signature_pde_kernel(X.astype(np.float64), dyadic_order=1)
  assert normwise(Kn, refn) < 1e-5


@pytest.mark.parametrize("difference", [True, False])
def test_pde_gram_input_and_linear(difference):
  rng = np.random.default_rng(3)
  M = rng.standard_normal((2, 3, 40, 70)) * 0.02
  for lam in (0, 1):
    got = ksig.algorithms.signature_kern_pde(M, difference=difference, dyadic_order=lam)
    ref = O.signature_kern_pde(M, difference, lam)
    assert normwise(got.cpu().numpy(), ref) < 1e-11
  X = rng.standard_normal((3, 20, 2)) * 0.2
  pk = ksig.kernels.SignaturePDEKernel(normalize=False, difference=difference,
                                       static_kernel=ksig.static.kernels.LinearKernel(0.7))
  got = pk(X)
  ref = O.signature_pde_kernel(X, normalize=False, difference=difference, kind="linear",
                               param=0.7)
  assert normwise(got, ref) < 1e-11


def test_pde_diag_and_ndim_check():
  rng = np.random.default_rng(4)
  X = rng.standard_normal((3, 15, 2)) * 0.3
  pk = ksig.kernels.SignaturePDEKernel(normalize=False)
  np.testing.assert_allclose(pk(X, diag=True), np.diag(pk(X)), rtol=1e-13)
  np.testing.assert_array_equal(ksig.kernels.SignaturePDEKernel()(X, diag=True), 1.0)
  with pytest.raises(ValueError):
    ksig.algorithms.signature_kern_pde(np.zeros((2, 3)))
