"""GPU parity of the matrix-core static Gram variants of the DP sweep
('fp32': float32 MFMA tiles, 'bf16x3': bf16 hi/lo split with three MFMA
products) against the FP64 oracle, float32 inputs, normwise error over each
pair's level vector.  Tolerances: 'fp32' 1e-5; 'bf16x3' 1e-5 for d >= 3 and
2e-5 at d = 1, where a single split product (~2^-16 relative error, no
averaging over features) feeds every Gram value."""

import math

import numpy as np
import pytest
import torch

import ksig_amd as ksig
from ksig_amd import _lib, _ops
from oracle import sigkern as O
from tests.gpu_util import dev, level_err, normwise, to_dev

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _levels(X, Y, n, order, kind, param, mode):
  st = (ksig.static.kernels.RBFKernel(param) if kind == "rbf"
        else ksig.static.kernels.LinearKernel(param))
  sk = ksig.kernels.SignatureKernel(n, order=order, normalize=False, static_kernel=st,
                                    gram_precision=mode)
  return sk._levels(to_dev(X, torch.float32), to_dev(Y, torch.float32)).cpu().numpy()


@pytest.mark.parametrize("mode", ["fp32", "bf16x3"])
@pytest.mark.parametrize("lx,ly,d,n,order", [
    (200, 200, 8, 5, 1),      # C1 shape
    (128, 128, 32, 6, 1),     # C4 shape
    (300, 257, 16, 8, 2),     # C3-like, order 2
    (65, 66, 5, 4, 1), (66, 65, 3, 3, 2), (3, 90, 8, 5, 1), (130, 17, 12, 5, 1),
    (129, 129, 1, 2, 1), (40, 300, 24, 6, 1)])
def test_mfma_modes_rbf(mode, lx, ly, d, n, order):
  rng = np.random.default_rng(lx * 3 + ly * 7 + d + n)
  X = O.brownian(rng, 2, lx, d) + 0.1 * rng.standard_normal((2, 1, d))
  Y = O.brownian(rng, 3, ly, d)
  sig = math.sqrt(d)
  got = _levels(X, Y, n, order, "rbf", sig, mode)
  X32 = X.astype(np.float32).astype(np.float64)
  Y32 = Y.astype(np.float32).astype(np.float64)
  ref = O.signature_kern(O.pairwise_gram(X32, Y32, "rbf", sig), n, order, True, True)
  err = level_err(got, ref)
  print(f"mode={mode} shape=({lx},{ly},{d}) n={n} order={order} err={err:.2e}")
  assert err < (2e-5 if (mode == "bf16x3" and d == 1) else TOL)


@pytest.mark.parametrize("mode", ["fp32", "bf16x3"])
def test_mfma_modes_linear(mode):
  rng = np.random.default_rng(5)
  X = rng.standard_normal((2, 70, 6)) * 0.2
  Y = rng.standard_normal((2, 90, 6)) * 0.2
  got = _levels(X, Y, 4, 1, "linear", 0.7, mode)
  ref = O.signature_kern(O.pairwise_gram(X.astype(np.float32).astype(np.float64),
                                         Y.astype(np.float32).astype(np.float64), "linear", 0.7),
                         4, 1, True, True)
  assert level_err(got, ref) < TOL


def test_mfma_matches_valu_path_normalised_gram():
  rng = np.random.default_rng(6)
  X = O.brownian(rng, 24, 200, 8).astype(np.float32)
  Ks = {}
  for mode in ("valu", "fp32", "bf16x3"):
    sk = ksig.kernels.SignatureKernel(5, static_kernel=ksig.static.kernels.RBFKernel(math.sqrt(8)),
                                      gram_precision=mode)
    Ks[mode] = sk(X)
  ref = O.signature_kernel(X.astype(np.float64), n_levels=5, param=math.sqrt(8))
  for mode, K in Ks.items():
    err = normwise(K, ref)
    print(f"normalised Gram mode={mode} err={err:.2e}")
    assert err < TOL


def test_mfma_ragged_pairs():
  rng = np.random.default_rng(7)
  lens_x = [40, 64, 65, 127, 300, 1023]
  lens_y = [1000, 65, 64, 129, 33, 128]
  xs = [O.brownian(rng, 1, L, 16)[0] for L in lens_x]
  ys = [O.brownian(rng, 1, L, 16)[0] for L in lens_y]
  Xs = _ops.SeqSet.ragged([to_dev(x, torch.float32) for x in xs])
  Ys = _ops.SeqSet.ragged([to_dev(y, torch.float32) for y in ys])
  pairs = torch.tensor([[i, i] for i in range(6)], device=dev(), dtype=torch.int32)
  for mode in (_lib.GRAM_MFMA_F32, _lib.GRAM_MFMA_BF16X3):
    got = _ops.signature_levels(Xs, Ys, pairs, 8, 2, True, _lib.STATIC_RBF, 4.0, mode).cpu().numpy()
    for p in range(6):
      ref = O.pair_levels(xs[p].astype(np.float32).astype(np.float64),
                          ys[p].astype(np.float32).astype(np.float64), 8, 2, True, "rbf", 4.0)
      err = level_err(got[p][:, None], ref[:, None])
      assert err < TOL, (mode, p, err)
