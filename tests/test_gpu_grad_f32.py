"""float32 DP adjoint (static Gram on MFMA, skewed LDS ring) against the
float64 adjoint, which is itself pinned to autograd through the torch float64
transcription and to central differences (test_gpu_grad.py).

Cases cover feature counts below, at and between the 8/16/32 buckets (vector
and clamped loads), lengths around the 64-row bands and the 16-column chunks,
both orientations (len_y > len_x swaps the sweep), RBF and linear, and 1 to 12
levels.  Tolerance: 2e-4 normwise relative error of dL/dx, dL/dy and dL/dparam
(float32 recursion against float64)."""

import numpy as np
import pytest
import torch

from ksig_amd import _lib, _ops, autograd
from oracle import sigkern as O
from tests.gpu_util import dev

pytestmark = pytest.mark.gpu
TOL = 2e-4


def _rel(a, b):
  a = a.detach().double().cpu()
  b = b.detach().double().cpu()
  return float((a - b).abs().max() / b.abs().max().clamp_min(1e-300))


def _grads(xs, ys, n_levels, kind, param, c, dtype):
  X = torch.tensor(xs, device=dev(), dtype=dtype).requires_grad_(True)
  Y = torch.tensor(ys, device=dev(), dtype=dtype).requires_grad_(True)
  p = torch.tensor(param, dtype=torch.float64, device=dev(), requires_grad=True)
  Xs, Ys = _ops.SeqSet.dense(X.detach()), _ops.SeqSet.dense(Y.detach())
  pairs = _ops.dense_pairs(xs.shape[0], ys.shape[0], dev())
  code = _lib.STATIC_RBF if kind == "rbf" else _lib.STATIC_LINEAR
  d = xs.shape[-1]
  lev = autograd.signature_levels(Xs, Ys, pairs, n_levels, code, p, X.reshape(-1, d),
                                  Y.reshape(-1, d))
  (lev * c).sum().backward()
  return lev.detach(), X.grad, Y.grad, p.grad


@pytest.mark.parametrize("d", [3, 8, 13, 16, 24, 32])
@pytest.mark.parametrize("lx,ly", [(2, 5), (17, 17), (65, 40), (40, 66), (80, 130), (130, 64)])
def test_float32_adjoint_vs_float64(d, lx, ly):
  rng = np.random.default_rng(1000 + 7 * d + lx + 3 * ly)
  xs = O.brownian(rng, 2, lx, d)
  ys = O.brownian(rng, 3, ly, d)
  n = 6
  c = torch.tensor(rng.standard_normal((6, n + 1)), device=dev())
  sig = float(np.sqrt(d))
  ref = _grads(xs, ys, n, "rbf", sig, c, torch.float64)
  got = _grads(xs, ys, n, "rbf", sig, c, torch.float32)
  assert _rel(got[0], ref[0]) < 1e-5
  assert _rel(got[1], ref[1]) < TOL
  assert _rel(got[2], ref[2]) < TOL
  assert abs(float(got[3]) - float(ref[3])) < TOL * max(1.0, abs(float(ref[3])))


@pytest.mark.parametrize("kind,param", [("rbf", 0.9), ("linear", 0.6)])
@pytest.mark.parametrize("n_levels", [1, 2, 4, 12])
def test_float32_adjoint_levels_and_kinds(kind, param, n_levels):
  rng = np.random.default_rng(2000 + n_levels)
  xs = O.brownian(rng, 2, 70, 5)
  ys = O.brownian(rng, 2, 90, 5)
  c = torch.tensor(rng.standard_normal((4, n_levels + 1)), device=dev())
  ref = _grads(xs, ys, n_levels, kind, param, c, torch.float64)
  got = _grads(xs, ys, n_levels, kind, param, c, torch.float32)
  assert _rel(got[1], ref[1]) < TOL
  assert _rel(got[2], ref[2]) < TOL
  assert abs(float(got[3]) - float(ref[3])) < TOL * max(1.0, abs(float(ref[3])))


def _grads_xy(xs, ys, n_levels, kind, param, c, dtype, param_grad):
  X = torch.tensor(xs, device=dev(), dtype=dtype).requires_grad_(True)
  Y = torch.tensor(ys, device=dev(), dtype=dtype).requires_grad_(True)
  p = torch.tensor(param, dtype=torch.float64, device=dev(), requires_grad=param_grad)
  Xs, Ys = _ops.SeqSet.dense(X.detach()), _ops.SeqSet.dense(Y.detach())
  pairs = _ops.dense_pairs(xs.shape[0], ys.shape[0], dev())
  code = _lib.STATIC_RBF if kind == "rbf" else _lib.STATIC_LINEAR
  d = xs.shape[-1]
  lev = autograd.signature_levels(Xs, Ys, pairs, n_levels, code, p, X.reshape(-1, d),
                                  Y.reshape(-1, d))
  (lev * c).sum().backward()
  return X.grad, Y.grad, p.grad


@pytest.mark.parametrize("kind,param", [("rbf", 1.7), ("linear", 0.4)])
@pytest.mark.parametrize("d,n_levels", [(32, 6), (8, 3), (16, 1)])
def test_adjoint_without_parameter_gradient(kind, param, d, n_levels):
  """With the bandwidth / scale not requiring a gradient the adjoint is
  launched without grad_param (the float32 adjoint then drops the argument
  flow): dL/dX and dL/dY equal those of the launch with it up to the order
  of the contraction's float atomics, and p.grad stays None."""
  rng = np.random.default_rng(3000 + d + n_levels)
  xs = O.brownian(rng, 3, 90, d)
  ys = O.brownian(rng, 2, 140, d)
  c = torch.tensor(rng.standard_normal((6, n_levels + 1)), device=dev())
  gx1, gy1, gp1 = _grads_xy(xs, ys, n_levels, kind, param, c, torch.float32, True)
  gx0, gy0, gp0 = _grads_xy(xs, ys, n_levels, kind, param, c, torch.float32, False)
  assert gp1 is not None and gp0 is None
  assert _rel(gx0, gx1) < 1e-6 and _rel(gy0, gy1) < 1e-6
  gx64, gy64, _ = _grads_xy(xs, ys, n_levels, kind, param, c, torch.float64, False)
  assert _rel(gx0, gx64) < TOL and _rel(gy0, gy64) < TOL
