"""The fused signature-PDE adjoint (chain rule through ΔΔ and the static
kernel inside the reverse sweep, ksig_signature_pde_bwd_fused) against the
two-call path (ksig_signature_pde_bwd_edges + ksig_signature_pde_grad_contract)
on the same float32 inputs, and against the float64 adjoint, on ragged sets
spanning one to three bands in both orientations, with the gradient flowing
to both sides (aliased for a Gram of one set)."""

import numpy as np
import pytest
import torch

from ksig_amd import _lib, _ops, autograd
from oracle import sigkern as O
from tests.gpu_util import dev

pytestmark = pytest.mark.gpu


def _rel(a, b):
  a = a.detach().double().cpu()
  b = b.detach().double().cpu()
  return float((a - b).abs().max() / b.abs().max().clamp_min(1e-300))


def _grads(Xs, Ys, pairs, code, param, lam, dtype, fused, same):
  """dL/d(x data), dL/d(y data), dL/dparam of sum_p c_p K_p."""
  old = autograd._FUSED_PAIRS
  autograd._FUSED_PAIRS = old if fused else 0
  try:
    xd = Xs.data.to(dtype).clone().requires_grad_(True)
    yd = xd if same else Ys.data.to(dtype).clone().requires_grad_(True)
    X = _ops.SeqSet(xd.detach(), Xs.off, Xs.len, Xs.max_len, Xs.n, Xs.host_len)
    Y = X if same else _ops.SeqSet(yd.detach(), Ys.off, Ys.len, Ys.max_len, Ys.n, Ys.host_len)
    p = torch.tensor(param, dtype=torch.float64, device=dev(), requires_grad=True)
    K = autograd.signature_pde(X, Y, pairs, code, p, lam, True, xd, yd)
    c = torch.linspace(-1.0, 1.5, K.shape[0], device=dev(), dtype=torch.float64)
    (K * c).sum().backward()
    return xd.grad, (None if same else yd.grad), p.grad
  finally:
    autograd._FUSED_PAIRS = old


@pytest.mark.parametrize("kind", ["rbf", "linear"])
@pytest.mark.parametrize("lam", [0, 1])
@pytest.mark.parametrize("d", [1, 3, 4, 5, 8])
def test_fused_vs_two_call_ragged(kind, lam, d):
  rng = np.random.default_rng(500 + 10 * d + lam + (kind == "rbf"))
  xs = [torch.tensor(O.brownian(rng, 1, L, d)[0], device=dev()) for L in (300, 131, 9)]
  ys = [torch.tensor(O.brownian(rng, 1, L, d)[0], device=dev()) for L in (90, 350, 2)]
  Xs, Ys = _ops.SeqSet.ragged(xs), _ops.SeqSet.ragged(ys)
  pairs = _ops.dense_pairs(3, 3, dev())
  code = _lib.STATIC_RBF if kind == "rbf" else _lib.STATIC_LINEAR
  param = 0.9 if kind == "rbf" else 0.6
  gx, gy, gp = _grads(Xs, Ys, pairs, code, param, lam, torch.float32, True, False)
  rx, ry, rp = _grads(Xs, Ys, pairs, code, param, lam, torch.float32, False, False)
  qx, qy, qp = _grads(Xs, Ys, pairs, code, param, lam, torch.float64, False, False)
  # the two float32 paths round differently (sums in the sweep vs MFMA tiles)
  assert _rel(gx, rx) < 2e-5 and _rel(gy, ry) < 2e-5
  assert abs(float(gp) - float(rp)) < 2e-5 * abs(float(rp))
  # both against the float64 adjoint
  assert _rel(gx, qx) < 1e-4 and _rel(gy, qy) < 1e-4
  assert abs(float(gp) - float(qp)) < 1e-4 * abs(float(qp))


@pytest.mark.parametrize("lam", [0, 1])
def test_fused_gram_of_one_set(lam):
  """K(X, X) on the upper triangle with the gradient to X through both
  sides of every pair (the two sides alias), C2-like shapes."""
  rng = np.random.default_rng(600 + lam)
  X = torch.tensor(O.brownian(rng, 5, 140, 4), device=dev())
  Xs = _ops.SeqSet.dense(X)
  pairs = _ops.triu_pairs(5, dev())
  gx, _, gp = _grads(Xs, Xs, pairs, _lib.STATIC_RBF, 1.1, lam, torch.float32, True, True)
  rx, _, rp = _grads(Xs, Xs, pairs, _lib.STATIC_RBF, 1.1, lam, torch.float32, False, True)
  qx, _, qp = _grads(Xs, Xs, pairs, _lib.STATIC_RBF, 1.1, lam, torch.float64, False, True)
  assert _rel(gx, rx) < 2e-5 and abs(float(gp) - float(rp)) < 2e-5 * abs(float(rp))
  assert _rel(gx, qx) < 1e-4 and abs(float(gp) - float(qp)) < 1e-4 * abs(float(qp))


def test_fused_abi_rejects_other_shapes():
  """float64, difference=0, dyadic order 2 and d > 8 are not taken by the
  fused entry (KSIG_EUNSUPPORTED, nothing launched); the workspace query is 0
  for them."""
  lib = _lib.load()
  assert lib.ksig_signature_pde_bwd_fused_workspace(_lib.F64, 4, 50, 50, 1, 1, 10) == 0
  assert lib.ksig_signature_pde_bwd_fused_workspace(_lib.F32, 9, 50, 50, 1, 1, 10) == 0
  assert lib.ksig_signature_pde_bwd_fused_workspace(_lib.F32, 4, 50, 50, 2, 1, 10) == 0
  assert lib.ksig_signature_pde_bwd_fused_workspace(_lib.F32, 4, 50, 50, 1, 0, 10) == 0
  assert lib.ksig_signature_pde_bwd_fused_workspace(_lib.F32, 4, 50, 50, 1, 1, 10) > 0
  null = None
  rc = lib.ksig_signature_pde_bwd_fused(
      _lib.F64, 1, 1, 1, 50, 1, 1, 1, 50, 4, 1, 1, _lib.STATIC_RBF, 1.0, 1, 1, 1, 1, 1,
      null, null, null, null, 0, null)
  assert rc == 1002
