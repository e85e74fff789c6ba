"""GPU parity of ksig_signature_pde_grad_contract (the signature-PDE adjoint's
chain rule through ΔΔ and the static kernel, fused with the contraction)
against a float64 numpy evaluation of the same sums:

  dL/dG = ΔΔ^T(dL/dM) (difference) or dL/dM,  W = dL/dG * G (RBF) or dL/dG,
  dL/dx_a = c_cross sum_b W y_b + c_self (sum_b W) x_a, likewise for y,
  dL/dsigma = sum W |x - y|^2 / sigma^3 (RBF),  dL/dscale = sum dL/dG <x, y>.

dL/dM is random, laid out as the adjoint writes it (rows = the longer
sequence of the pair).  Lengths cross the 64-point tiles, both orientations
occur, and one sequence is a single point (an empty dL/dM with difference).
Tolerance: 1e-5 normwise relative error (float32), 1e-12 (float64)."""

import numpy as np
import pytest
import torch

from ksig_amd import _lib, _ops
from tests.gpu_util import normwise, to_dev

pytestmark = pytest.mark.gpu


def _reference(xs, ys, pairs, gms, difference, kind, param):
  gx = [np.zeros_like(x) for x in xs]
  gy = [np.zeros_like(y) for y in ys]
  gp = 0.0
  sh = 1 if difference else 0
  for (i, j), gm in zip(pairs, gms):
    x, y = xs[i], ys[j]
    A, B = len(x), len(y)
    An, Bn = A - sh, B - sh
    if An > 0 and Bn > 0:
      nr, nc = max(An, Bn), min(An, Bn)
      M = gm.reshape(nr, nc)
      gM = M.T if B > A else M
    else:
      gM = np.zeros((max(An, 0), max(Bn, 0)))
    if difference:
      P = np.pad(gM, 1) if gM.size else np.zeros((A + 1, B + 1))
      gG = P[:-1, :-1] - P[:-1, 1:] - P[1:, :-1] + P[1:, 1:]
    else:
      gG = gM
    dot = x @ y.T
    if kind == "rbf":
      s = 1.0 / max(2.0 * param * param, 1e-12)
      d2 = (x * x).sum(1)[:, None] + (y * y).sum(1)[None, :] - 2 * dot
      W = gG * np.exp(-s * d2)
      gx[i] += 2 * s * (W @ y - W.sum(1)[:, None] * x)
      gy[j] += 2 * s * (W.T @ x - W.sum(0)[:, None] * y)
      gp += (W * d2).sum() / param ** 3
    else:
      W = gG
      gx[i] += param * W @ y
      gy[j] += param * W.T @ x
      gp += (W * dot).sum()
  return np.concatenate(gx), np.concatenate(gy), gp


@pytest.mark.parametrize("dtype,tol", [(torch.float32, 1e-5), (torch.float64, 1e-12)])
@pytest.mark.parametrize("kind,param", [("rbf", 1.3), ("linear", 0.7)])
@pytest.mark.parametrize("difference", [True, False])
@pytest.mark.parametrize("d", [4, 21, 40, 70])
def test_pde_grad_contract(dtype, tol, kind, param, difference, d):
  rng = np.random.default_rng(700 + d + 2 * int(difference))
  lx = [1, 40, 65, 130, 7]
  ly = [66, 3, 150, 64]
  xs = [rng.standard_normal((L, d)) * 0.3 for L in lx]
  ys = [rng.standard_normal((L, d)) * 0.3 for L in ly]
  X = _ops.SeqSet.ragged([to_dev(x, dtype) for x in xs])
  Y = _ops.SeqSet.ragged([to_dev(y, dtype) for y in ys])
  pairs = [(i, j) for i in range(len(lx)) for j in range(len(ly))]
  sh = 1 if difference else 0
  gms = [rng.standard_normal(max(lx[i] - sh, 0) * max(ly[j] - sh, 0)) for i, j in pairs]
  if dtype == torch.float32:  # the reference sees the float32-rounded inputs
    xs = [x.astype(np.float32).astype(np.float64) for x in xs]
    ys = [y.astype(np.float32).astype(np.float64) for y in ys]
    gms = [g.astype(np.float32).astype(np.float64) for g in gms]
  sizes = np.array([g.size for g in gms], dtype=np.int64)
  m_off = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
  gm_dev = to_dev(np.concatenate(gms + [np.zeros(1)]), dtype)
  gx = torch.zeros_like(X.data)
  gy = torch.zeros_like(Y.data)
  gp = torch.zeros((), device=gx.device, dtype=torch.float64)
  pr = to_dev(np.asarray(pairs, dtype=np.int32))
  lib = _lib.load()
  code = _lib.STATIC_RBF if kind == "rbf" else _lib.STATIC_LINEAR
  _lib.check(lib.ksig_signature_pde_grad_contract(
      _lib.dtype_code(dtype), X.data.data_ptr(), X.off.data_ptr(), X.len.data_ptr(), X.max_len,
      Y.data.data_ptr(), Y.off.data_ptr(), Y.len.data_ptr(), Y.max_len, d, pr.data_ptr(),
      len(pairs), code, param, int(difference), gm_dev.data_ptr(), to_dev(m_off).data_ptr(),
      gx.data_ptr(), gy.data_ptr(), gp.data_ptr(), _lib.stream_handle(gx.device)))
  torch.cuda.synchronize()
  rx, ry, rp = _reference(xs, ys, pairs, gms, difference, kind, param)
  assert normwise(gx.cpu().numpy(), rx) < tol
  assert normwise(gy.cpu().numpy(), ry) < tol
  assert abs(float(gp) - rp) < tol * max(1.0, abs(rp)) * 10
