"""Gradients of the signature-PDE kernel (HIP adjoint sweep) against autograd
through a PyTorch float64 transcription of the Goursat scheme and against
central finite differences (h = 1e-6) of the FP64 oracle, including the
unbiased MMD^2 with respect to the inputs and the RBF bandwidth."""

import math

import numpy as np
import pytest
import torch

import ksig_amd as ksig
from ksig_amd import _lib, _ops, autograd
from oracle import sigkern as O
from tests import torch_reference as TR
from tests.gpu_util import dev

pytestmark = pytest.mark.gpu


def _rel(a, b):
  a = a.detach().double().cpu()
  b = b.detach().double().cpu()
  return float((a - b).abs().max() / b.abs().max().clamp_min(1e-300))


@pytest.mark.parametrize("kind", ["rbf", "linear"])
@pytest.mark.parametrize("lx,ly,lam", [(6, 8, 0), (9, 5, 1), (12, 12, 2)])
def test_pde_gradients_float64_vs_autograd(kind, lx, ly, lam):
  rng = np.random.default_rng(lx * 7 + ly + lam)
  xs = torch.tensor(O.brownian(rng, 2, lx, 2))
  ys = torch.tensor(O.brownian(rng, 2, ly, 2))
  param = 0.6 if kind == "rbf" else 1.2
  c = torch.tensor(rng.standard_normal(4))
  xr = xs.clone().requires_grad_(True)
  yr = ys.clone().requires_grad_(True)
  pr = torch.tensor(param, dtype=torch.float64, requires_grad=True)
  ref = torch.stack([TR.pde(xr[a], yr[b], kind, pr, lam) for a in range(2) for b in range(2)])
  (ref * c).sum().backward()
  X = xs.to(dev()).requires_grad_(True)
  Y = ys.to(dev()).requires_grad_(True)
  p = torch.tensor(param, dtype=torch.float64, device=dev(), requires_grad=True)
  code = _lib.STATIC_RBF if kind == "rbf" else _lib.STATIC_LINEAR
  got = autograd.signature_pde(_ops.SeqSet.dense(X.detach()), _ops.SeqSet.dense(Y.detach()),
                               _ops.dense_pairs(2, 2, dev()), code, p, lam, True,
                               X.reshape(-1, 2), Y.reshape(-1, 2))
  assert _rel(got, ref) < 1e-11
  (got * c.to(dev())).sum().backward()
  assert _rel(X.grad, xr.grad) < 1e-8
  assert _rel(Y.grad, yr.grad) < 1e-8
  assert abs(float(p.grad) - float(pr.grad)) < 1e-8 * max(1.0, abs(float(pr.grad)))


def test_pde_mmd_gradient_vs_finite_differences():
  """Unbiased MMD^2 of the signature-PDE kernel (RBF, dyadic order 1) with
  respect to 8 random input coordinates and the bandwidth, against FP64
  central differences (h = 1e-6) of the oracle."""
  rng = np.random.default_rng(21)
  X = O.brownian(rng, 4, 24, 4)
  Y = O.brownian(rng, 3, 20, 4)
  sig, lam = 1.0, 1

  def mmd(Xn, s):
    kw = dict(normalize=False, param=s, dyadic_order=lam)
    return O.mmd2_unbiased(O.signature_pde_kernel(Xn, **kw), O.signature_pde_kernel(Y, **kw),
                           O.signature_pde_kernel(Xn, Y, **kw))

  Xt = torch.tensor(X, device=dev(), requires_grad=True)
  s = torch.tensor(sig, dtype=torch.float64, device=dev(), requires_grad=True)
  pk = ksig.kernels.SignaturePDEKernel(normalize=False, dyadic_order=lam,
                                       static_kernel=ksig.static.kernels.RBFKernel(s))
  Yt = torch.tensor(Y, device=dev())
  loss = autograd.mmd2_unbiased(pk(Xt, return_on_gpu=True), pk(Yt, return_on_gpu=True),
                                pk(Xt, Yt, return_on_gpu=True))
  assert abs(float(loss) - mmd(X, sig)) < 1e-12
  loss.backward()
  g = Xt.grad.cpu().numpy()
  h = 1e-6
  coords = [tuple(int(v) for v in rng.integers(0, [4, 24, 4])) for _ in range(8)]
  for (a, i, k) in coords:
    Xp, Xm = X.copy(), X.copy()
    Xp[a, i, k] += h
    Xm[a, i, k] -= h
    fd = (mmd(Xp, sig) - mmd(Xm, sig)) / (2 * h)
    assert abs(fd - g[a, i, k]) < 1e-6 * max(1.0, abs(fd)), (a, i, k, fd, g[a, i, k])
  fd = (mmd(X, sig + h) - mmd(X, sig - h)) / (2 * h)
  assert abs(fd - float(s.grad)) < 1e-6 * max(1.0, abs(fd))


def test_pde_gradients_float32_c2_shape_subset():
  """C2 shapes (L = 256, d = 4, dyadic 1) on 3 x 2 pairs: float32 adjoint
  against the float64 adjoint."""
  rng = np.random.default_rng(22)
  X = O.brownian(rng, 3, 256, 4)
  Y = O.brownian(rng, 2, 256, 4)
  grads = {}
  for dt in (torch.float64, torch.float32):
    Xt = torch.tensor(X, device=dev(), dtype=dt, requires_grad=True)
    s = torch.tensor(1.0, dtype=torch.float64, device=dev(), requires_grad=True)
    pk = ksig.kernels.SignaturePDEKernel(normalize=False, dyadic_order=1,
                                         static_kernel=ksig.static.kernels.RBFKernel(s))
    K = pk(Xt, torch.tensor(Y, device=dev(), dtype=dt), return_on_gpu=True)
    K.sum().backward()
    grads[dt] = (Xt.grad.double().cpu(), float(s.grad))
  assert _rel(grads[torch.float32][0], grads[torch.float64][0]) < 1e-4
  assert abs(grads[torch.float32][1] - grads[torch.float64][1]) < 1e-4 * abs(grads[torch.float64][1])


@pytest.mark.parametrize("lam,difference,kind", [
    (0, True, "rbf"), (1, True, "rbf"), (1, True, "linear"), (0, False, "linear"),
    (1, False, "linear")])
def test_pde_gradients_multiband_vs_finite_differences(lam, difference, kind):
  """Float64 adjoint on ragged sets whose grids span two and three bands of
  256 refined rows, in both orientations (lengths 300 and 131 against 90
  and 350), against central differences (h = 1e-6) of the float64 forward:
  the weighted sum of the four pair values with respect to 10 random input
  coordinates of each side and the static-kernel parameter; then the
  float32 adjoint on the same sets against the float64 one.  Without
  differencing the cells hold the static kernel itself, so the paths are
  scaled by 0.02 there (K up to about 340).  The adjoint rebuilds U backwards
  from the saved last row and column, which amplifies rounding with the
  growth of U: at scale 0.1 (K ~ 2e5) even the float64 rebuild is off by
  5e-6 in the gradient, which is a limit of the method, not of these
  tolerances."""
  rng = np.random.default_rng(40 + 2 * lam + int(difference))
  d = 3
  sc = 1.0 if difference else 0.02
  xs = [torch.tensor(sc * O.brownian(rng, 1, L, d)[0], device=dev()) for L in (300, 131)]
  ys = [torch.tensor(sc * O.brownian(rng, 1, L, d)[0], device=dev()) for L in (90, 350)]
  Xs, Ys = _ops.SeqSet.ragged(xs), _ops.SeqSet.ragged(ys)
  pairs = _ops.dense_pairs(2, 2, dev())
  code = _lib.STATIC_RBF if kind == "rbf" else _lib.STATIC_LINEAR
  param = 0.8 if kind == "rbf" else 0.7
  c = torch.tensor(rng.standard_normal(4), device=dev())
  xd = Xs.data.clone().requires_grad_(True)
  yd = Ys.data.clone().requires_grad_(True)
  p = torch.tensor(param, dtype=torch.float64, device=dev(), requires_grad=True)
  K = autograd.signature_pde(Xs, Ys, pairs, code, p, lam, difference, xd, yd)
  (K * c).sum().backward()

  def value(xdata, ydata, pv):
    Xn = _ops.SeqSet(xdata, Xs.off, Xs.len, Xs.max_len, Xs.n, Xs.host_len)
    Yn = _ops.SeqSet(ydata, Ys.off, Ys.len, Ys.max_len, Ys.n, Ys.host_len)
    return float((_ops.signature_pde(Xn, Yn, pairs, difference, code, pv, lam) * c).sum())

  h = 1e-6
  for side, data, grad in (("x", Xs.data, xd.grad), ("y", Ys.data, yd.grad)):
    rows = data.shape[0]
    for _ in range(10):
      r, k = int(rng.integers(0, rows)), int(rng.integers(0, d))
      dp, dm = data.clone(), data.clone()
      dp[r, k] += h
      dm[r, k] -= h
      if side == "x":
        fd = (value(dp, Ys.data, param) - value(dm, Ys.data, param)) / (2 * h)
      else:
        fd = (value(Xs.data, dp, param) - value(Xs.data, dm, param)) / (2 * h)
      g = float(grad[r, k])
      assert abs(fd - g) < 1e-6 * max(1.0, abs(fd)), (side, r, k, fd, g)
  fd = (value(Xs.data, Ys.data, param + h) - value(Xs.data, Ys.data, param - h)) / (2 * h)
  assert abs(fd - float(p.grad)) < 1e-6 * max(1.0, abs(fd)), (fd, float(p.grad))

  X32 = _ops.SeqSet(Xs.data.float(), Xs.off, Xs.len, Xs.max_len, Xs.n, Xs.host_len)
  Y32 = _ops.SeqSet(Ys.data.float(), Ys.off, Ys.len, Ys.max_len, Ys.n, Ys.host_len)
  xd32 = X32.data.clone().requires_grad_(True)
  yd32 = Y32.data.clone().requires_grad_(True)
  p32 = torch.tensor(param, dtype=torch.float64, device=dev(), requires_grad=True)
  K32 = autograd.signature_pde(X32, Y32, pairs, code, p32, lam, difference, xd32, yd32)
  (K32 * c).sum().backward()
  assert _rel(K32, K) < 1e-5
  assert _rel(xd32.grad, xd.grad) < 1e-4
  assert _rel(yd32.grad, yd.grad) < 1e-4
  assert abs(float(p32.grad) - float(p.grad)) < 1e-4 * max(1.0, abs(float(p.grad)))
