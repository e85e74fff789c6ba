"""Gradients of the DP signature kernel (HIP adjoint sweep) against autograd
through a PyTorch float64 transcription of the reference dataflow and against
central finite differences (h = 1e-6) of the FP64 oracle.  Tolerances:
float64 1e-8 relative (normwise), float32 2e-4."""

import math

import numpy as np
import pytest
import torch

import ksig_amd as ksig
from ksig_amd import _lib, _ops, autograd
from oracle import sigkern as O
from tests import torch_reference as TR
from tests.gpu_util import dev, to_dev

pytestmark = pytest.mark.gpu


def _rel(a, b):
  a = a.detach().double().cpu()
  b = b.detach().double().cpu()
  return float((a - b).abs().max() / b.abs().max().clamp_min(1e-300))


@pytest.mark.parametrize("kind", ["rbf", "linear"])
@pytest.mark.parametrize("lx,ly,n", [(7, 9, 3), (70, 66, 5), (130, 40, 4), (40, 130, 6)])
def test_level_gradients_float64(kind, lx, ly, n):
  rng = np.random.default_rng(lx + ly + n)
  xs = torch.tensor(O.brownian(rng, 2, lx, 3))
  ys = torch.tensor(O.brownian(rng, 3, ly, 3))
  param = 0.8 if kind == "rbf" else 1.3
  c = torch.tensor(rng.standard_normal((2 * 3, n + 1)))
  # reference: autograd through the torch transcription
  xr = xs.clone().requires_grad_(True)
  yr = ys.clone().requires_grad_(True)
  pr = torch.tensor(param, dtype=torch.float64, requires_grad=True)
  lev = torch.stack([TR.levels(xr[a], yr[b], n, kind, pr) for a in range(2) for b in range(3)])
  (lev * c).sum().backward()
  # HIP path
  X = xs.to(dev()).requires_grad_(True)
  Y = ys.to(dev()).requires_grad_(True)
  p = torch.tensor(param, dtype=torch.float64, device=dev(), requires_grad=True)
  Xs, Ys = _ops.SeqSet.dense(X.detach()), _ops.SeqSet.dense(Y.detach())
  pairs = _ops.dense_pairs(2, 3, dev())
  code = _lib.STATIC_RBF if kind == "rbf" else _lib.STATIC_LINEAR
  got = autograd.signature_levels(Xs, Ys, pairs, n, code, p, X.reshape(-1, 3), Y.reshape(-1, 3))
  assert _rel(got, lev) < 1e-11
  (got * c.to(dev())).sum().backward()
  assert _rel(X.grad, xr.grad) < 1e-8
  assert _rel(Y.grad, yr.grad) < 1e-8
  assert abs(float(p.grad) - float(pr.grad)) < 1e-8 * max(1.0, abs(float(pr.grad)))


def test_gradient_vs_finite_differences_float64():
  rng = np.random.default_rng(11)
  x = O.brownian(rng, 1, 40, 4)[0]
  y = O.brownian(rng, 1, 35, 4)[0]
  n, sig = 5, 2.0
  c = rng.standard_normal(n + 1)
  f = lambda xx, s=sig: float(np.dot(c, O.pair_levels(xx, y, n, 1, True, "rbf", s)))
  X = torch.tensor(x[None], device=dev(), requires_grad=True)
  Y = torch.tensor(y[None], device=dev())
  p = torch.tensor(sig, dtype=torch.float64, device=dev(), requires_grad=True)
  lev = autograd.signature_levels(_ops.SeqSet.dense(X.detach()), _ops.SeqSet.dense(Y),
                                  _ops.dense_pairs(1, 1, dev()), n, _lib.STATIC_RBF, p,
                                  X.reshape(-1, 4))
  (lev[0] * torch.tensor(c, device=dev())).sum().backward()
  g = X.grad[0].cpu().numpy()
  h = 1e-6
  for (i, k) in [(0, 0), (5, 1), (17, 3), (39, 2), (20, 0), (33, 1), (1, 2), (38, 3)]:
    xp, xm = x.copy(), x.copy()
    xp[i, k] += h
    xm[i, k] -= h
    fd = (f(xp) - f(xm)) / (2 * h)
    assert abs(fd - g[i, k]) < 1e-6 * max(1.0, abs(fd)), (i, k, fd, g[i, k])
  fd = (f(x, sig + h) - f(x, sig - h)) / (2 * h)
  assert abs(fd - float(p.grad)) < 1e-6 * max(1.0, abs(fd))


def test_normalised_kernel_gradient_through_class_api():
  rng = np.random.default_rng(12)
  xs = torch.tensor(O.brownian(rng, 3, 20, 2))
  ys = torch.tensor(O.brownian(rng, 2, 24, 2))
  W = torch.tensor(rng.standard_normal((3, 2)))
  xr = xs.clone().requires_grad_(True)
  (TR.signature_kernel(xr, ys, 4, "rbf", 0.7) * W).sum().backward()
  X = xs.to(dev()).requires_grad_(True)
  sk = ksig.kernels.SignatureKernel(4, static_kernel=ksig.static.kernels.RBFKernel(0.7))
  K = sk(X, ys.to(dev()), return_on_gpu=True)
  (K * W.to(dev())).sum().backward()
  assert _rel(X.grad, xr.grad) < 1e-8
  # symmetric Gram path: dK(X,X)/dX
  xr2 = xs.clone().requires_grad_(True)
  (TR.signature_kernel(xr2, xr2, 4, "rbf", 0.7) * W[:, :1]).sum().backward()
  X2 = xs.to(dev()).requires_grad_(True)
  (sk(X2, return_on_gpu=True) * W[:, :1].to(dev())).sum().backward()
  assert _rel(X2.grad, xr2.grad) < 1e-8


def test_mmd_gradients_float32_vs_float64():
  """C4-style MMD^2 (unnormalised, 6 levels, sigma = sqrt(d)) on a subset:
  float32 HIP gradients against the float64 HIP path and the torch reference."""
  rng = np.random.default_rng(13)
  d, L = 8, 30
  gen = rng.standard_normal((6, L, d)) * 0.5 / math.sqrt(L)
  tgt = O.brownian(rng, 5, L, d)
  sig = math.sqrt(d)
  grads = {}
  for dt in (torch.float64, torch.float32):
    inc = torch.tensor(gen, device=dev(), dtype=dt)
    scale = torch.full((d,), 0.5, device=dev(), dtype=dt, requires_grad=True)
    X = torch.cumsum(inc / 0.5 * scale, 1)
    Y = torch.tensor(tgt, device=dev(), dtype=dt)
    sk = ksig.kernels.SignatureKernel(6, normalize=False,
                                      static_kernel=ksig.static.kernels.RBFKernel(sig))
    loss = autograd.mmd2_unbiased(sk(X, return_on_gpu=True), sk(Y, return_on_gpu=True),
                                  sk(X, Y, return_on_gpu=True))
    loss.backward()
    grads[dt] = (float(loss), scale.grad.double().cpu())
  # torch reference on CPU float64
  inc = torch.tensor(gen)
  scale = torch.full((d,), 0.5, dtype=torch.float64, requires_grad=True)
  X = torch.cumsum(inc / 0.5 * scale, 1)
  Y = torch.tensor(tgt)
  Kxx = TR.signature_kernel(X, X, 6, "rbf", sig, normalize=False)
  Kyy = TR.signature_kernel(Y, Y, 6, "rbf", sig, normalize=False)
  Kxy = TR.signature_kernel(X, Y, 6, "rbf", sig, normalize=False)
  ref = autograd.mmd2_unbiased(Kxx, Kyy, Kxy)
  ref.backward()
  assert abs(grads[torch.float64][0] - float(ref)) < 1e-10
  assert _rel(grads[torch.float64][1], scale.grad) < 1e-8
  assert _rel(grads[torch.float32][1], scale.grad) < 2e-4


@pytest.mark.parametrize("d", [8, 13, 16, 32])
def test_training_forward_boundary_state_layout(d):
  """The float32 training forward (row stream for d in {8, 16, 32}) writes the
  same adjoint boundary state -- row-end prefixes then final column sums, in
  the rows-are-the-longer-sequence orientation -- as the float64 per-pair
  kernel, for uneven lengths (both orientations, several bands per pair)."""
  import ctypes
  from ksig_amd import _ops
  rng = np.random.default_rng(70 + d)
  lens_x = [3, 70, 140, 33, 300]
  lens_y = [90, 5, 140, 261]
  xs = [O.brownian(rng, 1, L, d)[0] for L in lens_x]
  ys = [O.brownian(rng, 1, L, d)[0] for L in lens_y]
  n_levels = 5
  pr = np.array([[a, b] for a in range(len(lens_x)) for b in range(len(lens_y))], np.int32)
  la = np.array([lens_x[a] for a, _ in pr])
  lb = np.array([lens_y[b] for _, b in pr])
  per = (n_levels - 1) * np.maximum(la + lb - 2, 0)
  off = np.concatenate([[0], np.cumsum(per)[:-1]]).astype(np.int64)
  lib = _lib.load()
  saves, outs = {}, {}
  # float64 per-pair kernel; float32 without a workspace (row stream) and, for
  # 8 < d <= 32, with one (the MFMA-Gram forward)
  for key in ("f64", "f32", "f32_ws"):
    dt = torch.float64 if key == "f64" else torch.float32
    X = _ops.SeqSet.ragged([torch.tensor(s, device=dev(), dtype=dt) for s in xs])
    Y = _ops.SeqSet.ragged([torch.tensor(s, device=dev(), dtype=dt) for s in ys])
    pairs = to_dev(pr)
    save = torch.full((int(per.sum()),), float("nan"), device=dev(), dtype=dt)
    so = to_dev(off)
    out = torch.empty((len(pr), n_levels + 1), device=dev(), dtype=torch.float64)
    wsb = (lib.ksig_signature_levels_fwd_save_workspace(
        _lib.dtype_code(dt), d, n_levels, X.max_len, Y.max_len, len(pr)) if key == "f32_ws" else 0)
    ws = torch.empty(max(wsb // 4, 1), device=dev(), dtype=torch.float32)
    _lib.check(lib.ksig_signature_levels_fwd_save(
        _lib.dtype_code(dt), X.data.data_ptr(), X.off.data_ptr(), X.len.data_ptr(), X.max_len,
        Y.data.data_ptr(), Y.off.data_ptr(), Y.len.data_ptr(), Y.max_len, d,
        pairs.data_ptr(), len(pr), _lib.STATIC_RBF, math.sqrt(d), n_levels, out.data_ptr(),
        save.data_ptr(), so.data_ptr(), ws.data_ptr() if wsb else None, wsb,
        _lib.stream_handle(dev())))
    saves[key] = save.double().cpu().numpy()
    outs[key] = out.cpu().numpy()
  for key in ("f32", "f32_ws"):
    assert not np.isnan(saves[key]).any()
    for p in range(len(pr)):
      a, b = off[p], off[p] + per[p]
      ref = saves["f64"][a:b]
      got = saves[key][a:b]
      if b > a:
        assert np.max(np.abs(got - ref)) <= 2e-5 * max(np.max(np.abs(ref)), 1e-30), (key, p)
    assert np.max(np.abs(outs[key] - outs["f64"])) <= 1e-5 * np.max(np.abs(outs["f64"]))
