"""GPU parity of ksig_static_grad_contract (the chain rule of a per-pair Gram
adjoint omega through the static kernel) against a float64 numpy evaluation
of the same sums:

  gx[x_i, a] += c_cross * sum_b W[a, b] y_b + c_self * (sum_b W[a, b]) x_a
  gy[y_j, b] += c_cross * sum_a W[a, b] x_a + c_self * (sum_a W[a, b]) y_b

Tolerance: 1e-5 normwise relative error for float32 (fp32 MFMA products and
atomic accumulation), 1e-12 for float64."""

import numpy as np
import pytest
import torch

from ksig_amd import _lib, _ops
from tests.gpu_util import normwise, to_dev

pytestmark = pytest.mark.gpu


def _sets(rng, lens, d, dtype):
  seqs = [rng.standard_normal((L, d)) for L in lens]
  S = _ops.SeqSet.ragged([to_dev(s, dtype) for s in seqs])
  return seqs, S


def _reference(xs, ys, pairs, Ws, cc, cs, want_x, want_y):
  gx = [np.zeros_like(x) for x in xs]
  gy = [np.zeros_like(y) for y in ys]
  for (i, j), W in zip(pairs, Ws):
    if want_x:
      gx[i] += cc * W @ ys[j] + cs * W.sum(1)[:, None] * xs[i]
    if want_y:
      gy[j] += cc * W.T @ xs[i] + cs * W.sum(0)[:, None] * ys[j]
  return np.concatenate(gx), np.concatenate(gy)


def _run(X, Y, pairs, Ws, cc, cs, gx, gy, dtype):
  flat = np.concatenate([W.ravel() for W in Ws]) if Ws else np.zeros(1)
  sizes = np.array([W.size for W in Ws], dtype=np.int64)
  off = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
  om = to_dev(flat, dtype)
  w_off = to_dev(off)
  pr = to_dev(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
  lib = _lib.load()
  _lib.check(lib.ksig_static_grad_contract(
      _lib.dtype_code(dtype), X.data.data_ptr(), X.off.data_ptr(), X.len.data_ptr(), X.max_len,
      Y.data.data_ptr(), Y.off.data_ptr(), Y.len.data_ptr(), Y.max_len, X.d, pr.data_ptr(),
      len(pairs), om.data_ptr(), w_off.data_ptr(), cc, cs,
      None if gx is None else gx.data_ptr(), None if gy is None else gy.data_ptr(),
      _lib.stream_handle(om.device)))
  torch.cuda.synchronize()


@pytest.mark.parametrize("dtype,tol", [(torch.float32, 1e-5), (torch.float64, 1e-12)])
@pytest.mark.parametrize("d", [1, 5, 16, 17, 32, 33, 48, 64, 100])
def test_ragged_pairs(dtype, tol, d):
  """Lengths around the 64-row tiles, unsorted pairs with repeats."""
  rng = np.random.default_rng(900 + d)
  lx = [1, 2, 63, 64, 65, 130, 7, 200]
  ly = [3, 64, 129, 1, 66, 40]
  xs, X = _sets(rng, lx, d, dtype)
  ys, Y = _sets(rng, ly, d, dtype)
  pairs = [(int(rng.integers(len(lx))), int(rng.integers(len(ly)))) for _ in range(30)]
  Ws = [rng.standard_normal((lx[i], ly[j])) for i, j in pairs]
  cc, cs = 0.37, -0.37
  gx = torch.zeros_like(X.data)
  gy = torch.zeros_like(Y.data)
  _run(X, Y, pairs, Ws, cc, cs, gx, gy, dtype)
  rx, ry = _reference(xs, ys, pairs, Ws, cc, cs, True, True)
  assert normwise(gx.cpu().numpy(), rx) < tol
  assert normwise(gy.cpu().numpy(), ry) < tol


@pytest.mark.parametrize("want", ["x", "y"])
def test_one_side_only(want):
  """A null gradient pointer skips that side (linear kernel: c_self = 0)."""
  rng = np.random.default_rng(7)
  lx, ly = [50, 90, 3], [70, 12]
  xs, X = _sets(rng, lx, 8, torch.float32)
  ys, Y = _sets(rng, ly, 8, torch.float32)
  pairs = [(0, 0), (0, 1), (1, 0), (2, 1), (1, 1)]
  Ws = [rng.standard_normal((lx[i], ly[j])) for i, j in pairs]
  g = torch.zeros_like(X.data if want == "x" else Y.data)
  _run(X, Y, pairs, Ws, 0.8, 0.0, g if want == "x" else None, g if want == "y" else None,
       torch.float32)
  rx, ry = _reference(xs, ys, pairs, Ws, 0.8, 0.0, want == "x", want == "y")
  assert normwise(g.cpu().numpy(), rx if want == "x" else ry) < 1e-5


def test_aliased_set_and_gradient():
  """x and y the same set with one shared gradient buffer (a symmetric Gram
  over upper-triangle pairs)."""
  rng = np.random.default_rng(11)
  lens = [20, 64, 100, 33]
  xs, X = _sets(rng, lens, 12, torch.float32)
  pairs = [(i, j) for i in range(4) for j in range(i, 4)]
  Ws = [rng.standard_normal((lens[i], lens[j])) for i, j in pairs]
  g = torch.zeros_like(X.data)
  _run(X, X, pairs, Ws, 0.5, -0.5, g, g, torch.float32)
  rx, ry = _reference(xs, xs, pairs, Ws, 0.5, -0.5, True, True)
  assert normwise(g.cpu().numpy(), rx + ry) < 1e-5


def test_many_pairs_share_x_rows():
  """An x-major dense pair list long enough that each workgroup takes runs of
  several consecutive pairs (x-side accumulation kept across pairs)."""
  rng = np.random.default_rng(12)
  lx = rng.integers(2, 10, 200).tolist()
  ly = rng.integers(2, 10, 100).tolist()
  xs, X = _sets(rng, lx, 3, torch.float32)
  ys, Y = _sets(rng, ly, 3, torch.float32)
  pairs = [(i, j) for i in range(200) for j in range(100)]
  Ws = [rng.standard_normal((lx[i], ly[j])) for i, j in pairs]
  gx = torch.zeros_like(X.data)
  gy = torch.zeros_like(Y.data)
  _run(X, Y, pairs, Ws, 1.1, -1.1, gx, gy, torch.float32)
  rx, ry = _reference(xs, ys, pairs, Ws, 1.1, -1.1, True, True)
  assert normwise(gx.cpu().numpy(), rx) < 1e-5
  assert normwise(gy.cpu().numpy(), ry) < 1e-5
