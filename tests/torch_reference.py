"""PyTorch float64 transcription of the reference dataflow, used only as an
autograd reference for gradient tests (ksig/static/kernels.py:248-260,
ksig/utils.py:63-90,138-157, ksig/algorithms.py:47-85,344-399,
ksig/kernels.py:180-214)."""

import torch


def gram(x, y, kind, param):
  if kind == "rbf":
    D2 = (x * x).sum(-1)[:, None] + (y * y).sum(-1)[None, :] - 2 * x @ y.T
    den = 2 * param**2
    den = torch.clamp(den, min=1e-12) if isinstance(den, torch.Tensor) else max(den, 1e-12)
    return torch.exp(-D2 / den)
  return param * (x @ y.T)


def excl_cumsum2(R):
  E = torch.cumsum(torch.cumsum(R[:-1, :-1], 0), 1)
  return torch.nn.functional.pad(E, (1, 0, 1, 0))


def levels(x, y, n_levels, kind="rbf", param=1.0):
  G = gram(x, y, kind, param)
  M = torch.diff(torch.diff(G, dim=0), dim=1)
  K = [torch.ones((), dtype=x.dtype), M.sum()]
  R = M
  for _ in range(1, n_levels):
    R = M * excl_cumsum2(R)
    K.append(R.sum())
  return torch.stack(K)


def pde(x, y, kind="rbf", param=1.0, dyadic_order=0):
  G = gram(x, y, kind, param)
  M = torch.diff(torch.diff(G, dim=0), dim=1)
  f = 1 << dyadic_order
  M = torch.repeat_interleave(torch.repeat_interleave(M, f, 0), f, 1) / (f * f)
  lx, ly = M.shape
  prev = [torch.ones((), dtype=x.dtype)] * (ly + 1)
  for i in range(lx):
    cur = [torch.ones((), dtype=x.dtype)]
    for j in range(ly):
      m = M[i, j]
      cur.append((prev[j + 1] + cur[j]) * (1 + m / 2 + m * m / 12) - prev[j] * (1 - m * m / 12))
    prev = cur
  return prev[ly]


def gram_matrix_levels(X, Y, n_levels, kind="rbf", param=1.0):
  return torch.stack([torch.stack([levels(X[a], Y[b], n_levels, kind, param)
                                   for b in range(Y.shape[0])], -1)
                      for a in range(X.shape[0])], -2)  # [n+1, nX, nY]


def signature_kernel(X, Y, n_levels, kind="rbf", param=1.0, normalize=True):
  K = gram_matrix_levels(X, Y, n_levels, kind, param)
  if not normalize:
    return K.sum(0)
  kx = torch.stack([levels(X[a], X[a], n_levels, kind, param) for a in range(X.shape[0])], -1)
  ky = torch.stack([levels(Y[b], Y[b], n_levels, kind, param) for b in range(Y.shape[0])], -1)
  sx = torch.clamp(torch.sqrt(torch.clamp(kx, min=1e-12)), min=1e-12)
  sy = torch.clamp(torch.sqrt(torch.clamp(ky, min=1e-12)), min=1e-12)
  return (K / (sx[:, :, None] * sy[:, None, :])).mean(0)
