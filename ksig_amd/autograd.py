"""Differentiable signature-kernel levels and MMD^2 (no reference counterpart:
ksig is forward-only).

``signature_levels`` returns the per-pair level values K_0..K_n of the
first-order DP signature kernel (RBF or linear static kernel, difference=True)
as a differentiable function of the sequences and of the bandwidth / scale.
The forward is the fused HIP sweep (it also leaves the O((Lx+Ly)·n) boundary
state per pair); the backward is the HIP adjoint sweep, which writes
omega = dL/dG * dG/dz per static-kernel entry, followed by the HIP contraction
of omega with the inputs (ksig_static_grad_contract, MFMA tiles per pair):

  RBF,    G = exp(-s |x-y|^2), s = 1/max(2 sigma^2, 1e-12):
    dL/dx_a = -2s (x_a sum_b omega_ab - sum_b omega_ab y_b)
    dL/dy_b =  2s (sum_a omega_ab x_a - y_b sum_a omega_ab)
    dL/dsigma = sum omega_ab |x_a - y_b|^2 / sigma^3
  linear, G = c <x, y>:
    dL/dx_a = c sum_b omega_ab y_b,  dL/dy_b = c sum_a omega_ab x_a,
    dL/dc = sum omega_ab G_ab / c
"""

from __future__ import annotations

import math
from typing import Optional

import torch

from . import _lib, _ops

LOG2E = 1.4426950408889634
# omega scratch per chunk of pairs: 8 GiB (fewer, larger adjoint launches
# fill the chip better; the 16,384-pair chunks of 1 GiB left a partial round
# of waves per launch)
_OMEGA_BYTES = 1 << 33


def _as_param(p, device) -> torch.Tensor:
  if isinstance(p, torch.Tensor):
    return p.to(device=device, dtype=torch.float64).reshape(())
  return torch.tensor(float(p), device=device, dtype=torch.float64)


def _uniform_lengths(Xs: _ops.SeqSet, Ys: _ops.SeqSet):
  """(len_x, len_y) when every sequence of each set has its set's maximum
  length (known on the host: no device synchronisation), else None."""
  if Xs.uniform and Ys.uniform:
    return Xs.max_len, Ys.max_len
  return None


class _SigLevels(torch.autograd.Function):
  """For sets of uniform length the forward and the adjoint size every buffer
  on the host and never synchronise with the device, so a whole training step
  can be captured in a graph."""

  @staticmethod
  def forward(ctx, xdata, ydata, param, pval, xs_meta, ys_meta, pairs, n_levels, kind):
    lib = _lib.load()
    Xs = _ops.SeqSet(xdata.detach().contiguous(), *xs_meta)
    Ys = _ops.SeqSet(ydata.detach().contiguous(), *ys_meta)
    if pval is None:
      pval = float(param.detach())
    n_pairs = pairs.shape[0]
    dev = xdata.device
    out = torch.empty((n_pairs, n_levels + 1), device=dev, dtype=torch.float64)
    uni = _uniform_lengths(Xs, Ys)
    if uni is not None:
      A, B = uni
      per_pair = (n_levels - 1) * max(A + B - 2, 0)
      save_off = torch.arange(n_pairs, device=dev, dtype=torch.int64) * per_pair
      total = n_pairs * per_pair
      la = lb = torch.empty(0, device=dev, dtype=torch.int64)
    else:
      la = Xs.len.to(torch.int64)[pairs[:, 0].long()]
      lb = Ys.len.to(torch.int64)[pairs[:, 1].long()]
      per = (n_levels - 1) * torch.clamp(la + lb - 2, min=0)
      save_off = torch.zeros(n_pairs, device=dev, dtype=torch.int64)
      if n_pairs > 1:
        save_off[1:] = torch.cumsum(per, 0)[:-1]
      total = int(per.sum().item()) if n_pairs else 0
    save = torch.empty(max(total, 1), device=dev, dtype=Xs.data.dtype)
    if n_pairs:
      ws_bytes = lib.ksig_signature_levels_fwd_save_workspace(
          _lib.dtype_code(Xs.data.dtype), Xs.d, n_levels, Xs.max_len, Ys.max_len, n_pairs)
      ws = torch.empty(max(ws_bytes // 4, 1), device=dev, dtype=torch.float32)
      _lib.check(lib.ksig_signature_levels_fwd_save(
          _lib.dtype_code(Xs.data.dtype), Xs.data.data_ptr(), Xs.off.data_ptr(),
          Xs.len.data_ptr(), Xs.max_len, Ys.data.data_ptr(), Ys.off.data_ptr(),
          Ys.len.data_ptr(), Ys.max_len, Xs.d, pairs.data_ptr(), n_pairs, kind, pval,
          n_levels, out.data_ptr(), save.data_ptr(), save_off.data_ptr(),
          ws.data_ptr(), ws_bytes, _lib.stream_handle(dev)))
    ctx.save_for_backward(xdata, ydata, pairs, save, save_off, la, lb)
    ctx.meta = (xs_meta, ys_meta, n_levels, kind, pval, uni)
    return out

  @staticmethod
  def backward(ctx, gout):
    lib = _lib.load()
    xdata, ydata, pairs, save, save_off, la, lb = ctx.saved_tensors
    xs_meta, ys_meta, n_levels, kind, pval, uni = ctx.meta
    Xs = _ops.SeqSet(xdata.detach().contiguous(), *xs_meta)
    Ys = _ops.SeqSet(ydata.detach().contiguous(), *ys_meta)
    dev = xdata.device
    dt = Xs.data.dtype
    gout = gout.to(torch.float64).contiguous()
    n_pairs = pairs.shape[0]
    gx = torch.zeros_like(Xs.data) if ctx.needs_input_grad[0] else None
    gy = torch.zeros_like(Ys.data) if ctx.needs_input_grad[1] else None
    gparam_total = torch.zeros((), device=dev, dtype=torch.float64)
    if n_pairs == 0:
      return gx, gy, gparam_total, None, None, None, None, None, None
    esz = Xs.data.element_size()
    s_rbf = 1.0 / max(2.0 * pval * pval, 1e-12)
    # chunk pairs so that omega stays within the scratch budget
    if uni is not None:
      A, B = uni
      step = max(1, _OMEGA_BYTES // (esz * max(A * B, 1)))
      bounds = list(range(0, n_pairs, step)) + [n_pairs]
    else:
      sizes = (la * lb)
      budget = max(_OMEGA_BYTES // esz, int(sizes.max().item()))
      chunk_id = ((torch.cumsum(sizes, 0) - sizes) // budget).cpu()
      change = torch.nonzero(chunk_id[1:] != chunk_id[:-1]).flatten() + 1
      bounds = [0] + change.tolist() + [n_pairs]
    for c0, c1 in zip(bounds[:-1], bounds[1:]):
      pc = pairs[c0:c1].contiguous()
      if uni is not None:
        w_off = torch.arange(c1 - c0, device=dev, dtype=torch.int64) * (A * B)
        omega = torch.empty(max((c1 - c0) * A * B, 1), device=dev, dtype=dt)
      else:
        sz = sizes[c0:c1]
        w_off = torch.zeros(c1 - c0, device=dev, dtype=torch.int64)
        if c1 - c0 > 1:
          w_off[1:] = torch.cumsum(sz, 0)[:-1]
        omega = torch.empty(int(sz.sum().item()), device=dev, dtype=dt)
      gp = torch.empty(c1 - c0, device=dev, dtype=torch.float64)
      ws_bytes = lib.ksig_signature_levels_bwd_workspace(
          _lib.dtype_code(dt), Xs.d, n_levels, Xs.max_len, Ys.max_len, c1 - c0)
      ws = torch.empty(max(ws_bytes // 4, 1), device=dev, dtype=torch.float32)
      _lib.check(lib.ksig_signature_levels_bwd(
          _lib.dtype_code(dt), Xs.data.data_ptr(), Xs.off.data_ptr(), Xs.len.data_ptr(),
          Xs.max_len, Ys.data.data_ptr(), Ys.off.data_ptr(), Ys.len.data_ptr(),
          Ys.max_len, Xs.d, pc.data_ptr(), c1 - c0, kind, pval, n_levels,
          gout[c0:c1].data_ptr(), save.data_ptr(), save_off[c0:c1].data_ptr(),
          omega.data_ptr(), w_off.data_ptr(),
          gp.data_ptr() if ctx.needs_input_grad[2] else None, ws.data_ptr(), ws_bytes,
          _lib.stream_handle(dev)))
      if ctx.needs_input_grad[2]:
        gparam_total = gparam_total + gp.sum()
      _contract_hip(omega, w_off, pc, Xs, Ys, kind, pval, s_rbf, gx, gy)
    if kind == _lib.STATIC_RBF:
      den = 2.0 * pval * pval
      if den > 1e-12:
        q = LOG2E / den
        gparam_total = gparam_total / (q * pval ** 3)
      else:
        gparam_total = gparam_total * 0.0
    else:
      gparam_total = gparam_total / pval
    if not ctx.needs_input_grad[2]:
      gparam_total = None
    return gx, gy, gparam_total, None, None, None, None, None, None


def _contract_hip(omega, w_off, pairs, Xs, Ys, kind, pval, s_rbf, gx, gy):
  """gx / gy += the chain rule of omega (per pair [len_x, len_y] at w_off)
  through the static kernel, on the device (ksig_static_grad_contract)."""
  if gx is None and gy is None:
    return
  if kind == _lib.STATIC_RBF:
    c_cross, c_self = 2.0 * s_rbf, -2.0 * s_rbf
  else:
    c_cross, c_self = float(pval), 0.0
  lib = _lib.load()
  _lib.check(lib.ksig_static_grad_contract(
      _lib.dtype_code(Xs.data.dtype), Xs.data.data_ptr(), Xs.off.data_ptr(),
      Xs.len.data_ptr(), Xs.max_len, Ys.data.data_ptr(), Ys.off.data_ptr(), Ys.len.data_ptr(),
      Ys.max_len, Xs.d, pairs.data_ptr(), pairs.shape[0], omega.data_ptr(), w_off.data_ptr(),
      c_cross, c_self, None if gx is None else gx.data_ptr(),
      None if gy is None else gy.data_ptr(), _lib.stream_handle(omega.device)))


def signature_levels(X: _ops.SeqSet, Y: _ops.SeqSet, pairs: torch.Tensor, n_levels: int,
                     kind: int, param, xdata: Optional[torch.Tensor] = None,
                     ydata: Optional[torch.Tensor] = None) -> torch.Tensor:
  """Differentiable level values [n_pairs, n_levels+1] (float64) of the
  first-order DP kernel with difference.  ``xdata`` / ``ydata`` are the packed
  [rows, d] tensors the gradient flows to (default: X.data / Y.data);
  ``param`` may be a tensor requiring grad."""
  xdata = X.data if xdata is None else xdata
  ydata = Y.data if ydata is None else ydata
  if isinstance(param, torch.Tensor):
    p, pval = _as_param(param, xdata.device), None
  else:  # a fixed parameter: its value stays on the host
    p, pval = None, float(param)
  xs_meta = (X.off, X.len, X.max_len, X.n, X.host_len)
  ys_meta = (Y.off, Y.len, Y.max_len, Y.n, Y.host_len)
  return _SigLevels.apply(xdata, ydata, p, pval, xs_meta, ys_meta, pairs, n_levels, kind)


def mmd2_unbiased(Kxx: torch.Tensor, Kyy: torch.Tensor, Kxy: torch.Tensor) -> torch.Tensor:
  """Unbiased MMD^2: mean_{i!=j} Kxx + mean_{i!=j} Kyy - 2 mean Kxy."""
  n, m = Kxx.shape[0], Kyy.shape[0]
  sxx = (Kxx.sum() - torch.diagonal(Kxx).sum()) / (n * (n - 1))
  syy = (Kyy.sum() - torch.diagonal(Kyy).sum()) / (m * (m - 1))
  return sxx + syy - 2.0 * Kxy.mean()


# -----------------------------------------------------------------------------
# Signature-PDE kernel
# -----------------------------------------------------------------------------

_GM_BYTES = 1 << 33  # dL/dM scratch per chunk of pairs (8 GiB, as omega)
# fused adjoint: at most this many pairs and this much FIFO workspace per
# launch (0 pairs: always the two-call path, as a reference for tests)
_FUSED_PAIRS = 1 << 30
_FUSED_WS_BYTES = 1 << 32


class _SigPDE(torch.autograd.Function):
  """The forward also leaves U of the last refined row and column of every
  pair (O(Lx + Ly) float64 values); the adjoint kernel rebuilds the rest of U
  from them in its reverse sweep and writes dL/dM per original cell, which
  ksig_signature_pde_grad_contract turns into input and parameter gradients
  (ΔΔ-transpose and static-kernel chain rule, fused with the contraction)."""

  @staticmethod
  def forward(ctx, xdata, ydata, param, xs_meta, ys_meta, pairs, kind, dyadic, difference):
    lib = _lib.load()
    Xs = _ops.SeqSet(xdata.detach().contiguous(), *xs_meta)
    Ys = _ops.SeqSet(ydata.detach().contiguous(), *ys_meta)
    pval = float(param.detach())
    n_pairs = pairs.shape[0]
    dev = xdata.device
    out = torch.empty((n_pairs,), device=dev, dtype=torch.float64)
    sh = 1 if difference else 0
    la = Xs.len.to(torch.int64)[pairs[:, 0].long()]
    lb = Ys.len.to(torch.int64)[pairs[:, 1].long()]
    nrr = torch.clamp(torch.maximum(la, lb) - sh, min=0) << dyadic
    ncr = torch.clamp(torch.minimum(la, lb) - sh, min=0) << dyadic
    esz = nrr + ncr
    e_off = torch.zeros(n_pairs, device=dev, dtype=torch.int64)
    if n_pairs > 1:
      e_off[1:] = torch.cumsum(esz, 0)[:-1]
    total = int(esz.sum().item()) if n_pairs else 0
    edges = torch.empty(max(total, 1), device=dev, dtype=torch.float64)
    if n_pairs:
      _lib.check(lib.ksig_signature_pde_fwd_edges(
          _lib.dtype_code(Xs.data.dtype), Xs.data.data_ptr(), Xs.off.data_ptr(),
          Xs.len.data_ptr(), Xs.max_len, Ys.data.data_ptr(), Ys.off.data_ptr(),
          Ys.len.data_ptr(), Ys.max_len, Xs.d, pairs.data_ptr(), n_pairs, kind, pval,
          int(dyadic), int(difference), out.data_ptr(), edges.data_ptr(), e_off.data_ptr(),
          _lib.stream_handle(dev)))
    ctx.save_for_backward(xdata, ydata, pairs, edges, e_off, la, lb)
    ctx.meta = (xs_meta, ys_meta, kind, pval, dyadic, difference)
    return out

  @staticmethod
  def backward(ctx, gout):
    lib = _lib.load()
    xdata, ydata, pairs, edges, e_off, la, lb = ctx.saved_tensors
    xs_meta, ys_meta, kind, pval, dyadic, difference = ctx.meta
    Xs = _ops.SeqSet(xdata.detach().contiguous(), *xs_meta)
    Ys = _ops.SeqSet(ydata.detach().contiguous(), *ys_meta)
    dev, dt = xdata.device, Xs.data.dtype
    gout = gout.to(torch.float64).contiguous()
    n_pairs = pairs.shape[0]
    gx = torch.zeros_like(Xs.data) if ctx.needs_input_grad[0] else None
    gy = torch.zeros_like(Ys.data) if ctx.needs_input_grad[1] else None
    gparam = torch.zeros((), device=dev, dtype=torch.float64)
    if n_pairs == 0:
      return gx, gy, gparam, None, None, None, None, None, None
    stream = _lib.stream_handle(dev)
    # the fused adjoint (chain rule in the sweep, no dL/dM array) where it is
    # built: float32, difference, dyadic order 0 / 1, d <= 8
    fused_pairs = _FUSED_PAIRS if dt == torch.float32 else 0
    ws_pp = lib.ksig_signature_pde_bwd_fused_workspace(
        _lib.dtype_code(dt), Xs.d, Xs.max_len, Ys.max_len, int(dyadic), int(difference), 1)
    if ws_pp > 0 and fused_pairs > 0:
      need_gp = ctx.needs_input_grad[2]
      step = max(1, min(fused_pairs, _FUSED_WS_BYTES // ws_pp))
      for c0 in range(0, n_pairs, step):
        c1 = min(n_pairs, c0 + step)
        P = c1 - c0
        ws = torch.empty(max(ws_pp * P // 4, 1), device=dev, dtype=torch.float32)
        gpp = torch.empty(P, device=dev, dtype=torch.float64) if need_gp else None
        _lib.check(lib.ksig_signature_pde_bwd_fused(
            _lib.dtype_code(dt), Xs.data.data_ptr(), Xs.off.data_ptr(), Xs.len.data_ptr(),
            Xs.max_len, Ys.data.data_ptr(), Ys.off.data_ptr(), Ys.len.data_ptr(), Ys.max_len,
            Xs.d, pairs[c0:c1].contiguous().data_ptr(), P, kind, pval, int(dyadic),
            int(difference), gout[c0:c1].data_ptr(), edges.data_ptr(),
            e_off[c0:c1].data_ptr(), None if gx is None else gx.data_ptr(),
            None if gy is None else gy.data_ptr(), None if gpp is None else gpp.data_ptr(),
            ws.data_ptr(), ws_pp * P, stream))
        if need_gp:
          gparam = gparam + gpp.sum()
      return gx, gy, gparam, None, None, None, None, None, None
    sh = 1 if difference else 0
    nr = torch.clamp(torch.maximum(la, lb) - sh, min=0)   # kernel orientation
    nc = torch.clamp(torch.minimum(la, lb) - sh, min=0)
    msz = nr * nc
    budget = max(_GM_BYTES // Xs.data.element_size(), int(msz.max().item()))
    chunk_id = ((torch.cumsum(msz, 0) - msz) // budget).cpu()
    change = torch.nonzero(chunk_id[1:] != chunk_id[:-1]).flatten() + 1
    bounds = [0] + change.tolist() + [n_pairs]
    for c0, c1 in zip(bounds[:-1], bounds[1:]):
      pc = pairs[c0:c1].contiguous()
      P = c1 - c0
      m_off = torch.zeros(P, device=dev, dtype=torch.int64)
      if P > 1:
        m_off[1:] = torch.cumsum(msz[c0:c1], 0)[:-1]
      gm = torch.zeros(max(int(msz[c0:c1].sum()), 1), device=dev, dtype=dt)
      ws_bytes = lib.ksig_signature_pde_bwd_workspace(
          _lib.dtype_code(dt), Xs.max_len, Ys.max_len, int(dyadic), int(difference), P)
      ws = torch.empty(max(ws_bytes // 4, 1), device=dev, dtype=torch.float32)
      _lib.check(lib.ksig_signature_pde_bwd_edges(
          _lib.dtype_code(dt), Xs.data.data_ptr(), Xs.off.data_ptr(), Xs.len.data_ptr(),
          Xs.max_len, Ys.data.data_ptr(), Ys.off.data_ptr(), Ys.len.data_ptr(), Ys.max_len,
          Xs.d, pc.data_ptr(), P, kind, pval, int(dyadic), int(difference),
          gout[c0:c1].data_ptr(), edges.data_ptr(), e_off[c0:c1].data_ptr(), gm.data_ptr(),
          m_off.data_ptr(), ws.data_ptr(), ws_bytes, stream))
      _lib.check(lib.ksig_signature_pde_grad_contract(
          _lib.dtype_code(dt), Xs.data.data_ptr(), Xs.off.data_ptr(), Xs.len.data_ptr(),
          Xs.max_len, Ys.data.data_ptr(), Ys.off.data_ptr(), Ys.len.data_ptr(), Ys.max_len,
          Xs.d, pc.data_ptr(), P, kind, pval, int(difference), gm.data_ptr(),
          m_off.data_ptr(), None if gx is None else gx.data_ptr(),
          None if gy is None else gy.data_ptr(),
          gparam.data_ptr() if ctx.needs_input_grad[2] else None, stream))
    return gx, gy, gparam, None, None, None, None, None, None


def signature_pde(X: _ops.SeqSet, Y: _ops.SeqSet, pairs: torch.Tensor, kind: int, param,
                  dyadic_order: int = 0, difference: bool = True,
                  xdata: Optional[torch.Tensor] = None,
                  ydata: Optional[torch.Tensor] = None) -> torch.Tensor:
  """Differentiable signature-PDE values [n_pairs] (float64)."""
  xdata = X.data if xdata is None else xdata
  ydata = Y.data if ydata is None else ydata
  p = _as_param(param, xdata.device)
  xs_meta = (X.off, X.len, X.max_len, X.n)
  ys_meta = (Y.off, Y.len, Y.max_len, Y.n)
  return _SigPDE.apply(xdata, ydata, p, xs_meta, ys_meta, pairs, kind, int(dyadic_order),
                       bool(difference))
