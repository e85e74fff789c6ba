"""Host-side launchers: torch device tensors -> C-ABI calls -> torch results.

Every function here hands device pointers of contiguous torch tensors to
libksig_hip.so and returns freshly allocated float64 device tensors; nothing
is computed on the host.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib


@dataclass
class SeqSet:
  """Packed sequence set: rows [total, d], per-sequence offsets and lengths."""
  data: torch.Tensor     # [total_rows, d], contiguous
  off: torch.Tensor      # int64 [n]
  len: torch.Tensor      # int32 [n]
  max_len: int
  n: int
  host_len: Optional[np.ndarray] = field(default=None, repr=False)  # ragged sets only

  @property
  def d(self) -> int:
    return self.data.shape[-1]

  @staticmethod
  def dense(X: torch.Tensor) -> "SeqSet":
    """From a dense [n, L, d] device tensor."""
    n, L, d = X.shape
    data = X.reshape(n * L, d).contiguous()
    off = torch.arange(n, device=X.device, dtype=torch.int64) * L
    lens = torch.full((n,), L, device=X.device, dtype=torch.int32)
    return SeqSet(data, off, lens, L, n)

  @staticmethod
  def ragged(seqs: Sequence[torch.Tensor], device=None, dtype=None) -> "SeqSet":
    """From a list of [L_a, d] tensors of differing lengths."""
    device = device or seqs[0].device
    dtype = dtype or seqs[0].dtype
    lens_list = [int(s.shape[0]) for s in seqs]
    data = torch.cat([s.to(device=device, dtype=dtype) for s in seqs], 0).contiguous()
    offs = [0]
    for L in lens_list[:-1]:
      offs.append(offs[-1] + L)
    off = torch.tensor(offs, device=device, dtype=torch.int64)
    lens = torch.tensor(lens_list, device=device, dtype=torch.int32)
    return SeqSet(data, off, lens, max(lens_list) if lens_list else 0, len(seqs),
                  np.asarray(lens_list, dtype=np.int64))

  def lengths(self) -> np.ndarray:
    """Sequence lengths on the host."""
    if self.host_len is not None:
      return self.host_len
    return np.full(self.n, self.max_len, dtype=np.int64)

  @property
  def uniform(self) -> bool:
    return self.host_len is None or bool(np.all(self.host_len == self.max_len))


def _scalar(p) -> float:
  return float(p.detach()) if isinstance(p, torch.Tensor) else float(p)


def static_code(static_kernel, raw: bool = False) -> tuple:
  """(kind, param) of a fused static kernel, or None for a generic one.
  With ``raw`` the parameter is returned as stored (it may be a tensor that
  requires grad)."""
  from .static.kernels import LinearKernel, RBFKernel
  if type(static_kernel) is RBFKernel:
    p = static_kernel.bandwidth
    return _lib.STATIC_RBF, (p if raw else _scalar(p))
  if type(static_kernel) is LinearKernel:
    p = static_kernel.scale
    return _lib.STATIC_LINEAR, (p if raw else _scalar(p))
  return None


def dense_pairs(nx: int, ny: int, device) -> torch.Tensor:
  ix = torch.arange(nx, device=device, dtype=torch.int32)
  iy = torch.arange(ny, device=device, dtype=torch.int32)
  return torch.stack(torch.meshgrid(ix, iy, indexing="ij"), -1).reshape(-1, 2).contiguous()


def triu_pairs(n: int, device) -> torch.Tensor:
  return torch.triu_indices(n, n, device=device).t().to(torch.int32).contiguous()


def diag_pairs(n: int, device) -> torch.Tensor:
  i = torch.arange(n, device=device, dtype=torch.int32)
  return torch.stack([i, i], -1).contiguous()


def _check_sets(X: SeqSet, Y: SeqSet):
  if X.data.dtype != Y.data.dtype:
    raise TypeError("X and Y must share a dtype")
  if X.d != Y.d:
    raise ValueError("X and Y must have the same number of features")
  if X.data.device != Y.data.device:
    raise ValueError("X and Y must live on the same device")


# Largest d of the fused sweeps (features in registers / LDS per lane).
FUSED_MAX_D = 32
# Dot products per GEMM chunk of the GEMM route (bytes).
GEMM_CHUNK_BYTES = 1 << 31


def _norm_order(n_levels: int, order: int) -> int:
  return n_levels if (order <= 0 or order >= n_levels) else order


def fused_dp_ok(d: int, n_levels: int, order: int) -> bool:
  """True when the fused sweeps take this shape: d <= 32 and (n_levels,
  order) in the compiled table; otherwise the GEMM route runs."""
  lib = _lib.load()
  return d <= FUSED_MAX_D and bool(lib.ksig_dp_supported(n_levels, _norm_order(n_levels, order)))


def _static_factors(kind: int, param: float):
  """(scale folded into the row operand, norm factor h or None) of the
  static kernel: RBF exp2(2h<x,y> - h|x|^2 - h|y|^2), h = log2(e) /
  max(2 sigma^2, 1e-12) (ksig/static/kernels.py:258-260); linear scale<x,y>."""
  if kind == _lib.STATIC_RBF:
    if not param > 0:
      raise ValueError(f"static kernel parameter must be positive (got {param})")
    h = 1.4426950408889634 / max(2.0 * param * param, 1e-12)
    return 2.0 * h, h
  if kind == _lib.STATIC_LINEAR:
    if not param > 0:
      raise ValueError(f"static kernel parameter must be positive (got {param})")
    return float(param), None
  raise ValueError(f"unknown static kernel kind {kind}")


def signature_levels_dots(dots: torch.Tensor, mat_off: torch.Tensor, ld: int,
                          lens: Optional[torch.Tensor], lx: int, ly: int, max_ly: int,
                          kind: int, hx: Optional[torch.Tensor], hx_off: Optional[torch.Tensor],
                          hy: Optional[torch.Tensor], hy_off: Optional[torch.Tensor],
                          n_levels: int, order: int, difference: bool) -> torch.Tensor:
  """Level values [n_mats, n_levels+1] (float64) of ksig_signature_levels_dots:
  matrix p at element offset mat_off[p] of ``dots`` with row stride ``ld``,
  static-kernel values formed on read (RBF: exp2(v - hx - hy))."""
  lib = _lib.load()
  n = mat_off.shape[0]
  out = torch.empty((n, n_levels + 1), device=dots.device, dtype=torch.float64)
  if n == 0:
    return out
  ptr = lambda t: t.data_ptr() if t is not None else None
  _lib.check(lib.ksig_signature_levels_dots(
      _lib.dtype_code(dots.dtype), dots.data_ptr(), mat_off.data_ptr(), int(ld), ptr(lens),
      int(lx), int(ly), int(max_ly), int(kind), ptr(hx), ptr(hx_off), ptr(hy), ptr(hy_off), n,
      n_levels, order, int(bool(difference)), out.data_ptr(), _lib.stream_handle(dots.device)))
  return out


def signature_levels_gemm(X: SeqSet, Y: SeqSet, pairs: torch.Tensor, n_levels: int,
                          order: int, difference: bool, kind: int, param: float
                          ) -> torch.Tensor:
  """Level values [n_pairs, n_levels+1] (float64) with the static kernel's
  inner products from library GEMMs on the matrix cores (K = d) and the
  recursion in the HIP kernels of ksig_signature_levels_dots: any d, any
  (n_levels, order).  Dense sets take one GEMM per block of x sequences
  against the y range its pairs need, read in place; other pair lists gather
  their sequences and take a batched GEMM per chunk.  The pair orientation
  puts the longer sequences on the rows (the levels are symmetric in x and
  y), so the LDS column state holds the shorter side."""
  _check_sets(X, Y)
  n_pairs = pairs.shape[0]
  dev = X.data.device
  out = torch.empty((n_pairs, n_levels + 1), device=dev, dtype=torch.float64)
  if n_pairs == 0:
    return out
  if Y.max_len > X.max_len:  # rows = the longer side
    return signature_levels_gemm(Y, X, pairs.flip(-1).contiguous(), n_levels, order,
                                 difference, kind, param)
  gdot, h = _static_factors(kind, param)
  Xd = X.data * gdot
  hx = (h * (X.data * X.data).sum(-1)).contiguous() if h is not None else None
  hy = (h * (Y.data * Y.data).sum(-1)).contiguous() if h is not None else None
  esz = X.data.element_size()
  pr = pairs.detach().cpu().numpy().astype(np.int64)
  xoff, yoff = X.off.detach().cpu().numpy(), Y.off.detach().cpu().numpy()
  lxs, lys = X.lengths()[pr[:, 0]], Y.lengths()[pr[:, 1]]

  def run(dots, idx, mat_off, ld, lens, lx, ly, max_ly):
    t = lambda a, dt=torch.int64: torch.as_tensor(np.ascontiguousarray(a), device=dev, dtype=dt)
    hxo = t(xoff[pr[idx, 0]]) if h is not None else None
    hyo = t(yoff[pr[idx, 1]]) if h is not None else None
    lv = signature_levels_dots(dots, t(mat_off), ld, None if lens is None else t(lens,
                               torch.int32), lx, ly, max_ly, kind, hx, hxo, hy, hyo, n_levels,
                               order, difference)
    out[torch.as_tensor(idx, device=dev)] = lv

  done = np.zeros(n_pairs, dtype=bool)
  if X.uniform and Y.uniform:
    # block GEMMs: x sequences [a0, a1) against y sequences [b0, b1]
    Lx, Ly = X.max_len, Y.max_len
    by_a = np.argsort(pr[:, 0], kind="stable")
    a_sorted = pr[by_a, 0]
    a0 = 0
    while a0 < X.n:
      lo = np.searchsorted(a_sorted, a0)
      if lo >= n_pairs:
        break
      a0 = int(a_sorted[lo])
      # grow the block while the dot products stay within the chunk
      nb = 1
      while True:
        hi = np.searchsorted(a_sorted, a0 + 2 * nb)
        bsel = pr[by_a[lo:hi], 1]
        span = int(bsel.max() - bsel.min() + 1)
        if a0 + 2 * nb > X.n or 2 * nb * Lx * span * Ly * esz > GEMM_CHUNK_BYTES:
          break
        nb *= 2
      hi = np.searchsorted(a_sorted, a0 + nb)
      idx = by_a[lo:hi]
      b = pr[idx, 1]
      b0, b1 = int(b.min()), int(b.max())
      span = b1 - b0 + 1
      if len(idx) * 4 >= nb * span:  # the block GEMM's pairs are mostly needed
        A = Xd[int(xoff[a0]):int(xoff[a0]) + nb * Lx]
        B = Y.data[int(yoff[b0]):int(yoff[b0]) + span * Ly]
        dots = A @ B.T
        ld = span * Ly
        mat_off = (pr[idx, 0] - a0) * Lx * ld + (b - b0) * Ly
        run(dots, idx, mat_off, ld, None, Lx, Ly, Ly)
        done[idx] = True
      a0 += nb
  rest = np.nonzero(~done)[0]
  # gathered pairs: batched GEMMs over chunks of pairs, padded to the chunk's
  # longest sides (the kernel reads each pair within its own lengths)
  pos = 0
  d = X.d
  while pos < len(rest):
    Lxm, Lym = int(lxs[rest[pos]]), int(lys[rest[pos]])
    end = pos + 1
    while end < len(rest):
      nx, ny = max(Lxm, int(lxs[rest[end]])), max(Lym, int(lys[rest[end]]))
      if (end - pos + 1) * (nx * ny + (nx + ny) * d) * esz > GEMM_CHUNK_BYTES:
        break
      Lxm, Lym = nx, ny
      end += 1
    idx = rest[pos:end]
    ar = np.arange(max(Lxm, Lym))
    ix = xoff[pr[idx, 0]][:, None] + np.minimum(ar[None, :Lxm], lxs[idx][:, None] - 1)
    iy = yoff[pr[idx, 1]][:, None] + np.minimum(ar[None, :Lym], lys[idx][:, None] - 1)
    Xp = Xd[torch.as_tensor(ix, device=dev)]
    Yp = Y.data[torch.as_tensor(iy, device=dev)]
    dots = torch.bmm(Xp, Yp.transpose(1, 2).contiguous())
    mat_off = np.arange(len(idx), dtype=np.int64) * Lxm * Lym
    lens = np.stack([lxs[idx], lys[idx]], -1).astype(np.int32)
    run(dots, idx, mat_off, Lym, lens, 0, 0, int(lys[idx].max()))
    pos = end
  return out


def static_gram_chunks(X: SeqSet, Y: SeqSet, pairs: torch.Tensor, kind: int, param: float):
  """Static-kernel matrices of a pair list over uniform-length sets, in
  chunks of pairs: yields (pair indices, M [chunk, Lx, Ly]) with M = k(x_i,
  y_j) formed from a batched library GEMM (the inner products of
  ksig/utils.py:105-157) and one elementwise pass (RBF exp2(2h<x,y> -
  h|x|^2 - h|y|^2), ksig/static/kernels.py:248-260; linear scale<x,y>).  The
  route of the sweeps over caller matrices for d beyond the fused kernels."""
  if not (X.uniform and Y.uniform):
    raise NotImplementedError(
        f"d = {X.d} > {FUSED_MAX_D} is taken for sets of one sequence length only")
  gdot, h = _static_factors(kind, param)
  Lx, Ly, d = X.max_len, Y.max_len, X.d
  X3 = X.data.view(X.n, Lx, d)
  Y3 = Y.data.view(Y.n, Ly, d)
  esz = X.data.element_size()
  chunk = max(1, GEMM_CHUNK_BYTES // ((Lx * Ly + (Lx + Ly) * d) * esz))
  pl = pairs.long()
  for p0 in range(0, pairs.shape[0], chunk):
    pa, pb = pl[p0:p0 + chunk, 0], pl[p0:p0 + chunk, 1]
    Xp, Yp = X3[pa], Y3[pb]
    dots = torch.bmm(Xp * gdot, Yp.transpose(1, 2).contiguous())
    if h is not None:
      dots = torch.exp2(dots - h * (Xp * Xp).sum(-1)[:, :, None]
                        - h * (Yp * Yp).sum(-1)[:, None, :])
    yield p0, dots


def signature_levels(X: SeqSet, Y: SeqSet, pairs: torch.Tensor, n_levels: int,
                     order: int, difference: bool, kind: int, param: float,
                     gram_mode: int = -1) -> torch.Tensor:
  """Level values [n_pairs, n_levels+1] (float64) of the fused DP kernel, or
  of the GEMM route for shapes the fused kernels do not take."""
  lib = _lib.load()
  _check_sets(X, Y)
  n_pairs = pairs.shape[0]
  if kind in (_lib.STATIC_RBF, _lib.STATIC_LINEAR) and not fused_dp_ok(X.d, n_levels, order):
    return signature_levels_gemm(X, Y, pairs, n_levels, order, difference, kind, param)
  if (n_pairs > 0 and X.data.dtype == torch.float32 and difference and gram_mode in (-1, 0)
      and kind in (_lib.STATIC_RBF, _lib.STATIC_LINEAR) and not (X.uniform and Y.uniform)):
    # uneven lengths: length-bucketed schedule for the row stream
    from . import tune
    try:
      sched = tune.tuned_schedule(X, Y, pairs, kind, n_levels, order)
    except NotImplementedError:
      sched = None
    if sched is not None:
      return signature_levels_batched(X, Y, pairs, n_levels, order, kind, param, sched)
  out = torch.empty((n_pairs, n_levels + 1), device=X.data.device, dtype=torch.float64)
  if n_pairs == 0:
    return out
  eff_order = n_levels if (order <= 0 or order >= n_levels) else order
  if (X.data.dtype == torch.float32 and difference and gram_mode in (-1, 0) and eff_order == 1
      and kind in (_lib.STATIC_RBF, _lib.STATIC_LINEAR)):
    # first order, float32: the training forward's entry without the state
    # picks the MFMA-Gram forward when it is faster (d > 16) given a workspace
    ws_bytes = lib.ksig_signature_levels_fwd_save_workspace(
        _lib.F32, X.d, n_levels, X.max_len, Y.max_len, n_pairs)
    if ws_bytes > 0:
      ws = torch.empty(ws_bytes // 4, device=X.data.device, dtype=torch.float32)
      _lib.check(lib.ksig_signature_levels_fwd_save(
          _lib.F32, X.data.data_ptr(), X.off.data_ptr(), X.len.data_ptr(), X.max_len,
          Y.data.data_ptr(), Y.off.data_ptr(), Y.len.data_ptr(), Y.max_len, X.d,
          pairs.data_ptr(), n_pairs, kind, float(param), n_levels, out.data_ptr(), None, None,
          ws.data_ptr(), ws_bytes, _lib.stream_handle(X.data.device)))
      return out
  _lib.check(lib.ksig_signature_levels(
      _lib.dtype_code(X.data.dtype), X.data.data_ptr(), X.off.data_ptr(),
      X.len.data_ptr(), X.max_len, Y.data.data_ptr(), Y.off.data_ptr(),
      Y.len.data_ptr(), Y.max_len, X.d, pairs.data_ptr(), n_pairs, kind,
      float(param), n_levels, order, int(bool(difference)), int(gram_mode),
      out.data_ptr(), _lib.stream_handle(X.data.device)))
  return out


def stream_slots(X: SeqSet, Y: SeqSet, kind: int, n_levels: int, order: int,
                 max_slice: int = 0, global_fifo: bool = False):
  """(resident waves, max pairs per wave, workspace bytes per wave) of the
  row-stream kernel for this problem shape, slice bound and FIFO placement;
  resident waves is 0 when no row-stream kernel fits."""
  lib = _lib.load()
  s, m, w = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_size_t(0)
  _lib.check(lib.ksig_stream_slots(X.d, kind, n_levels, order, X.max_len, Y.max_len,
                                   int(max_slice), int(bool(global_fifo)), ctypes.byref(s),
                                   ctypes.byref(m), ctypes.byref(w)))
  return s.value, m.value, w.value


# Two-wave slices (global FIFO): a slice whose sweep is longer than a factor
# (default TWO_WAVE_FACTOR) x the mean work per resident wave, with at least
# TWO_WAVE_MIN_COLS columns and TWO_WAVE_MIN_ROWS rows, is split band by band
# between the two waves of one block.
TWO_WAVE_FACTOR = 0.8
TWO_WAVE_MIN_COLS = 384
TWO_WAVE_MIN_ROWS = 512


def stream_schedule(lx: np.ndarray, ly: np.ndarray, slots: int, max_per_wave: int,
                    min_cols: int = 68, two_wave: float = 0.0, resident: int = 0):
  """Length-bucketed schedule of a pair list for the row stream (host side).

  Pairs are ordered by their shorter side, longest first (a wave sweeps all
  its pairs with one column count), then cut into contiguous slices of about
  equal work (rows x columns), one slice per resident wave and round, each
  slice at most ``max_per_wave`` pairs.  The slices are then listed by their
  sweep length, longest first (the hardware starts waves in launch order and
  hands each later slice to the first wave slot that frees up, so the slices
  that set the critical path start at once and the short ones fill the
  gaps).

  ``two_wave`` > 0 (global-FIFO launches, two waves per block: block b runs
  slices 2b and 2b + 1; True means TWO_WAVE_FACTOR): the slices longer than
  that factor x the mean work per resident wave become two-wave slices, each followed by
  an empty slice that marks it, so that the block's second wave sweeps every
  other band of it; the remaining slices are paired into blocks in the same
  longest-first order, and blocks are listed by their sweep length.
  ``resident`` is the resident wave count the mean work is taken over
  (default ``slots``).  Returns (order, wave_beg) as int32 numpy arrays."""
  lx = np.asarray(lx, dtype=np.int64)
  ly = np.asarray(ly, dtype=np.int64)
  n = lx.shape[0]
  cols = np.minimum(lx, ly)
  rows = np.maximum(lx, ly)
  order = np.argsort(-cols, kind="stable")
  live = cols[order] >= 2
  cost = np.where(live, rows[order] * np.maximum(cols[order], min_cols), 0).astype(np.float64)
  slots = max(int(slots), 1)
  rounds = max(1, -(-n // (slots * max_per_wave)))
  n_target = min(n, slots * rounds)
  cum = np.cumsum(cost)
  total = cum[-1] if n else 0.0
  if total > 0 and n_target > 1:
    targets = total * np.arange(1, n_target) / n_target
    b = np.searchsorted(cum, targets, side="left") + 1
  else:
    b = np.zeros(0, dtype=np.int64)
  b = np.unique(np.concatenate([[0], np.clip(b, 0, n), [n]]))
  sizes = np.diff(b)
  parts = np.maximum(1, -(-sizes // max_per_wave))
  base = np.repeat(b[:-1], parts)
  size_r = np.repeat(sizes, parts)
  part_r = np.repeat(parts, parts)
  t = np.arange(parts.sum()) - np.repeat(np.cumsum(parts) - parts, parts)
  starts = base + (t * size_r) // part_r
  wave_beg = np.concatenate([starts, [n]]).astype(np.int64)
  if len(wave_beg) > 2:
    # sweep steps of a slice ~ its padded rows x its widest pair's columns
    so_rows = np.where(live, rows[order], 0)
    so_cols = np.maximum(cols[order], min_cols)
    slice_rows = np.add.reduceat(so_rows, wave_beg[:-1])
    slice_cols = so_cols[wave_beg[:-1]]
    steps = slice_rows * slice_cols
    perm = np.argsort(-steps, kind="stable")
    # (block length, slices) with None for the empty slice of a two-wave block
    tail = []
    if two_wave:
      factor = TWO_WAVE_FACTOR if two_wave is True else float(two_wave)
      mean = steps.sum() / max(int(resident) if resident > 0 else slots, 1)
      tw = ((steps[perm] > factor * mean)
            & (slice_cols[perm] >= TWO_WAVE_MIN_COLS) & (slice_rows[perm] >= TWO_WAVE_MIN_ROWS))
      solo = [int(k) for k, t in zip(perm, tw) if not t]
      blocks = [(steps[k] / 2, [int(k), None]) for k, t in zip(perm, tw) if t]
      blocks += [(steps[solo[i]], solo[i:i + 2]) for i in range(0, len(solo) - 1, 2)]
      tail = solo[len(solo) - len(solo) % 2:]  # an unpaired slice goes last
      blocks.sort(key=lambda b: -b[0])
    else:
      blocks = [(steps[k], [int(k)]) for k in perm]
    seq = [k for _, ks in blocks for k in ks] + tail
    idx = np.concatenate([np.arange(wave_beg[k], wave_beg[k + 1]) for k in seq if k is not None])
    sizes = np.array([0 if k is None else wave_beg[k + 1] - wave_beg[k] for k in seq])
    order = order[idx]
    wave_beg = np.concatenate([[0], np.cumsum(sizes)])
  return order.astype(np.int32), wave_beg.astype(np.int32)


def signature_levels_batched(X: SeqSet, Y: SeqSet, pairs: torch.Tensor, n_levels: int,
                             order: int, kind: int, param: float,
                             schedule: Optional[tuple] = None) -> torch.Tensor:
  """Level values [n_pairs, n_levels+1] of float32 pairs of uneven lengths in
  one persistent row-stream launch.  ``schedule`` = (order, wave_beg[,
  max_slice[, global_fifo]]) with device int32 tensors (order may be None),
  or None for the length-bucketed schedule of ``make_schedule``."""
  lib = _lib.load()
  _check_sets(X, Y)
  if X.data.dtype != torch.float32:
    raise TypeError("the batched row stream takes float32 inputs")
  n_pairs = pairs.shape[0]
  dev = X.data.device
  out = torch.empty((n_pairs, n_levels + 1), device=dev, dtype=torch.float64)
  if n_pairs == 0:
    return out
  if schedule is None:
    schedule = make_schedule(X, Y, pairs, kind, n_levels, order)
  so, wb = schedule[0], schedule[1]
  max_slice = schedule[2] if len(schedule) > 2 else 128
  global_fifo = bool(schedule[3]) if len(schedule) > 3 else False
  n_waves = wb.shape[0] - 1
  ws, ws_bytes = None, 0
  if global_fifo:
    norm_order = n_levels if (order <= 0 or order >= n_levels) else order
    _, _, per_wave = stream_slots(X, Y, kind, n_levels, norm_order, max_slice, True)
    ws_bytes = per_wave * n_waves
    ws = torch.empty(max(ws_bytes, 16), device=dev, dtype=torch.uint8)
  _lib.check(lib.ksig_signature_levels_batched(
      X.data.data_ptr(), X.off.data_ptr(), X.len.data_ptr(), X.max_len,
      Y.data.data_ptr(), Y.off.data_ptr(), Y.len.data_ptr(), Y.max_len, X.d,
      pairs.data_ptr(), n_pairs, kind, float(param), n_levels, order,
      so.data_ptr() if so is not None else None, wb.data_ptr(), n_waves,
      int(max_slice), ws.data_ptr() if ws is not None else None, ws_bytes,
      out.data_ptr(), _lib.stream_handle(dev)))
  return out


def make_schedule(X: SeqSet, Y: SeqSet, pairs: torch.Tensor, kind: int, n_levels: int,
                  order: int, global_fifo: Optional[bool] = None, max_slice: int = 0,
                  oversub: int = 1, two_wave: Optional[float] = None):
  """(order, wave_beg, max_slice, global_fifo) of the length-bucketed schedule
  (device int32 tensors, the largest slice and the FIFO placement).

  ``global_fifo`` None picks the FIFO placement with more resident waves;
  ``max_slice`` (0 = the kernel's 128) bounds the pairs of one wave, which
  sizes its LDS pair table; ``oversub`` cuts that many slices per resident
  wave (the hardware hands the later, lighter slices to waves that finish
  early); ``two_wave`` (global FIFO only; the factor of ``stream_schedule``,
  0 = off, None = TWO_WAVE_FACTOR) splits the longest slices between the two
  waves of a block.
  The autotuner (ksig_amd.tune) searches these four."""
  norm_order = n_levels if (order <= 0 or order >= n_levels) else order
  if global_fifo is None:
    slots, maxp, _ = stream_slots(X, Y, kind, n_levels, norm_order, max_slice)
    g_slots, _, _ = stream_slots(X, Y, kind, n_levels, norm_order, max_slice, True)
    # the FIFO goes to global memory when an LDS FIFO would cost resident waves
    global_fifo = g_slots > slots
    slots = max(slots, g_slots)
  else:
    slots, maxp, _ = stream_slots(X, Y, kind, n_levels, norm_order, max_slice,
                                  bool(global_fifo))
  if slots <= 0:
    raise NotImplementedError("no row-stream kernel fits this problem shape")
  if max_slice > 0:
    maxp = min(maxp, int(max_slice))
  pr = pairs.detach().cpu().numpy()
  lx = X.lengths()[pr[:, 0]]
  ly = Y.lengths()[pr[:, 1]]
  so, wb = stream_schedule(lx, ly, slots * max(int(oversub), 1), maxp,
                           min_cols=96 if global_fifo else 68,
                           two_wave=(TWO_WAVE_FACTOR if two_wave is None else float(two_wave))
                           if global_fifo else 0.0,
                           resident=slots)
  dev = X.data.device
  return (torch.as_tensor(so, device=dev), torch.as_tensor(wb, device=dev),
          int(np.diff(wb).max()) if len(wb) > 1 else 1, bool(global_fifo))


def signature_levels_gram(M: torch.Tensor, n_levels: int, order: int,
                          difference: bool) -> torch.Tensor:
  """Level values [prod(lead), n_levels+1] for matrices M [..., lx, ly]."""
  lib = _lib.load()
  lx, ly = M.shape[-2:]
  Mc = M.reshape(-1, lx, ly).contiguous()
  n = Mc.shape[0]
  out = torch.empty((n, n_levels + 1), device=M.device, dtype=torch.float64)
  if n == 0:
    return out
  _lib.check(lib.ksig_signature_levels_gram(
      _lib.dtype_code(Mc.dtype), Mc.data_ptr(), n, lx, ly, n_levels, order,
      int(bool(difference)), out.data_ptr(), _lib.stream_handle(M.device)))
  return out


def signature_pde(X: SeqSet, Y: SeqSet, pairs: torch.Tensor, difference: bool,
                  kind: int, param: float, dyadic_order: int) -> torch.Tensor:
  """Signature-PDE values [n_pairs] (float64)."""
  lib = _lib.load()
  _check_sets(X, Y)
  n_pairs = pairs.shape[0]
  out = torch.empty((n_pairs,), device=X.data.device, dtype=torch.float64)
  if n_pairs == 0:
    return out
  if X.d > FUSED_MAX_D:
    for p0, M in static_gram_chunks(X, Y, pairs, kind, param):
      out[p0:p0 + M.shape[0]] = signature_pde_gram(M, difference, dyadic_order)
    return out
  _lib.check(lib.ksig_signature_pde(
      _lib.dtype_code(X.data.dtype), X.data.data_ptr(), X.off.data_ptr(),
      X.len.data_ptr(), X.max_len, Y.data.data_ptr(), Y.off.data_ptr(),
      Y.len.data_ptr(), Y.max_len, X.d, pairs.data_ptr(), n_pairs, kind,
      float(param), int(dyadic_order), int(bool(difference)), out.data_ptr(),
      _lib.stream_handle(X.data.device)))
  return out


def signature_pde_gram(M: torch.Tensor, difference: bool, dyadic_order: int
                       ) -> torch.Tensor:
  lib = _lib.load()
  lx, ly = M.shape[-2:]
  Mc = M.reshape(-1, lx, ly).contiguous()
  n = Mc.shape[0]
  out = torch.empty((n,), device=M.device, dtype=torch.float64)
  if n == 0:
    return out
  _lib.check(lib.ksig_signature_pde_gram(
      _lib.dtype_code(Mc.dtype), Mc.data_ptr(), n, lx, ly, int(dyadic_order),
      int(bool(difference)), out.data_ptr(), _lib.stream_handle(M.device)))
  return out


def global_align_log(X: SeqSet, Y: SeqSet, pairs: torch.Tensor, kind: int,
                     param: float) -> torch.Tensor:
  """log GA values [n_pairs] (float64) with the static kernel fused."""
  lib = _lib.load()
  _check_sets(X, Y)
  n_pairs = pairs.shape[0]
  out = torch.empty((n_pairs,), device=X.data.device, dtype=torch.float64)
  if n_pairs == 0:
    return out
  if X.d > FUSED_MAX_D:
    for p0, M in static_gram_chunks(X, Y, pairs, kind, param):
      out[p0:p0 + M.shape[0]] = global_align_log_gram(M)
    return out
  _lib.check(lib.ksig_global_align_log(
      _lib.dtype_code(X.data.dtype), X.data.data_ptr(), X.off.data_ptr(),
      X.len.data_ptr(), X.max_len, Y.data.data_ptr(), Y.off.data_ptr(),
      Y.len.data_ptr(), Y.max_len, X.d, pairs.data_ptr(), n_pairs, kind,
      float(param), out.data_ptr(), _lib.stream_handle(X.data.device)))
  return out


def global_align_log_gram(M: torch.Tensor) -> torch.Tensor:
  """log GA values [prod(lead)] (float64) for matrices M [..., lx, ly]."""
  lib = _lib.load()
  lx, ly = M.shape[-2:]
  Mc = M.reshape(-1, lx, ly).contiguous()
  n = Mc.shape[0]
  out = torch.empty((n,), device=M.device, dtype=torch.float64)
  if n == 0:
    return out
  _lib.check(lib.ksig_global_align_log_gram(
      _lib.dtype_code(Mc.dtype), Mc.data_ptr(), n, lx, ly, out.data_ptr(),
      _lib.stream_handle(M.device)))
  return out


def rws_groups(warp_lens: np.ndarray, n_seq: int, target_waves: int = 4096) -> np.ndarray:
  """Split points [n_groups + 1] of the warping series into groups of about
  equal rows: enough groups that n_seq x n_groups waves fill the chip, but
  at least one 64-row band of rows per group where there are enough rows."""
  lens = np.asarray(warp_lens, dtype=np.int64)
  n_w = lens.shape[0]
  total = int(lens.sum())
  if n_w == 0:
    return np.zeros(1, dtype=np.int32)
  g = -(-int(target_waves) // max(int(n_seq), 1))
  g = max(1, min(g, n_w, max(1, total // 64)))
  cum = np.cumsum(lens)
  cuts = np.searchsorted(cum, total * np.arange(1, g) / g, side="left") + 1
  b = np.unique(np.concatenate([[0], np.clip(cuts, 1, n_w - 1) if g > 1 else [], [n_w]]))
  return b.astype(np.int32)


def random_warping_series(X: SeqSet, warp: torch.Tensor, warp_lens: np.ndarray
                          ) -> torch.Tensor:
  """DTW values [n_seq, n_warps] (float64) of every sequence against every
  warping series (points ``warp`` [sum of lengths, d], lengths ``warp_lens``)."""
  lib = _lib.load()
  if warp.dtype != X.data.dtype or warp.device != X.data.device:
    raise ValueError("warping series must share the sequences' dtype and device")
  if warp.shape[-1] != X.d:
    raise ValueError("warping series must have the sequences' number of features")
  lens = np.asarray(warp_lens, dtype=np.int64)
  n_w = lens.shape[0]
  dev = X.data.device
  out = torch.empty((X.n, n_w), device=dev, dtype=torch.float64)
  if X.n == 0 or n_w == 0:
    return out
  if X.d > FUSED_MAX_D:
    # distances from a library GEMM (ksig/kernels.py:466), then the sweep
    if not X.uniform:
      raise NotImplementedError(
          f"d = {X.d} > {FUSED_MAX_D} is taken for sets of one sequence length only")
    Xf = X.data
    w = warp.contiguous()
    D = ((Xf * Xf).sum(-1)[:, None] + (w * w).sum(-1)[None, :] - 2 * Xf @ w.T)
    return random_warping_series_gram(D.view(X.n, X.max_len, -1), lens)
  off = torch.as_tensor(np.concatenate([[0], np.cumsum(lens)]).astype(np.int32), device=dev)
  groups = rws_groups(lens, X.n)
  g = torch.as_tensor(groups, device=dev)
  w = warp.contiguous()
  _lib.check(lib.ksig_random_warping_series(
      _lib.dtype_code(X.data.dtype), X.data.data_ptr(), X.off.data_ptr(), X.len.data_ptr(),
      X.max_len, X.n, X.d, w.data_ptr(), off.data_ptr(), n_w, int(lens.sum()),
      g.data_ptr(), len(groups) - 1, out.data_ptr(), _lib.stream_handle(dev)))
  return out


def random_warping_series_gram(D: torch.Tensor, warp_lens: np.ndarray) -> torch.Tensor:
  """DTW values [n_seq, n_warps] (float64) from distances D [n_seq, lx, sum of
  lengths]."""
  lib = _lib.load()
  lens = np.asarray(warp_lens, dtype=np.int64)
  n_w = lens.shape[0]
  n, lx, W = D.shape
  if W != int(lens.sum()):
    raise ValueError("the last axis of D must be the total length of the warping series")
  dev = D.device
  out = torch.empty((n, n_w), device=dev, dtype=torch.float64)
  if n == 0 or n_w == 0:
    return out
  Dc = D.contiguous()
  off = torch.as_tensor(np.concatenate([[0], np.cumsum(lens)]).astype(np.int32), device=dev)
  groups = rws_groups(lens, n)
  g = torch.as_tensor(groups, device=dev)
  _lib.check(lib.ksig_random_warping_series_gram(
      _lib.dtype_code(Dc.dtype), Dc.data_ptr(), n, lx, W, off.data_ptr(), n_w,
      g.data_ptr(), len(groups) - 1, out.data_ptr(), _lib.stream_handle(dev)))
  return out


def lowrank_levels(Z: torch.Tensor, order: int = 1) -> torch.Tensor:
  """Low-rank level features [n_seq, n_levels, n_comp] (float64) from the
  projected increments Z [n_levels, n_seq, steps, n_comp], embedding order
  ``order`` (<= 0 or >= n_levels: all orders)."""
  lib = _lib.load()
  Zc = Z.contiguous()
  n_levels, n, steps, P = Zc.shape
  out = torch.empty((n, n_levels, P), device=Z.device, dtype=torch.float64)
  if n == 0 or P == 0:
    return out
  _lib.check(lib.ksig_lowrank_levels(_lib.dtype_code(Zc.dtype), Zc.data_ptr(), n_levels,
                                     int(order), n, steps, P, out.data_ptr(),
                                     _lib.stream_handle(Z.device)))
  return out


def lowrank_levels_diag(dU: torch.Tensor) -> torch.Tensor:
  """RFSF-DP level features [n_seq, sum_i 2^(i+1) * half] (float64) from the
  differenced random Fourier features dU [n_levels, n_seq, steps, 2 * half]."""
  lib = _lib.load()
  dc = dU.contiguous()
  n_levels, n, steps, width = dc.shape
  half = width // 2
  tot = (2 << n_levels) - 2
  out = torch.empty((n, tot * half), device=dU.device, dtype=torch.float64)
  if n == 0 or half == 0:
    return out
  _lib.check(lib.ksig_lowrank_levels_diag(_lib.dtype_code(dc.dtype), dc.data_ptr(), n_levels, n,
                                          steps, half, out.data_ptr(),
                                          _lib.stream_handle(dU.device)))
  return out


def lowrank_levels_diagonal(dU: torch.Tensor, internal_size: int = 2,
                            order: int = 1) -> torch.Tensor:
  """Level features [n_seq, sum_i q^(i+1) * half] (float64) of the low-rank
  recursion with a DiagonalProjection of internal size q = internal_size at
  embedding order ``order`` (<= 0 or >= n_levels: all orders), from the
  differenced lifted paths dU [n_levels, n_seq, steps, q * half].  Shapes
  the HIP scan is not built for raise NotImplementedError (nothing is
  launched)."""
  lib = _lib.load()
  dc = dU.contiguous()
  n_levels, n, steps, width = dc.shape
  q = int(internal_size)
  if q < 1 or width % q:
    raise ValueError(f'cannot split {width} features into internal_size={q} parts')
  half = width // q
  tot = sum(q ** (i + 1) for i in range(n_levels))
  out = torch.empty((n, tot * half), device=dU.device, dtype=torch.float64)
  _lib.check(lib.ksig_lowrank_levels_diagonal(_lib.dtype_code(dc.dtype), dc.data_ptr(), n_levels,
                                              int(order), q, n, steps, half, out.data_ptr(),
                                              _lib.stream_handle(dU.device)))
  return out


def lowrank_levels_sketch(dU: torch.Tensor, hash_idx: torch.Tensor, hash_bit: torch.Tensor,
                          n_comp: int, order: int = 1) -> torch.Tensor:
  """Level features [n_seq, n_levels, n_comp] (float64) of the low-rank
  recursion with TensorSketch projections from the lifted paths dU
  [n_levels, n_seq, steps, width], each level's hashes hash_idx [n_levels,
  width] (int32 after conversion) and signs hash_bit [n_levels, width],
  embedding order ``order``.  Shapes the HIP scan is not built for raise
  NotImplementedError (nothing is launched)."""
  lib = _lib.load()
  dc = dU.contiguous()
  n_levels, n, steps, width = dc.shape
  hi = hash_idx.to(device=dU.device, dtype=torch.int32).contiguous()
  hb = hash_bit.to(device=dU.device, dtype=dc.dtype).contiguous()
  assert tuple(hi.shape) == (n_levels, width) and tuple(hb.shape) == (n_levels, width)
  out = torch.empty((n, n_levels, int(n_comp)), device=dU.device, dtype=torch.float64)
  _lib.check(lib.ksig_lowrank_levels_sketch(_lib.dtype_code(dc.dtype), dc.data_ptr(), n_levels,
                                            int(order), n, steps, width, hi.data_ptr(),
                                            hb.data_ptr(), int(n_comp), out.data_ptr(),
                                            _lib.stream_handle(dU.device)))
  return out


def rff_increments(proj: torch.Tensor, difference: bool = True,
                   out: Optional[torch.Tensor] = None) -> torch.Tensor:
  """[sin, cos](proj) / sqrt(half), differenced along time, from projections
  proj [n_seq, L, half] in one pass.  ``proj`` may be a column slice of a
  wider GEMM output (unit stride along its last axis); ``out`` optionally
  receives the result [n_seq, L - 1 (or L), 2 half] (contiguous)."""
  lib = _lib.load()
  n, L, half = proj.shape
  if proj.stride(2) != 1 or proj.stride(0) != L * proj.stride(1):
    proj = proj.contiguous()
  ld = proj.stride(1)
  steps = L - 1 if difference else L
  if out is None:
    out = torch.empty((n, max(steps, 0), 2 * half), device=proj.device, dtype=proj.dtype)
  if out.numel() == 0:
    return out
  assert out.is_contiguous() and tuple(out.shape) == (n, steps, 2 * half)
  _lib.check(lib.ksig_rff_increments(_lib.dtype_code(proj.dtype), proj.data_ptr(), n, L, half,
                                     ld, int(bool(difference)), out.data_ptr(),
                                     _lib.stream_handle(proj.device)))
  return out
