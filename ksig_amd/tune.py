"""Autotuner for the row-stream launch of a pair list.

The row stream (one persistent launch over a length-bucketed pair list) has
three launch choices whose best setting depends on the length distribution,
d, the level count and the order, and which no closed-form model ranks
reliably:

* ``global_fifo`` -- the band FIFO in LDS, or in a caller workspace in
  global memory (frees LDS for more resident waves, adds a global round trip
  per band);
* ``max_slice`` -- the most pairs one wave owns; the LDS pair table is sized
  to it (32 B per pair), so smaller slices can raise the resident waves per
  CU, at the price of more pipeline fills;
* ``oversub`` -- slices cut per resident wave: 1 runs every wave once, more
  lets the dispatcher hand the later (shorter) slices to waves that finish
  early, which evens out the tail set by the largest pairs;
* ``two_wave`` (global FIFO only) -- the slices longer than this factor x
  the mean work per resident wave (0: none) are split band by band between
  the two waves of a block, which halves the sweeps that set the critical
  path, at the price of a second pipeline fill and a step-count hand-shake
  between the waves.

``autotune`` times every feasible candidate on the caller's own data with
events on the launch stream (the kernels run on torch's current stream) and
keeps the fastest per problem signature; ``signature_levels`` then uses the
kept choice for every later call with the same signature.  Candidates only
change how pairs are grouped into waves, never the arithmetic per pair, so
every candidate computes the same values (up to the float32 rounding of
per-pair partial sums, far below 1e-6 relative).  The kept choices can be
saved to and loaded from a JSON file (``save`` / ``load``, or the
``KSIG_TUNE_CACHE`` environment variable, read on first use).
"""

from __future__ import annotations

import json
import os
from dataclasses import asdict, dataclass
from typing import Callable, Dict, List, Optional, Tuple

import numpy as np
import torch

from . import _ops

MAX_SLICES = (128, 64, 32)
OVERSUB = (1, 2, 4, 8)
TWO_WAVE = (0.8, 1.25)


@dataclass(frozen=True)
class StreamConfig:
  global_fifo: bool
  max_slice: int
  oversub: int
  two_wave: float = 0.0


_cache: Dict[str, StreamConfig] = {}
_loaded = False


def _norm_order(n_levels: int, order: int) -> int:
  return n_levels if (order <= 0 or order >= n_levels) else order


def signature_key(lx: np.ndarray, ly: np.ndarray, d: int, kind: int, n_levels: int,
                  order: int) -> str:
  """Problem signature the tuned choice is kept under: d, static kind, levels,
  order, the pair count (to a power of two) and the length distribution (the
  quartiles and maximum of the shorter and the longer side, to 32 points)."""
  lx = np.asarray(lx, dtype=np.int64)
  ly = np.asarray(ly, dtype=np.int64)
  if lx.size == 0:
    return f"d{d}-k{kind}-n{n_levels}-o{_norm_order(n_levels, order)}-p0"
  lo = np.minimum(lx, ly)
  hi = np.maximum(lx, ly)
  q = lambda v: "/".join(str(int(x) // 32) for x in np.percentile(v, [25, 50, 75, 100]))
  return (f"d{d}-k{kind}-n{n_levels}-o{_norm_order(n_levels, order)}"
          f"-p{int(np.log2(max(lx.size, 1)))}-c{q(lo)}-r{q(hi)}")


def _pair_lengths(X: _ops.SeqSet, Y: _ops.SeqSet, pairs: torch.Tensor):
  pr = pairs.detach().cpu().numpy()
  return X.lengths()[pr[:, 0]], Y.lengths()[pr[:, 1]]


def candidates(X: _ops.SeqSet, Y: _ops.SeqSet, kind: int, n_levels: int,
               order: int) -> List[StreamConfig]:
  """Every launch choice for which a row-stream kernel fits (resident waves
  > 0); only the slice bounds that change the resident waves are kept."""
  o = _norm_order(n_levels, order)
  out = []
  for gf in (False, True):
    seen = set()
    for ms in MAX_SLICES:
      slots, _, _ = _ops.stream_slots(X, Y, kind, n_levels, o, ms, gf)
      if slots <= 0 or slots in seen:
        continue
      seen.add(slots)
      for ov in OVERSUB:
        out.append(StreamConfig(gf, ms, ov))
        if gf:
          out += [StreamConfig(gf, ms, ov, f) for f in TWO_WAVE]
  return out


def _event_timer(fn: Callable[[], object], reps: int) -> float:
  """Mean milliseconds of ``fn`` over ``reps`` launches after one warm-up,
  from events on torch's current stream."""
  fn()
  stream = torch.cuda.current_stream()
  e0 = torch.cuda.Event(enable_timing=True)
  e1 = torch.cuda.Event(enable_timing=True)
  e0.record(stream)
  for _ in range(reps):
    fn()
  e1.record(stream)
  e1.synchronize()
  return e0.elapsed_time(e1) / reps


def autotune(X: _ops.SeqSet, Y: _ops.SeqSet, pairs: torch.Tensor, n_levels: int,
             order: int, kind: int, param: float, reps: int = 3,
             timer: Optional[Callable[[Callable[[], object], int], float]] = None,
             configs: Optional[List[StreamConfig]] = None
             ) -> Tuple[StreamConfig, Dict[StreamConfig, float]]:
  """Time every candidate launch of this pair list and keep the fastest.

  Returns (best config, {config: ms}).  ``timer(fn, reps)`` defaults to
  events on the current stream; ``configs`` defaults to ``candidates``."""
  _ensure_loaded()
  if X.data.dtype != torch.float32:
    raise TypeError("the row stream takes float32 inputs")
  configs = candidates(X, Y, kind, n_levels, order) if configs is None else configs
  if not configs:
    raise NotImplementedError("no row-stream kernel fits this problem shape")
  timer = timer or _event_timer
  times: Dict[StreamConfig, float] = {}
  for cfg in configs:
    sched = _ops.make_schedule(X, Y, pairs, kind, n_levels, order, cfg.global_fifo,
                               cfg.max_slice, cfg.oversub, cfg.two_wave)
    times[cfg] = float(timer(lambda: _ops.signature_levels_batched(
        X, Y, pairs, n_levels, order, kind, param, sched), reps))
  best = min(times, key=times.get)
  lx, ly = _pair_lengths(X, Y, pairs)
  _cache[signature_key(lx, ly, X.d, kind, n_levels, order)] = best
  _generation[0] += 1
  return best, times


def lookup(X: _ops.SeqSet, Y: _ops.SeqSet, pairs: torch.Tensor, kind: int, n_levels: int,
           order: int) -> Optional[StreamConfig]:
  """The kept choice for this problem signature, or None."""
  _ensure_loaded()
  if not _cache:
    return None
  lx, ly = _pair_lengths(X, Y, pairs)
  return _cache.get(signature_key(lx, ly, X.d, kind, n_levels, order))


# Schedules already built, keyed by the identity of the pair list and the
# sets' host length arrays (a repeated call with the same tensors skips the
# host planning and the device copies), and by the content of the pair
# lengths (a new tensor with the same lengths costs one device-to-host copy
# of the pair list and a hash).
_SCHED_MAX = 64
_sched_by_id: Dict[tuple, tuple] = {}
_sched_by_content: Dict[tuple, tuple] = {}
_generation = [0]   # bumped whenever the kept launch choices change


def _remember(table: dict, key, value) -> None:
  if len(table) >= _SCHED_MAX:
    table.pop(next(iter(table)))
  table[key] = value


def tuned_schedule(X: _ops.SeqSet, Y: _ops.SeqSet, pairs: torch.Tensor, kind: int,
                   n_levels: int, order: int):
  """The schedule of the kept choice, or the default length-bucketed schedule
  when this signature has not been tuned; built schedules are cached."""
  _ensure_loaded()
  base = (X.d, kind, n_levels, _norm_order(n_levels, order), X.max_len, Y.max_len,
          str(pairs.device), _generation[0])
  idkey = base + (id(pairs), id(X.host_len), id(Y.host_len))
  hit = _sched_by_id.get(idkey)
  if (hit is not None and hit[0] is pairs and hit[1] == pairs._version
      and hit[2] is X.host_len and hit[3] is Y.host_len):
    return hit[4]
  lx, ly = _pair_lengths(X, Y, pairs)
  ckey = base + (hash(lx.tobytes()), hash(ly.tobytes()), lx.size)
  sched = _sched_by_content.get(ckey)
  if sched is None:
    cfg = _cache.get(signature_key(lx, ly, X.d, kind, n_levels, order)) if _cache else None
    if cfg is None:
      sched = _ops.make_schedule(X, Y, pairs, kind, n_levels, order)
    else:
      sched = _ops.make_schedule(X, Y, pairs, kind, n_levels, order, cfg.global_fifo,
                                 cfg.max_slice, cfg.oversub, cfg.two_wave)
    _remember(_sched_by_content, ckey, sched)
  # the entry holds the pair tensor and the length arrays, so their ids stay
  # theirs while it lives
  _remember(_sched_by_id, idkey, (pairs, pairs._version, X.host_len, Y.host_len, sched))
  return sched


def clear() -> None:
  _cache.clear()
  _generation[0] += 1
  _sched_by_id.clear()
  _sched_by_content.clear()


def save(path: str) -> None:
  with open(path, "w") as f:
    json.dump({k: asdict(v) for k, v in sorted(_cache.items())}, f, indent=1)


def load(path: str) -> None:
  with open(path) as f:
    raw = json.load(f)
  for k, v in raw.items():
    _cache[k] = StreamConfig(bool(v["global_fifo"]), int(v["max_slice"]), int(v["oversub"]),
                             float(v.get("two_wave", 0.0)))
  _generation[0] += 1


def _ensure_loaded() -> None:
  global _loaded
  if _loaded:
    return
  _loaded = True
  path = os.environ.get("KSIG_TUNE_CACHE")
  if path and os.path.exists(path):
    load(path)
