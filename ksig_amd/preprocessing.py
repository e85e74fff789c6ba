"""Sequence preprocessing transformers: the ksig.preprocessing counterparts
(ksig/preprocessing.py:15-138).

They take NumPy arrays (or lists of them) on the host or torch tensors on the
GPU and return the same kind, with the reference's steps: interpolation to a
common length, normalisation, lead-lag, time and basepoint augmentation.
"""

from __future__ import annotations

from typing import List, Optional, Union

import numpy as np
import torch
from sklearn.base import BaseEstimator, TransformerMixin

from .utils import check_positive_value


def _interp(t_new, t_old, v):
  """Piecewise-linear interpolation of the columns of v [n, c] from the
  increasing grid t_old [n] to t_new [m] (np.interp per column)."""
  if isinstance(v, np.ndarray):
    return np.stack([np.interp(t_new, t_old, v[:, c]) for c in range(v.shape[1])], axis=1)
  i = torch.clamp(torch.searchsorted(t_old, t_new, right=True), 1, t_old.shape[0] - 1)
  t0, t1 = t_old[i - 1], t_old[i]
  w = ((t_new - t0) / torch.clamp(t1 - t0, min=torch.finfo(v.dtype).tiny))[:, None]
  return v[i - 1] * (1 - w) + v[i] * w


def _linspace(like, n):
  if isinstance(like, np.ndarray):
    return np.linspace(0., 1., n)
  return torch.linspace(0., 1., n, device=like.device, dtype=like.dtype)


class SequenceTabulator(BaseEstimator, TransformerMixin):
  """Tabulates sequences to one length (ksig/preprocessing.py:15-68): rows
  holding a NaN are dropped and every channel is linearly interpolated to
  ``max_len_`` points whenever lengths differ, a NaN is present or a
  sequence exceeds ``max_len_``."""

  def __init__(self, max_len: Optional[int] = None):
    self.max_len = (check_positive_value(max_len, 'max_len')
                    if max_len is not None else None)

  def fit(self, X_seq, y=None) -> 'SequenceTabulator':
    max_seq_len = int(np.max([x.shape[0] for x in X_seq]))
    self.max_len_ = (min(self.max_len, max_seq_len) if self.max_len is not None
                     else max_seq_len)
    return self

  def transform(self, X_seq: Union[np.ndarray, torch.Tensor, List]):
    is_np = isinstance(X_seq[0], np.ndarray)
    isnan = np.isnan if is_np else torch.isnan
    needs = any(x.shape[0] != X_seq[0].shape[0] or bool(isnan(x).any())
                or x.shape[0] > self.max_len_ for x in X_seq)
    if needs:
      out = []
      for x in X_seq:
        x = x[~isnan(x).any(-1)]
        out.append(_interp(_linspace(x, self.max_len_), _linspace(x, x.shape[0]), x))
      X_seq = out
    if isinstance(X_seq, list):
      X_seq = np.stack(X_seq, 0) if is_np else torch.stack(X_seq, 0)
    return X_seq


class SequenceAugmentor(BaseEstimator, TransformerMixin):
  """Normalisation, lead-lag, time and basepoint augmentation of sequences
  [n, L, d] (ksig/preprocessing.py:71-138), in the reference's order, with a
  final interpolation to ``max_len`` points if the result is longer."""

  def __init__(self, add_time: bool = True, lead_lag: bool = True,
               basepoint: bool = True, normalize: bool = True,
               max_time: float = 1., max_len: Optional[int] = None):
    self.add_time = add_time
    self.lead_lag = lead_lag
    self.basepoint = basepoint
    self.normalize = normalize
    self.max_time = max_time
    self.max_len = max_len

  def fit(self, X_seq, y=None) -> 'SequenceAugmentor':
    if self.normalize:
      self.scale_ = X_seq.max()
    return self

  def transform(self, X_seq):
    is_np = isinstance(X_seq, np.ndarray)
    if self.normalize:
      # the reference divides in place (ksig/preprocessing.py:116-117)
      X_seq /= self.scale_
    if self.lead_lag:
      if is_np:
        X_seq = np.repeat(X_seq, 2, axis=1)
        X_seq = np.concatenate((X_seq[:, 1:], X_seq[:, :-1]), axis=-1)
      else:
        X_seq = torch.repeat_interleave(X_seq, 2, dim=1)
        X_seq = torch.cat((X_seq[:, 1:], X_seq[:, :-1]), dim=-1)
    if self.add_time and self.max_time > 1e-6:
      n, L = X_seq.shape[:2]
      if is_np:
        t = np.linspace(0., self.max_time, L)
        X_seq = np.concatenate((np.tile(t[None, :, None], [n, 1, 1]), X_seq), axis=-1)
      else:
        t = torch.linspace(0., self.max_time, L, device=X_seq.device, dtype=X_seq.dtype)
        X_seq = torch.cat((t[None, :, None].expand(n, L, 1), X_seq), dim=-1)
    if self.basepoint:
      if is_np:
        X_seq = np.concatenate((np.zeros_like(X_seq[:, :1]), X_seq), axis=1)
      else:
        X_seq = torch.cat((torch.zeros_like(X_seq[:, :1]), X_seq), dim=1)
    if self.max_len is not None and X_seq.shape[1] > self.max_len:
      cur = _linspace(X_seq, X_seq.shape[1])
      tgt = _linspace(X_seq, self.max_len)
      stack = np.stack if is_np else torch.stack
      X_seq = stack([_interp(tgt, cur, x) for x in X_seq], 0)
    return X_seq
