"""Shared helpers of the GPU parity tests."""

import numpy as np
import torch


def dev():
  return torch.device("cuda", 0)


def to_dev(a, dtype=None):
  t = torch.as_tensor(np.asarray(a), device=dev())
  return t if dtype is None else t.to(dtype)


def normwise(got, ref):
  got = np.asarray(got, dtype=np.float64)
  ref = np.asarray(ref, dtype=np.float64)
  return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300))


def level_err(got, ref):
  """max over pairs of the normwise error of the level vector."""
  got = np.asarray(got, dtype=np.float64)
  ref = np.asarray(ref, dtype=np.float64)
  num = np.max(np.abs(got - ref), axis=0)
  den = np.maximum(np.max(np.abs(ref), axis=0), 1e-300)
  return float(np.max(num / den))
