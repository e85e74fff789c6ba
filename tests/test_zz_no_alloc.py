"""GPU test of the allocation-free entry point
ksig_signature_levels_batched_noalloc (the batched row stream with its band
FIFO in the library's static device scratch), called through the C ABI.

Its output must equal that of ksig_signature_levels_batched with a caller
workspace on the same schedule (the same kernel on the same inputs: exactly
equal values), on two calls in a row, the second with a smaller input, and
match the FP64 oracle within the float32 bar of the DP kernels (1e-5
normwise per pair).  A schedule whose FIFO exceeds the scratch is refused
with ValueError before anything is launched."""

import ctypes

import numpy as np
import pytest
import torch

from ksig_amd import _lib, _ops
from oracle import sigkern as O
from tests.gpu_util import level_err, to_dev

pytestmark = pytest.mark.gpu
TOL32 = 1e-5


def _ragged(rng, lens, d):
  seqs = [O.brownian(rng, 1, int(L), d)[0].astype(np.float32) for L in lens]
  return seqs, _ops.SeqSet.ragged([to_dev(s) for s in seqs])


def _problem(seed, n_seq, n_pairs, lmin, lmax, d):
  rng = np.random.default_rng(seed)
  lx = rng.integers(lmin, lmax + 1, n_seq)
  ly = rng.integers(lmin, lmax + 1, n_seq)
  xs, X = _ragged(rng, lx, d)
  ys, Y = _ragged(rng, ly, d)
  pr = np.stack([rng.integers(0, n_seq, n_pairs), rng.integers(0, n_seq, n_pairs)],
                -1).astype(np.int32)
  return xs, ys, X, Y, pr


def _call(entry, X, Y, pairs, sched, n_levels, order, sigma, workspace=None):
  lib = _lib.load()
  so, wb, max_slice, gf = sched
  assert gf
  n_waves = wb.shape[0] - 1
  out = torch.full((pairs.shape[0], n_levels + 1), float("nan"), device=X.data.device,
                   dtype=torch.float64)
  args = [X.data.data_ptr(), X.off.data_ptr(), X.len.data_ptr(), X.max_len,
          Y.data.data_ptr(), Y.off.data_ptr(), Y.len.data_ptr(), Y.max_len, X.d,
          pairs.data_ptr(), pairs.shape[0], _lib.STATIC_RBF, sigma, n_levels, order,
          so.data_ptr(), wb.data_ptr(), n_waves, max_slice]
  if entry == "batched":
    args += [workspace.data_ptr(), workspace.numel()]
  args += [out.data_ptr(), _lib.stream_handle(X.data.device)]
  fn = (lib.ksig_signature_levels_batched if entry == "batched"
        else lib.ksig_signature_levels_batched_noalloc)
  _lib.check(fn(*args))
  return out


@pytest.mark.parametrize("nlev,order", [(8, 2), (5, 1)])
def test_noalloc_matches_workspace_entry_twice(nlev, order):
  lib = _lib.load()
  assert lib.ksig_noalloc_scratch_bytes() == 64 << 20
  sigma = 4.0
  # call 1: 160 pairs, lengths up to 300; call 2: a smaller input (40 pairs,
  # lengths up to 120), so stale FIFO contents of call 1 sit in the scratch
  for seed, n_seq, n_pairs, lmax in ((31, 60, 160, 300), (32, 20, 40, 120)):
    xs, ys, X, Y, pr = _problem(seed, n_seq, n_pairs, 2, lmax, 16)
    pairs = to_dev(pr)
    sched = _ops.make_schedule(X, Y, pairs, _lib.STATIC_RBF, nlev, order, global_fifo=True,
                               max_slice=32, oversub=2)
    _, _, per_wave = _ops.stream_slots(X, Y, _lib.STATIC_RBF, nlev,
                                       nlev if order >= nlev else order, sched[2], True)
    n_waves = sched[1].shape[0] - 1
    assert per_wave * n_waves <= lib.ksig_noalloc_scratch_bytes()
    ws = torch.empty(per_wave * n_waves, device=X.data.device, dtype=torch.uint8)
    ref_out = _call("batched", X, Y, pairs, sched, nlev, order, sigma, ws)
    got = _call("noalloc", X, Y, pairs, sched, nlev, order, sigma)
    torch.cuda.synchronize()
    assert torch.equal(got, ref_out), float((got - ref_out).abs().max())
    sample = range(0, n_pairs, max(1, n_pairs // 12))
    ref64 = np.stack([O.pair_levels(xs[pr[p, 0]].astype(np.float64),
                                    ys[pr[p, 1]].astype(np.float64), nlev, order, True, "rbf",
                                    sigma) for p in sample])
    assert level_err(got.cpu().numpy()[list(sample)].T, ref64.T) < TOL32


def test_noalloc_refuses_oversized_fifo():
  lib = _lib.load()
  xs, ys, X, Y, pr = _problem(33, 8, 8, 2, 40, 16)
  pairs = to_dev(pr)
  # a schedule of one pair per wave over a long column side: the FIFO of
  # 4096 waves x 1000 columns exceeds 64 MiB, so the call must fail before
  # launching (the kernel never runs: the pair list holds 8 pairs, the rest
  # of the waves would find empty slices)
  wb = to_dev(np.minimum(np.arange(4097), 8).astype(np.int32))
  out = torch.zeros((8, 6), device=X.data.device, dtype=torch.float64)
  rc = lib.ksig_signature_levels_batched_noalloc(
      X.data.data_ptr(), X.off.data_ptr(), X.len.data_ptr(), 1000,
      Y.data.data_ptr(), Y.off.data_ptr(), Y.len.data_ptr(), 1000, 16,
      pairs.data_ptr(), 8, _lib.STATIC_RBF, 4.0, 5, 1, None, wb.data_ptr(), 4096, 1,
      out.data_ptr(), _lib.stream_handle(X.data.device))
  assert rc == 1001
  assert b"scratch" in lib.ksig_last_error()
  torch.cuda.synchronize()
  assert float(out.abs().max()) == 0.0
