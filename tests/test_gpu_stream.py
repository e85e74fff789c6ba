"""GPU parity of the row-stream DP kernel (float32, difference=True): many
pairs stacked in one wave, caller schedules, uneven lengths, every feature
bucket, against the FP64 oracle on the float32-rounded inputs.

Tolerance: 1e-5 normwise relative error over each pair's level vector (the
float32 bar of the DP kernels)."""

import math

import numpy as np
import pytest
import torch

import ksig_amd as ksig
from ksig_amd import _lib, _ops
from oracle import sigkern as O
from tests.gpu_util import dev, level_err, to_dev

pytestmark = pytest.mark.gpu
TOL32 = 1e-5


def _ragged(rng, lens, d):
  seqs = [O.brownian(rng, 1, max(L, 1), d)[0].astype(np.float32) for L in lens]
  return seqs, _ops.SeqSet.ragged([to_dev(s) for s in seqs])


def _oracle(xs, ys, pairs, n_levels, order, kind, param):
  out = []
  for a, b in pairs:
    xa = xs[a].astype(np.float64)
    yb = ys[b].astype(np.float64)
    out.append(O.pair_levels(xa, yb, n_levels, order, True, kind, param))
  return np.stack(out)


def _check(got, ref):
  got = np.asarray(got)
  # K_0 = 1 exactly; empty grids give exact zeros
  np.testing.assert_array_equal(got[:, 0], 1.0)
  assert level_err(got.T, ref.T) < TOL32


@pytest.mark.parametrize("kind,param", [("rbf", 1.5), ("linear", 0.7)])
@pytest.mark.parametrize("nlev,order", [(5, 1), (4, 2)])
def test_schedules_with_many_pairs_per_wave(kind, param, nlev, order):
  rng = np.random.default_rng(21)
  lx = rng.integers(1, 41, 40)
  ly = rng.integers(1, 41, 30)
  lx[:3] = [1, 2, 3]
  ly[:2] = [1, 2]
  xs, X = _ragged(rng, lx, 3)
  ys, Y = _ragged(rng, ly, 3)
  n_pairs = 600
  pr = np.stack([rng.integers(0, 40, n_pairs), rng.integers(0, 30, n_pairs)], -1).astype(np.int32)
  pairs = to_dev(pr)
  code = _lib.STATIC_RBF if kind == "rbf" else _lib.STATIC_LINEAR
  ref = _oracle(xs, ys, pr, nlev, order, kind, param)
  # 1) length-bucketed schedule from the capacity query
  got = _ops.signature_levels_batched(X, Y, pairs, nlev, order, code, param)
  _check(got.cpu().numpy(), ref)
  # 2) a random processing order in full 128-pair slices
  perm = rng.permutation(n_pairs).astype(np.int32)
  wb = np.arange(0, n_pairs + 128, 128).clip(max=n_pairs).astype(np.int32)
  got = _ops.signature_levels_batched(X, Y, pairs, nlev, order, code, param,
                                      (to_dev(perm), to_dev(wb)))
  _check(got.cpu().numpy(), ref)
  # 3) identity order, ragged slice sizes including single pairs
  wb = np.unique(np.concatenate([[0, 1, 2, 130, 131], rng.integers(0, n_pairs, 12),
                                 np.arange(0, n_pairs, 100), [n_pairs]])).astype(np.int32)
  assert np.diff(wb).max() <= 128
  got = _ops.signature_levels_batched(X, Y, pairs, nlev, order, code, param,
                                      (None, to_dev(wb)))
  _check(got.cpu().numpy(), ref)


@pytest.mark.parametrize("d", [1, 5, 8, 13, 16, 24, 32])
def test_uneven_lengths_across_bands(d):
  """Lengths that cross the 68-column minimum band, 64R-row bands and R-row
  padding, stacked several to a wave."""
  rng = np.random.default_rng(300 + d)
  lens = [2, 3, 67, 68, 69, 70, 129, 130, 150, 257, 5, 40]
  xs, X = _ragged(rng, lens, d)
  n = len(lens)
  pr = np.array([[a, b] for a in range(n) for b in range(n) if (a + 2 * b) % 3 == 0],
                dtype=np.int32)
  sig = math.sqrt(d)
  ref = _oracle(xs, xs, pr, 5, 1, "rbf", sig)
  for wb in (None, np.array([0, 7, 8, 40, len(pr)], dtype=np.int32)):
    sched = None if wb is None else (None, to_dev(wb))
    got = _ops.signature_levels_batched(X, X, to_dev(pr), 5, 1, _lib.STATIC_RBF, sig, sched)
    _check(got.cpu().numpy(), ref)


@pytest.mark.parametrize("d,L,n", [(8, 24, 70), (16, 20, 60), (32, 18, 55)])
def test_dense_gram_more_pairs_than_resident_waves(d, L, n):
  """Full Gram through SignatureKernel: more pairs than resident waves, so
  every wave stacks several pairs (equal slices chosen by the library)."""
  rng = np.random.default_rng(40 + d)
  X = O.brownian(rng, n, L, d).astype(np.float32)
  sig = math.sqrt(d)
  su = ksig.kernels.SignatureKernel(5, normalize=False,
                                    static_kernel=ksig.static.kernels.RBFKernel(sig))
  lev = su._levels(to_dev(X), to_dev(X)).cpu().numpy()
  ref = O.signature_kern(O.pairwise_gram(X.astype(np.float64), None, "rbf", sig), 5, 1,
                         True, True)
  assert level_err(lev, ref) < TOL32
  slots, _, _ = _ops.stream_slots(_ops.SeqSet.dense(to_dev(X)), _ops.SeqSet.dense(to_dev(X)),
                               _lib.STATIC_RBF, 5, 1)
  assert 0 < slots < n * n


def test_ragged_signature_kernel_uses_schedule():
  """Uneven float32 sets through the host op pick the batched row stream and
  agree with the per-pair oracle."""
  rng = np.random.default_rng(77)
  lens_x = [33, 90, 7, 64, 120]
  lens_y = [50, 2, 75]
  xs, X = _ragged(rng, lens_x, 4)
  ys, Y = _ragged(rng, lens_y, 4)
  pairs = _ops.dense_pairs(len(lens_x), len(lens_y), dev())
  got = _ops.signature_levels(X, Y, pairs, 6, 2, True, _lib.STATIC_RBF, 1.3)
  ref = _oracle(xs, ys, pairs.cpu().numpy(), 6, 2, "rbf", 1.3)
  _check(got.cpu().numpy(), ref)


@pytest.mark.parametrize("d,nlev,order", [(8, 5, 1), (16, 8, 2), (32, 6, 1), (16, 4, 4)])
def test_global_memory_fifo(d, nlev, order):
  """Band hand-over through the global-memory workspace (the long-sequence
  layout): lengths below and above the 96-column minimum band, several pairs
  per wave, random order, against the oracle and the LDS-FIFO layout."""
  rng = np.random.default_rng(500 + d + nlev)
  lens = [2, 5, 40, 95, 96, 97, 150, 200, 260, 31]
  xs, X = _ragged(rng, lens, d)
  n = len(lens)
  pr = np.array([[a, b] for a in range(n) for b in range(n)], dtype=np.int32)
  perm = rng.permutation(len(pr)).astype(np.int32)
  sig = math.sqrt(d)
  ref = _oracle(xs, xs, pr, nlev, order, "rbf", sig)
  pairs = to_dev(pr)
  wb = np.array([0, 1, 9, 30, 64, 99, len(pr)], dtype=np.int32)
  sched_g = (to_dev(perm), to_dev(wb), int(np.diff(wb).max()), True)
  got_g = _ops.signature_levels_batched(X, X, pairs, nlev, order, _lib.STATIC_RBF, sig, sched_g)
  _check(got_g.cpu().numpy(), ref)
  sched_l = (to_dev(perm), to_dev(wb), int(np.diff(wb).max()), False)
  got_l = _ops.signature_levels_batched(X, X, pairs, nlev, order, _lib.STATIC_RBF, sig, sched_l)
  _check(got_l.cpu().numpy(), ref)
  assert level_err(got_g.cpu().numpy().T, got_l.cpu().numpy().T) < 1e-6


def test_c3_shape_sample_uses_workspace_fifo():
  """A 48-pair sample of the uneven-length workload (lengths in [32, 1024],
  d = 16, 8 levels, order 2, sigma = 4) through the length-bucketed schedule;
  long columns select the global-memory FIFO."""
  rng = np.random.default_rng(3)
  n = 48
  lx = rng.integers(32, 1025, n)
  ly = rng.integers(32, 1025, n)
  xs = [(O.brownian(rng, 1, L, 16)[0] + 0.1 * rng.standard_normal(16)).astype(np.float32)
        for L in lx]
  ys = [(O.brownian(rng, 1, L, 16)[0] + 0.1 * rng.standard_normal(16)).astype(np.float32)
        for L in ly]
  X = _ops.SeqSet.ragged([to_dev(s) for s in xs])
  Y = _ops.SeqSet.ragged([to_dev(s) for s in ys])
  pr = np.stack([np.arange(n), np.arange(n)], -1).astype(np.int32)
  sched = _ops.make_schedule(X, Y, to_dev(pr), _lib.STATIC_RBF, 8, 2)
  assert sched[3] is True
  got = _ops.signature_levels_batched(X, Y, to_dev(pr), 8, 2, _lib.STATIC_RBF, 4.0, sched)
  idx = rng.choice(n, 6, replace=False)
  ref = _oracle([xs[i] for i in idx], [ys[i] for i in idx],
                [(k, k) for k in range(len(idx))], 8, 2, "rbf", 4.0)
  _check(got.cpu().numpy()[idx], ref)


def test_autotuned_launch_choices_agree_with_oracle():
  """Every candidate the autotuner times (FIFO placement, slice bound,
  slices per resident wave) computes the same levels, and the kept choice
  is the one signature_levels uses afterwards."""
  from ksig_amd import tune
  rng = np.random.default_rng(8)
  n = 96
  lx = rng.integers(2, 300, n)
  ly = rng.integers(2, 300, n)
  xs, X = _ragged(rng, lx, 16)
  ys, Y = _ragged(rng, ly, 16)
  pr = np.stack([np.arange(n), rng.permutation(n)], -1).astype(np.int32)
  pairs = to_dev(pr)
  cands = tune.candidates(X, Y, _lib.STATIC_RBF, 4, 2)
  assert len(cands) >= 3
  results = []
  for cfg in cands:
    sched = _ops.make_schedule(X, Y, pairs, _lib.STATIC_RBF, 4, 2, cfg.global_fifo,
                               cfg.max_slice, cfg.oversub, cfg.two_wave)
    assert sched[3] == cfg.global_fifo and sched[2] <= cfg.max_slice
    results.append(_ops.signature_levels_batched(X, Y, pairs, 4, 2, _lib.STATIC_RBF, 2.0,
                                                 sched).cpu().numpy())
  idx = rng.choice(n, 8, replace=False)
  ref = _oracle(xs, ys, pr[idx], 4, 2, "rbf", 2.0)
  for got in results:
    _check(got[idx], ref)
    assert level_err(got.T, results[0].T) < 1e-6
  tune.clear()
  best, times = tune.autotune(X, Y, pairs, 4, 2, _lib.STATIC_RBF, 2.0, reps=1)
  assert best in times and set(times) == set(cands)
  assert tune.lookup(X, Y, pairs, _lib.STATIC_RBF, 4, 2) == best
  got = _ops.signature_levels(X, Y, pairs, 4, 2, True, _lib.STATIC_RBF, 2.0).cpu().numpy()
  _check(got[idx], ref)
  tune.clear()


def _sched(order, wave_beg, global_fifo=True):
  wb = np.asarray(wave_beg, np.int32)
  return (to_dev(np.asarray(order, np.int32)), to_dev(wb), int(np.diff(wb).max()), global_fifo)


@pytest.mark.parametrize("kind,param,d,nlev,order", [("rbf", 4.0, 16, 8, 2),
                                                     ("linear", 0.5, 8, 5, 1),
                                                     ("rbf", 3.0, 32, 3, 3)])
def test_two_wave_slices_match_one_wave(kind, param, d, nlev, order):
  """Slices split band by band between the two waves of a block (an empty
  slice after a slice marks it): stacked pairs of uneven lengths, pairs that
  straddle band boundaries, empty grids, odd and even band counts, a slice
  whose columns are short of the two-wave minimum (padded), and a final
  block with one wave, against the same slices swept by one wave and
  against the FP64 oracle."""
  rng = np.random.default_rng(70 + d)
  lx = np.array([700, 300, 2, 451, 260, 240, 90, 1, 333, 517, 128, 64, 300, 250])
  ly = np.array([500, 610, 400, 260, 700, 300, 40, 5, 400, 290, 230, 64, 128, 250])
  n = len(lx)
  xs, X = _ragged(rng, lx, d)
  ys, Y = _ragged(rng, ly, d)
  pr = np.stack([np.arange(n), np.arange(n)], -1).astype(np.int32)
  pairs = to_dev(pr)
  code = _lib.STATIC_RBF if kind == "rbf" else _lib.STATIC_LINEAR
  # two-wave slices: [0, 3), [3, 6), [12, 14); one-wave blocks: [6, 8) + [8, 10),
  # [10, 12) alone at the end (odd slice count: the last block has one wave)
  seq = [(0, 3), None, (3, 6), None, (12, 14), None, (6, 8), (8, 10), (10, 12)]
  o2, wb2 = [], [0]
  for sl in seq:
    if sl is not None:
      o2 += list(range(*sl))
    wb2.append(len(o2))
  one = [0]
  for sl in seq:
    if sl is not None:
      one.append(one[-1] + sl[1] - sl[0])
  got2 = _ops.signature_levels_batched(X, Y, pairs, nlev, order, code, param,
                                       _sched(o2, wb2)).cpu().numpy()
  got1 = _ops.signature_levels_batched(X, Y, pairs, nlev, order, code, param,
                                       _sched(o2, one)).cpu().numpy()
  assert np.all(np.isfinite(got2))
  assert level_err(got2.T, got1.T) < 1e-6
  ref = _oracle(xs, ys, pr, nlev, order, kind, param)
  _check(got2, ref)
  # the schedule builder makes two-wave slices of the longest sweeps
  so, wb = _ops.stream_schedule(lx, ly, 2048, 128, min_cols=96, two_wave=True)
  sizes = np.diff(wb)
  assert np.any(sizes == 0) and np.all(np.nonzero(sizes == 0)[0] % 2 == 1)
  got3 = _ops.signature_levels_batched(X, Y, pairs, nlev, order, code, param,
                                       _sched(so, wb)).cpu().numpy()
  _check(got3, ref)


def test_two_wave_launch_replays_from_a_graph():
  """The batched row stream with two-wave slices and its global FIFO
  workspace captured in a HIP graph: every replay (with eager launches of
  the same op in between) equals the eager result bit for bit."""
  rng = np.random.default_rng(90)
  n = 40
  lx = rng.integers(200, 1025, n)
  ly = rng.integers(200, 1025, n)
  xs = [(O.brownian(rng, 1, L, 16)[0]).astype(np.float32) for L in lx]
  ys = [(O.brownian(rng, 1, L, 16)[0]).astype(np.float32) for L in ly]
  X = _ops.SeqSet.ragged([to_dev(s) for s in xs])
  Y = _ops.SeqSet.ragged([to_dev(s) for s in ys])
  pairs = to_dev(np.stack([np.arange(n), np.arange(n)], -1).astype(np.int32))
  sched = _ops.make_schedule(X, Y, pairs, _lib.STATIC_RBF, 8, 2)
  assert sched[3] and int((np.diff(sched[1].cpu().numpy()) == 0).sum()) > 0
  run = lambda: _ops.signature_levels_batched(X, Y, pairs, 8, 2, _lib.STATIC_RBF, 4.0, sched)
  eager = run()
  torch.cuda.synchronize()
  stream = torch.cuda.current_stream()
  side = torch.cuda.Stream()
  side.wait_stream(stream)
  with torch.cuda.stream(side):
    run()
  stream.wait_stream(side)
  g = torch.cuda.CUDAGraph()
  with torch.cuda.graph(g):
    out = run()
  for _ in range(3):
    g.replay()
    torch.cuda.synchronize()
    assert torch.equal(out, eager)
    again = run()
    torch.cuda.synchronize()
    assert torch.equal(again, eager)
  idx = [0, 7, 19]
  ref = _oracle([xs[i] for i in idx], [ys[i] for i in idx],
                [(k, k) for k in range(len(idx))], 8, 2, "rbf", 4.0)
  _check(eager.cpu().numpy()[idx], ref)
