"""Host-side length-bucketed schedule of the row-stream kernel (no GPU)."""

import numpy as np
import pytest

from ksig_amd._ops import stream_schedule


def _check(lx, ly, slots, maxp):
  order, wb = stream_schedule(lx, ly, slots, maxp)
  n = len(lx)
  assert order.dtype == np.int32 and wb.dtype == np.int32
  np.testing.assert_array_equal(np.sort(order), np.arange(n))
  assert wb[0] == 0 and wb[-1] == n
  sizes = np.diff(wb)
  assert sizes.min() >= 1 and sizes.max() <= maxp
  cols = np.minimum(lx, ly)[order]
  rows = np.maximum(lx, ly)[order]
  for b, e in zip(wb[:-1], wb[1:]):
    assert np.all(np.diff(cols[b:e]) <= 0)  # shorter side, longest first
  # slices listed by sweep length (rows x widest pair's columns), longest first
  steps = [int(np.sum(np.where(cols[b:e] >= 2, rows[b:e], 0))) * max(int(cols[b]), 68)
           for b, e in zip(wb[:-1], wb[1:])]
  assert all(steps[k] >= steps[k + 1] for k in range(len(steps) - 1))
  return order, wb


def test_uniform_lengths_one_round_of_equal_slices():
  n = 131328
  order, wb = _check(np.full(n, 200), np.full(n, 200), 1792, 128)
  sizes = np.diff(wb)
  assert len(sizes) == 1792
  assert sizes.max() - sizes.min() <= 1


def test_more_pairs_than_table_capacity_uses_balanced_rounds():
  n = 300000
  order, wb = _check(np.full(n, 20), np.full(n, 20), 1024, 128)
  sizes = np.diff(wb)
  assert len(sizes) % 1024 == 0
  assert sizes.max() - sizes.min() <= 1


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_uneven_lengths_balance_work(seed):
  rng = np.random.default_rng(seed)
  n = 4096
  lx = rng.integers(32, 1025, n)
  ly = rng.integers(32, 1025, n)
  order, wb = _check(lx, ly, 512, 128)
  cost = (np.maximum(lx, ly) * np.maximum(np.minimum(lx, ly), 68))[order].astype(float)
  per = np.add.reduceat(cost, wb[:-1])
  # slices hold whole pairs: each is within one pair of the mean work
  assert per.max() <= per.mean() + cost.max()
  assert len(per) <= 512


def test_degenerate_and_tiny_lists():
  _check(np.array([1]), np.array([5]), 64, 128)
  _check(np.array([3, 1, 2]), np.array([2, 2, 9]), 1, 128)
  order, wb = stream_schedule(np.zeros(0, np.int64), np.zeros(0, np.int64), 8, 128)
  assert len(order) == 0 and list(wb) == [0]


@pytest.mark.parametrize("seed", [0, 1])
def test_two_wave_schedule_marks_long_slices(seed):
  """Two-wave slices: the longest sweeps (above the mean work per resident
  wave by the factor, with enough columns and rows) each take a block,
  marked by an empty slice right after them at an odd index; the other
  slices pair up in blocks; every pair appears once."""
  from ksig_amd import _ops
  rng = np.random.default_rng(seed)
  n = 4096
  lx = rng.integers(32, 1025, n)
  ly = rng.integers(32, 1025, n)
  for resident in (2048, 256):
    order, wb = stream_schedule(lx, ly, 4 * resident, 128, min_cols=96, two_wave=True,
                                resident=resident)
    np.testing.assert_array_equal(np.sort(order), np.arange(n))
    sizes = np.diff(wb)
    empty = np.nonzero(sizes == 0)[0]
    assert (len(empty) > 0) == (resident == 2048)  # 256 waves: no sweep is that long
    assert np.all(empty % 2 == 1) and np.all(sizes[empty - 1] > 0)
    cols = np.minimum(lx, ly)[order]
    rows = np.maximum(lx, ly)[order]
    steps = np.array([rows[b:e].sum() * max(int(cols[b]), 96) if e > b else 0
                      for b, e in zip(wb[:-1], wb[1:])])
    mean = steps.sum() / resident
    tw = empty - 1
    assert np.all(steps[tw] > _ops.TWO_WAVE_FACTOR * mean)
    assert np.all(cols[wb[tw]] >= _ops.TWO_WAVE_MIN_COLS)
    solo = np.setdiff1d(np.arange(len(sizes)), np.concatenate([tw, empty]))
    long_solo = (steps[solo] > _ops.TWO_WAVE_FACTOR * mean) & \
        (cols[wb[solo]] >= _ops.TWO_WAVE_MIN_COLS) & \
        (np.add.reduceat(rows, wb[:-1])[solo] >= _ops.TWO_WAVE_MIN_ROWS)
    assert not np.any(long_solo)
  # without the flag nothing changes
  o1, w1 = stream_schedule(lx, ly, 8192, 128, min_cols=96)
  assert np.all(np.diff(w1) > 0)
