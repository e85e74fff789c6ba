"""Host-side parts of the row-stream autotuner (no GPU): the problem
signature its choices are kept under, the JSON cache, and the oversubscribed
schedule it times."""

import numpy as np

from ksig_amd import tune
from ksig_amd._ops import stream_schedule


def test_signature_key_separates_problems_and_ignores_pair_order():
  rng = np.random.default_rng(0)
  lx = rng.integers(32, 1025, 4096)
  ly = rng.integers(32, 1025, 4096)
  k = tune.signature_key(lx, ly, 16, 0, 8, 2)
  perm = rng.permutation(4096)
  assert tune.signature_key(lx[perm], ly[perm], 16, 0, 8, 2) == k
  # X/Y swapped per pair: same shorter / longer sides
  assert tune.signature_key(ly, lx, 16, 0, 8, 2) == k
  assert tune.signature_key(lx, ly, 8, 0, 8, 2) != k
  assert tune.signature_key(lx, ly, 16, 1, 8, 2) != k
  assert tune.signature_key(lx, ly, 16, 0, 7, 2) != k
  assert tune.signature_key(lx, ly, 16, 0, 8, 1) != k
  assert tune.signature_key(lx[:1000], ly[:1000], 16, 0, 8, 2) != k
  assert tune.signature_key(lx // 2, ly // 2, 16, 0, 8, 2) != k
  # order normalisation: order >= n_levels and order <= 0 are the same problem
  assert tune.signature_key(lx, ly, 16, 0, 8, 8) == tune.signature_key(lx, ly, 16, 0, 8, -1)
  assert tune.signature_key(np.zeros(0), np.zeros(0), 16, 0, 8, 2).endswith("p0")


def test_cache_round_trip(tmp_path):
  tune.clear()
  cfg = tune.StreamConfig(True, 64, 2)
  tune._cache["k1"] = cfg
  tune._cache["k2"] = tune.StreamConfig(False, 128, 1)
  path = str(tmp_path / "tune.json")
  tune.save(path)
  tune.clear()
  tune.load(path)
  assert tune._cache["k1"] == cfg
  assert tune._cache["k2"] == tune.StreamConfig(False, 128, 1)
  tune.clear()


def test_oversubscribed_schedule_covers_every_pair_once():
  rng = np.random.default_rng(4)
  n = 4096
  lx = rng.integers(32, 1025, n)
  ly = rng.integers(32, 1025, n)
  for ov in (1, 2, 4):
    order, wb = stream_schedule(lx, ly, 512 * ov, 64)
    np.testing.assert_array_equal(np.sort(order), np.arange(n))
    sizes = np.diff(wb)
    assert sizes.min() >= 1 and sizes.max() <= 64
    assert len(sizes) >= min(n // 64, 512 * ov) and len(sizes) <= 512 * ov + n // 64 + 1


def test_rws_groups_partition_the_series():
  from ksig_amd._ops import rws_groups
  rng = np.random.default_rng(5)
  for n_seq, n_w in [(512, 100), (1, 300), (4096, 100), (3, 1), (7, 2)]:
    lens = rng.integers(1, 33, n_w)
    g = rws_groups(lens, n_seq)
    assert g[0] == 0 and g[-1] == n_w and np.all(np.diff(g) > 0)
    rows = np.diff(np.concatenate([[0], np.cumsum(lens)])[g])
    n_g = len(g) - 1
    # enough groups to fill the chip where the rows allow, at least a band each
    assert n_g == 1 or rows.min() >= 1
    assert n_g <= max(1, lens.sum() // 64) and n_g <= n_w
  assert list(rws_groups(np.zeros(0, np.int64), 4)) == [0]
