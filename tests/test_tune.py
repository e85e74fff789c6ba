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


def test_schedule_cache_skips_repeated_planning(monkeypatch):
  """A repeated call with the same pair tensor reuses the built schedule
  without touching the pair list; a new tensor with the same pair lengths
  reuses it after one content hash; changed lengths rebuild it."""
  import torch
  from ksig_amd import _ops
  tune.clear()
  builds = []

  def fake_make_schedule(X, Y, pairs, kind, n_levels, order, *cfg):
    builds.append(cfg)
    return ("schedule", len(builds))
  monkeypatch.setattr(_ops, "make_schedule", fake_make_schedule)
  rng = np.random.default_rng(6)
  lx = rng.integers(2, 50, 30)
  ly = rng.integers(2, 50, 20)
  X = _ops.SeqSet(torch.zeros((int(lx.sum()), 4)), torch.zeros(30, dtype=torch.int64),
                  torch.as_tensor(lx, dtype=torch.int32), int(lx.max()), 30, lx)
  Y = _ops.SeqSet(torch.zeros((int(ly.sum()), 4)), torch.zeros(20, dtype=torch.int64),
                  torch.as_tensor(ly, dtype=torch.int32), int(ly.max()), 20, ly)
  pairs = torch.as_tensor(np.stack([rng.integers(0, 30, 100), rng.integers(0, 20, 100)], -1),
                          dtype=torch.int32)
  s1 = tune.tuned_schedule(X, Y, pairs, 0, 5, 1)
  s2 = tune.tuned_schedule(X, Y, pairs, 0, 5, 1)
  assert s1 == s2 and len(builds) == 1
  s3 = tune.tuned_schedule(X, Y, pairs.clone(), 0, 5, 1)   # same content
  assert s3 == s1 and len(builds) == 1
  pairs[:, 0] = (pairs[:, 0] + 1) % 30                     # in place: new lengths
  tune.tuned_schedule(X, Y, pairs, 0, 5, 1)
  assert len(builds) == 2
  tune.tuned_schedule(X, Y, pairs, 0, 6, 1)                # another problem
  assert len(builds) == 3
  tune.clear()                                              # kept choices changed
  tune.tuned_schedule(X, Y, pairs, 0, 6, 1)
  assert len(builds) == 4
  tune.clear()
