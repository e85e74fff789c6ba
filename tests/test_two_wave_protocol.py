"""Host model of the two-wave slice hand-over of the row stream
(ksig_amd/csrc/dp_stream.h, coop_sync): two waves sweep the even and the odd
bands of one slice; lane 63 of each writes its band output to its own FIFO
region (slot = column, at step band*Pc + 63 + column), lane 0 of the other
reads it two steps ahead; every K steps a wave publishes its step count and
waits until the other's published count reaches

    wave 0: t + K + 65 - Pc        wave 1: t + K + 65

The model runs both waves with random step times and checks, from the counts
each wave has seen published (all it may rely on), that every FIFO slot is
read after it was written (RAW), and written again only after it was read
(WAR), and that the two never wait for each other at once — for every band
pass of at least 2K + 130 slots.  Below that bound the model finds the
deadlock the bound excludes."""

import numpy as np
import pytest

K = 32          # kStreamCoopSync
MIN_COLS = 224  # kStreamCoopMinCols
INF = 1 << 40


def _need(v, t, pc):
  return t + K + 65 - pc if v == 0 else t + K + 65


def _simulate(nb, pc, rng, stall=0.0):
  """Returns None on success, or a string naming the first violation."""
  nbw = [(nb + 1) // 2, nb // 2]
  total = [n * pc + 63 for n in nbw]
  pub = [0, 0]            # published counts
  seen = [-1, -1]         # partner count seen at the last completed wait
  t = [0, 0]              # next step to run
  started = [False, False]  # prologue (sync at 0 + reads of steps 0, 1) done
  clock = [0.0, 0.0]
  waiting = [False, False]

  def step_time():
    d = rng.uniform(0.5, 1.5)
    if stall and rng.random() < stall:
      d += rng.uniform(10, 200)
    return d

  def try_sync(v, tv):
    pub[v] = tv
    if pub[1 - v] >= _need(v, tv, pc):
      seen[v] = pub[1 - v]
      return True
    return False

  def read(v, s):
    # lane 0 of wave v loads the input of its data step s
    k, c = divmod(s, pc)
    if k >= nbw[v]:
      return None
    if v == 0 and k == 0:
      return None  # the slice's first band: its first rows start a pair
    kp = k - 1 if v == 0 else k
    produced = kp * pc + 63 + c
    if seen[v] <= produced:
      return f"RAW: wave {v} step {s} reads slot {c} before step {produced} of the other"
    return None

  def write(v, tv):
    if tv < 63 or tv - 63 >= nbw[v] * pc:
      return None
    k, c = divmod(tv - 63, pc)
    if k == 0:
      return None
    # the previous content (band k-1) is read by the other wave at data step s
    s = (k - 1) * pc + c if v == 0 else k * pc + c
    if (k - 1 if v == 0 else k) >= nbw[1 - v]:
      return None  # nobody reads it
    read_step = max(s - 2, 0)
    if seen[v] <= read_step:
      return f"WAR: wave {v} step {tv} overwrites slot {c} before the other read it"
    return None

  while t[0] < total[0] or t[1] < total[1]:
    live = [v for v in (0, 1) if t[v] < total[v]]
    # a waiting wave retries when the other publishes
    for v in live:
      if waiting[v]:
        ok = try_sync(v, t[v] if started[v] else 0)
        waiting[v] = not ok
    runnable = [v for v in live if not waiting[v]]
    if not runnable:
      return "deadlock"
    v = min(runnable, key=lambda u: clock[u])
    if not started[v]:
      if not try_sync(v, 0):
        waiting[v] = True
        continue
      started[v] = True
      for s in (0, 1):
        err = read(v, s)
        if err:
          return err
      continue
    tv = t[v]
    if tv > 0 and tv % K == 0 and not waiting[v]:
      if not try_sync(v, tv):
        waiting[v] = True
        continue
    err = read(v, tv + 2) or write(v, tv)
    if err:
      return err
    clock[v] += step_time()
    t[v] += 1
    if t[v] == total[v]:
      pub[v] = INF
  return None


@pytest.mark.parametrize("pc", [MIN_COLS, 300, 1024])
@pytest.mark.parametrize("nb", [2, 3, 5, 8])
def test_hand_over_is_ordered_and_deadlock_free(nb, pc):
  rng = np.random.default_rng(nb * 7919 + pc)
  for trial in range(6):
    assert _simulate(nb, pc, rng, stall=0.02 * trial) is None


def test_bound_is_needed():
  """Below 2K + 130 slots both waves can wait for each other."""
  rng = np.random.default_rng(1)
  results = {_simulate(4, 2 * K + 100, rng) for _ in range(4)}
  assert "deadlock" in results
  assert MIN_COLS >= 2 * K + 130


def test_model_catches_a_short_lag(monkeypatch):
  """The waits are tight: wave 1 trailing by K + 62 instead of K + 65 steps
  (or wave 0 leading by 3 more) lets a read overtake its write."""
  import tests.test_two_wave_protocol as m
  rng = np.random.default_rng(3)
  monkeypatch.setattr(m, "_need", lambda v, t, pc: t + K + 65 - pc if v == 0 else t + K + 62)
  assert any((m._simulate(5, MIN_COLS, rng, 0.02) or "").startswith("RAW") for _ in range(5))
  monkeypatch.setattr(m, "_need", lambda v, t, pc: t + K + 62 - pc if v == 0 else t + K + 65)
  assert any((m._simulate(5, MIN_COLS, rng, 0.02) or "").startswith("RAW") for _ in range(5))
