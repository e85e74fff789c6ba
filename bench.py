"""Benchmark of the signature-kernel hot path on one MI355X.

Default workload (the headline line): the batched uneven DP signature kernel,
4096 sequence pairs whose lengths are drawn independently from U{32..1024},
d = 16, float32 random walks (cumsum of N(0, 1/L) increments plus a
per-sequence per-channel offset of 0.1 N(0, 1)), truncated signature kernel
by dynamic programming with n_levels = 8, order = 2, RBF bandwidth 4 =
sqrt(d).  A step is one persistent launch of the fused HIP row stream over all
4096 pairs (length-bucketed slices, launch choices picked by the autotuner
before the timed region), inputs already resident in HBM.

Metric: DP cell-levels per second = sum over pairs of (Lx-1)(Ly-1) x n_levels / s.

The JSON line (rank 0) also carries: the per-pair error against the FP64
oracle; the reference dataflow restated in plain FP32 GPU tensor ops (the
comparator of the >= 5x target); the matrix-core Gram variants (bf16 hi/lo
split and float32 MFMA, per-pair kernel) with their errors; a roofline object
for the row-stream kernel (HIP-event timing on the launch stream); the FP64
and FP32 NumPy oracle timed over a process pool on the host cores; the MMD^2
training step (1024 vs 1024 paths, L = 128, d = 32, 6 levels) as forward+
adjoint over forward; and the 512 x L = 200 x d = 8 symmetric Gram (5 levels,
order 1) as a secondary figure.  Other workloads: --workload c1|c2|c4|gak|rws|rfsf.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 (vector and matrix)
PEAK_HBM_GBS = 8000.0

# HBM bytes per launch of the dominant kernel at the default workload sizes,
# from two rocprofv3 --pmc passes (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the
# factors an elementwise copy of 1 GiB reads on the same counters); recorded
# with the kernel version and schedule they were measured on (c3: the
# autotuned schedule with two-wave slices, 3680 wave slots in 1840 blocks).
MEASURED_TRAFFIC_BYTES = {"c1": 38658 * 1024, "c3": (666357 * 2 + 600020) * 1024,
                          "gak": (21130 * 2 + 4104) * 1024, "rws": (13451 * 2 + 3200) * 1024,
                          "rfsf": (230276 * 2 + 3906) * 1024}


def brownian(gen: torch.Generator, n, L, d, device, dtype=torch.float32):
  inc = torch.randn((n, L, d), generator=gen, device=device, dtype=torch.float64) / math.sqrt(L)
  return torch.cumsum(inc, 1).to(dtype)


def plain_fp32_levels(X, Y, n_levels, sigma):
  """The reference dataflow (ksig/kernels.py:124-164, ksig/static/kernels.py:
  248-260, ksig/algorithms.py:47-85) restated with FP32 torch ops: materialise
  M, ΔΔ, then per level an exclusive 2-D cumsum, a multiply and a sum."""
  Xn = (X * X).sum(-1)
  Yn = (Y * Y).sum(-1)
  dot = torch.einsum("ail,bjl->abij", X, Y)
  D2 = Xn[:, None, :, None] + Yn[None, :, None, :] - 2 * dot
  M = torch.exp(-D2 / max(2 * sigma**2, 1e-12))
  M = torch.diff(torch.diff(M, dim=-2), dim=-1)
  K = [M.sum((-2, -1))]
  R = M
  for _ in range(1, n_levels):
    E = torch.cumsum(torch.cumsum(R[..., :-1, :-1], -2), -1)
    E = torch.nn.functional.pad(E, (1, 0, 1, 0))
    R = M * E
    K.append(R.sum((-2, -1)))
  return torch.stack(K, -1)


def recursion_flops(n_levels: int, order: int) -> int:
  """Multiply/add count of one cell of the level recursion (the reference's
  R-matrix updates, ksig/algorithms.py:88-140, of which order 1 is the
  special case :47-85), excluding the static kernel and ΔΔ."""
  f = 0
  for m in range(1, n_levels):
    dn = min(m + 1, order)
    f += 1 + 4 * (dn - 1) + 2 * (dn - 1) ** 2      # new R-matrix entries
  for m in range(1, n_levels + 1):
    dm = min(m, order)
    f += dm * dm                                    # level sum + column flow
    if m < n_levels:
      dn = min(m + 1, order)
      f += 1 + 2 * (dn - 1) * dm                    # prefix and side sums
  return f


def excl_cumsum(M: torch.Tensor, dims) -> torch.Tensor:
  """Exclusive cumulative sum over `dims` (ksig/utils.py:63-90)."""
  sl = [slice(None)] * M.dim()
  pad = [0, 0] * M.dim()
  for a in dims:
    a = a % M.dim()
    sl[a] = slice(0, -1)
  R = M[tuple(sl)]
  for a in dims:
    R = torch.cumsum(R, a)
  for a in dims:
    a = a % M.dim()
    pad[2 * (M.dim() - 1 - a)] = 1
  return torch.nn.functional.pad(R, pad)


def plain_fp32_levels_ho(X, Y, n_levels, order, sigma):
  """Paired batches X [B, Lx, d], Y [B, Ly, d]: the reference dataflow of the
  higher-order recursion (ksig/algorithms.py:88-140) on the materialised
  ΔΔ RBF Gram (ksig/static/kernels.py:248-260, ksig/algorithms.py:106) in
  FP32 torch ops."""
  Xn = (X * X).sum(-1)
  Yn = (Y * Y).sum(-1)
  dot = torch.bmm(X, Y.transpose(1, 2).contiguous())
  M = torch.exp(-(Xn[:, :, None] + Yn[:, None, :] - 2 * dot) / max(2 * sigma**2, 1e-12))
  M = torch.diff(torch.diff(M, dim=-2), dim=-1)
  levels = [torch.ones(M.shape[0], device=M.device, dtype=M.dtype), M.sum((-2, -1))]
  R = {(0, 0): M}
  for i in range(1, n_levels):
    dd = min(i + 1, order)
    Rn = {(0, 0): M * excl_cumsum(sum(R.values()), (-2, -1))}
    for r in range(1, dd):
      Rn[(0, r)] = M * excl_cumsum(sum(v for (a, b), v in R.items() if b == r - 1), (-2,)) / (r + 1)
      Rn[(r, 0)] = M * excl_cumsum(sum(v for (a, b), v in R.items() if a == r - 1), (-1,)) / (r + 1)
      for s2 in range(1, dd):
        Rn[(r, s2)] = M * R[(r - 1, s2 - 1)] / ((r + 1) * (s2 + 1)) if (r - 1, s2 - 1) in R \
            else torch.zeros_like(M)
    R = Rn
    levels.append(sum(v.sum((-2, -1)) for v in R.values()))
  return torch.stack(levels, -1)


# Work of one process-pool CPU baseline, set before the pool forks (the
# workers inherit it; nothing large is pickled).
_CPU_JOB: dict = {}


def _cpu_worker(wid: int):
  """One pool worker: evaluate randomly drawn units of the workload with the
  NumPy oracle on one thread until the job's time budget ends."""
  job = _CPU_JOB
  try:
    from threadpoolctl import threadpool_limits
    threadpool_limits(1)
  except Exception:  # pragma: no cover
    pass
  rng = np.random.default_rng(job["seed"] + 7919 * wid)
  t0 = time.perf_counter()
  units = 0
  work = 0.0
  done = []
  while time.perf_counter() - t0 < job["seconds"]:
    u = job["pick"](rng)
    val = job["eval"](u)
    work += job["work"](u)
    units += 1
    if job.get("keep"):
      done.append((u, val))
  return work, units, time.perf_counter() - t0, done


def host_cpu_model() -> str:
  try:
    with open("/proc/cpuinfo") as f:
      for line in f:
        if line.startswith("model name"):
          return line.split(":", 1)[1].strip()
  except OSError:  # pragma: no cover
    pass
  return "unknown"


def pool_cores() -> int:
  """Worker processes of the CPU baseline: every host core this process may
  use, at most 16 (the host share of one GPU)."""
  try:
    n = len(os.sched_getaffinity(0))
  except AttributeError:  # pragma: no cover
    n = os.cpu_count() or 1
  return max(1, min(16, n))


def cpu_pool_run(pick, evaluate, work, seconds: float, seed: int, cores: int,
                 keep: bool = False):
  """Run the oracle over a fork pool of ``cores`` single-threaded workers for
  ``seconds`` each.  Returns (work per second of the whole pool, units done,
  work per second of one worker on average, kept (unit, value) results).
  Must be called before this process initialises the GPU."""
  import multiprocessing as mp
  _CPU_JOB.clear()
  _CPU_JOB.update(pick=pick, eval=evaluate, work=work, seconds=seconds, seed=seed, keep=keep)
  ctx = mp.get_context("fork")
  t0 = time.perf_counter()
  with ctx.Pool(cores) as pool:
    parts = pool.map(_cpu_worker, range(cores))
  wall = time.perf_counter() - t0
  tot_work = sum(p[0] for p in parts)
  units = sum(p[1] for p in parts)
  per_core = float(np.mean([p[0] / p[2] for p in parts]))
  kept = [x for p in parts for x in p[3]]
  span = max(p[2] for p in parts)
  return tot_work / max(span, 1e-9), units, per_core, kept, wall


def cpu_baseline_pool(unit: str, sample: str, runs: dict, cores: int) -> dict:
  """The cpu_baseline object: FP64 over the pool is the value; the FP32 run
  and the mean per-core rate are reported beside it."""
  f64 = runs["f64"]
  res = {"value": f64[0], "unit": unit, "cores": cores, "kind": "port",
         "cpu_model": host_cpu_model(), "threads_per_worker": 1,
         "sample": sample + f"; FP64: {f64[1]} units in {f64[4]:.1f} s wall"}
  res["fp64_per_core"] = f64[2]
  if "f32" in runs:
    res["fp32_value"] = runs["f32"][0]
    res["fp32_per_core"] = runs["f32"][2]
    res["sample"] += f"; FP32: {runs['f32'][1]} units in {runs['f32'][4]:.1f} s wall"
  return res


def main():
  ap = argparse.ArgumentParser()
  ap.add_argument("--workload", default="c3",
                  choices=["c1", "c2", "c3", "c4", "gak", "rws", "rfsf"],
                  help="c3 (default, the headline line): 4096 uneven-length pairs, L in "
                       "[32, 1024], d=16, 8 levels, order 2, with the c4 forward+adjoint ratio "
                       "and the c1 figure attached; c1: 512 x L=200 x d=8 symmetric Gram; "
                       "c4: MMD^2 training step, 1024 x 1024 paths, L=128, d=32, 6 levels; "
                       "c2: signature-PDE MMD^2 of 256 vs 256 paths, L=256, d=4, dyadic 1")
  ap.add_argument("--gpus", type=int, default=1)
  ap.add_argument("--steps", type=int, default=20)
  ap.add_argument("--warmup", type=int, default=3)
  ap.add_argument("--cpu-seconds", type=float, default=10.0,
                  help="time budget of each CPU-baseline pool run (FP64, then FP32)")
  ap.add_argument("--no-cpu", action="store_true")
  ap.add_argument("--no-plain", action="store_true")
  ap.add_argument("--no-extras", action="store_true",
                  help="c3: leave out the attached c4 and c1 figures")
  ap.add_argument("--no-tune", action="store_true",
                  help="c3: skip the autotuner and time the default schedule")
  ap.add_argument("--gram", default="valu", choices=["valu", "fp32", "bf16x3"],
                  help="c1: static-kernel Gram evaluation inside the sweep")
  ap.add_argument("--no-variants", action="store_true")
  ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                  help="only run an elementwise copy of MIB MiB (5 times): the "
                       "reference point for scaling HBM byte counters")
  ap.add_argument("--kernel-probe", default=None,
                  choices=["fwd", "fwd_save", "adjoint", "adjoint_gp", "contract", "pde_fwd",
                           "pde_fwd_edges",
                           "pde_adjoint", "pde_contract", "pde_fused"],
                  help="only time one kernel of the C4 training step on one 16384-pair "
                       "chunk (128 x 128 pairs, L=128, d=32, 6 levels), --steps launches: "
                       "the unit that counter collection and the per-chunk costs refer to")
  args = ap.parse_args()
  if args.gpus != 1:
    print(f"bench.py measures one GPU only (the signature-kernel path shards by "
          f"pairs with no collective in the timed loop); --gpus {args.gpus} is not "
          f"supported in this build", file=sys.stderr)
    sys.exit(2)

  if args.copy_probe:
    dev = torch.device("cuda", 0)
    n = args.copy_probe * (1 << 20) // 4
    src = torch.ones(n, device=dev, dtype=torch.float32)
    dst = torch.empty_like(src)
    for _ in range(5):   # an elementwise kernel: plain vector loads and stores
      torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4,
                      "calls": 5}), flush=True)
    return

  if args.kernel_probe:
    if args.kernel_probe.startswith("pde_"):
      return pde_kernel_probe(args)
    return kernel_probe(args)
  if args.workload == "c3":
    return bench_c3(args)
  if args.workload == "c4":
    print(json.dumps(run_c4(args)), flush=True)
    return
  if args.workload == "c2":
    return bench_c2(args)
  if args.workload == "gak":
    return bench_gak(args)
  if args.workload == "rws":
    return bench_rws(args)
  if args.workload == "rfsf":
    return bench_rfsf(args)
  cpu = None if args.no_cpu else c1_cpu_baseline(args.cpu_seconds)
  res = run_c1(args)
  if cpu is not None:
    res["cpu_baseline"] = cpu
  print(json.dumps(res), flush=True)


C1 = dict(n=512, L=200, d=8, n_levels=5, seed=1234)


def c1_cpu_baseline(seconds: float) -> dict:
  """The FP64 and FP32 NumPy oracle over a process pool on the host cores, on
  random pairs of the 512 x L=200 x d=8 workload (its own seeded paths with
  the same distribution; the GPU is not touched)."""
  from oracle import sigkern as O
  n, L, d, nl = C1["n"], C1["L"], C1["d"], C1["n_levels"]
  sigma = math.sqrt(d)
  rng = np.random.default_rng(C1["seed"])
  X64 = np.cumsum(rng.standard_normal((n, L, d)) / math.sqrt(L), 1)
  X32 = X64.astype(np.float32)
  batch = 8
  cores = pool_cores()
  runs = {}
  for tag, X in (("f64", X64), ("f32", X32)):
    runs[tag] = cpu_pool_run(
        lambda r: (r.integers(0, n, batch), r.integers(0, n, batch)),
        lambda u, X=X: O.signature_kern(O.rbf_gram(X[u[0]], X[u[1]], sigma), nl, 1, True, True),
        lambda u: batch * (L - 1) ** 2 * nl, seconds, 7, cores)
  return cpu_baseline_pool(
      "DP cell-levels/s",
      f"random pairs of the workload (L={L}, d={d}, {nl} levels, order 1), NumPy oracle "
      f"vectorised over {batch} pairs per call, {cores} single-threaded worker processes, "
      f"{seconds:.0f} s per precision", runs, cores)


def run_c1(args) -> dict:
  """512 Brownian paths, L=200, d=8, 5 levels, order 1, RBF sigma=sqrt(8):
  the symmetric Gram as its 131,328 upper-triangle pairs in one launch."""
  import ksig_amd
  from ksig_amd import _lib, _ops
  from oracle import sigkern as O

  dev = torch.device("cuda", 0)
  torch.cuda.set_device(dev)
  n, L, d, n_levels = C1["n"], C1["L"], C1["d"], C1["n_levels"]
  sigma = math.sqrt(d)
  gen = torch.Generator(device=dev).manual_seed(C1["seed"])
  X = brownian(gen, n, L, d, dev)
  Xs = _ops.SeqSet.dense(X)
  pairs = _ops.triu_pairs(n, dev)
  n_pairs = pairs.shape[0]
  cells = n_pairs * (L - 1) ** 2
  cell_levels = cells * n_levels

  mode = _lib.GRAM_MODES[args.gram]

  def step(m=mode):
    return _ops.signature_levels(Xs, Xs, pairs, n_levels, 1, True, _lib.STATIC_RBF, sigma, m)

  for _ in range(args.warmup):
    step()
  torch.cuda.synchronize()
  stream = torch.cuda.current_stream(dev)
  starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
  ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  for k in range(args.steps):
    starts[k].record(stream)
    out = step()
    ends[k].record(stream)
  torch.cuda.synchronize()
  wall = time.perf_counter() - t0
  kern_ms = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))
  ms_per_step = wall * 1e3 / args.steps

  # accuracy of this run's output vs the FP64 oracle on a sample of pairs
  X64 = X.double().cpu().numpy()
  idx = np.random.default_rng(3).choice(n_pairs, 24, replace=False)
  pr = pairs.cpu().numpy()[idx]
  got = out.cpu().numpy()[idx]
  ref = np.stack([O.pair_levels(X64[a], X64[b], n_levels, 1, True, "rbf", sigma)
                  for a, b in pr])
  err = float(np.max(np.abs(got - ref).max(1) / np.abs(ref).max(1)))

  def time_mode(m, reps=5):
    step(m)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
      o = step(m)
    e1.record(stream)
    torch.cuda.synchronize()
    gotm = o.cpu().numpy()[idx]
    errm = float(np.max(np.abs(gotm - ref).max(1) / np.abs(ref).max(1)))
    return e0.elapsed_time(e1) / reps, errm

  variants = {}
  if not args.no_variants:
    for name in ("valu", "fp32", "bf16x3"):
      ms, e = time_mode(_lib.GRAM_MODES[name])
      variants[name] = {"kernel_ms": ms, "cell_levels_per_s": cell_levels / (ms * 1e-3),
                        "max_rel_err_vs_fp64": e}

  flops_cell = 4 * n_levels + 2 * d   # recursion + static Gram
  achieved = cells * flops_cell / (kern_ms * 1e-3) / 1e12
  res = {
      "metric": "DP-cells/sec (pairs x (L-1)^2 x levels), fused fwd, fp32",
      "value": cell_levels / (ms_per_step * 1e-3),
      "unit": "DP cell-levels/s",
      "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
      "ms_per_step": ms_per_step,
      "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
      "dtype": "f32", "data": "synthetic (seeded Brownian paths)",
      "config": {"workload": "signature kernel DP, 512 x L=200 x d=8 Brownian, "
                             "5 levels, order 1, RBF sigma=sqrt(8), symmetric Gram "
                             "(131328 upper-triangle pairs evaluated)",
                 "pairs_evaluated": n_pairs, "L": L, "d": d, "n_levels": n_levels,
                 "full_gram_equivalent_cell_levels_per_s":
                     n * n * (L - 1) ** 2 * n_levels / (ms_per_step * 1e-3)},
      "max_rel_err_vs_fp64": err,
      "gram": args.gram,
      "gram_variants": variants,
      "roofline": {"bound": "valu" if args.gram == "valu" else "mfma", "achieved": achieved,
                   "peak": PEAK_FP32_TFLOPS,
                   "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                   "traffic": MEASURED_TRAFFIC_BYTES["c1"] if args.gram == "valu" else None,
                   "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE)",
                   "kernel": ("dp_mfma_kernel<5,1,8,%d>" % (0 if args.gram == "fp32" else 1)
                              if args.gram != "valu" else "dp_stream_kernel<5,1,8,4,0,true,false>"),
                   "kernel_ms": kern_ms,
                   "flops_per_cell": flops_cell},
  }
  if not args.no_plain:
    # plain FP32 reference dataflow on 8 x 512 pairs (a slice of the same Gram)
    Xa = X[:8]
    torch.cuda.synchronize()
    plain_levels = plain_fp32_levels(Xa, X, n_levels, sigma)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    reps = 3
    for _ in range(reps):
      plain_fp32_levels(Xa, X, n_levels, sigma)
    torch.cuda.synchronize()
    tp = (time.perf_counter() - t1) / reps
    pv = 8 * n * (L - 1) ** 2 * n_levels / tp
    res["plain_fp32"] = {"value": pv, "unit": "DP cell-levels/s",
                         "speedup_fused_over_plain": res["value"] / pv,
                         "sample": "8 x 512 pairs per call, torch FP32 ops"}
    del plain_levels
  return res


# per cell: static kernel (d multiply-adds, norm terms, exp), the transform
# log(M / (2 - M)) (sub, log, sub, max), the log-sum-exp of three (2 max,
# 3 sub, 3 exp, 2 add, log) and the two final adds
GAK_FLOPS_PER_CELL_BASE = 4 + 4 + 11 + 2


def plain_gak_log(X, Y, sigma):
  """The reference dataflow of ksig/algorithms.py:514-559 in torch FP32 ops on
  the GPU: materialise M [b, Lx, Ly] for pairs (X[k], Y[k]), transform and
  take logs, then sweep the antidiagonals with one batched log-sum-exp per
  antidiagonal over every pair and cell on it."""
  b, Lx, _ = X.shape
  Ly = Y.shape[1]
  D2 = (X * X).sum(-1)[:, :, None] + (Y * Y).sum(-1)[:, None, :] - 2 * X @ Y.transpose(1, 2)
  M = torch.exp(-D2 / max(2 * sigma ** 2, 1e-12))
  lg = torch.log(torch.clamp(M / (2 - M), min=1e-12))
  K = torch.full((b, Lx + 1, Ly + 1), float("-inf"), device=X.device, dtype=X.dtype)
  K[:, 0, 0] = 0.0
  for t in range(Lx + Ly - 1):
    i = torch.arange(max(0, t - Ly + 1), min(t, Lx - 1) + 1, device=X.device)
    j = t - i
    up, left, dg = K[:, i, j + 1], K[:, i + 1, j], K[:, i, j]
    m = torch.maximum(torch.maximum(up, left), dg)
    K[:, i + 1, j + 1] = lg[:, i, j] + m + torch.log(
        torch.exp(up - m) + torch.exp(left - m) + torch.exp(dg - m))
  return K[:, Lx, Ly]


def bench_gak(args):
  """Global alignment kernel (log space), fused HIP sweep: 512 Brownian paths,
  L=200, d=8, RBF sigma=sqrt(8), symmetric Gram as 131,328 upper-triangle
  pairs (the shape of the DP signature-kernel workload)."""
  from ksig_amd import _lib, _ops
  from oracle import sigkern as O
  dev = torch.device("cuda", 0)
  torch.cuda.set_device(dev)
  n, L, d = 512, 200, 8
  sigma = math.sqrt(d)
  g = torch.Generator(device=dev).manual_seed(0)
  X = brownian(g, n, L, d, dev)
  S = _ops.SeqSet.dense(X)
  pairs = _ops.triu_pairs(n, dev)
  n_pairs = pairs.shape[0]
  cells = float(n_pairs) * L * L

  def step():
    return _ops.global_align_log(S, S, pairs, _lib.STATIC_RBF, sigma)
  for _ in range(args.warmup):
    step()
  torch.cuda.synchronize()
  stream = torch.cuda.current_stream(dev)
  e0 = torch.cuda.Event(enable_timing=True)
  e1 = torch.cuda.Event(enable_timing=True)
  t0 = time.perf_counter()
  e0.record(stream)
  for _ in range(args.steps):
    out = step()
  e1.record(stream)
  torch.cuda.synchronize()
  wall = time.perf_counter() - t0
  kern_ms = e0.elapsed_time(e1) / args.steps
  ms_per_step = wall * 1e3 / args.steps

  X64 = X.double().cpu().numpy()
  idx = np.random.default_rng(3).choice(n_pairs, 16, replace=False)
  pr = pairs.cpu().numpy()[idx]
  got = out.cpu().numpy()[idx]
  ref = np.array([O.pair_gak_log(X64[a], X64[b], "rbf", sigma) for a, b in pr])
  err = float(np.max(np.abs(got - ref) / np.maximum(1.0, np.abs(ref))))

  flops_cell = GAK_FLOPS_PER_CELL_BASE + 2 * d
  achieved = cells * flops_cell / (kern_ms * 1e-3) / 1e12
  res = {
      "metric": "GA-cells/sec (pairs x Lx x Ly), fused log-space fwd, fp32 inputs",
      "value": cells / (ms_per_step * 1e-3),
      "unit": "GA cells/s",
      "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
      "ms_per_step": ms_per_step,
      "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
      "dtype": "f32 inputs, f64 log accumulation", "data": "synthetic (seeded Brownian paths)",
      "config": {"workload": "global alignment kernel, 512 x L=200 x d=8 Brownian, RBF "
                             "sigma=sqrt(8), symmetric Gram (131328 upper-triangle pairs)",
                 "pairs_evaluated": n_pairs, "L": L, "d": d},
      "max_abs_log_err_vs_fp64": err,
      "err_sample": "16 random pairs, |log GA - log GA_64| / max(1, |log GA_64|)",
      "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                   "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                   "traffic": MEASURED_TRAFFIC_BYTES["gak"],
                   "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE)",
                   "kernel": "gak_forward_kernel<float,8,RBF,4>", "kernel_ms": kern_ms,
                   "flops_per_cell": flops_cell},
  }
  if not args.no_plain:
    nb = 256
    a = torch.arange(nb, device=dev) % n
    b = (torch.arange(nb, device=dev) * 7 + 3) % n
    Xa, Xb = X[a].contiguous(), X[b].contiguous()
    pl = plain_gak_log(Xa, Xb, sigma)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    plain_gak_log(Xa, Xb, sigma)
    torch.cuda.synchronize()
    tp = time.perf_counter() - t1
    pv = nb * L * L / tp
    chk = _ops.global_align_log(S, S, torch.stack([a, b], -1).to(torch.int32).contiguous(),
                                _lib.STATIC_RBF, sigma)
    pdiff = float(torch.max(torch.abs(pl.double() - chk) / torch.clamp(chk.abs(), min=1.0)))
    res["plain_fp32"] = {"value": pv, "unit": "GA cells/s",
                         "speedup_fused_over_plain": res["value"] / pv,
                         "sample": f"{nb} pairs per call: M materialised, one batched "
                                   "log-sum-exp per antidiagonal, torch FP32 ops",
                         "max_abs_log_diff_vs_fused": pdiff}
  if not args.no_cpu:
    try:
      from threadpoolctl import threadpool_limits
      lim = threadpool_limits(1)
    except Exception:  # pragma: no cover
      lim = None
    rng = np.random.default_rng(7)
    batch, done = 64, 0
    t2 = time.perf_counter()
    while time.perf_counter() - t2 < args.cpu_seconds:
      ia = rng.integers(0, n, batch)
      ib = rng.integers(0, n, batch)
      O.global_align_kern_log(O.rbf_gram(X64[ia], X64[ib], sigma))
      done += batch
    dt = time.perf_counter() - t2
    res["cpu_baseline"] = {"value": done * L * L / dt, "unit": "GA cells/s", "cores": 1,
                           "kind": "port",
                           "sample": f"{done} random pairs of the workload, FP64 NumPy "
                                     f"oracle (vectorised over {batch} pairs), {dt:.1f} s "
                                     "on 1 host core"}
  print(json.dumps(res), flush=True)


def plain_rws(X, W, lens):
  """The reference dataflow of ksig/algorithms.py:630-659 and
  ksig/kernels.py:466 in torch FP32 ops on the GPU: materialise the distances
  D [n, L, sum of lengths], then per warping series sweep the antidiagonals
  with one batched min over every sequence and cell on it."""
  n, L, _ = X.shape
  D = (X * X).sum(-1)[:, :, None] + (W * W).sum(-1)[None, None, :] - 2 * X @ W.t()
  out = []
  off = 0
  for lb in lens:
    Db = D[:, :, off:off + lb]
    off += lb
    P = torch.full((n, L + 1, lb + 1), float("inf"), device=X.device, dtype=X.dtype)
    P[:, 0, 0] = 0.0
    for t in range(L + lb - 1):
      i = torch.arange(max(0, t - lb + 1), min(t, L - 1) + 1, device=X.device)
      j = t - i
      P[:, i + 1, j + 1] = Db[:, i, j] + torch.minimum(
          torch.minimum(P[:, i, j + 1], P[:, i + 1, j]), P[:, i, j])
    out.append(P[:, L, lb])
  return torch.stack(out, -1)


def bench_rws(args):
  """Random Warping Series features: 4096 Brownian paths, L=200, d=8,
  against 100 warping series of lengths U{1..32} (the class defaults),
  DTW values of every (sequence, series) pair in one launch."""
  from ksig_amd import _ops
  from oracle import sigkern as O
  dev = torch.device("cuda", 0)
  torch.cuda.set_device(dev)
  n, L, d, n_w, max_warp = 4096, 200, 8, 100, 32
  g = torch.Generator(device=dev).manual_seed(0)
  X = brownian(g, n, L, d, dev)
  rng = np.random.RandomState(0)
  lens = rng.randint(1, max_warp + 1, size=n_w)
  W = torch.as_tensor(rng.standard_normal((int(lens.sum()), d)) * 0.5, device=dev,
                      dtype=torch.float32)
  S = _ops.SeqSet.dense(X)
  cells = float(n) * L * float(lens.sum())

  def step():
    return _ops.random_warping_series(S, W, lens)
  for _ in range(args.warmup):
    step()
  torch.cuda.synchronize()
  stream = torch.cuda.current_stream(dev)
  e0 = torch.cuda.Event(enable_timing=True)
  e1 = torch.cuda.Event(enable_timing=True)
  t0 = time.perf_counter()
  e0.record(stream)
  for _ in range(args.steps):
    out = step()
  e1.record(stream)
  torch.cuda.synchronize()
  wall = time.perf_counter() - t0
  kern_ms = e0.elapsed_time(e1) / args.steps
  ms_per_step = wall * 1e3 / args.steps

  X64 = X.double().cpu().numpy()
  W64 = W.double().cpu().numpy()
  idx = np.random.default_rng(3).choice(n, 4, replace=False)
  ref = O.random_warping_series(O.squared_euclid_dist(X64[idx], W64[None]), lens)
  got = out.cpu().numpy()[idx]
  err = float(np.max(np.abs(got - ref).max(-1) / np.abs(ref).max(-1)))
  flops_cell = 3 * d + 4   # distance (d mul-adds, norm terms) + min of three + add
  achieved = cells * flops_cell / (kern_ms * 1e-3) / 1e12
  res = {
      "metric": "DTW-cells/sec (sequences x L x total warping-series length), fused fwd, fp32",
      "value": cells / (ms_per_step * 1e-3),
      "unit": "DTW cells/s",
      "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
      "ms_per_step": ms_per_step,
      "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
      "dtype": "f32", "data": "synthetic (seeded Brownian paths, Gaussian warping series)",
      "config": {"workload": "random warping series, 4096 x L=200 x d=8 Brownian against 100 "
                             "series of lengths U{1..32}",
                 "n_seq": n, "L": L, "d": d, "n_warps": n_w, "warp_rows": int(lens.sum())},
      "max_rel_err_vs_fp64": err,
      "err_sample": "4 random sequences, normwise over their 100 DTW values",
      "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                   "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                   "traffic": MEASURED_TRAFFIC_BYTES["rws"],
                   "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE)",
                   "kernel": "rws_forward_kernel<float,8,false,4>", "kernel_ms": kern_ms,
                   "flops_per_cell": flops_cell},
  }
  if not args.no_plain:
    nb, nwb = 256, 25
    Xs = X[:nb].contiguous()
    Ws = W[:int(lens[:nwb].sum())].contiguous()
    pl = plain_rws(Xs, Ws, lens[:nwb])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    plain_rws(Xs, Ws, lens[:nwb])
    torch.cuda.synchronize()
    tp = time.perf_counter() - t1
    pv = nb * L * float(lens[:nwb].sum()) / tp
    pdiff = float(torch.max(torch.abs(pl.double() - out[:nb, :nwb]) / out[:nb, :nwb].abs()))
    res["plain_fp32"] = {"value": pv, "unit": "DTW cells/s",
                         "speedup_fused_over_plain": res["value"] / pv,
                         "sample": f"{nb} sequences x {nwb} series per call: distances "
                                   "materialised, one batched min per antidiagonal, torch "
                                   "FP32 ops",
                         "max_rel_diff_vs_fused": pdiff}
  if not args.no_cpu:
    try:
      from threadpoolctl import threadpool_limits
      lim = threadpool_limits(1)
    except Exception:  # pragma: no cover
      lim = None
    rng2 = np.random.default_rng(7)
    batch, done = 64, 0
    t2 = time.perf_counter()
    while time.perf_counter() - t2 < args.cpu_seconds:
      ia = rng2.integers(0, n, batch)
      O.random_warping_series(O.squared_euclid_dist(X64[ia], W64[None]), lens[:10])
      done += batch
    dt = time.perf_counter() - t2
    res["cpu_baseline"] = {"value": done * L * float(lens[:10].sum()) / dt,
                           "unit": "DTW cells/s", "cores": 1, "kind": "port",
                           "sample": f"{done} random sequences against the first 10 series, "
                                     f"FP64 NumPy oracle (vectorised over {batch} sequences), "
                                     f"{dt:.1f} s on 1 host core"}
  print(json.dumps(res), flush=True)


def bench_rfsf(args):
  """Random Fourier signature features with a tensorized random projection
  (RFSF-TRP), the reference README's example shape: 1000 sequences, L=200,
  d=100, 5 levels, 100 RFF components (200 features), 100 projection
  components.  Times the whole transform (RFF and projection GEMMs in
  hipBLASLt, the HIP level scan) and the level scan alone."""
  import ksig_amd as ksig
  from ksig_amd import _ops
  from oracle import sigkern as O
  dev = torch.device("cuda", 0)
  torch.cuda.set_device(dev)
  n, L, d, nl, D, P = 1000, 200, 100, 5, 100, 100
  g = torch.Generator(device=dev).manual_seed(0)
  X = torch.randn((n, L, d), generator=g, device=dev, dtype=torch.float32)
  sf = ksig.kernels.SignatureFeatures(
      n_levels=nl, static_features=ksig.static.features.RandomFourierFeatures(
          n_components=D, random_state=0),
      projection=ksig.projections.TensorizedRandomProjection(n_components=P, random_state=1))
  sf.fit(X)

  def step():
    return sf.transform(X, return_on_gpu=True)
  for _ in range(args.warmup):
    step()
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  for _ in range(args.steps):
    F = step()
  torch.cuda.synchronize()
  ms_per_step = (time.perf_counter() - t0) * 1e3 / args.steps
  # the level scan alone on the projected increments of this input
  U = [torch.diff(f.transform(X.reshape(-1, d), return_on_gpu=True).reshape(n, L, -1), dim=1)
       for f in sf.static_features_]
  Z = [sf.projections_[0]._project(U[0])]
  for u, p in zip(U[1:], sf.projections_[1:]):
    Z.append(p.scaling_ * (u.reshape(-1, u.shape[-1]) @ p.components_).reshape(n, L - 1, -1))
  Z = torch.stack(Z, 0).contiguous()
  stream = torch.cuda.current_stream(dev)
  e0 = torch.cuda.Event(enable_timing=True)
  e1 = torch.cuda.Event(enable_timing=True)
  _ops.lowrank_levels(Z)
  e0.record(stream)
  for _ in range(args.steps):
    _ops.lowrank_levels(Z)
  e1.record(stream)
  torch.cuda.synchronize()
  scan_ms = e0.elapsed_time(e1) / args.steps
  scan_bytes = Z.numel() * 4 + n * nl * P * 8
  # features of 4 sequences vs the FP64 oracle with the same random components
  w = [f.random_weights_.double().cpu().numpy() for f in sf.static_features_]
  trp = [(p.components_.double().cpu().numpy(), p.scaling_) for p in sf.projections_]
  idx = [0, 1, 500, 999]
  ref = O.signature_features(X[idx].double().cpu().numpy(), nl, w, trp)
  err = float(np.max(np.abs(F[idx].double().cpu().numpy() - ref).max(-1) / np.abs(ref).max(-1)))
  achieved = scan_bytes / (scan_ms * 1e-3) / 1e9
  res = {
      "metric": "RFSF-TRP feature throughput (sequences x L / s), fp32, whole transform",
      "value": n * L / (ms_per_step * 1e-3),
      "unit": "sequence points/s",
      "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
      "ms_per_step": ms_per_step,
      "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
      "dtype": "f32", "data": "synthetic (seeded N(0,1) sequences)",
      "config": {"workload": "RFSF-TRP: 1000 x L=200 x d=100, 5 levels, 100 RFF components, "
                             "100 projection components (501 features)",
                 "features": int(F.shape[1])},
      "max_rel_err_vs_fp64": err,
      "err_sample": "4 sequences, normwise over their normalised feature vectors",
      "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                   "frac": achieved / PEAK_HBM_GBS, "traffic": MEASURED_TRAFFIC_BYTES["rfsf"],
                   "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE)",
                   "kernel": "lowrank_scan_kernel<float,5,1>", "kernel_ms": scan_ms,
                   "bytes_per_launch": scan_bytes},
  }
  # RFSF-DP on the same input: diagonal projection of the RFF increments
  dp = ksig.kernels.SignatureFeatures(
      n_levels=nl, static_features=ksig.static.features.RandomFourierFeatures(
          n_components=D, random_state=0),
      projection=ksig.projections.DiagonalProjection())
  dp.fit(X)
  for _ in range(args.warmup):
    dp.transform(X, return_on_gpu=True)
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  for _ in range(args.steps):
    Fd = dp.transform(X, return_on_gpu=True)
  torch.cuda.synchronize()
  dp_ms = (time.perf_counter() - t0) * 1e3 / args.steps
  dU = torch.stack([torch.diff(f.transform(X.reshape(-1, d), return_on_gpu=True).reshape(n, L, -1),
                               dim=1) for f in dp.static_features_], 0).contiguous()
  _ops.lowrank_levels_diag(dU)
  e0.record(stream)
  for _ in range(args.steps):
    _ops.lowrank_levels_diag(dU)
  e1.record(stream)
  torch.cuda.synchronize()
  dscan_ms = e0.elapsed_time(e1) / args.steps
  dscan_bytes = dU.numel() * 4 + n * ((2 << nl) - 2) * D * 8
  wd = [f.random_weights_.double().cpu().numpy() for f in dp.static_features_]
  refd = O.signature_features(X[idx].double().cpu().numpy(), nl, wd, 'diag')
  errd = float(np.max(np.abs(Fd[idx].double().cpu().numpy() - refd).max(-1) /
                      np.abs(refd).max(-1)))
  res["dp"] = {"ms_per_transform": dp_ms, "features": int(Fd.shape[1]),
               "max_rel_err_vs_fp64": errd,
               "scan_kernel": "lowrank_diag_scan_kernel<float,2,5,1>",
               "scan_ms": dscan_ms, "scan_bytes": dscan_bytes,
               "scan_GBps": dscan_bytes / (dscan_ms * 1e-3) / 1e9}
  # RFSF-TS on the same input: TensorSketch of the RFF increments, 100
  # sketch components
  ts = ksig.kernels.SignatureFeatures(
      n_levels=nl, static_features=ksig.static.features.RandomFourierFeatures(
          n_components=D, random_state=0),
      projection=ksig.projections.TensorSketch(n_components=P, random_state=2))
  ts.fit(X)
  for _ in range(args.warmup):
    ts.transform(X, return_on_gpu=True)
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  for _ in range(args.steps):
    Ft = ts.transform(X, return_on_gpu=True)
  torch.cuda.synchronize()
  ts_ms = (time.perf_counter() - t0) * 1e3 / args.steps
  hidx = torch.stack([p.hash_idx_ for p in ts.projections_], 0)
  hbit = torch.stack([p.hash_bit_.float() for p in ts.projections_], 0)
  _ops.lowrank_levels_sketch(dU, hidx, hbit, P)
  e0.record(stream)
  for _ in range(args.steps):
    _ops.lowrank_levels_sketch(dU, hidx, hbit, P)
  e1.record(stream)
  torch.cuda.synchronize()
  tscan_ms = e0.elapsed_time(e1) / args.steps
  # the scan convolves each level's running sum with the step's count
  # sketch: from the sparse sketch (one multiply-add per input feature and
  # component, 2 * width * P flop per level-step) or, with fewer components
  # than features, from the dense sketch (width adds, then 2 * P * P flop)
  width = dU.shape[-1]
  dense = P < width
  per = (width + 2.0 * P * P) if dense else 2.0 * width * P
  tscan_flop = float(n) * (L - 1) * nl * per
  wt = [f.random_weights_.double().cpu().numpy() for f in ts.static_features_]
  skt = [(p.hash_idx_.cpu().numpy(), p.hash_bit_.double().cpu().numpy(), P)
         for p in ts.projections_]
  Ut = [O.rff_features(X[idx].double().cpu().numpy(), wi) for wi in wt]
  reft = O.low_rank_features_normalised(O.signature_kern_low_rank_sketch(Ut, nl, skt), nl)
  errt = float(np.max(np.abs(Ft[idx].double().cpu().numpy() - reft).max(-1) /
                      np.abs(reft).max(-1)))
  res["ts"] = {"ms_per_transform": ts_ms, "features": int(Ft.shape[1]),
               "max_rel_err_vs_fp64": errt,
               "scan_kernel": f"lowrank_sketch_kernel<float,5,{'true' if dense else 'false'}>",
               "scan_mode": "dense count sketch" if dense else "sparse count sketch",
               "scan_ms": tscan_ms, "scan_flop": tscan_flop,
               "scan_TFLOPs": tscan_flop / (tscan_ms * 1e-3) / 1e12,
               "flop_note": "float64 flop of the convolutions as evaluated: width adds and "
                            "2 n_comp^2 per level-step from the dense sketch, or 2 width n_comp "
                            "from the sparse one"}
  print(json.dumps(res), flush=True)


def kernel_probe(args):
  """One C4 kernel on one chunk of the training step: 128 Brownian paths
  (L=128, d=32) against themselves, 16384 pairs, 6 levels, RBF sigma=sqrt(32);
  the forward without (fwd) or with (fwd_save) the boundary state, the
  adjoint (omega; adjoint_gp also the bandwidth gradient, which the C4 step
  does not take) or the contraction of omega with the inputs, each launched
  --steps times after one untimed launch, timed with HIP events on the
  library's stream."""
  from ksig_amd import _lib, _ops
  dev = torch.device("cuda", 0)
  lib = _lib.load()
  st = _lib.stream_handle(dev)
  n, L, d, nl = 128, 128, 32, 6
  g = torch.Generator(device=dev).manual_seed(1)
  X = brownian(g, n, L, d, dev).reshape(-1, d).contiguous()
  S = _ops.SeqSet.dense(X.reshape(n, L, d))
  pairs = _ops.dense_pairs(n, n, dev)
  P = pairs.shape[0]
  per = (nl - 1) * (2 * L - 2)
  save_off = torch.arange(P, device=dev, dtype=torch.int64) * per
  save = torch.empty(P * per, device=dev)
  out = torch.empty((P, nl + 1), device=dev, dtype=torch.float64)
  sig = math.sqrt(d)
  xa = (S.data.data_ptr(), S.off.data_ptr(), S.len.data_ptr(), L)
  fwsb = lib.ksig_signature_levels_fwd_save_workspace(0, d, nl, L, L, P)
  fws = torch.empty(max(fwsb // 4, 1), device=dev)
  gout = torch.randn((P, nl + 1), device=dev, dtype=torch.float64, generator=g)
  omega = torch.empty(P * L * L, device=dev)
  w_off = torch.arange(P, device=dev, dtype=torch.int64) * (L * L)
  gp = torch.empty(P, device=dev, dtype=torch.float64)
  wsb = lib.ksig_signature_levels_bwd_workspace(0, d, nl, L, L, P)
  ws = torch.empty(max(wsb // 4, 1), device=dev)
  gx = torch.zeros_like(X)

  def fwd_save():
    _lib.check(lib.ksig_signature_levels_fwd_save(
        0, *xa, *xa, d, pairs.data_ptr(), P, _lib.STATIC_RBF, sig, nl, out.data_ptr(),
        save.data_ptr(), save_off.data_ptr(), fws.data_ptr(), fwsb, st))

  def fwd():  # the same kernel without the boundary state (the forward-only path)
    _lib.check(lib.ksig_signature_levels_fwd_save(
        0, *xa, *xa, d, pairs.data_ptr(), P, _lib.STATIC_RBF, sig, nl, out.data_ptr(),
        None, save_off.data_ptr(), fws.data_ptr(), fwsb, st))

  def adjoint(with_gp=False):  # as the C4 step runs it: no bandwidth gradient
    _lib.check(lib.ksig_signature_levels_bwd(
        0, *xa, *xa, d, pairs.data_ptr(), P, _lib.STATIC_RBF, sig, nl, gout.data_ptr(),
        save.data_ptr(), save_off.data_ptr(), omega.data_ptr(), w_off.data_ptr(),
        gp.data_ptr() if with_gp else None, ws.data_ptr(), wsb, st))

  def contract():
    c = 1.0 / (sig * sig)
    _lib.check(lib.ksig_static_grad_contract(
        0, *xa, *xa, d, pairs.data_ptr(), P, omega.data_ptr(), w_off.data_ptr(), c, -c,
        gx.data_ptr(), gx.data_ptr(), st))

  fwd_save()
  if args.kernel_probe not in ("fwd_save", "fwd"):
    adjoint()
  f = {"fwd": fwd, "fwd_save": fwd_save, "adjoint": adjoint,
       "adjoint_gp": lambda: adjoint(True),
       "contract": contract}[args.kernel_probe]
  f()
  torch.cuda.synchronize()
  e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
  e0.record()
  for _ in range(args.steps):
    f()
  e1.record()
  torch.cuda.synchronize()
  ms = e0.elapsed_time(e1) / args.steps
  print(json.dumps({"kernel_probe": args.kernel_probe, "pairs": P, "L": L, "d": d,
                    "n_levels": nl, "launches": args.steps, "ms_per_launch": ms}), flush=True)


def pde_kernel_probe(args):
  """One kernel of the signature-PDE training step on one chunk of the C2
  shape: 128 Brownian paths (L=256, d=4) against themselves, 16384 pairs, RBF
  sigma=1, dyadic order 1; the forward (levels only), the forward that also
  saves the edges, the adjoint (dL/dM per original cell) or the contraction of
  dL/dM with the inputs, each launched --steps times after one untimed
  launch, timed with HIP events on the library's stream."""
  from ksig_amd import _lib, _ops
  dev = torch.device("cuda", 0)
  lib = _lib.load()
  st = _lib.stream_handle(dev)
  n, L, d, lam = 128, 256, 4, 1
  g = torch.Generator(device=dev).manual_seed(1)
  X = brownian(g, n, L, d, dev).reshape(-1, d).contiguous()
  S = _ops.SeqSet.dense(X.reshape(n, L, d))
  pairs = _ops.dense_pairs(n, n, dev)
  P = pairs.shape[0]
  xa = (S.data.data_ptr(), S.off.data_ptr(), S.len.data_ptr(), L)
  out = torch.empty(P, device=dev, dtype=torch.float64)
  per = 2 * ((L - 1) << lam)
  e_off = torch.arange(P, device=dev, dtype=torch.int64) * per
  edges = torch.empty(P * per, device=dev, dtype=torch.float64)
  gout = torch.randn(P, device=dev, dtype=torch.float64, generator=g)
  m_off = torch.arange(P, device=dev, dtype=torch.int64) * ((L - 1) ** 2)
  gm = torch.zeros(P * (L - 1) ** 2, device=dev)
  wsb = lib.ksig_signature_pde_bwd_workspace(0, L, L, lam, 1, P)
  ws = torch.empty(max(wsb // 4, 1), device=dev)
  gx = torch.zeros_like(X)
  gparam = torch.zeros((), device=dev, dtype=torch.float64)

  def fwd():
    _lib.check(lib.ksig_signature_pde(0, *xa, *xa, d, pairs.data_ptr(), P, _lib.STATIC_RBF,
                                      1.0, lam, 1, out.data_ptr(), st))

  def fwd_edges():
    _lib.check(lib.ksig_signature_pde_fwd_edges(
        0, *xa, *xa, d, pairs.data_ptr(), P, _lib.STATIC_RBF, 1.0, lam, 1, out.data_ptr(),
        edges.data_ptr(), e_off.data_ptr(), st))

  def adjoint():
    _lib.check(lib.ksig_signature_pde_bwd_edges(
        0, *xa, *xa, d, pairs.data_ptr(), P, _lib.STATIC_RBF, 1.0, lam, 1, gout.data_ptr(),
        edges.data_ptr(), e_off.data_ptr(), gm.data_ptr(), m_off.data_ptr(), ws.data_ptr(),
        wsb, st))

  def contract():
    _lib.check(lib.ksig_signature_pde_grad_contract(
        0, *xa, *xa, d, pairs.data_ptr(), P, _lib.STATIC_RBF, 1.0, 1, gm.data_ptr(),
        m_off.data_ptr(), gx.data_ptr(), gx.data_ptr(), gparam.data_ptr(), st))

  fwsb = lib.ksig_signature_pde_bwd_fused_workspace(0, d, L, L, lam, 1, P)
  fws = torch.empty(max(fwsb // 4, 1), device=dev)
  gpp = torch.empty(P, device=dev, dtype=torch.float64)

  def fused():
    _lib.check(lib.ksig_signature_pde_bwd_fused(
        0, *xa, *xa, d, pairs.data_ptr(), P, _lib.STATIC_RBF, 1.0, lam, 1, gout.data_ptr(),
        edges.data_ptr(), e_off.data_ptr(), gx.data_ptr(), gx.data_ptr(), gpp.data_ptr(),
        fws.data_ptr(), fwsb, st))

  probes = {"pde_fwd": fwd, "pde_fwd_edges": fwd_edges, "pde_adjoint": adjoint,
            "pde_contract": contract, "pde_fused": fused}
  if args.kernel_probe in ("pde_adjoint", "pde_contract", "pde_fused"):
    fwd_edges()
  if args.kernel_probe in ("pde_adjoint", "pde_contract"):
    adjoint()
  f = probes[args.kernel_probe]
  f()
  torch.cuda.synchronize()
  e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
  e0.record()
  for _ in range(args.steps):
    f()
  e1.record()
  torch.cuda.synchronize()
  ms = e0.elapsed_time(e1) / args.steps
  print(json.dumps({"kernel_probe": args.kernel_probe, "pairs": P, "L": L, "d": d,
                    "dyadic_order": lam, "launches": args.steps, "ms_per_launch": ms,
                    "refined_cells_per_s": P * ((L - 1) << lam) ** 2 / (ms * 1e-3)}),
        flush=True)


def ragged_walks_host(rng: np.random.Generator, lens, d: int, offset_scale: float):
  """Random walks of the given lengths (cumsum of N(0, 1/L) increments plus a
  per-sequence per-channel offset), packed row-wise in float64: (walk
  [sum of lengths, d], start row of every sequence)."""
  lens = np.asarray(lens, dtype=np.int64)
  total = int(lens.sum())
  seg = np.repeat(np.arange(len(lens)), lens)
  inc = rng.standard_normal((total, d)) / np.sqrt(lens[seg])[:, None]
  csum = np.cumsum(inc, 0)
  starts = np.concatenate([[0], np.cumsum(lens)[:-1]])
  base = np.where((starts > 0)[:, None], csum[np.maximum(starts - 1, 0)], 0.0)
  walk = csum - base[seg] + offset_scale * rng.standard_normal((len(lens), d))[seg]
  return walk, starts


def seqset_from_host(walk: np.ndarray, starts: np.ndarray, lens, dev):
  """The float32 ragged SeqSet of packed host rows."""
  from ksig_amd import _ops
  lens = np.asarray(lens, dtype=np.int64)
  data = torch.as_tensor(walk.astype(np.float32), device=dev)
  off = torch.as_tensor(starts, device=dev, dtype=torch.int64)
  ln = torch.as_tensor(lens, device=dev, dtype=torch.int32)
  return _ops.SeqSet(data, off, ln, int(lens.max()), len(lens), lens)


def ragged_walks(rng: np.random.Generator, lens, d: int, offset_scale: float, dev):
  """ragged_walks_host moved to the device as a float32 SeqSet: (set, walk)."""
  walk, starts = ragged_walks_host(rng, lens, d, offset_scale)
  return seqset_from_host(walk, starts, lens, dev), walk


def c3_host_problem(seed: int = 2024) -> dict:
  """The batched uneven workload on the host: 4096 pairs (x_p, y_p), each
  length drawn from U{32..1024}, d=16, 8 levels, order 2, RBF sigma=4."""
  n, d = 4096, 16
  rng = np.random.default_rng(seed)
  lx = rng.integers(32, 1025, n)
  ly = rng.integers(32, 1025, n)
  xw, xs = ragged_walks_host(rng, lx, d, 0.1)
  yw, ys = ragged_walks_host(rng, ly, d, 0.1)
  return dict(n=n, d=d, n_levels=8, order=2, sigma=4.0, lx=lx, ly=ly, xw=xw, yw=yw, xs=xs,
              ys=ys)


def c3_pair(P: dict, p: int, dtype):
  """Pair p of the workload as the GPU sees it (float32 values) in ``dtype``."""
  x = P["xw"][P["xs"][p]:P["xs"][p] + P["lx"][p]].astype(np.float32).astype(dtype)
  y = P["yw"][P["ys"][p]:P["ys"][p] + P["ly"][p]].astype(np.float32).astype(dtype)
  return x, y


def c3_cpu_baseline(P: dict, seconds: float):
  """The FP64 and FP32 NumPy oracle (the reference algorithm and dataflow:
  materialised Gram, ΔΔ, per-level exclusive cumsums of the higher-order
  recursion) over a process pool on the host cores, on uniformly drawn pairs
  of the workload.  Returns (cpu_baseline object, {pair: FP64 levels})."""
  from oracle import sigkern as O
  nl, order, sigma = P["n_levels"], P["order"], P["sigma"]
  cores = pool_cores()
  runs, ref = {}, {}
  for tag, dt in (("f64", np.float64), ("f32", np.float32)):
    def evaluate(p, dt=dt):
      x, y = c3_pair(P, p, dt)
      return O.pair_levels(x, y, nl, order, True, "rbf", sigma)
    runs[tag] = cpu_pool_run(
        lambda r: int(r.integers(0, P["n"])), evaluate,
        lambda p: float((P["lx"][p] - 1) * (P["ly"][p] - 1) * nl), seconds,
        17 if tag == "f64" else 29, cores, keep=(tag == "f64"))
    if tag == "f64":
      ref = {int(p): v for p, v in runs[tag][3]}
  obj = cpu_baseline_pool(
      "DP cell-levels/s",
      f"uniformly drawn pairs of the workload ({nl} levels, order {order}, d={P['d']}, "
      f"lengths U{{32..1024}}), NumPy oracle one pair per call, {cores} single-threaded "
      f"worker processes, {seconds:.0f} s per precision", runs, cores)
  return obj, ref


def run_c3(args, P: dict, ref64: dict) -> dict:
  """4096 pairs of uneven lengths in one persistent row-stream launch; the
  error is taken against the FP64 levels in ``ref64`` (pair -> levels), or
  against 16 pairs evaluated here when it is empty."""
  from ksig_amd import _lib, _ops
  from oracle import sigkern as O
  dev = torch.device("cuda", 0)
  torch.cuda.set_device(dev)
  n, d, n_levels, order, sigma = P["n"], P["d"], P["n_levels"], P["order"], P["sigma"]
  lx, ly = P["lx"], P["ly"]
  X = seqset_from_host(P["xw"], P["xs"], lx, dev)
  Y = seqset_from_host(P["yw"], P["ys"], ly, dev)
  pairs = torch.stack([torch.arange(n, device=dev, dtype=torch.int32)] * 2, -1).contiguous()
  cells = float(np.sum((lx - 1) * (ly - 1)))
  cell_levels = cells * n_levels
  t0 = time.perf_counter()
  sched = _ops.make_schedule(X, Y, pairs, _lib.STATIC_RBF, n_levels, order)
  plan_ms = (time.perf_counter() - t0) * 1e3
  tuned = None
  if not args.no_tune:
    # launch choices (FIFO placement, slice bound, slices per resident wave)
    # timed on this pair list by the autotuner, before the timed region
    from ksig_amd import tune
    t0 = time.perf_counter()
    best, times = tune.autotune(X, Y, pairs, n_levels, order, _lib.STATIC_RBF, sigma, reps=3)
    tune_s = time.perf_counter() - t0
    default_cfg = tune.StreamConfig(bool(sched[3]), 128, 1,
                                    _ops.TWO_WAVE_FACTOR if sched[3] else 0.0)
    tuned = {"best": {"global_fifo": best.global_fifo, "max_slice": best.max_slice,
                      "oversub": best.oversub, "two_wave": best.two_wave},
             "candidates_ms": {f"{'gfifo' if c.global_fifo else 'lds'}/s{c.max_slice}/x{c.oversub}"
                               f"{f'/2w{c.two_wave:g}' if c.two_wave else ''}":
                               round(t, 4) for c, t in times.items()},
             "default_ms": times.get(default_cfg), "tune_s": tune_s}
    sched = tune.tuned_schedule(X, Y, pairs, _lib.STATIC_RBF, n_levels, order)

  def step():
    return _ops.signature_levels_batched(X, Y, pairs, n_levels, order, _lib.STATIC_RBF, sigma,
                                         sched)
  for _ in range(args.warmup):
    step()
  torch.cuda.synchronize()
  stream = torch.cuda.current_stream(dev)
  e0 = torch.cuda.Event(enable_timing=True)
  e1 = torch.cuda.Event(enable_timing=True)
  t1 = time.perf_counter()
  e0.record(stream)
  for _ in range(args.steps):
    out = step()
  e1.record(stream)
  torch.cuda.synchronize()
  wall = time.perf_counter() - t1
  kern_ms = e0.elapsed_time(e1) / args.steps
  ms_per_step = wall * 1e3 / args.steps

  # host planning of a repeated ragged call (the schedule cache) and the
  # end-to-end time of the public op on this pair list (outside the timed
  # region: it includes the planning)
  from ksig_amd import tune as _tune
  _tune.tuned_schedule(X, Y, pairs, _lib.STATIC_RBF, n_levels, order)
  t2 = time.perf_counter()
  for _ in range(5):
    _tune.tuned_schedule(X, Y, pairs, _lib.STATIC_RBF, n_levels, order)
  plan_repeat_ms = (time.perf_counter() - t2) * 1e3 / 5
  torch.cuda.synchronize()
  t2 = time.perf_counter()
  for _ in range(5):
    _ops.signature_levels(X, Y, pairs, n_levels, order, True, _lib.STATIC_RBF, sigma)
  torch.cuda.synchronize()
  call_ms = (time.perf_counter() - t2) * 1e3 / 5

  # per-pair error vs the FP64 oracle (inputs rounded to float32 as on the GPU)
  if not ref64:
    for p in np.random.default_rng(5).choice(n, 16, replace=False):
      x, y = c3_pair(P, int(p), np.float64)
      ref64 = dict(ref64)
      ref64[int(p)] = O.pair_levels(x, y, n_levels, order, True, "rbf", sigma)
  got = out.cpu().numpy()

  def pair_errors(vals):
    return np.array([float(np.max(np.abs(vals[p] - r)) / np.max(np.abs(r)))
                     for p, r in ref64.items()])
  errs = pair_errors(got)

  res = {
      "metric": "DP-cells/sec (pairs×L²×levels) fwd & fwd+bwd vs plain FP32, at ≤1e-5 rel err",
      "value": cell_levels / (ms_per_step * 1e-3),
      "unit": "DP cell-levels/s",
      "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
      "ms_per_step": ms_per_step,
      "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
      "dtype": "f32", "data": "synthetic (seeded random walks with channel offsets)",
      "config": {"workload": "batched uneven DP signature kernel: 4096 pairs, L ~ U{32..1024} "
                             "per sequence, d=16, 8 levels, order 2, RBF sigma=4, one persistent "
                             "launch (fused FP32 row stream)",
                 "cells": cells, "cell_levels": cell_levels,
                 "n_waves": int(sched[1].shape[0] - 1),
                 "global_fifo": bool(sched[3]), "plan_ms": plan_ms,
                 "plan_ms_repeat_call": plan_repeat_ms, "public_op_call_ms": call_ms,
                 "autotune": tuned},
      "max_rel_err_vs_fp64": float(errs.max()),
      "median_rel_err_vs_fp64": float(np.median(errs)),
      "err_sample": f"{len(errs)} uniformly drawn pairs, normwise over each pair's level "
                    "vector (K_0..K_8) against the FP64 oracle on the float32-rounded inputs",
  }
  flops_cell = recursion_flops(n_levels, order) + 3 + 2 * d + 3
  achieved = cells * flops_cell / (kern_ms * 1e-3) / 1e12
  res["roofline"] = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": MEASURED_TRAFFIC_BYTES["c3"],
                     "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE)",
                     "kernel": "dp_stream_kernel<8,2,16,2,0,true,%s>" %
                               ("true" if sched[3] else "false"),
                     "kernel_ms": kern_ms, "flops_per_cell": flops_cell,
                     "peak_note": "157.3 TFLOP/s counts packed FP32 FMAs; this build compiles "
                                  "with packed FP32 disabled, so the issue ceiling of its "
                                  "instructions is half of it"}
  variants = {"valu_row_stream": {"kernel_ms": kern_ms,
                                  "cell_levels_per_s": cell_levels / (kern_ms * 1e-3),
                                  "max_rel_err_vs_fp64": float(errs.max())}}
  if not args.no_variants:
    # the static Gram on the matrix cores (bf16 hi/lo split with three
    # products, or float32 MFMA): the per-pair MFMA kernel, one launch over
    # all 4096 pairs, recursion in FP32 on the VALU
    for name in ("bf16x3", "fp32"):
      mode = _lib.GRAM_MODES[name]

      def vstep(mode=mode):
        return _ops.signature_levels(X, Y, pairs, n_levels, order, True, _lib.STATIC_RBF,
                                     sigma, mode)
      vo = vstep()
      torch.cuda.synchronize()
      e0.record(stream)
      reps = 3
      for _ in range(reps):
        vo = vstep()
      e1.record(stream)
      torch.cuda.synchronize()
      vms = e0.elapsed_time(e1) / reps
      verr = pair_errors(vo.cpu().numpy())
      variants[name + "_mfma_per_pair"] = {
          "kernel_ms": vms, "cell_levels_per_s": cell_levels / (vms * 1e-3),
          "max_rel_err_vs_fp64": float(verr.max()),
          "kernel": "dp_mfma_kernel<8,2,16,%d>" % (1 if name == "bf16x3" else 0)}
  res["gram_variants"] = variants
  if not args.no_plain:
    # plain FP32: the reference dataflow over length buckets (pairs padded to
    # the bucket's lengths by repeating their last point: ΔΔ of the padding
    # is exactly zero), timed on a sample of pairs, counted in true cells
    prng = np.random.default_rng(11)
    sub = prng.choice(n, 512, replace=False)
    bl = 256
    keyx = -(-lx[sub] // bl) * bl
    keyy = -(-ly[sub] // bl) * bl
    groups = {}
    for p, kx, ky in zip(sub, keyx, keyy):
      groups.setdefault((int(kx), int(ky)), []).append(int(p))

    def padded(walk, starts, lens, idx, L):
      o = np.empty((len(idx), L, d), dtype=np.float32)
      for k, p in enumerate(idx):
        seq = walk[starts[p]:starts[p] + lens[p]]
        o[k, :lens[p]] = seq
        o[k, lens[p]:] = seq[-1]
      return torch.as_tensor(o, device=dev)
    batches = [(padded(P["xw"], P["xs"], lx, idx, kx), padded(P["yw"], P["ys"], ly, idx, ky),
                idx) for (kx, ky), idx in groups.items()]
    plain_out = {}
    for bx, by, idx in batches:   # warm-up + results
      lev = plain_fp32_levels_ho(bx, by, n_levels, order, sigma)
      for k, p in enumerate(idx):
        plain_out[p] = lev[k]
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for bx, by, idx in batches:
      plain_fp32_levels_ho(bx, by, n_levels, order, sigma)
    torch.cuda.synchronize()
    tp = time.perf_counter() - t2
    sub_cells = float(np.sum((lx[sub] - 1) * (ly[sub] - 1))) * n_levels
    pv = sub_cells / tp
    perr = max(float(torch.max(torch.abs(plain_out[p].cpu() - out[p].cpu())) /
                     torch.max(torch.abs(out[p].cpu()))) for p in sub[:8])
    res["plain_fp32"] = {"value": pv, "unit": "DP cell-levels/s",
                         "speedup_fused_over_plain": res["value"] / pv,
                         "sample": "512 random pairs batched in length buckets of 256 "
                                   "(about 32 pairs per batch), torch FP32 ops",
                         "max_rel_diff_vs_fused": perr}
    for v in variants.values():
      v["speedup_over_plain_fp32"] = v["cell_levels_per_s"] / pv
  return res


def bench_c3(args):
  """The headline line: C3, with the C4 forward+adjoint ratio and the C1
  figure attached (unless --no-extras)."""
  P = c3_host_problem()
  cpu, ref64 = (None, {}) if args.no_cpu else c3_cpu_baseline(P, args.cpu_seconds)
  res = run_c3(args, P, ref64)
  if cpu is not None:
    res["cpu_baseline"] = cpu
  if not args.no_extras:
    c4 = run_c4(args)
    res["fwd_bwd_over_fwd"] = c4["bwd_over_fwd"]
    res["c4_training_step"] = {
        k: c4[k] for k in ("config", "ms_per_step", "value", "unit", "fwd_ms",
                           "fwd_bwd_eager_ms", "fwd_bwd_graph_ms", "bwd_over_fwd",
                           "graph_replay_max_rel_diff_vs_eager",
                           "grad_scale_max_rel_err_vs_fp64", "loss_rel_err_vs_fp64",
                           "grad_check", "plain_fp32") if k in c4}
    c1 = run_c1(args)
    res["c1_symmetric_gram"] = {
        k: c1[k] for k in ("config", "ms_per_step", "value", "unit", "max_rel_err_vs_fp64",
                           "roofline", "plain_fp32", "gram_variants") if k in c1}
  print(json.dumps(res), flush=True)


def plain_levels_fo(X, Y, n_levels, sigma):
  """All pairs of X [n, L, d] and Y [m, L, d]: the reference dataflow of the
  first-order recursion (ksig/algorithms.py:47-85) on the materialised ΔΔ RBF
  Gram, in the inputs' dtype, differentiable by torch autograd."""
  Xn = (X * X).sum(-1)
  Yn = (Y * Y).sum(-1)
  dot = torch.einsum("ail,bjl->abij", X, Y)
  M = torch.exp(-(Xn[:, None, :, None] + Yn[None, :, None, :] - 2 * dot) /
                max(2 * sigma**2, 1e-12))
  M = torch.diff(torch.diff(M, dim=-2), dim=-1)
  K = [torch.ones(M.shape[:2], device=M.device, dtype=M.dtype), M.sum((-2, -1))]
  R = M
  for _ in range(1, n_levels):
    R = M * excl_cumsum(R, (-2, -1))
    K.append(R.sum((-2, -1)))
  return torch.stack(K, -1)


def robust_sqrt(v, eps=1e-12):
  """ksig/utils.py:173-182 with the level normalisation of ksig/kernels.py:180-214."""
  return torch.clamp(torch.sqrt(torch.clamp(v, min=eps)), min=eps)


def flat_sum(t: torch.Tensor) -> torch.Tensor:
  """Sum in two per-row stages (rows of 1024), so no reduction spans several
  blocks: such reductions record memset nodes, which misbehave in replayed
  HIP graphs."""
  v = t.reshape(-1)
  pad = (-v.numel()) % 1024
  if pad:
    v = torch.nn.functional.pad(v, (0, pad))
  return v.view(-1, 1024).sum(1).sum()


def run_c4(args) -> dict:
  """MMD^2 training step: generated paths with a learnable per-channel scale
  against fixed targets; forward + adjoint, eager and HIP-graph captured."""
  from ksig_amd import _lib, _ops, autograd
  dev = torch.device("cuda", 0)
  torch.cuda.set_device(dev)
  n, L, d, n_levels = 1024, 128, 32, 6
  sigma = math.sqrt(d)
  gen = torch.Generator(device=dev).manual_seed(44)
  Z = torch.randn((n, L, d), generator=gen, device=dev)
  Y = brownian(gen, n, L, d, dev)
  scale = torch.full((d,), 0.5, device=dev, dtype=torch.float32, requires_grad=True)
  Ys = _ops.SeqSet.dense(Y)
  triu = _ops.triu_pairs(n, dev)
  full = _ops.dense_pairs(n, n, dev)
  ti, tj = triu[:, 0].long(), triu[:, 1].long()
  diag_pos = torch.nonzero(ti == tj).flatten()
  off_x = torch.arange(n, device=dev, dtype=torch.int64) * L
  len_x = torch.full((n,), L, device=dev, dtype=torch.int32)
  with torch.no_grad():
    lev_yy = _ops.signature_levels(Ys, Ys, triu, n_levels, 1, True, _lib.STATIC_RBF, sigma)
    dy = lev_yy[diag_pos]
    kyy = (lev_yy / (robust_sqrt(dy[ti]) * robust_sqrt(dy[tj]))).mean(-1)
    syy = 2.0 * kyy[ti != tj].sum() / (n * (n - 1))
  offmask = (ti != tj).to(torch.float64)  # masks, not boolean indexing: capturable
  # The normalised MMD^2 terms as per-level bilinear forms u_m^T K_m v_m over
  # dense level matrices (u = 1 / sqrt(diag K_XX), v = 1 / sqrt(diag K_YY)):
  # their backward is a gather and batched products, where per-pair
  # normalisation through index_select would need an atomic index_add_ of
  # every pair into n diagonal entries.
  tri_flat = ti * n + tj
  v_y = (1.0 / robust_sqrt(dy)).t().contiguous().unsqueeze(-1)  # [levels+1, n, 1]
  nl1 = n_levels + 1

  def loss_fn(fwd_only=False):
    Xg = torch.cumsum(scale * Z, 1)
    xd = Xg.reshape(-1, d)
    Xs = _ops.SeqSet(xd, off_x, len_x, L, n)
    if fwd_only:
      lev_xx = _ops.signature_levels(Xs, Xs, triu, n_levels, 1, True, _lib.STATIC_RBF, sigma)
      lev_xy = _ops.signature_levels(Xs, Ys, full, n_levels, 1, True, _lib.STATIC_RBF, sigma)
    else:
      lev_xx = autograd.signature_levels(Xs, Xs, triu, n_levels, _lib.STATIC_RBF, sigma,
                                         xdata=xd, ydata=xd)
      lev_xy = autograd.signature_levels(Xs, Ys, full, n_levels, _lib.STATIC_RBF, sigma,
                                         xdata=xd, ydata=Ys.data)
    u = (1.0 / robust_sqrt(torch.index_select(lev_xx, 0, diag_pos))).t().contiguous()
    dxx = torch.zeros((n * n, nl1), device=dev, dtype=torch.float64).index_put(
        (tri_flat,), lev_xx * offmask[:, None])
    kxx_m = dxx.view(n, n, nl1).permute(2, 0, 1).contiguous()     # [levels+1, n, n]
    kxy_m = lev_xy.view(n, n, nl1).permute(2, 0, 1).contiguous()
    ut = u.unsqueeze(1)                                              # [levels+1, 1, n]
    sxx = torch.bmm(torch.bmm(ut, kxx_m), u.unsqueeze(-1)).sum() * (2.0 / (n * (n - 1) * nl1))
    sxy = torch.bmm(torch.bmm(ut, kxy_m), v_y).sum() * (1.0 / (n * n * nl1))
    return sxx + syy - 2.0 * sxy

  def train_step():
    scale.grad = None
    loss = loss_fn()
    loss.backward()
    return loss

  cells = (triu.shape[0] + full.shape[0]) * (L - 1) ** 2
  cell_levels = cells * n_levels
  stream = torch.cuda.current_stream(dev)

  def timed(fn, reps):
    fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(reps):
      fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / reps, e0.elapsed_time(e1) / reps

  with torch.no_grad():
    fwd_wall, fwd_ms = timed(lambda: loss_fn(True), max(2, args.steps // 2))
  for _ in range(args.warmup):
    train_step()
  eager_wall, eager_ms = timed(train_step, args.steps)
  loss_eager = float(train_step())
  grad_eager = scale.grad.detach().clone()

  # HIP graph of the whole step (forward, adjoint kernels, contractions)
  side = torch.cuda.Stream(dev)
  side.wait_stream(stream)
  with torch.cuda.stream(side):
    for _ in range(2):
      train_step()
  stream.wait_stream(side)
  graph = torch.cuda.CUDAGraph()
  scale.grad = None
  with torch.cuda.graph(graph):
    static_loss = loss_fn()
    static_loss.backward()
  static_grad = scale.grad
  replay_err = 0.0
  for _ in range(3):   # replays against the eager step, eager work in between
    graph.replay()
    torch.cuda.synchronize()
    replay_err = max(replay_err, abs(float(static_loss) - loss_eager) / abs(loss_eager),
                     float((static_grad - grad_eager).abs().max() / grad_eager.abs().max()))
    _ = torch.cumsum(Z[:4], 1).sum().item()
  graph_wall, graph_ms = timed(graph.replay, args.steps)

  # gradients vs autograd through the FP64 transcription, 16 x 16 subset
  ns = 16
  sc64 = scale.detach().double().clone().requires_grad_(True)
  Xg64 = torch.cumsum(sc64 * Z[:ns].double(), 1)
  kxx64 = plain_levels_fo(Xg64, Xg64, n_levels, sigma)
  kyy64 = plain_levels_fo(Y[:ns].double(), Y[:ns].double(), n_levels, sigma)
  kxy64 = plain_levels_fo(Xg64, Y[:ns].double(), n_levels, sigma)

  def mmd_sub(kxx, kyy, kxy):
    dxs = torch.diagonal(kxx, 0, 0, 1).t()
    dys = torch.diagonal(kyy, 0, 0, 1).t()
    nx = (kxx / (robust_sqrt(dxs[:, None]) * robust_sqrt(dxs[None]))).mean(-1)
    ny = (kyy / (robust_sqrt(dys[:, None]) * robust_sqrt(dys[None]))).mean(-1)
    nxy = (kxy / (robust_sqrt(dxs[:, None]) * robust_sqrt(dys[None]))).mean(-1)
    return autograd.mmd2_unbiased(nx, ny, nxy)
  l64 = mmd_sub(kxx64, kyy64, kxy64)
  l64.backward()
  sc32 = scale.detach().clone().requires_grad_(True)
  Xg32 = torch.cumsum(sc32 * Z[:ns], 1)
  Xs32 = _ops.SeqSet.dense(Xg32.detach())
  Ys32 = _ops.SeqSet.dense(Y[:ns].contiguous())
  pp = _ops.dense_pairs(ns, ns, dev)
  xd32 = Xg32.reshape(-1, d)
  k_xx = autograd.signature_levels(Xs32, Xs32, pp, n_levels, _lib.STATIC_RBF, sigma,
                                   xdata=xd32, ydata=xd32).view(ns, ns, -1)
  k_yy = _ops.signature_levels(Ys32, Ys32, pp, n_levels, 1, True, _lib.STATIC_RBF,
                               sigma).view(ns, ns, -1)
  k_xy = autograd.signature_levels(Xs32, Ys32, pp, n_levels, _lib.STATIC_RBF, sigma,
                                   xdata=xd32, ydata=Ys32.data).view(ns, ns, -1)
  l32 = mmd_sub(k_xx, k_yy, k_xy)
  l32.backward()
  grad_err = float((sc32.grad.double() - sc64.grad).abs().max() / sc64.grad.abs().max())
  loss_err = abs(float(l32) - float(l64)) / abs(float(l64))

  res = {
      "metric": "DP-cells/sec (pairs x (L-1)^2 x levels), MMD^2 training step fwd+bwd, fp32",
      "value": cell_levels / (graph_wall * 1e-3),
      "unit": "DP cell-levels/s",
      "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
      "ms_per_step": graph_wall,
      "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
      "dtype": "f32", "data": "synthetic (seeded Gaussian increments, learnable scale 0.5)",
      "config": {"workload": "MMD^2 training step: 1024 generated (L=128, d=32) vs 1024 "
                             "Brownian targets, DP kernel 6 levels, order 1, RBF sigma=sqrt(32), "
                             "normalised levels; K_XX on the upper triangle, K_YY constant",
                 "pairs": int(triu.shape[0] + full.shape[0]), "cells": cells},
      "fwd_ms": fwd_wall, "fwd_bwd_eager_ms": eager_wall, "fwd_bwd_graph_ms": graph_wall,
      "bwd_over_fwd": graph_wall / fwd_wall,
      "graph_replay_max_rel_diff_vs_eager": replay_err,
      "grad_scale_max_rel_err_vs_fp64": grad_err, "loss_rel_err_vs_fp64": loss_err,
      "grad_check": "16 x 16 subset, autograd through the FP64 torch transcription",
  }
  if not args.no_plain:
    # plain FP32 fallback: the reference dataflow with torch autograd, on an
    # 8 x 1024 slice of K_XY
    sc = scale.detach().clone().requires_grad_(True)

    def plain_step():
      sc.grad = None
      Xg = torch.cumsum(sc * Z[:8], 1)
      lv = plain_levels_fo(Xg, Y, n_levels, sigma)
      lv.mean().backward()
    plain_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(2):
      plain_step()
    torch.cuda.synchronize()
    tp = (time.perf_counter() - t0) / 2
    pv = 8 * n * (L - 1) ** 2 * n_levels / tp
    res["plain_fp32"] = {"value": pv, "unit": "DP cell-levels/s (fwd+bwd)",
                         "speedup_fused_over_plain": res["value"] / pv,
                         "sample": "8 x 1024 pairs of K_XY, torch FP32 ops + autograd"}
  return res


def bench_c2(args):
  """Signature-PDE kernel (dyadic order 1): forward Grams XX, YY, XY of 256 vs
  256 Brownian paths and the adjoint gradient of the unbiased MMD^2 with
  respect to X and the RBF bandwidth, checked against FP64 central finite
  differences."""
  from ksig_amd import _lib, _ops, autograd
  dev = torch.device("cuda", 0)
  torch.cuda.set_device(dev)
  n, L, d, dyadic = 256, 256, 4, 1
  gen = torch.Generator(device=dev).manual_seed(2)
  X = brownian(gen, n, L, d, dev)
  Y = brownian(gen, n, L, d, dev)
  triu = _ops.triu_pairs(n, dev)
  full = _ops.dense_pairs(n, n, dev)
  ti, tj = triu[:, 0].long(), triu[:, 1].long()
  offmask = (ti != tj).to(torch.float64)
  refined = ((L - 1) << dyadic) ** 2
  cells = (2 * triu.shape[0] + full.shape[0]) * refined

  def mmd(Xd, Yd, sig, grad):
    """MMD^2 over sequences X [n, L, d], Y (float32 or float64)."""
    Xs, Ys = _ops.SeqSet.dense(Xd.detach()), _ops.SeqSet.dense(Yd.detach())
    if grad:
      xd, yd = Xd.reshape(-1, d), Yd.reshape(-1, d)
      kxx = autograd.signature_pde(Xs, Xs, triu, _lib.STATIC_RBF, sig, dyadic, True, xd, xd)
      kyy = autograd.signature_pde(Ys, Ys, triu, _lib.STATIC_RBF, sig, dyadic, True, yd, yd)
      kxy = autograd.signature_pde(Xs, Ys, full, _lib.STATIC_RBF, sig, dyadic, True, xd, yd)
    else:
      s = float(sig)
      kxx = _ops.signature_pde(Xs, Xs, triu, True, _lib.STATIC_RBF, s, dyadic)
      kyy = _ops.signature_pde(Ys, Ys, triu, True, _lib.STATIC_RBF, s, dyadic)
      kxy = _ops.signature_pde(Xs, Ys, full, True, _lib.STATIC_RBF, s, dyadic)
    norm = n * (n - 1)
    return (2.0 * (kxx * offmask).sum() / norm + 2.0 * (kyy * offmask).sum() / norm
            - 2.0 * kxy.mean())

  stream = torch.cuda.current_stream(dev)

  def timed(fn, reps):
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
      fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / reps

  fwd_ms = timed(lambda: mmd(X, Y, 1.0, False), args.steps)
  Xp = X.clone().requires_grad_(True)
  sig = torch.tensor(1.0, device=dev, dtype=torch.float64, requires_grad=True)

  def fwd_bwd():
    Xp.grad = None
    sig.grad = None
    loss = mmd(Xp, Y, sig, True)
    loss.backward()
    return loss
  fb_ms = timed(fwd_bwd, args.steps)
  loss = fwd_bwd()
  gX, gS = Xp.grad.detach().clone(), float(sig.grad)

  # FP64 central finite differences (h = 1e-6) on 8 random coordinates and
  # the bandwidth, with the float64 instantiation of the same kernels
  X64, Y64 = X.double(), Y.double()
  rng = np.random.default_rng(8)
  coords = [(int(rng.integers(n)), int(rng.integers(L)), int(rng.integers(d))) for _ in range(8)]
  h = 1e-6
  fd_err = 0.0
  rows = []
  for (a, t, f) in coords:
    Xh = X64.clone()
    Xh[a, t, f] += h
    lp = float(mmd(Xh, Y64, 1.0, False))
    Xh[a, t, f] -= 2 * h
    lm = float(mmd(Xh, Y64, 1.0, False))
    fd = (lp - lm) / (2 * h)
    rows.append((fd, float(gX[a, t, f])))
  scale_fd = max(abs(r[0]) for r in rows)
  fd_err = max(abs(fd - g) for fd, g in rows) / scale_fd
  lp = float(mmd(X64, Y64, 1.0 + h, False))
  lm = float(mmd(X64, Y64, 1.0 - h, False))
  fd_s = (lp - lm) / (2 * h)
  sig_err = abs(fd_s - gS) / abs(fd_s)
  loss64 = float(mmd(X64, Y64, 1.0, False))

  res = {
      "metric": "PDE-cells/sec (pairs x (2^l (L-1))^2), signature-PDE fwd+bwd, fp32",
      "value": cells / (fb_ms * 1e-3),
      "unit": "PDE refined cells/s",
      "n_gpus": 1, "steps": args.steps, "warmup": 1,
      "ms_per_step": fb_ms,
      "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
      "dtype": "f32", "data": "synthetic (seeded Brownian paths)",
      "config": {"workload": "signature-PDE MMD^2: 256 vs 256 Brownian paths, L=256, d=4, "
                             "RBF sigma=1, dyadic order 1; K_XX, K_YY on the upper triangle",
                 "pairs": int(2 * triu.shape[0] + full.shape[0]), "refined_cells": cells},
      "fwd_ms": fwd_ms, "fwd_bwd_ms": fb_ms, "bwd_over_fwd": fb_ms / fwd_ms,
      "fwd_cells_per_s": cells / (fwd_ms * 1e-3),
      "loss": float(loss), "loss_rel_err_vs_fp64": abs(float(loss) - loss64) / abs(loss64),
      "grad_x_max_rel_err_vs_fd": fd_err, "grad_sigma_rel_err_vs_fd": sig_err,
      "fd": "FP64 central differences, h=1e-6, 8 random coordinates of X and sigma, with "
            "the float64 kernels (themselves checked against the FP64 oracle in tests)",
  }
  print(json.dumps(res), flush=True)


if __name__ == "__main__":
  main()
