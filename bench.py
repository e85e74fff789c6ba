"""Benchmark of the signature-kernel hot path on one MI355X.

Workload (default): 512 Brownian paths (cumsum of N(0, 1/L) increments),
L = 200, d = 8, float32; truncated signature kernel by dynamic programming,
n_levels = 5, order = 1, RBF bandwidth sqrt(8); full symmetric Gram plus
diagonal, evaluated as the n(n+1)/2 = 131,328 upper-triangle pairs (the lower
triangle is the mirror image).  A step is one pass of the fused HIP sweep over
all those pairs with the inputs already resident in HBM.

Metric: DP cell-levels per second = evaluated pairs x (L-1)^2 x n_levels / s.

Prints ONE JSON line (rank 0) with the measured value, a roofline object for
the sweep kernel (HIP-event timing on the launch stream), the FP64 CPU oracle
timed on a bounded sample of the same workload, and the reference dataflow
restated in plain FP32 GPU tensor ops as a comparator.
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


def cpu_baseline(X64: np.ndarray, n_levels: int, sigma: float, seconds: float):
  """FP64 NumPy oracle (the reference algorithm and dataflow) on one host
  core, over seeded pairs of the same workload until `seconds` elapse."""
  from oracle import sigkern as O
  try:
    from threadpoolctl import threadpool_limits
    lim = threadpool_limits(1)
  except Exception:  # pragma: no cover
    lim = None
  rng = np.random.default_rng(7)
  n = X64.shape[0]
  L = X64.shape[1]
  batch = 8
  cells = 0
  pairs = 0
  t0 = time.perf_counter()
  while time.perf_counter() - t0 < seconds:
    a = rng.integers(0, n, batch)
    b = rng.integers(0, n, batch)
    M = O.rbf_gram(X64[a], X64[b], sigma)
    O.signature_kern(M, n_levels, 1, True, True)
    cells += batch * (L - 1) ** 2 * n_levels
    pairs += batch
  dt = time.perf_counter() - t0
  if lim is not None:
    lim.unregister() if hasattr(lim, "unregister") else None
  return {"value": cells / dt, "unit": "DP cell-levels/s", "cores": 1, "kind": "port",
          "sample": f"{pairs} random pairs of the workload (L={L}, d={X64.shape[2]}, "
                    f"{n_levels} levels), FP64 NumPy oracle, {dt:.1f} s on 1 host core"}


def main():
  ap = argparse.ArgumentParser()
  ap.add_argument("--gpus", type=int, default=1)
  ap.add_argument("--steps", type=int, default=20)
  ap.add_argument("--warmup", type=int, default=3)
  ap.add_argument("--cpu-seconds", type=float, default=12.0)
  ap.add_argument("--no-cpu", action="store_true")
  ap.add_argument("--no-plain", action="store_true")
  ap.add_argument("--gram", default="valu", choices=["valu", "fp32", "bf16x3"],
                  help="static-kernel Gram evaluation inside the sweep")
  ap.add_argument("--no-variants", action="store_true")
  ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                  help="only run a device-to-device copy of MIB MiB (5 times): the "
                       "reference point for scaling HBM byte counters")
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
    for _ in range(5):
      dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4,
                      "calls": 5}), flush=True)
    return

  import ksig_amd
  from ksig_amd import _lib, _ops
  from oracle import sigkern as O

  dev = torch.device("cuda", 0)
  torch.cuda.set_device(dev)
  n, L, d, n_levels = 512, 200, 8, 5
  sigma = math.sqrt(d)
  gen = torch.Generator(device=dev).manual_seed(1234)
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
      "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                   "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                   "traffic": None, "kernel": ("dp_mfma_kernel<5,1,8,%d>" % (0 if args.gram == "fp32" else 1)
                              if args.gram != "valu" else "dp_stream_kernel<5,1,8,4,0,true>"),
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
  if not args.no_cpu:
    res["cpu_baseline"] = cpu_baseline(X64, n_levels, sigma, args.cpu_seconds)
  print(json.dumps(res), flush=True)


if __name__ == "__main__":
  main()
