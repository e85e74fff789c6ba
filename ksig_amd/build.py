"""Build recipe for libksig_hip.so (gfx950).

Compiles every ``csrc/*.hip`` translation unit with hipcc in parallel and links
them into ``ksig_amd/libksig_hip.so`` (in-tree, so it travels with the source
tree).  Usage: ``python -m ksig_amd.build`` or ``ksig_amd.build.build()``.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libksig_hip.so"
ARCH = os.environ.get("KSIG_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 VALU forms are disabled on purpose for this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", str(ROOT / "include"), "-I", str(CSRC),
    "-Wno-unused-result",
]


def _compile(src: Path) -> Path:
  obj = OBJ / (src.stem + ".o")
  cmd = [HIPCC, *CFLAGS, "-c", str(src), "-o", str(obj)]
  r = subprocess.run(cmd, capture_output=True, text=True)
  if r.returncode != 0:
    raise RuntimeError(f"hipcc failed on {src.name}:\n{r.stderr[-8000:]}")
  return obj


def build(verbose: bool = True) -> Path:
  """Compile all HIP sources and link libksig_hip.so; returns its path."""
  OBJ.mkdir(exist_ok=True)
  srcs = sorted(CSRC.glob("*.hip"))
  workers = min(len(srcs), max(1, (os.cpu_count() or 2)))
  with cf.ThreadPoolExecutor(workers) as ex:
    objs = list(ex.map(_compile, srcs))
  tmp = LIB.with_suffix(".so.tmp")
  cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
         *map(str, objs)]
  r = subprocess.run(cmd, capture_output=True, text=True)
  if r.returncode != 0:
    raise RuntimeError(f"link failed:\n{r.stderr[-8000:]}")
  os.replace(tmp, LIB)
  if verbose:
    print(f"built {LIB} from {len(srcs)} sources", file=sys.stderr)
  return LIB


if __name__ == "__main__":
  build()
