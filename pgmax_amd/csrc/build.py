"""Builds libpgx.so (the HIP engine and its C ABI) in-tree for gfx950.

Usage: python -m pgmax_amd.csrc.build   (or __graft_entry__.build())
"""

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
LIB = os.path.join(PKG, "libpgx.so")

SOURCES = ["pgx_engine.hip", "pgx_plan.cpp", "pgx_mpnn.hip"]

# -ffp-contract=off: no fused multiply-adds, the reference rounds every op.
# -packed-fp32-ops off: packed FP32 instructions are avoided on these GPUs.
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc() -> str:
  for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
    if cand and os.path.exists(cand):
      return cand
  raise RuntimeError("hipcc not found; the HIP engine cannot be built")


def build(verbose: bool = False) -> str:
  """Compiles every source into pgmax_amd/libpgx.so and returns its path."""
  srcs = [os.path.join(HERE, s) for s in SOURCES if os.path.exists(os.path.join(HERE, s))]
  tmp = LIB + ".tmp"
  cmd = [hipcc()] + FLAGS + ["-I", os.path.join(ROOT, "include"), "-o", tmp] + srcs
  if verbose:
    print(" ".join(cmd), flush=True)
  res = subprocess.run(cmd, capture_output=True, text=True)
  if res.returncode != 0:
    sys.stderr.write(res.stdout + res.stderr)
    raise RuntimeError("hipcc failed building libpgx.so")
  if verbose and res.stderr.strip():
    sys.stderr.write(res.stderr)
  os.replace(tmp, LIB)
  return LIB


if __name__ == "__main__":
  print(build(verbose=True))
