"""Builds libpgx.so (the HIP engine and its C ABI) in-tree for gfx950.

Usage: python -m pgmax_amd.csrc.build   (or __graft_entry__.build())
"""

import concurrent.futures
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
LIB = os.path.join(PKG, "libpgx.so")

# the kernels are split over several files so they compile in parallel
SOURCES = ["pgx_engine.hip", "pgx_phase.hip", "pgx_task_t0.hip", "pgx_task_t1.hip",
           "pgx_bip.hip", "pgx_plan.cpp", "pgx_mpnn.hip", "pgx_vjp.hip"]

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
  """Compiles every source (in parallel) and links pgmax_amd/libpgx.so; returns its path."""
  srcs = [os.path.join(HERE, s) for s in SOURCES if os.path.exists(os.path.join(HERE, s))]
  compile_flags = [f for f in FLAGS if f != "-shared"]
  with tempfile.TemporaryDirectory(prefix="pgx_build_") as tmpdir:
    jobs = []
    for src in srcs:
      obj = os.path.join(tmpdir, os.path.basename(src) + ".o")
      cmd = [hipcc()] + compile_flags + ["-I", os.path.join(ROOT, "include"), "-c", src, "-o", obj]
      if verbose:
        print(" ".join(cmd), flush=True)
      jobs.append((obj, cmd))
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(jobs)) as ex:
      results = list(ex.map(lambda j: subprocess.run(j[1], capture_output=True, text=True), jobs))
    for (_, cmd), res in zip(jobs, results):
      if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed on {cmd[-3]}")
      if verbose and res.stderr.strip():
        sys.stderr.write(res.stderr)
    tmp = LIB + ".tmp"
    link = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + [o for o, _ in jobs]
    res = subprocess.run(link, capture_output=True, text=True)
    if res.returncode != 0:
      sys.stderr.write(res.stdout + res.stderr)
      raise RuntimeError("hipcc failed linking libpgx.so")
  os.replace(tmp, LIB)
  return LIB


if __name__ == "__main__":
  print(build(verbose=True))
