"""Benchmark: batched binary-RBM loopy BP on one MI355X.

Workload (one "step"): BP.run over a batch of 64 binary RBMs with 500
visible x 100 hidden units (50,000 pairwise EnumFactors each, one W ~ N(0,
0.1^2) draw and bias draw per batch element), max-product (temperature 0),
damping 0.5, 200 iterations, FP32. Inputs are resident in GPU memory before
the timed region. The metric is edge-state message updates per second:
E_s * batch * iterations / time, E_s = 200,000.

Also reported: the same workload at temperature 1 (sum-product), the
roofline of the BP kernel (algorithmic bytes per iteration from
HIP-event-timed kernel time) and a CPU baseline (the NumPy FP32 oracle on
one batch element, bounded to ~10 s).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

N_VISIBLE, N_HIDDEN, BATCH, ITERS = 500, 100, 64, 200
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_per_iter(num_es, num_lp, num_edges, num_vs, batch):
  """8*E_s + 4*C + 12*V_s per batch element, plus 4*E of shared topology.

  Messages are read and written once (8 B per edge state), log-potentials
  read once, evidence read once and variable sums written and read once
  (12 B per variable state); the edge->variable topology is shared by the
  batch and counted once.
  """
  return batch * (8 * num_es + 4 * num_lp + 12 * num_vs) + 4 * num_edges


def cpu_baseline(fg, lp, ev, budget_s=10.0):
  """NumPy FP32 oracle on one batch element, single core, ~budget_s seconds."""
  from oracle import lbp  # pylint: disable=g-import-not-at-top
  from pgmax_amd.infer.inferer import InfererContext  # pylint: disable=g-import-not-at-top
  ctx = InfererContext(fg.bp_state)
  g = lbp.graph_from_context(ctx)
  es = ctx.var_states_for_edge_states.shape[0]
  ftov = np.zeros(es, np.float32)
  lbp.run_with_diffs(g, lp, ftov, ev, 1, 0.5, 0.0)  # warm-up
  iters, t0 = 0, time.perf_counter()
  while time.perf_counter() - t0 < budget_s or iters < 2:
    lbp.run_with_diffs(g, lp, ftov, ev, 2, 0.5, 0.0)
    iters += 2
  dt = time.perf_counter() - t0
  return {
      "value": es * iters / dt,
      "unit": "edge-state msg updates/s",
      "cores": 1,
      "kind": "port",
      "sample": f"NumPy FP32 oracle, 1 RBM {N_VISIBLE}x{N_HIDDEN}, max-product, "
                f"{iters} iterations in {dt:.1f} s",
  }


def main():
  ap = argparse.ArgumentParser()
  ap.add_argument("--gpus", type=int, default=1)
  ap.add_argument("--steps", type=int, default=10)
  ap.add_argument("--warmup", type=int, default=3)
  ap.add_argument("--no-cpu", action="store_true")
  args = ap.parse_args()
  if args.gpus != 1:
    print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported "
          "(independent graphs would shard across GPUs with no per-iteration "
          "communication; this build measures the one-GPU kernel).", file=sys.stderr)
    sys.exit(2)

  import torch  # pylint: disable=g-import-not-at-top
  import graphs  # pylint: disable=g-import-not-at-top
  from pgmax_amd import _lib  # pylint: disable=g-import-not-at-top
  from pgmax_amd import infer  # pylint: disable=g-import-not-at-top

  fg, _, _, lp_np, ev_np = graphs.rbm(N_VISIBLE, N_HIDDEN, batch=BATCH, seed=0)
  bp = infer.BP(fg.bp_state, temperature=0.0)
  base = bp.init()
  lp = torch.as_tensor(lp_np, dtype=torch.float32, device="cuda")
  ev = torch.as_tensor(ev_np, dtype=torch.float32, device="cuda")
  arrays = infer.BPArrays(log_potentials=lp, ftov_msgs=base.ftov_msgs, evidence=ev)
  g = bp.context.engine.graph
  num_es, num_lp, num_vs, num_edges = g.num_es, g.num_lp, g.num_vs, g.num_edges

  def timed(temperature, steps, warmup):
    for _ in range(warmup):
      bp.run(arrays, ITERS, 0.5, temperature)
    torch.cuda.synchronize()
    _lib.profile_read()
    _lib.set_option("profile", 1)
    t0 = time.perf_counter()
    for _ in range(steps):
      out = bp.run(arrays, ITERS, 0.5, temperature)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    _lib.set_option("profile", 0)
    kernel_ms, runs = _lib.profile_read()
    return dt, kernel_ms / max(runs, 1), out

  dt0, k0_ms, _ = timed(0.0, args.steps, args.warmup)
  dt1, k1_ms, _ = timed(1.0, max(args.steps // 2, 1), 1)
  updates = num_es * BATCH * ITERS
  value = updates * args.steps / dt0
  alg_bytes = algorithmic_bytes_per_iter(num_es, num_lp, num_edges, num_vs, BATCH) * ITERS
  achieved = alg_bytes / (k0_ms * 1e-3) / 1e9
  line = {
      "metric": "edge-state msg updates/s",
      "value": value,
      "unit": "edge-state msg updates/s",
      "n_gpus": 1,
      "steps": args.steps,
      "warmup": args.warmup,
      "ms_per_step": dt0 * 1e3 / args.steps,
      "higher_is_better": True,
      "scaling": "weak",
      "vs_baseline": None,
      "dtype": "f32",
      "data": "synthetic",
      "config": {
          "workload": f"rbm_{N_VISIBLE}v_{N_HIDDEN}h_batch{BATCH}_maxproduct_{ITERS}it",
          "num_edge_states": num_es, "batch": BATCH, "iters": ITERS,
          "damping": 0.5, "temperature": 0.0,
          "sumproduct_value": updates * max(args.steps // 2, 1) / dt1,
          "sumproduct_kernel_ms": k1_ms,
      },
      "roofline": {
          "bound": "hbm",
          "achieved": achieved,
          "peak": HBM_PEAK_GBS,
          "unit": "GB/s",
          "frac": achieved / HBM_PEAK_GBS,
          "traffic": None,
          "kernel_ms": k0_ms,
          "alg_bytes_per_launch": alg_bytes,
      },
  }
  if not args.no_cpu:
    line["cpu_baseline"] = cpu_baseline(fg, lp_np[0].astype(np.float32),
                                        ev_np[0].astype(np.float32))
  print(json.dumps(line), flush=True)


if __name__ == "__main__":
  main()
