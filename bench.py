"""Benchmark: batched binary-RBM loopy BP on one MI355X.

Workload (one "step"): BP.run over a batch of 64 binary RBMs with 500
visible x 100 hidden units (50,000 pairwise EnumFactors each, one W ~ N(0,
0.1^2) draw and bias draw per batch element), max-product (temperature 0),
damping 0.5, 200 iterations, FP32. Inputs are resident in GPU memory before
the timed region. The metric is edge-state message updates per second:
E_s * batch * iterations / time, E_s = 200,000.

Also reported:
  * the same workload at temperature 1 (sum-product);
  * roofline of the dominant kernel: its algorithmic bytes per launch over
    its average launch time (HIP events on the run's stream around every
    launch; around the 200 back-to-back launches of the bipartite iteration
    as a whole), and its HBM traffic per launch from two rocprofv3 --pmc passes
    (FETCH_SIZE, WRITE_SIZE), each scaled by the same counter's reading for a
    elementwise multiply of known size;
  * a CPU baseline: the C/OpenMP restatement of the reference's Enum
    iteration (oracle/lbp_cpu.c) on the box's host cores, bounded to ~10-20 s.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-traffic]
       python bench.py --workload glass|potts|deconv [--steps K] [--warmup W]

The other workloads (engine shape chosen by pgx_autotune first; no CPU
baseline or traffic passes) print one JSON line of the same form:
  glass   1,024 spin-glass tori of 128 x 128 (J ~ N(0,1), unaries ~ N(0,0.5)),
          sum-product, 100 iterations per step;
  potts   2,048 Potts grids, side ~ U{16..128}, 2..8 states, beta ~ U(0.2, 1.5),
          sum-product, 100 iterations per step (one union graph);
  deconv  binary deconvolution, 64 images 40 x 40, 5 features 6 x 6, OR/AND
          factors, max-product, 100 iterations per step.
"""

import argparse
import csv
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

N_VISIBLE, N_HIDDEN, BATCH, ITERS = 500, 100, 64, 200
HBM_PEAK_GBS = 8000.0
COPY_BYTES = 256 << 20  # calibration copy for the traffic counters


def algorithmic_bytes_per_iter(num_es, num_lp, num_edges, num_vs, batch):
  """8*E_s + 4*C + 12*V_s per batch element, plus 4*E of shared topology.

  Messages are read and written once (8 B per edge state), log-potentials
  read once, evidence read once and variable sums written and read once
  (12 B per variable state); the edge->variable topology is shared by the
  batch and counted once.
  """
  return batch * (8 * num_es + 4 * num_lp + 12 * num_vs) + 4 * num_edges


def stored_bytes(name, num_es, num_lp, num_vs, batch):
  """Bytes the kernel's own data layout moves per launch, where it differs
  from the algorithmic count: the bipartite iteration stores one float per
  binary edge (both states recoverable), so messages cost 4 bytes per edge
  state read + written instead of 8."""
  if name == "bipartite":
    slots = rbm_chunk_slots(N_VISIBLE, N_HIDDEN)
    return batch * (4 * num_es + 4 * num_lp + 4 * num_vs + 8 * slots) + 8 * (num_vs // 2)
  return None


def rbm_chunk_slots(n_visible, n_hidden):
  """Chunk sums per batch element of a binary RBM: each unit's edges in runs of 32, 2 states."""
  def runs(n):
    return (n + 31) // 32
  return 2 * (n_hidden * runs(n_visible) + n_visible * runs(n_hidden))


def kernel_bytes(name, num_es, num_lp, num_edges, num_vs, batch, iters):
  """Algorithmic bytes of one launch of a kernel of the BP iteration."""
  if name == "bipartite":
    # messages read + written (4 bytes per edge state each way, the reference's
    # representation), log-potentials, evidence read, chunk sums read +
    # written, per-unit (state, chunk slot) table
    slots = rbm_chunk_slots(N_VISIBLE, N_HIDDEN)
    return batch * (8 * num_es + 4 * num_lp + 4 * num_vs + 8 * slots) + 8 * (num_vs // 2)
  if name.startswith("fac_phase"):
    # messages read + written, log-potentials, variable sums read, edge table
    return batch * (8 * num_es + 4 * num_lp + 4 * num_vs) + 4 * num_edges
  if name == "fac_fused":
    # messages read + written, log-potentials, evidence read, edge tables
    return batch * (8 * num_es + 4 * num_lp + 4 * num_vs) + 8 * num_edges
  if name == "var_phase":
    # messages read, evidence read, variable sums written, adjacency
    return batch * (4 * num_es + 8 * num_vs) + 4 * num_edges
  if name == "task_loop":
    return algorithmic_bytes_per_iter(num_es, num_lp, num_edges, num_vs, batch) * iters
  return None


ROCPROF_KERNEL = {
    "bipartite": "k_bipartite<true, 32, 256>",
    "fac_phase_enum2x2": "k_fac_phase<true, 0>",
    "var_phase": "k_var_phase",
    "fac_fused": "k_fac_fused",
    "task_loop": "k_task_loop",
}


def cpu_baseline(lp, ev, budget_s=15.0):
  """C/OpenMP restatement of the Enum iteration over the batch, all host cores."""
  from oracle import cpu  # pylint: disable=g-import-not-at-top
  from oracle import lbp  # pylint: disable=g-import-not-at-top
  import graphs  # pylint: disable=g-import-not-at-top
  from pgmax_amd.infer.inferer import InfererContext  # pylint: disable=g-import-not-at-top
  if not os.path.exists(cpu.LIB):
    cpu.build()
  fg, _, _, _, _ = graphs.rbm(N_VISIBLE, N_HIDDEN, batch=1, seed=0)
  g = lbp.graph_from_context(InfererContext(fg.bp_state))
  es = g["var_states_for_edge_states"].shape[0]
  try:
    cores = len(os.sched_getaffinity(0))
  except AttributeError:  # pragma: no cover
    cores = os.cpu_count() or 1
  cores = max(1, min(cores, 16))
  threads = cpu.set_threads(cores)
  ft0 = np.zeros(es, np.float32)
  probe_iters = 4
  t0 = time.perf_counter()
  cpu.run_enum(g, lp, ft0, ev, probe_iters, 0.5, 0.0)
  per_iter = (time.perf_counter() - t0) / probe_iters
  iters = int(max(4, min(ITERS, budget_s / max(per_iter, 1e-6))))
  t0 = time.perf_counter()
  cpu.run_enum(g, lp, ft0, ev, iters, 0.5, 0.0)
  dt = time.perf_counter() - t0
  return {
      "value": es * lp.shape[0] * iters / dt,
      "unit": "edge-state msg updates/s",
      "cores": threads,
      "kind": "port",
      "sample": f"C/OpenMP restatement of the reference Enum iteration (oracle/lbp_cpu.c), "
                f"the same {lp.shape[0]} RBMs {N_VISIBLE}x{N_HIDDEN}, max-product, "
                f"{iters} of {ITERS} iterations in {dt:.1f} s on {threads} threads",
  }


def _pmc_pass(counter, kernel_key, tmp):
  """KiB per dispatch of `counter` for the BP kernel and for the copy."""
  out_dir = os.path.join(tmp, counter)
  cmd = ["timeout", "-s", "KILL", "150", "rocprofv3", "--pmc", counter, "--output-format", "csv",
         "-d", out_dir, "-o", "pmc", "--", sys.executable, os.path.abspath(__file__), "--inner"]
  res = subprocess.run(cmd, cwd=tmp, capture_output=True, text=True)
  files = glob.glob(os.path.join(out_dir, "**", "*counter_collection.csv"), recursive=True)
  if not files:
    raise RuntimeError(f"rocprofv3 --pmc {counter} wrote no counters (rc {res.returncode}): "
                       f"{res.stderr[-300:]}")
  vals, copy_vals = [], []
  with open(files[0]) as f:
    for row in csv.DictReader(f):
      name = row.get("Kernel_Name", "")
      if row.get("Counter_Name", counter) != counter:
        continue
      v = float(row["Counter_Value"])
      if kernel_key in name:
        vals.append(v)
      elif "elementwise" in name.lower() and "Mul" in name:
        copy_vals.append(v)
  if not vals or not copy_vals:
    raise RuntimeError(f"no {counter} rows for {kernel_key} or the calibration copy")
  return float(np.mean(vals)), float(max(copy_vals))


def measure_traffic(kernel_name):
  """HBM bytes per launch of the dominant kernel from two rocprofv3 --pmc passes."""
  key = ROCPROF_KERNEL.get(kernel_name)
  if key is None or shutil.which("rocprofv3") is None:
    return None, "rocprofv3 or kernel mapping unavailable"
  tmp = tempfile.mkdtemp(prefix="pgx_pmc_")
  try:
    fetch, fetch_copy = _pmc_pass("FETCH_SIZE", key, tmp)
    write, write_copy = _pmc_pass("WRITE_SIZE", key, tmp)
  except Exception as e:  # pylint: disable=broad-except
    return None, str(e)[:300]
  finally:
    shutil.rmtree(tmp, ignore_errors=True)
  kib = COPY_BYTES / 1024.0
  f_scale, w_scale = kib / fetch_copy, kib / write_copy
  traffic = (fetch * f_scale + write * w_scale) * 1024.0
  note = (f"FETCH_SIZE {fetch:.0f} KiB x {f_scale:.3f} + WRITE_SIZE {write:.0f} KiB x "
          f"{w_scale:.3f} per dispatch (scales from a {COPY_BYTES >> 20} MiB elementwise "
          f"multiply, which reads and writes {COPY_BYTES >> 20} MiB)")
  return traffic, note


def build_workload():
  import torch  # pylint: disable=g-import-not-at-top
  import graphs  # pylint: disable=g-import-not-at-top
  from pgmax_amd import infer  # pylint: disable=g-import-not-at-top
  fg, _, _, lp_np, ev_np = graphs.rbm(N_VISIBLE, N_HIDDEN, batch=BATCH, seed=0)
  bp = infer.BP(fg.bp_state, temperature=0.0)
  base = bp.init()
  lp = torch.as_tensor(lp_np, dtype=torch.float32, device="cuda")
  ev = torch.as_tensor(ev_np, dtype=torch.float32, device="cuda")
  arrays = infer.BPArrays(log_potentials=lp, ftov_msgs=base.ftov_msgs, evidence=ev)
  return bp, arrays, lp_np, ev_np


def inner():
  """One run of the workload plus the calibration copy (under rocprofv3 --pmc)."""
  import torch  # pylint: disable=g-import-not-at-top
  bp, arrays, _, _ = build_workload()
  bp.run(arrays, ITERS, 0.5, 0.0)
  src = torch.empty(COPY_BYTES // 4, dtype=torch.float32, device="cuda").fill_(1.0)
  dst = torch.empty_like(src)
  torch.mul(src, 2.0, out=dst)  # one elementwise dispatch: reads and writes COPY_BYTES
  torch.cuda.synchronize()


def other_workload(name):
  """(tuned shape, run(iters) callable, edge states per iteration, T, description)."""
  import torch  # pylint: disable=g-import-not-at-top
  import graphs  # pylint: disable=g-import-not-at-top
  from pgmax_amd import batch as pbatch  # pylint: disable=g-import-not-at-top
  from pgmax_amd import engine as pengine  # pylint: disable=g-import-not-at-top
  from pgmax_amd import infer  # pylint: disable=g-import-not-at-top
  rng = np.random.default_rng(1)
  if name == "glass":
    side, b = 128, 1024
    fg, _, _ = graphs.ising_torus(side, seed=0)
    bp = infer.BP(fg.bp_state, temperature=1.0)
    eng = bp.context.engine
    nf = 2 * side * side
    j = torch.as_tensor(rng.normal(size=(b, nf)).astype(np.float32), device="cuda")
    lp = torch.stack([j, -j, -j, j], -1).reshape(b, -1)
    ev = torch.as_tensor(rng.normal(0, 0.5, size=(b, eng.graph.num_vs)).astype(np.float32),
                         device="cuda")
    ft = torch.zeros(eng.graph.num_es, device="cuda")
    tuned = eng.autotune(lp, ft, ev, 0.5, 1.0, probe_iters=5)
    return (tuned, lambda n: eng.run(lp, ft, ev, n, 0.5, 1.0), eng.graph.num_es * b, 1.0,
            f"glass_{b}x{side}x{side}_sumproduct")
  if name == "potts":
    n = 2048
    side = rng.integers(16, 129, size=n)
    states = rng.integers(2, 9, size=n)
    beta = rng.uniform(0.2, 1.5, size=n)
    ws = [pbatch.grid_wiring(int(a), int(a), int(k), full_tables=True) for a, k in zip(side, states)]
    bb = pbatch.BatchedBP(ws, temperature=1.0)
    lp = torch.cat([torch.eye(int(k), device="cuda").mul(float(bt)).reshape(-1).repeat(
        w.num_log_potentials // (int(k) ** 2)) for w, k, bt in zip(ws, states, beta)])
    ev = torch.randn(bb.union.num_var_states, generator=torch.Generator(device="cuda").manual_seed(1),
                     device="cuda")
    tuned = bb.autotune(lp, None, ev, probe_iters=3)
    return (tuned, lambda it: bb.run(lp, None, ev, it), bb.union.num_edge_states, 1.0,
            f"potts_{n}_uneven_sumproduct")
  if name == "deconv":
    flat, meta = graphs.deconvolution_wiring(64, 1, 40, 40, 5, 6, 6)
    x = graphs.deconvolution_problem(64, 40, 40, n_feat_true=5, feat_true=6, seed=0)
    ev = torch.as_tensor(graphs.deconvolution_evidence(meta, x), device="cuda")
    eng = pengine.Engine.from_wiring(flat.to_engine_wiring())
    lp = torch.zeros(1, device="cuda")[:0]
    ft = torch.zeros(eng.graph.num_es, device="cuda")
    tuned = eng.autotune(lp, ft, ev, 0.5, 0.0, probe_iters=3)
    return (tuned, lambda it: eng.run(lp, ft, ev, it, 0.5, 0.0), eng.graph.num_es, 0.0,
            "deconv_64x40x40_5x6x6_maxproduct")
  raise ValueError(name)


def run_other(args):
  """One JSON line for a secondary workload (after pgx_autotune)."""
  import torch  # pylint: disable=g-import-not-at-top
  from pgmax_amd import _lib  # pylint: disable=g-import-not-at-top
  iters = 100
  tuned, run, es_per_iter, temperature, desc = other_workload(args.workload)
  for _ in range(args.warmup):
    run(iters)
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  for _ in range(args.steps):
    run(iters)
  torch.cuda.synchronize()
  dt = time.perf_counter() - t0
  _lib.profile_kernels()
  _lib.set_option("profile", 2)
  run(iters)
  torch.cuda.synchronize()
  _lib.set_option("profile", 0)
  _lib.profile_read()
  kernels = _lib.profile_kernels()
  print(json.dumps({
      "metric": "edge-state msg updates/s",
      "value": es_per_iter * iters * args.steps / dt,
      "unit": "edge-state msg updates/s",
      "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
      "ms_per_step": dt * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
      "vs_baseline": None, "dtype": "f32", "data": "synthetic",
      "config": {"workload": desc, "iters": iters, "damping": 0.5, "temperature": temperature,
                 "edge_states_per_iteration": es_per_iter,
                 "iteration_us": dt * 1e6 / args.steps / iters,
                 "tuned_shape": tuned,
                 "kernels_us_per_launch": {k: round(ms * 1e3 / n, 2)
                                           for k, (ms, n) in kernels.items()}},
  }), flush=True)


def main():
  ap = argparse.ArgumentParser()
  ap.add_argument("--gpus", type=int, default=1)
  ap.add_argument("--steps", type=int, default=10)
  ap.add_argument("--warmup", type=int, default=3)
  ap.add_argument("--no-cpu", action="store_true")
  ap.add_argument("--no-traffic", action="store_true")
  ap.add_argument("--inner", action="store_true", help=argparse.SUPPRESS)
  ap.add_argument("--workload", default="rbm", choices=["rbm", "glass", "potts", "deconv"])
  args = ap.parse_args()
  if args.gpus != 1:
    print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported "
          "(independent graphs would shard across GPUs with no per-iteration "
          "communication; this build measures the one-GPU kernel).", file=sys.stderr)
    sys.exit(2)
  if args.inner:
    inner()
    return
  if args.workload != "rbm":
    run_other(args)
    return

  import torch  # pylint: disable=g-import-not-at-top
  from pgmax_amd import _lib  # pylint: disable=g-import-not-at-top

  bp, arrays, lp_np, ev_np = build_workload()
  g = bp.context.engine.graph
  num_es, num_lp, num_vs, num_edges = g.num_es, g.num_lp, g.num_vs, g.num_edges

  def timed(temperature, steps, warmup):
    for _ in range(warmup):
      bp.run(arrays, ITERS, 0.5, temperature)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
      out = bp.run(arrays, ITERS, 0.5, temperature)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, out

  dt0, _ = timed(0.0, args.steps, args.warmup)
  steps1 = max(args.steps // 2, 1)
  dt1, _ = timed(1.0, steps1, 1)
  print(f"timed: {args.steps} max-product runs in {dt0:.3f} s, {steps1} sum-product runs "
        f"in {dt1:.3f} s", file=sys.stderr, flush=True)

  # per-kernel GPU time over the same runs (events around every launch)
  _lib.profile_kernels()
  _lib.set_option("profile", 2)
  for _ in range(max(args.steps, 1)):
    bp.run(arrays, ITERS, 0.5, 0.0)
  torch.cuda.synchronize()
  _lib.set_option("profile", 0)
  _lib.profile_read()
  kernels = _lib.profile_kernels()
  dom = max(kernels, key=lambda k: kernels[k][0])
  dom_ms = kernels[dom][0] / kernels[dom][1]
  dom_bytes = kernel_bytes(dom, num_es, num_lp, num_edges, num_vs, BATCH, ITERS)
  achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
  iter_ms = sum(ms for ms, _ in kernels.values()) / max(args.steps, 1) / ITERS

  updates = num_es * BATCH * ITERS
  value = updates * args.steps / dt0
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
          "sumproduct_value": updates * steps1 / dt1,
          "iteration_gpu_us": iter_ms * 1e3,
          "iteration_alg_GBps": algorithmic_bytes_per_iter(
              num_es, num_lp, num_edges, num_vs, BATCH) / (iter_ms * 1e-3) / 1e9,
          "kernels_us_per_launch": {k: round(ms * 1e3 / n, 2) for k, (ms, n) in kernels.items()},
      },
      "roofline": {
          "bound": "hbm",
          "kernel": dom,
          "achieved": achieved,
          "peak": HBM_PEAK_GBS,
          "unit": "GB/s",
          "frac": achieved / HBM_PEAK_GBS,
          "traffic": None,
          "alg_bytes_per_launch": dom_bytes,
          "avg_launch_us": dom_ms * 1e3,
      },
  }
  sb = stored_bytes(dom, num_es, num_lp, num_vs, BATCH)
  if sb is not None:
    line["roofline"]["stored_bytes_per_launch"] = sb
    line["roofline"]["stored_GBps"] = sb / (dom_ms * 1e-3) / 1e9
  if not args.no_traffic:
    traffic, note = measure_traffic(dom)
    line["roofline"]["traffic"] = traffic
    line["roofline"]["traffic_note"] = note
  if not args.no_cpu:
    line["cpu_baseline"] = cpu_baseline(lp_np.astype(np.float32), ev_np.astype(np.float32))
  print(json.dumps(line), flush=True)


if __name__ == "__main__":
  main()
