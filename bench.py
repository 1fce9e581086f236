"""Benchmark: loopy BP (the reference's BP.run) on one MI355X.

Default workload (one "step" = one BP.run): 2,048 uneven open Potts grids in
one disjoint union -- side ~ U{16..128}, states ~ U{2..8}, 4-neighbour
pairwise EnumFactors beta * I_S with beta ~ U(0.2, 1.5), unaries ~ N(0, 1)
(``pgmax_amd.workloads.potts_batch``, seed 0) -- sum-product (temperature
1), damping 0.5, 500 iterations, FP32, engine shape chosen by the project's
autotuner. Inputs are resident in GPU memory before the timed region. The
metric is edge-state message updates per second: E_s * iterations / time.

The JSON line also carries:
  * roofline of the dominant kernel: its algorithmic bytes per launch (the
    messages read and written at 8 bytes per edge state, log-potentials,
    variable sums and edge table it must touch) over its average launch time
    from HIP events on the run's stream, plus its HBM traffic per launch
    from two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE), each scaled by
    the same counter's reading for an elementwise multiply of known size;
  * the whole iteration's algorithmic bandwidth (8 E_s + 4 C + 12 V_s + 4 E
    bytes per iteration);
  * time to max |message change| < 1e-4 per graph (pgmax/infer/bp.py:136
    semantics, per graph) for plain batched LBP;
  * a CPU baseline: the C/OpenMP restatement of the reference's Enum
    iteration (oracle/lbp_cpu.c) on a bounded sample of the same graphs.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-traffic]
       python bench.py --workload rbm|glass|deconv|glass_grad [--steps K] [--warmup W]
       python bench.py --warm-predict   (the warm-start prediction alone)

Secondary workloads (same line format):
  rbm     64 binary RBMs 500 visible x 100 hidden (one W draw each),
          max-product, 200 iterations (working set inside the 256 MB
          Infinity Cache: not an HBM measurement);
  glass   1,024 toroidal spin glasses of 128 x 128, J ~ N(0,1), unaries
          ~ N(0, 0.5), sum-product, 100 iterations;
  deconv  binary deconvolution, 64 images 40 x 40, 5 features 6 x 6, OR/AND
          factors, max-product, 100 iterations;
  glass_grad  a gradient step: 256 spin glasses 128 x 128, 100 iterations
          forward, then the reverse-mode product of the final messages.
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

HBM_PEAK_GBS = 8000.0
COPY_BYTES = 256 << 20  # calibration multiply for the traffic counters
CPU_BUDGET_S = 15.0

FAC_KINDS = {"fac_phase_enum2x2": 0, "fac_phase_enum_pair": 1, "fac_phase_enum": 2,
             "fac_phase_or": 3, "fac_phase_and": 4, "fac_phase_pool": 5}


def rocprof_key(kernel, temperature, lanes=False):
  """Substring of the rocprofv3 name of a kernel id (pgmax_amd._lib.KERNEL_NAMES);
  lanes: square pairwise tables ran on lane groups (k_pair_lanes)."""
  t0 = "true" if temperature == 0.0 else "false"
  if kernel == "fac_phase_enum_pair" and lanes:
    return f"k_pair_lanes<{t0}>"
  if kernel == "fac_fused" and lanes:
    return f"k_pair_lanes_fused<{t0}>"
  if kernel in FAC_KINDS:
    return f"k_fac_phase<{t0}, {FAC_KINDS[kernel]}>"
  return {"bipartite": f"k_bipartite<{t0}, ", "var_phase": "k_var_phase",
          "fac_fused": f"k_fac_fused<{t0}, ", "task_loop": "k_task_loop<",
          "vjp_pair_bwd": "k_pair_bwd<", "vjp_pair_fwd": "k_pair_fwd<",
          "vjp_var_sums": "k_var_sums", "vjp_var_bwd": "k_gvs"}.get(kernel)


def alg_bytes_iter(s):
  """8 E_s + 4 C + 12 V_s per batch element, plus 4 E of shared topology.

  Messages are read and written once (8 B per edge state), log-potentials
  read once, evidence read once and variable sums written and read once
  (12 B per variable state); the edge -> variable table is counted once.
  """
  return s["batch"] * (8 * s["es"] + 4 * s["lp"] + 12 * s["vs"]) + 4 * s["edges"]


def kernel_bytes(name, s):
  """Algorithmic bytes of one launch of a kernel of the BP iteration.

  ``s`` holds the sizes of the part of the graph the kernel covers (the
  factor-phase kernels of one block kind cover that kind's factors and
  their variables).
  """
  b = s["batch"]
  if name == "bipartite":
    # one launch is one whole iteration: its algorithmic bytes (the kernel's
    # own chunk-sum scratch and unit tables are traffic, not algorithmic)
    return alg_bytes_iter(s)
  if name.startswith("fac_phase"):
    # messages read + written, log-potentials, variable sums read, edge table
    return b * (8 * s["es"] + 4 * s["lp"] + 4 * s["vs"]) + 4 * s["edges"]
  if name == "fac_fused":
    return b * (8 * s["es"] + 4 * s["lp"] + 4 * s["vs"]) + 8 * s["edges"]
  if name == "var_phase":
    # messages read, evidence read, variable sums written, adjacency
    return b * (4 * s["es"] + 8 * s["vs"]) + 4 * s["edges"]
  if name == "task_loop":
    return alg_bytes_iter(s) * s["iters"]
  # reverse mode on pairwise graphs (one thread per factor)
  if name == "vjp_pair_bwd":
    # messages read; message and vt cotangents read + written; log-potentials
    # read, their cotangents read + written; variable sums and their
    # cotangents read; the edge-state -> variable-state table
    return b * (20 * s["es"] + 12 * s["lp"] + 8 * s["vs"]) + 4 * s["es"]
  if name == "vjp_pair_fwd":
    return b * (8 * s["es"] + 4 * s["lp"] + 4 * s["vs"]) + 4 * s["es"]
  if name == "vjp_var_sums":
    # messages read, evidence read, sums written, adjacency
    return b * (4 * s["es"] + 8 * s["vs"]) + 4 * (s["es"] + s["vs"])
  if name == "vjp_var_bwd":
    # vt cotangents read, sums written, evidence cotangents read + written
    return b * (4 * s["es"] + 12 * s["vs"]) + 4 * (s["es"] + s["vs"])
  return None


def host_cores():
  try:
    n = len(os.sched_getaffinity(0))
  except AttributeError:  # pragma: no cover
    n = os.cpu_count() or 1
  return max(1, min(n, 16))


# ---------------------------------------------------------------------------
# learned warm start vs plain LBP (spin-glass tori)
# ---------------------------------------------------------------------------

WARM_TRAIN_SIDES = (32, 48, 64, 80, 96, 112, 128)


def predict_alg_bytes(num_edges, num_vars, batch):
  """Fewest HBM bytes one prediction can move: the input layer reads 16 bytes
  of log-potentials per factor (8 per edge) and 8 of evidence per variable
  and writes a 128-byte BF16 row per edge; each of the three layers reads
  its input rows once and writes 128 bytes per edge, the last one 8 bytes of
  messages per edge instead."""
  rows, var_rows = batch * num_edges, batch * num_vars
  return rows * (8 + 128 + 3 * 128 + 2 * 128 + 8) + var_rows * 8


def predict_timing(net, lp_t, ev_t, reps=5):
  """Steady-state prediction time (CUDA events; buffers sized by a first
  call) and its rate in algorithmic bytes."""
  import torch  # pylint: disable=g-import-not-at-top
  net.predict(lp_t, ev_t)
  a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
  a.record()
  for _ in range(reps):
    net.predict(lp_t, ev_t)
  b.record()
  b.synchronize()
  ms = a.elapsed_time(b) / reps
  g = net.graph
  batch = lp_t.shape[0] if lp_t.dim() > 1 else 1
  alg = predict_alg_bytes(g.num_edges, g.num_vars, batch)
  return {"predict_ms": ms, "alg_bytes": alg, "alg_GBps": alg / (ms * 1e-3) / 1e9,
          "alg_floor_ms_at_8TBps": alg / HBM_PEAK_GBS / 1e6}


def warm_predict_bench(grids=1024, side=128):
  """Prediction alone on grids toroidal spin glasses side x side (weights
  from the seed-0 initialisation; the cost does not depend on them)."""
  import torch  # pylint: disable=g-import-not-at-top
  from pgmax_amd import engine  # pylint: disable=g-import-not-at-top
  from pgmax_amd import warmstart  # pylint: disable=g-import-not-at-top
  from pgmax_amd import workloads  # pylint: disable=g-import-not-at-top
  w, lp, ev = workloads.spin_glass_batch(side, grids, seed=99)
  g = engine.Graph(w.to_engine_wiring())
  net = warmstart.WarmStart(g, seed=0)
  res = predict_timing(net, torch.as_tensor(lp, device="cuda"), torch.as_tensor(ev, device="cuda"))
  res.update({"grids": grids, "side": side})
  net.close()
  g.close()
  return res


def warm_start_eval(train_grids=20000, eval_grids=1024, eval_side=128, budget=1000, tol=1e-4,
                    train_steps=600, damping=0.5, temperature=1.0):
  """Time to max |message change| <= tol, learned warm start + batched LBP
  against plain batched LBP from zero messages.

  Training set: train_grids toroidal spin glasses (J ~ N(0,1), fields
  ~ N(0, 0.5)) spread evenly over the sides WARM_TRAIN_SIDES; plain LBP runs
  each to tol within `budget` iterations, and the graphs that get there give
  the training targets (their converged messages). The network (one set of
  weights for every side) trains with the project's own kernels. Evaluation:
  eval_grids held-out grids of side eval_side (a different seed), both
  methods to tol with the same budget; a warm-started graph that misses it is
  rerun from zeros (warmstart.solve), and its iterations and time count.
  """
  import torch  # pylint: disable=g-import-not-at-top
  from pgmax_amd import engine  # pylint: disable=g-import-not-at-top
  from pgmax_amd import warmstart  # pylint: disable=g-import-not-at-top
  from pgmax_amd import workloads  # pylint: disable=g-import-not-at-top
  t_start = time.perf_counter()
  per = -(-train_grids // len(WARM_TRAIN_SIDES))
  engines, nets, samples, conv_by_side = {}, [], [], {}
  for k, side in enumerate(WARM_TRAIN_SIDES):
    w, lp, ev = workloads.spin_glass_batch(side, per, seed=1000 + k)
    eng = engine.Engine.from_wiring(w.to_engine_wiring())
    engines[side] = eng
    lp_t = torch.as_tensor(lp, device="cuda")
    ev_t = torch.as_tensor(ev, device="cuda")
    zero = torch.zeros((per, w.num_edge_states), device="cuda")
    out, deltas, iters = eng.run_converge(lp_t, zero, ev_t, budget, tol, damping, temperature)
    last = deltas[:, 0].gather(-1, (iters[:, 0].long() - 1).clamp(min=0)[:, None])[:, 0]
    keep = torch.nonzero(last <= tol).reshape(-1)
    conv_by_side[side] = int(keep.numel())
    print(f"warm start: side {side}: {keep.numel()} of {per} training grids reach {tol} "
          f"within {budget} iterations", file=sys.stderr, flush=True)
    batches = [(lp_t[keep[i:i + 32]], ev_t[keep[i:i + 32]], out[keep[i:i + 32]])
               for i in range(0, int(keep.numel()), 32)]
    nets.append(warmstart.WarmStart(eng, seed=0))
    samples.append(batches)
    del out, deltas, zero
  n_conv = sum(conv_by_side.values())
  t_train = time.perf_counter()
  losses = warmstart.fit_shared(nets, samples, train_steps if n_conv else 0, lr=3e-3)
  train_s = time.perf_counter() - t_train
  net = nets[WARM_TRAIN_SIDES.index(eval_side)]
  # held-out evaluation grids
  w, lp, ev = workloads.spin_glass_batch(eval_side, eval_grids, seed=99)
  eng = engines[eval_side]
  lp_t = torch.as_tensor(lp, device="cuda")
  ev_t = torch.as_tensor(ev, device="cuda")
  eng.run(lp_t, torch.zeros((eval_grids, w.num_edge_states), device="cuda"), ev_t, 2, damping,
          temperature)  # warm the shapes up
  res = {}
  for name, which in (("plain", None), ("warm", net)):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = warmstart.solve(eng, which, lp_t, ev_t, max_iters=budget, tol=tol, damping=damping,
                        temperature=temperature)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    it = r.iterations.float().cpu().numpy().reshape(-1)
    fell = r.fell_back.cpu().numpy().reshape(-1)
    res[name] = {"seconds": dt, "iters_mean": float(it.mean()), "predict_ms": r.predict_ms,
                 "fallbacks": int(fell.sum()), "ftov": r.ftov}
  steady = predict_timing(net, lp_t, ev_t)
  # graphs that reach tol: plain runs from zero; warm runs without a fallback
  d_plain = eng.run_converge(lp_t, torch.zeros_like(res["plain"]["ftov"]), ev_t, budget, tol,
                             damping, temperature)
  it_p, dl_p = d_plain[2][:, 0], d_plain[1][:, 0]
  conv_plain = int((dl_p.gather(-1, (it_p.long() - 1).clamp(min=0)[:, None])[:, 0] <= tol).sum())
  # final beliefs of the two methods (equal where the warm start fell back)
  b_p = eng.beliefs(ev_t, res["plain"]["ftov"]).reshape(eval_grids, -1, 2)
  b_w = eng.beliefs(ev_t, res["warm"]["ftov"]).reshape(eval_grids, -1, 2)
  belief_diff = float(((b_p[..., 1] - b_p[..., 0]) - (b_w[..., 1] - b_w[..., 0])).abs().max())
  out = {
      "tol": tol, "budget_iters": budget, "damping": damping, "temperature": temperature,
      "train": {"grids": per * len(WARM_TRAIN_SIDES), "sides": list(WARM_TRAIN_SIDES),
                "converged_targets": n_conv, "converged_by_side": conv_by_side,
                "steps": int(losses.size), "seconds": round(train_s, 1),
                "loss_first": float(losses[:10].mean()) if losses.size else None,
                "loss_last": float(losses[-10:].mean()) if losses.size else None},
      "eval": {"grids": eval_grids, "side": eval_side, "seed": 99,
               "plain_converged": conv_plain,
               "warm_converged_without_fallback": eval_grids - res["warm"]["fallbacks"],
               "plain_seconds": res["plain"]["seconds"], "warm_seconds": res["warm"]["seconds"],
               "plain_iters_mean": res["plain"]["iters_mean"],
               "warm_iters_mean": res["warm"]["iters_mean"],
               "warm_predict_ms": res["warm"]["predict_ms"],
               "predict_steady": steady,
               "speedup_plain_over_warm": res["plain"]["seconds"] / res["warm"]["seconds"],
               "max_belief_logratio_diff": belief_diff},
      "seconds_total": round(time.perf_counter() - t_start, 1),
  }
  for eng_k in engines.values():
    eng_k.graph.close()
  for n in nets:
    n.close()
  return out


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------


class Potts:
  """2,048 uneven open Potts grids (one union graph), sum-product, 500 iterations."""

  name = "potts"
  iters = 500
  temperature = 1.0
  damping = 0.5

  def __init__(self):
    import torch  # pylint: disable=g-import-not-at-top
    from pgmax_amd import batch  # pylint: disable=g-import-not-at-top
    from pgmax_amd import workloads  # pylint: disable=g-import-not-at-top
    t0 = time.perf_counter()
    self.pb = workloads.potts_batch(2048, seed=0)
    self.bb = batch.BatchedBP(self.pb.wirings, temperature=self.temperature)
    lay = self.bb.layout
    # every graph is Enum-only: union arrays are the graphs' arrays in order
    assert all(ix.size == 0 or ix[0] + ix.size - 1 == ix[-1] for ix in lay.lp_index)
    self.lp = torch.as_tensor(self.pb.log_potentials, device="cuda")
    self.ev = torch.as_tensor(self.pb.evidence, device="cuda")
    self.ft = torch.zeros(self.bb.union.num_edge_states, device="cuda")
    self.setup_s = time.perf_counter() - t0
    g = self.bb.graph
    st = self.pb.states.astype(np.int64)
    sd = self.pb.sides.astype(np.int64)
    nf = 2 * sd * (sd - 1)
    part = lambda m: {"batch": 1, "es": int((2 * nf * st)[m].sum()),
                      "lp": int((nf * st * st)[m].sum()), "vs": int((sd * sd * st)[m].sum()),
                      "edges": int((2 * nf)[m].sum()), "iters": self.iters}
    self.sizes = {"batch": 1, "es": g.num_es, "lp": g.num_lp, "vs": g.num_vs,
                  "edges": g.num_edges, "iters": self.iters}
    self.kind_sizes = {"fac_phase_enum2x2": part(st == 2), "fac_phase_enum_pair": part(st > 2)}
    self.tuned = None

  def autotune(self):
    self.tuned = self.bb.autotune(self.lp, self.ft, self.ev, self.damping, self.temperature,
                                  probe_iters=3)
    return self.tuned

  def run(self, iters=None):
    return self.bb.run(self.lp, self.ft, self.ev, iters or self.iters, self.damping,
                       self.temperature)

  def updates_per_step(self):
    return self.sizes["es"] * self.iters

  def sizes_for(self, kernel):
    return self.kind_sizes.get(kernel, self.sizes)

  def config(self):
    return {"workload": f"potts_2048_uneven_side16-128_states2-8_sumproduct_{self.iters}it",
            "graphs": 2048, "seed": 0, "num_edge_states": self.sizes["es"],
            "num_log_potentials": self.sizes["lp"], "num_var_states": self.sizes["vs"],
            "iters": self.iters, "damping": self.damping, "temperature": self.temperature,
            "setup_s": round(self.setup_s, 1), "tuned_shape": self.tuned}

  def time_to_tol(self, tol=1e-4, max_iters=500):
    """Plain batched LBP until every graph's max |change| <= tol (pgx_run_converge)."""
    import torch  # pylint: disable=g-import-not-at-top
    self.bb.run_converge(self.lp, self.ft, self.ev, 2, tol, self.damping, self.temperature)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, deltas, iters = self.bb.run_converge(self.lp, self.ft, self.ev, max_iters, tol,
                                            self.damping, self.temperature)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    it = iters.cpu().numpy()
    last = deltas.cpu().numpy()[np.arange(it.size), np.maximum(it - 1, 0)]
    conv = (last <= tol)
    return {"tol": tol, "max_iters": max_iters, "seconds": dt,
            "graphs_converged": int(conv.sum()), "graphs": int(it.size),
            "iters_mean": float(it.mean()), "iters_max": int(it.max()),
            "method": "plain batched LBP from zero messages, per-graph early exit"}

  def cpu_baseline(self):
    from oracle import cpu  # pylint: disable=g-import-not-at-top
    from oracle import lbp  # pylint: disable=g-import-not-at-top
    if not os.path.exists(cpu.LIB):
      cpu.build()
    cores = host_cores()
    threads = cpu.set_threads(cores)
    # a sample: the first graphs of the batch, about 6 M edge states
    pb, lay = self.pb, self.bb.layout
    k, es = 0, 0
    while k < pb.num_graphs and es < 6_000_000:
      es += pb.wirings[k].num_edge_states
      k += 1
    graphs = [lbp.graph_from_flat(w) for w in pb.wirings[:k]]
    lps = [pb.log_potentials[lay.lp_index[g][0]:lay.lp_index[g][-1] + 1] for g in range(k)]
    evs = [pb.evidence[lay.vs_offsets[g]:lay.vs_offsets[g + 1]] for g in range(k)]
    fts = [np.zeros(w.num_edge_states, np.float32) for w in pb.wirings[:k]]
    t0 = time.perf_counter()
    cpu.run_enum_graphs(graphs, lps, evs, fts, 2, self.damping, self.temperature, cores)
    per_iter = (time.perf_counter() - t0) / 2
    iters = int(max(2, min(self.iters, CPU_BUDGET_S / max(per_iter, 1e-6))))
    t0 = time.perf_counter()
    cpu.run_enum_graphs(graphs, lps, evs, fts, iters, self.damping, self.temperature, cores)
    dt = time.perf_counter() - t0
    return {"value": es * iters / dt, "unit": "edge-state msg updates/s", "cores": threads,
            "kind": "port",
            "sample": f"C/OpenMP restatement of the reference Enum iteration (oracle/lbp_cpu.c), "
                      f"graphs 0..{k - 1} of the same batch ({es} edge states), sum-product, "
                      f"{iters} of {self.iters} iterations in {dt:.1f} s on {threads} threads "
                      f"(parallel over graphs)"}

  def close(self):
    self.bb.close()


class Rbm:
  """64 binary RBMs 500 x 100, max-product, 200 iterations (cache-resident)."""

  name = "rbm"
  iters = 200
  temperature = 0.0
  damping = 0.5
  n_visible, n_hidden, batch = 500, 100, 64

  def __init__(self):
    import torch  # pylint: disable=g-import-not-at-top
    import graphs  # pylint: disable=g-import-not-at-top
    from pgmax_amd import infer  # pylint: disable=g-import-not-at-top
    t0 = time.perf_counter()
    fg, _, _, lp_np, ev_np = graphs.rbm(self.n_visible, self.n_hidden, batch=self.batch, seed=0)
    self.bp = infer.BP(fg.bp_state, temperature=self.temperature)
    base = self.bp.init()
    self.lp_np, self.ev_np = lp_np.astype(np.float32), ev_np.astype(np.float32)
    lp = torch.as_tensor(self.lp_np, device="cuda")
    ev = torch.as_tensor(self.ev_np, device="cuda")
    self.arrays = infer.BPArrays(log_potentials=lp, ftov_msgs=base.ftov_msgs, evidence=ev)
    self.setup_s = time.perf_counter() - t0
    g = self.bp.context.engine.graph
    self.sizes = {"batch": self.batch, "es": g.num_es, "lp": g.num_lp, "vs": g.num_vs,
                  "edges": g.num_edges, "iters": self.iters}
    self.tuned = None

  def autotune(self):
    self.tuned = self.bp.context.engine.autotune(self.arrays.log_potentials,
                                                 self.arrays.ftov_msgs, self.arrays.evidence,
                                                 self.damping, self.temperature, probe_iters=5)
    return self.tuned

  def run(self, iters=None):
    return self.bp.run(self.arrays, iters or self.iters, self.damping, self.temperature)

  def updates_per_step(self):
    return self.sizes["es"] * self.batch * self.iters

  def sizes_for(self, kernel):
    return self.sizes

  def config(self):
    return {"workload": f"rbm_{self.n_visible}v_{self.n_hidden}h_batch{self.batch}_maxproduct_"
                        f"{self.iters}it_cache_resident",
            "num_edge_states": self.sizes["es"], "batch": self.batch, "iters": self.iters,
            "damping": self.damping, "temperature": self.temperature,
            "tuned_shape": self.tuned}

  def time_to_tol(self):
    return None

  def cpu_baseline(self):
    from oracle import cpu  # pylint: disable=g-import-not-at-top
    from oracle import lbp  # pylint: disable=g-import-not-at-top
    if not os.path.exists(cpu.LIB):
      cpu.build()
    g = lbp.graph_from_context(self.bp.context)
    es = g["var_states_for_edge_states"].shape[0]
    threads = cpu.set_threads(host_cores())
    ft0 = np.zeros(es, np.float32)
    t0 = time.perf_counter()
    cpu.run_enum(g, self.lp_np, ft0, self.ev_np, 4, self.damping, self.temperature)
    per_iter = (time.perf_counter() - t0) / 4
    iters = int(max(4, min(self.iters, CPU_BUDGET_S / max(per_iter, 1e-6))))
    t0 = time.perf_counter()
    cpu.run_enum(g, self.lp_np, ft0, self.ev_np, iters, self.damping, self.temperature)
    dt = time.perf_counter() - t0
    return {"value": es * self.batch * iters / dt, "unit": "edge-state msg updates/s",
            "cores": threads, "kind": "port",
            "sample": f"C/OpenMP restatement (oracle/lbp_cpu.c), the same {self.batch} RBMs, "
                      f"max-product, {iters} of {self.iters} iterations in {dt:.1f} s on "
                      f"{threads} threads"}

  def close(self):
    self.bp.context.engine.graph.close()


class Glass:
  """1,024 toroidal spin glasses 128 x 128, sum-product, 100 iterations."""

  name = "glass"
  iters = 100
  temperature = 1.0
  damping = 0.5

  def __init__(self):
    import torch  # pylint: disable=g-import-not-at-top
    from pgmax_amd import engine  # pylint: disable=g-import-not-at-top
    from pgmax_amd import workloads  # pylint: disable=g-import-not-at-top
    w, lp, ev = workloads.spin_glass_batch(128, 1024, seed=0)
    self.eng = engine.Engine.from_wiring(w.to_engine_wiring())
    self.lp = torch.as_tensor(lp, device="cuda")
    self.ev = torch.as_tensor(ev, device="cuda")
    self.ft = torch.zeros(w.num_edge_states, device="cuda")
    g = self.eng.graph
    self.sizes = {"batch": 1024, "es": g.num_es, "lp": g.num_lp, "vs": g.num_vs,
                  "edges": g.num_edges, "iters": self.iters}
    self.tuned = None

  def autotune(self):
    self.tuned = self.eng.autotune(self.lp, self.ft, self.ev, self.damping, self.temperature,
                                   probe_iters=5)
    return self.tuned

  def run(self, iters=None):
    return self.eng.run(self.lp, self.ft, self.ev, iters or self.iters, self.damping,
                        self.temperature)

  def updates_per_step(self):
    return self.sizes["es"] * self.sizes["batch"] * self.iters

  def sizes_for(self, kernel):
    return self.sizes

  def config(self):
    return {"workload": f"glass_1024x128x128_sumproduct_{self.iters}it",
            "iters": self.iters, "damping": self.damping, "temperature": self.temperature,
            "tuned_shape": self.tuned}

  def time_to_tol(self):
    return None

  def cpu_baseline(self):
    return None

  def close(self):
    self.eng.graph.close()


class Deconv:
  """Binary deconvolution, 64 images 40 x 40, OR/AND factors, max-product."""

  name = "deconv"
  iters = 100
  temperature = 0.0
  damping = 0.5

  def __init__(self):
    import torch  # pylint: disable=g-import-not-at-top
    from pgmax_amd import engine  # pylint: disable=g-import-not-at-top
    from pgmax_amd import workloads  # pylint: disable=g-import-not-at-top
    flat, meta = workloads.deconvolution_wiring(64, 1, 40, 40, 5, 6, 6)
    x = workloads.deconvolution_problem(64, 40, 40, n_feat_true=5, feat_true=6, seed=0)
    self.ev = torch.as_tensor(workloads.deconvolution_evidence(meta, x), device="cuda")
    self.eng = engine.Engine.from_wiring(flat.to_engine_wiring())
    self.lp = torch.zeros(1, device="cuda")[:0]
    self.ft = torch.zeros(self.eng.graph.num_es, device="cuda")
    g = self.eng.graph
    self.sizes = {"batch": 1, "es": g.num_es, "lp": g.num_lp, "vs": g.num_vs,
                  "edges": g.num_edges, "iters": self.iters}
    self.tuned = None

  def autotune(self):
    self.tuned = self.eng.autotune(self.lp, self.ft, self.ev, self.damping, self.temperature,
                                   probe_iters=3)
    return self.tuned

  def run(self, iters=None):
    return self.eng.run(self.lp, self.ft, self.ev, iters or self.iters, self.damping,
                        self.temperature)

  def updates_per_step(self):
    return self.sizes["es"] * self.iters

  def sizes_for(self, kernel):
    return self.sizes

  def config(self):
    return {"workload": f"deconv_64x40x40_5x6x6_maxproduct_{self.iters}it",
            "iters": self.iters, "damping": self.damping, "temperature": self.temperature,
            "tuned_shape": self.tuned}

  def time_to_tol(self):
    return None

  def cpu_baseline(self):
    return None

  def close(self):
    self.eng.graph.close()


class GlassGrad:
  """Gradient step of BP: 256 toroidal spin glasses 128 x 128, sum-product,
  damping 0.5, 100 iterations forward (the engine's run), then the
  vector-Jacobian product of the final messages with respect to the
  log-potentials, evidence and initial messages (pgx_run_vjp: forward
  replay on its tape and the reverse pass) for a fixed random cotangent --
  what jax.grad through BP.run computes in the reference's learning example
  (examples/grid_mrf.ipynb cells 15-16). The value counts the forward run's
  edge-state updates per second of the whole step."""

  name = "glass_grad"
  iters = 100
  temperature = 1.0
  damping = 0.5
  batch = 256

  def __init__(self):
    import torch  # pylint: disable=g-import-not-at-top
    from pgmax_amd import engine  # pylint: disable=g-import-not-at-top
    from pgmax_amd import workloads  # pylint: disable=g-import-not-at-top
    from pgmax_amd.infer import vjp  # pylint: disable=g-import-not-at-top
    w, lp, ev = workloads.spin_glass_batch(128, self.batch, seed=0)
    self.eng = engine.Engine.from_wiring(w.to_engine_wiring())
    self.rv = vjp.RunVJP.from_wiring(w, self.eng)
    self.lp = torch.as_tensor(lp, device="cuda")
    self.ev = torch.as_tensor(ev, device="cuda")
    self.ft = torch.zeros(w.num_edge_states, device="cuda")
    rng = np.random.default_rng(1)
    self.g_out = torch.as_tensor(
        rng.normal(size=(self.batch, w.num_edge_states)).astype(np.float32), device="cuda")
    g = self.eng.graph
    self.sizes = {"batch": self.batch, "es": g.num_es, "lp": g.num_lp, "vs": g.num_vs,
                  "edges": g.num_edges, "iters": self.iters}
    self.tuned = None
    self.split_ms = None

  def autotune(self):
    self.tuned = self.eng.autotune(self.lp, self.ft, self.ev, self.damping, self.temperature,
                                   probe_iters=5)
    return self.tuned

  def forward(self, iters=None):
    return self.eng.run(self.lp, self.ft, self.ev, iters or self.iters, self.damping,
                        self.temperature)

  def backward(self, iters=None):
    return self.rv(self.lp, self.ev, self.ft, iters or self.iters, self.damping,
                   self.temperature, self.g_out)

  def run(self, iters=None):
    out = self.forward(iters)
    return out, self.backward(iters)

  def measure_split(self, reps=3):
    """(forward ms, reverse-mode product ms), CUDA events on the current stream."""
    import torch  # pylint: disable=g-import-not-at-top
    res = []
    for fn in (self.forward, self.backward):
      fn()
      a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
      a.record()
      for _ in range(reps):
        fn()
      b.record()
      torch.cuda.synchronize()
      res.append(a.elapsed_time(b) / reps)
    self.split_ms = res
    return res

  def updates_per_step(self):
    return self.sizes["es"] * self.sizes["batch"] * self.iters

  def alg_bytes_iteration(self):
    """One forward iteration plus one iteration of the product (its replay
    and its reverse step)."""
    return alg_bytes_iter(self.sizes) + sum(
        kernel_bytes(k, self.sizes)
        for k in ("vjp_var_sums", "vjp_pair_fwd", "vjp_pair_bwd", "vjp_var_bwd"))

  def sizes_for(self, kernel):
    return self.sizes

  def config(self):
    cfg = {"workload": f"glass_grad_{self.batch}x128x128_sumproduct_{self.iters}it",
           "step": "forward run + reverse-mode product of the final messages",
           "iters": self.iters, "damping": self.damping, "temperature": self.temperature,
           "tuned_shape": self.tuned}
    if self.split_ms is None:
      self.measure_split()
    cfg["forward_ms"], cfg["vjp_ms"] = self.split_ms
    cfg["vjp_over_forward"] = self.split_ms[1] / self.split_ms[0]
    return cfg

  def time_to_tol(self):
    return None

  def cpu_baseline(self):
    return None

  def close(self):
    self.rv.close()
    self.eng.graph.close()


WORKLOADS = {"potts": Potts, "rbm": Rbm, "glass": Glass, "deconv": Deconv,
             "glass_grad": GlassGrad}


# ---------------------------------------------------------------------------
# HBM traffic from rocprofv3 counters
# ---------------------------------------------------------------------------


def _pmc_pass(counter, kernel_key, workload, shape, tmp):
  """KiB per dispatch of `counter` for the kernel and for the calibration multiply."""
  out_dir = os.path.join(tmp, counter)
  cmd = ["timeout", "-s", "KILL", "240", "rocprofv3", "--pmc", counter, "--output-format", "csv",
         "-d", out_dir, "-o", "pmc", "--", sys.executable, os.path.abspath(__file__), "--inner",
         "--workload", workload, "--shape", shape]
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
    raise RuntimeError(f"no {counter} rows for {kernel_key} or the calibration multiply")
  return float(np.mean(vals)), float(max(copy_vals)), len(vals)


def measure_traffic(kernel_key, workload, shape):
  """HBM bytes per launch of the dominant kernel from two rocprofv3 --pmc passes
  (the workload run with the timed run's tuned shape, no autotuning probes)."""
  if kernel_key is None or shutil.which("rocprofv3") is None:
    return None, "rocprofv3 or kernel mapping unavailable"
  tmp = tempfile.mkdtemp(prefix="pgx_pmc_")
  try:
    fetch, fetch_copy, n = _pmc_pass("FETCH_SIZE", kernel_key, workload, shape, tmp)
    write, write_copy, _ = _pmc_pass("WRITE_SIZE", kernel_key, workload, shape, tmp)
  except Exception as e:  # pylint: disable=broad-except
    return None, str(e)[:300]
  finally:
    shutil.rmtree(tmp, ignore_errors=True)
  kib = COPY_BYTES / 1024.0
  f_scale, w_scale = kib / fetch_copy, kib / write_copy
  traffic = (fetch * f_scale + write * w_scale) * 1024.0
  note = (f"FETCH_SIZE {fetch:.0f} KiB x {f_scale:.3f} + WRITE_SIZE {write:.0f} KiB x "
          f"{w_scale:.3f} per dispatch, mean over {n} dispatches whose name holds "
          f"'{kernel_key}' (scales from a {COPY_BYTES >> 20} MiB elementwise multiply, "
          f"which reads and writes {COPY_BYTES >> 20} MiB)")
  return traffic, note


SHAPE_KEYS = ("mode", "block", "fuse", "vec", "tile", "lanes", "xcd")


def shape_arg(tuned):
  return ",".join(str(int(tuned[k])) for k in SHAPE_KEYS)


def inner(workload, shape):
  """A short run of the workload in a given shape plus the calibration multiply
  (under rocprofv3 --pmc)."""
  import torch  # pylint: disable=g-import-not-at-top
  from pgmax_amd import _lib  # pylint: disable=g-import-not-at-top
  w = WORKLOADS[workload]()
  if shape:
    for k, v in zip(SHAPE_KEYS, shape.split(",")):
      _lib.set_option(k, int(v))
  w.run(max(2, min(w.iters, 20)))
  src = torch.empty(COPY_BYTES // 4, dtype=torch.float32, device="cuda").fill_(1.0)
  dst = torch.empty_like(src)
  torch.mul(src, 2.0, out=dst)  # one elementwise dispatch: reads and writes COPY_BYTES
  torch.cuda.synchronize()
  w.close()


# ---------------------------------------------------------------------------
# main
# ---------------------------------------------------------------------------


def main():
  ap = argparse.ArgumentParser()
  ap.add_argument("--gpus", type=int, default=1)
  ap.add_argument("--steps", type=int, default=10)
  ap.add_argument("--warmup", type=int, default=2)
  ap.add_argument("--no-cpu", action="store_true")
  ap.add_argument("--no-traffic", action="store_true")
  ap.add_argument("--no-tol", action="store_true")
  ap.add_argument("--no-warm", action="store_true")
  ap.add_argument("--inner", action="store_true", help=argparse.SUPPRESS)
  ap.add_argument("--shape", default="", help=argparse.SUPPRESS)
  ap.add_argument("--workload", default="potts", choices=sorted(WORKLOADS))
  ap.add_argument("--warm-predict", action="store_true",
                  help="time the warm-start prediction alone (1,024 128x128 tori)")
  args = ap.parse_args()
  if args.gpus != 1:
    print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported "
          "(independent graphs would shard across GPUs by whole graphs with no per-iteration "
          "communication; this build measures the one-GPU kernel).", file=sys.stderr)
    sys.exit(2)
  if args.inner:
    inner(args.workload, args.shape)
    return
  if args.warm_predict:
    print(json.dumps({"warm_predict": warm_predict_bench()}), flush=True)
    return

  import torch  # pylint: disable=g-import-not-at-top
  from pgmax_amd import _lib  # pylint: disable=g-import-not-at-top

  w = WORKLOADS[args.workload]()
  print(f"{w.name}: set up in {getattr(w, 'setup_s', 0.0):.1f} s", file=sys.stderr, flush=True)
  w.autotune()
  print(f"{w.name}: tuned shape {w.tuned}", file=sys.stderr, flush=True)
  for _ in range(args.warmup):
    w.run()
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  for _ in range(args.steps):
    w.run()
  torch.cuda.synchronize()
  dt = time.perf_counter() - t0
  print(f"{w.name}: {args.steps} runs in {dt:.3f} s", file=sys.stderr, flush=True)

  # per-kernel GPU time (HIP events around every launch on the run's stream)
  _lib.profile_kernels()
  _lib.set_option("profile", 2)
  prof_runs = max(1, min(args.steps, 3))
  for _ in range(prof_runs):
    w.run()
  torch.cuda.synchronize()
  _lib.set_option("profile", 0)
  _lib.profile_read()
  kernels = _lib.profile_kernels()
  dom = max(kernels, key=lambda k: kernels[k][0])
  dom_ms = kernels[dom][0] / kernels[dom][1]
  dom_bytes = kernel_bytes(dom, w.sizes_for(dom))
  achieved = dom_bytes / (dom_ms * 1e-3) / 1e9 if dom_bytes else None
  iter_ms = sum(ms for ms, _ in kernels.values()) / prof_runs / w.iters
  value = w.updates_per_step() * args.steps / dt
  cfg = w.config()
  it_bytes = (w.alg_bytes_iteration() if hasattr(w, "alg_bytes_iteration")
              else alg_bytes_iter(w.sizes))
  cfg.update({
      "iteration_gpu_us": iter_ms * 1e3,
      "iteration_wall_us": dt * 1e6 / args.steps / w.iters,
      "iteration_alg_bytes": it_bytes,
      "iteration_alg_GBps": it_bytes / (dt / args.steps / w.iters) / 1e9,
      "kernels_us_per_launch": {k: round(ms * 1e3 / n, 2) for k, (ms, n) in kernels.items()},
  })
  line = {
      "metric": "edge-state msg updates/s; time to \u0394<1e-4: warm-start vs plain LBP",
      "value": value,
      "unit": "edge-state msg updates/s",
      "n_gpus": 1,
      "steps": args.steps,
      "warmup": args.warmup,
      "ms_per_step": dt * 1e3 / args.steps,
      "higher_is_better": True,
      "scaling": "weak",
      "vs_baseline": None,
      "dtype": "f32",
      "data": "synthetic",
      "config": cfg,
      "roofline": {
          "bound": "hbm",
          "kernel": dom,
          "achieved": achieved,
          "peak": HBM_PEAK_GBS,
          "unit": "GB/s",
          "frac": achieved / HBM_PEAK_GBS if achieved else None,
          "traffic": None,
          "alg_bytes_per_launch": dom_bytes,
          "avg_launch_us": dom_ms * 1e3,
      },
  }
  if not args.no_tol:
    ttt = w.time_to_tol()
    if ttt is not None:
      line["time_to_tol"] = ttt
      print(f"{w.name}: time to tol {ttt}", file=sys.stderr, flush=True)
  if not args.no_warm and w.name == "potts":
    w_eval = warm_start_eval()
    line["warm_start"] = w_eval
    print(f"warm start: {w_eval}", file=sys.stderr, flush=True)
  if not args.no_traffic:
    traffic, note = measure_traffic(rocprof_key(dom, w.temperature, w.tuned.get("lanes", 0)),
                                    w.name, shape_arg(w.tuned))
    line["roofline"]["traffic"] = traffic
    line["roofline"]["traffic_note"] = note
    if traffic and dom_bytes:
      line["roofline"]["traffic_over_alg"] = traffic / dom_bytes
  if not args.no_cpu:
    cb = w.cpu_baseline()
    if cb is not None:
      line["cpu_baseline"] = cb
  w.close()
  print(json.dumps(line), flush=True)


if __name__ == "__main__":
  main()
