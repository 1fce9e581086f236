"""Full-size parity against the C restatement of the oracle, bit for bit.

oracle/lbp_cpu.c computes what the NumPy FP32 oracle does, bit for bit
(tests/test_oracle.py checks it), multi-threaded over graphs and batch
elements, so it checks the benchmark workloads at their real sizes:

* the RBM batch: 64 draws of a 500 x 100 binary RBM (50,000 pairwise
  factors each), 200 iterations, max-product and sum-product, in both the
  bias-as-evidence form and the reference's own form with unary Enum
  factors for the biases (benchmark/rbm_lib.py:140-168);
* the 2,048 uneven Potts grids: 500 iterations in the tuned shape, and the
  per-graph convergence run to max |change| <= 1e-4, on a sample of graphs
  (graphs are disjoint: each graph's messages depend on its own inputs);
* the 1,024 spin-glass tori of 128 x 128 of the warm-start evaluation,
  1,000 iterations, on a sample of batch elements;
* the binary deconvolution: 64 images of 40 x 40, 5 features of 6 x 6
  (14.1M AND and 102,400 OR factors, 1.13e8 edge states), max-product, in
  the tuned shape, against the C restatement of max-product OR / AND BP.
"""

import os

import numpy as np
import pytest

import graphs
from oracle import cpu
from oracle import lbp
from pgmax_amd import batch
from pgmax_amd import infer
from pgmax_amd import workloads

pytestmark = pytest.mark.gpu

THREADS = 16  # the host share of one GPU


def _np(x):
  return x.detach().cpu().numpy() if hasattr(x, "detach") else np.asarray(x)


@pytest.fixture(scope="module", autouse=True)
def _c_oracle():
  if not os.path.exists(cpu.LIB):
    cpu.build()


@pytest.mark.parametrize("unary_factors", [False, True])
@pytest.mark.parametrize("temperature", [0.0, 1.0])
def test_rbm_64_full_size(gpu, temperature, unary_factors):
  fg, _, _, lp, ev = graphs.rbm(500, 100, batch=64, seed=0, unary_factors=unary_factors)
  bp = infer.BP(fg.bp_state, temperature=temperature)
  lp32, ev32 = lp.astype(np.float32), ev.astype(np.float32)
  arrays = infer.BPArrays(log_potentials=lp32, ftov_msgs=bp.init().ftov_msgs, evidence=ev32)
  print(f"RBM 64 x 500 x 100, unary factors {unary_factors}, T={temperature}", flush=True)
  out, deltas = bp.run_with_diffs(arrays, num_iters=200, damping=0.5)
  msgs, deltas = _np(out.ftov_msgs), _np(deltas)
  g = lbp.graph_from_context(bp.context)
  ft0 = np.zeros(msgs.shape[1], np.float32)
  ref, ref_d = cpu.run_enum(g, lp32, ft0, ev32, 200, 0.5, temperature, threads=THREADS)
  np.testing.assert_array_equal(msgs, ref)
  np.testing.assert_array_equal(deltas, ref_d)


def _sample(pb):
  """The first and last graph, and among the graphs whose oracle run costs at
  most a quarter of the largest one's (the C oracle runs one graph on one
  thread): the largest, the largest with 2 and with 8 states, and 10 more
  drawn at random."""
  cost = pb.sides.astype(np.int64) ** 2 * pb.states.astype(np.int64) ** 2
  ok = np.flatnonzero(cost <= cost.max() // 4)
  sample = {0, pb.num_graphs - 1, int(ok[np.argmax(cost[ok])])}
  for s in (2, 8):
    sel = ok[pb.states[ok] == s]
    if sel.size:
      sample.add(int(sel[np.argmax(cost[sel])]))
  sample.update(np.random.default_rng(5).choice(ok, 10, replace=False).tolist())
  return sorted(sample)


def _graph_inputs(pb, lay, g):
  w = pb.wirings[g]
  lp_g = pb.log_potentials[lay.lp_index[g][0]:lay.lp_index[g][-1] + 1]
  ev_g = pb.evidence[lay.vs_offsets[g]:lay.vs_offsets[g + 1]]
  return lbp.graph_from_flat(w), lp_g, ev_g, np.zeros(w.num_edge_states, np.float32)


def test_potts_2048_full_size_500_iterations(gpu):
  import torch  # pylint: disable=g-import-not-at-top
  pb = workloads.potts_batch(2048, seed=0)
  bb = batch.BatchedBP(pb.wirings, temperature=1.0)
  lp = torch.as_tensor(pb.log_potentials, device="cuda")
  ev = torch.as_tensor(pb.evidence, device="cuda")
  print("Potts 2048, tuned shape, 500 iterations", flush=True)
  bb.autotune(lp, None, ev, 0.5, 1.0, probe_iters=2)
  out, _ = bb.run(lp, None, ev, 500, 0.5, 1.0)
  out = _np(out)
  lay = bb.layout
  sample = _sample(pb)
  inputs = [_graph_inputs(pb, lay, g) for g in sample]
  refs, ref_d = cpu.run_enum_graphs([x[0] for x in inputs], [x[1] for x in inputs],
                                [x[2] for x in inputs], [x[3] for x in inputs], 500, 0.5, 1.0,
                                threads=THREADS)
  for g, ref in zip(sample, refs):
    np.testing.assert_array_equal(out[lay.es_index[g]], ref,
                                  err_msg=f"graph {g}: side {pb.sides[g]}, {pb.states[g]} states")

  print("Potts 2048, per-graph convergence to 1e-4", flush=True)
  out_c, task_deltas, iters = bb.run_converge(lp, None, ev, 500, tol=1e-4, damping=0.5,
                                              temperature=1.0)
  out_c, task_deltas, iters = _np(out_c), _np(task_deltas), _np(iters)
  stopped = []
  for k, g in enumerate(sample):
    below = np.flatnonzero(ref_d[k] <= np.float32(1e-4))
    n = int(below[0]) + 1 if below.size else 500
    assert int(iters[g]) == n, f"graph {g}: {int(iters[g])} iterations, oracle {n}"
    np.testing.assert_array_equal(task_deltas[g, :n], ref_d[k, :n], err_msg=f"graph {g}")
    assert not np.any(task_deltas[g, n:])
    stopped.append(n)
  # the messages a stopped graph holds are those of a run of its own length
  for k in np.argsort(stopped)[:3]:
    g = sample[int(k)]
    x = inputs[int(k)]
    ref, _ = cpu.run_enum(x[0], x[1], x[3], x[2], stopped[int(k)], 0.5, 1.0, threads=1)
    np.testing.assert_array_equal(out_c[lay.es_index[g]], ref, err_msg=f"graph {g}")
  bb.close()


def test_deconvolution_64_full_size(gpu):
  """The deconvolution workload at its stated size (the bench's inputs:
  priors of log-odds -2, the images' pixels as evidence, zero initial
  messages; its many exact ties included), 80 max-product iterations with
  damping 0.5 in the tuned shape: messages and per-iteration deltas bit for
  bit."""
  import torch  # pylint: disable=g-import-not-at-top
  from pgmax_amd import engine  # pylint: disable=g-import-not-at-top
  n_it = 80
  flat, meta = workloads.deconvolution_wiring(64, 1, 40, 40, 5, 6, 6)
  x = workloads.deconvolution_problem(64, 40, 40, n_feat_true=5, feat_true=6, seed=0)
  ev = workloads.deconvolution_evidence(meta, x).astype(np.float32)
  print(f"deconvolution 64 x 40 x 40: {flat.num_edge_states} edge states, {n_it} iterations",
        flush=True)
  eng = engine.Engine.from_wiring(flat.to_engine_wiring())
  lp = torch.zeros(1, device="cuda")[:0]
  ft = torch.zeros(flat.num_edge_states, device="cuda")
  ev_t = torch.as_tensor(ev, device="cuda")
  eng.autotune(lp, ft, ev_t, 0.5, 0.0, probe_iters=2)
  out, d = eng.run(lp, ft, ev_t, n_it, 0.5, 0.0)
  got, got_d = _np(out), _np(d)
  eng.graph.close()
  ref, ref_d = cpu.run_logical_t0(lbp.graph_from_flat(flat), np.zeros(flat.num_edge_states),
                                  ev, n_it, 0.5, threads=THREADS)
  np.testing.assert_array_equal(got_d, ref_d)
  np.testing.assert_array_equal(got, ref)


def test_spin_glass_1024_full_size(gpu):
  """The warm-start evaluation's graphs at their stated size: 1,024 toroidal
  spin glasses of 128 x 128 (J ~ N(0, 1), unaries ~ N(0, 0.5), seed 0),
  sum-product, damping 0.5, 1,000 iterations (the evaluation's budget) in
  the tuned shape; 12 sampled batch elements (independent of each other)
  against the C restatement, messages and deltas bit for bit."""
  import torch  # pylint: disable=g-import-not-at-top
  from pgmax_amd import engine  # pylint: disable=g-import-not-at-top
  n_it = 1000
  w, lp, ev = workloads.spin_glass_batch(128, 1024, seed=0)
  print(f"spin glass 1024 x 128 x 128, {n_it} iterations", flush=True)
  eng = engine.Engine.from_wiring(w.to_engine_wiring())
  lp_t, ev_t = torch.as_tensor(lp, device="cuda"), torch.as_tensor(ev, device="cuda")
  ft = torch.zeros(w.num_edge_states, device="cuda")
  eng.autotune(lp_t, ft, ev_t, 0.5, 1.0, probe_iters=2)
  out, d = eng.run(lp_t, ft, ev_t, n_it, 0.5, 1.0)
  sample = np.sort(np.random.default_rng(8).choice(1024, 12, replace=False))
  sample[0], sample[-1] = 0, 1023
  got, got_d = _np(out)[sample], _np(d)[sample]
  eng.graph.close()
  ref, ref_d = cpu.run_enum(lbp.graph_from_flat(w), lp[sample], np.zeros(w.num_edge_states),
                            ev[sample], n_it, 0.5, 1.0, threads=THREADS)
  np.testing.assert_array_equal(got_d, ref_d)
  np.testing.assert_array_equal(got, ref)
