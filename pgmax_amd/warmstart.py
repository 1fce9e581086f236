"""Learned warm start for loopy BP on pairwise binary graphs.

A 3-layer message-passing network over the factor graph (hidden width 64,
BF16 activations, FP32 accumulation; kernels in csrc/pgx_mpnn.hip) predicts
the initial ``ftov_msgs`` of BP. Plain BP then runs from that start to the
same convergence test as from zeros, so it alone decides the final
beliefs; when the warm-started run misses its iteration budget, the graph is
rerun from zero messages (the fallback).

No counterpart in the reference: the prediction enters through the
``ftov_msgs`` of BPArrays (pgmax/infer/bp_state.py:172-255), exactly where a
user's own initial messages would.

Typical use::

    net = WarmStart(graph)                 # engine.Graph or batch.BatchedBP
    net.fit(samples, steps=2000)           # samples: (lp, ev, converged ftov)
    ftov0 = net.predict(lp, ev)
    out = solve(engine_or_batch, net, lp, ev, max_iters=500, tol=1e-4)
"""

import ctypes
import dataclasses
from typing import Optional, Sequence, Tuple

import numpy as np

from pgmax_amd import _lib
from pgmax_amd import engine

HIDDEN = 64
IN_FEATURES = 8
LAYERS = 3
NUM_PARAMS = _lib.PGX_MPNN_NUM_PARAMS


def init_params(seed: int = 0) -> np.ndarray:
  """He-uniform weights and zero biases in the flat layout of include/pgx.h."""
  rng = np.random.default_rng(seed)
  parts = []
  lim0 = np.sqrt(6.0 / IN_FEATURES)
  parts += [rng.uniform(-lim0, lim0, size=(IN_FEATURES, HIDDEN)).ravel(), np.zeros(HIDDEN)]
  lim = np.sqrt(6.0 / (3 * HIDDEN)) / np.sqrt(3.0)
  for _ in range(LAYERS):
    parts += [rng.uniform(-lim, lim, size=(3 * HIDDEN, HIDDEN)).ravel(), np.zeros(HIDDEN)]
  parts += [rng.uniform(-1, 1, size=HIDDEN) * np.sqrt(3.0 / HIDDEN), np.zeros(1)]
  p = np.concatenate(parts).astype(np.float32)
  assert p.size == NUM_PARAMS
  return p


def _graph_handle(target):
  g = getattr(target, "graph", target)
  if isinstance(g, engine.Graph):
    return g
  raise TypeError("WarmStart needs an engine.Engine, engine.Graph or batch.BatchedBP")


class WarmStart(_lib.NativeHandle):
  """The network's parameters and GPU state for one compiled graph."""

  _destroy_symbol = "pgx_mpnn_destroy"

  def __init__(self, target, params=None, seed: int = 0):
    torch = _lib.require_gpu()
    self.graph = _graph_handle(target)
    lib = _lib.load()
    h = ctypes.c_void_p()
    _lib.check(lib.pgx_mpnn_create(self.graph.handle, ctypes.byref(h)), "pgx_mpnn_create")
    self._adopt(h)
    p = init_params(seed) if params is None else np.asarray(
        params.detach().cpu().numpy() if hasattr(params, "detach") else params, np.float32)
    if p.size != NUM_PARAMS:
      raise ValueError(f"expected {NUM_PARAMS} parameters, got {p.size}")
    self.params = torch.as_tensor(p.reshape(-1), device="cuda").contiguous()
    self._m = torch.zeros_like(self.params)
    self._v = torch.zeros_like(self.params)
    self.step = 0

  # ---- inputs -------------------------------------------------------------
  def _inputs(self, log_potentials, evidence):
    g = self.graph
    lp, lp_lead = engine._as_batched(log_potentials, g.num_lp, "log_potentials")  # pylint: disable=protected-access
    ev, ev_lead = engine._as_batched(evidence, g.num_vs, "evidence")  # pylint: disable=protected-access
    leads = {l for l in (lp_lead, ev_lead) if l is not None}
    if len(leads) > 1:
      raise ValueError(f"inconsistent batch shapes {sorted(leads)}")
    lead = leads.pop() if leads else None
    batch = int(np.prod(lead)) if lead is not None else 1
    return lp, (g.num_lp if lp_lead else 0), ev, (g.num_vs if ev_lead else 0), batch, lead

  def predict(self, log_potentials, evidence):
    """Predicted normalised ftov messages, shaped like ``evidence``'s batch."""
    torch = _lib.require_gpu()
    lp, lps, ev, evs, batch, lead = self._inputs(log_potentials, evidence)
    out = torch.empty((batch, self.graph.num_es), device="cuda", dtype=torch.float32)
    _lib.check(_lib.load().pgx_mpnn_predict(
        self.handle, _lib.ptr(self.params), _lib.ptr(lp), lps, _lib.ptr(ev), evs,
        _lib.ptr(out), batch, _lib.current_stream()), "pgx_mpnn_predict")
    return out[0] if lead is None else out.reshape(lead + (self.graph.num_es,))

  def loss_and_grad(self, log_potentials, evidence, target_ftov):
    """(loss [1], gradient [NUM_PARAMS]) of the log-ratio mean squared error."""
    torch = _lib.require_gpu()
    lp, lps, ev, evs, batch, _ = self._inputs(log_potentials, evidence)
    tgt, t_lead = engine._as_batched(target_ftov, self.graph.num_es, "target_ftov")  # pylint: disable=protected-access
    if (t_lead is None and batch != 1) or (t_lead is not None and tgt.shape[0] != batch):
      raise ValueError("target_ftov must have one row per batch element")
    grads = torch.empty_like(self.params)
    loss = torch.empty((1,), device="cuda", dtype=torch.float32)
    _lib.check(_lib.load().pgx_mpnn_grad(
        self.handle, _lib.ptr(self.params), _lib.ptr(lp), lps, _lib.ptr(ev), evs,
        _lib.ptr(tgt), self.graph.num_es if t_lead is not None else 0,
        _lib.ptr(grads), _lib.ptr(loss), batch, _lib.current_stream()), "pgx_mpnn_grad")
    return loss, grads

  def train_step(self, log_potentials, evidence, target_ftov, lr: float = 1e-3):
    """One Adam step on a batch; returns the loss before the step (device)."""
    loss, grads = self.loss_and_grad(log_potentials, evidence, target_ftov)
    self.step += 1
    _lib.check(_lib.load().pgx_mpnn_adam(
        _lib.ptr(self.params), _lib.ptr(grads), _lib.ptr(self._m), _lib.ptr(self._v),
        NUM_PARAMS, self.step, float(lr), 0.9, 0.999, 1e-8, _lib.current_stream()),
               "pgx_mpnn_adam")
    return loss

  def fit(self, samples: Sequence[Tuple], steps: int, lr: float = 1e-3,
          lr_final: Optional[float] = None, seed: int = 0, log_every: int = 0):
    """Adam over ``samples`` (each (lp, ev, converged ftov) for this graph,
    batched or not), cycling in a seeded random order; cosine-decays the
    learning rate to ``lr_final``. Returns the per-step losses (host)."""
    torch = _lib.require_gpu()
    rng = np.random.default_rng(seed)
    lr_final = lr * 0.05 if lr_final is None else lr_final
    losses = []
    order = []
    for s in range(steps):
      if not order:
        order = list(rng.permutation(len(samples)))
      lp, ev, tgt = samples[order.pop()]
      cur = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * s / max(steps - 1, 1)))
      losses.append(self.train_step(lp, ev, tgt, cur))
      if log_every and (s + 1) % log_every == 0:
        print(f"warm-start fit step {s + 1}/{steps} loss {float(losses[-1]):.5f}", flush=True)
    return torch.cat(losses).cpu().numpy()


def fit_shared(nets: Sequence[WarmStart], samples: Sequence[Sequence[Tuple]], steps: int,
               lr: float = 1e-3, lr_final: Optional[float] = None, seed: int = 0):
  """Trains one set of parameters over several graphs: nets[k] is a
  WarmStart for graph k and samples[k] its (lp, ev, converged ftov) batches.
  Every net is made to share nets[0]'s parameters and Adam state; each step
  draws a graph with probability proportional to its sample count, then one
  of its batches. Returns the per-step losses (host)."""
  torch = _lib.require_gpu()
  head = nets[0]
  for n in nets[1:]:
    n.params, n._m, n._v = head.params, head._m, head._v  # pylint: disable=protected-access
  weights = np.array([len(s) for s in samples], np.float64)
  if weights.sum() == 0:
    return np.zeros(0, np.float32)
  rng = np.random.default_rng(seed)
  lr_final = lr * 0.05 if lr_final is None else lr_final
  losses = []
  for s in range(steps):
    k = int(rng.choice(len(nets), p=weights / weights.sum()))
    lp, ev, tgt = samples[k][int(rng.integers(len(samples[k])))]
    cur = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * s / max(steps - 1, 1)))
    nets[k].step = head.step  # one Adam step count for the shared state
    losses.append(nets[k].train_step(lp, ev, tgt, cur))
    head.step = nets[k].step
  return torch.cat(losses).cpu().numpy()


@dataclasses.dataclass
class SolveResult:
  ftov: object          # final messages (device)
  iterations: object    # BP iterations per (batch element, task), incl. fallback reruns
  fell_back: object     # bool per (batch element, task): warm start missed the budget
  predict_ms: float     # GPU time of the prediction


def solve(target, net: Optional[WarmStart], log_potentials, evidence, max_iters: int,
          tol: float = 1e-4, damping: float = 0.5, temperature: float = 1.0) -> SolveResult:
  """BP to max |delta| <= tol from the network's prediction (zeros if net is
  None). Every graph whose warm-started run misses the budget is rerun from
  zero messages and reports the iterations of both runs.

  ``target`` is an engine.Engine whose graph is one connected component
  (batched inputs allowed) or a batch.BatchedBP (one union, any number of
  graphs).
  """
  torch = _lib.require_gpu()
  g = _graph_handle(target)
  is_batch = hasattr(target, "layout")
  if not is_batch and g.num_tasks != 1:
    raise ValueError("solve on an Engine needs a connected graph; use batch.BatchedBP")
  _, ev_lead = engine._as_batched(evidence, g.num_vs, "evidence")  # pylint: disable=protected-access
  shape = (ev_lead or ()) + (g.num_es,)
  pred_ms = 0.0
  if net is None:
    ftov0 = torch.zeros(shape, device="cuda", dtype=torch.float32)
  else:
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    ftov0 = net.predict(log_potentials, evidence)
    e1.record()
    e1.synchronize()
    pred_ms = e0.elapsed_time(e1)
  out, deltas, iters = _converge(target, log_potentials, ftov0, evidence, max_iters, tol,
                                 damping, temperature)
  last = deltas.gather(-1, (iters.long() - 1).clamp(min=0).unsqueeze(-1)).squeeze(-1)
  fell = torch.zeros(iters.shape, dtype=torch.bool, device="cuda")
  if net is not None:
    fell = ~(last <= tol)
    if bool(fell.any()):
      zero = torch.zeros(shape, device="cuda", dtype=torch.float32)
      out0, _, iters0 = _converge(target, log_potentials, zero, evidence, max_iters, tol,
                                  damping, temperature)
      if is_batch:
        mask = fell[torch.as_tensor(_es_task(target), device="cuda")]
      else:
        mask = fell.unsqueeze(-1)  # one task per batch element
      out = torch.where(mask, out0, out)
      iters = torch.where(fell, iters + iters0, iters)
  return SolveResult(out, iters, fell, pred_ms)


def _es_task(bb) -> np.ndarray:
  """Task (in BatchedBP.run_converge's graph order) of every union edge state."""
  cached = getattr(bb, "_warm_es_task", None)
  if cached is not None:
    return cached
  es_task = np.empty(bb.layout.num_es, np.int64)
  for gi, idx in enumerate(bb.layout.es_index):
    es_task[idx] = gi
  bb._warm_es_task = es_task  # pylint: disable=protected-access
  return es_task


def _converge(target, lp, ftov, ev, max_iters, tol, damping, temperature):
  if hasattr(target, "layout"):  # batch.BatchedBP: per-graph results in graph order
    return target.run_converge(lp, ftov, ev, max_iters, tol, damping, temperature)
  out, deltas, iters = target.run_converge(lp, ftov, ev, max_iters, tol, damping, temperature)
  return out, deltas[..., 0, :], iters[..., 0:1].squeeze(-1)
