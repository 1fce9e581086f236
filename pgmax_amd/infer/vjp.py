"""Reverse-mode BP: gradients through ``run`` (row f3).

The reference differentiates BP with JAX (``jax.grad`` through
``bp.run``, examples/grid_mrf.ipynb cells 15-16). Here the vector-Jacobian
product of the final messages is a native call (``pgx_run_vjp``), and
``run_differentiable`` wraps run + product as a torch autograd function, so a
loss on beliefs or marginals built from its output with torch operations
differentiates back to the log-potentials, the evidence and the initial
messages. Graphs of Enum, OR, AND and Pool factors.
"""

import ctypes

import numpy as np

from pgmax_amd import _lib
from pgmax_amd import batch as batch_mod
from pgmax_amd import engine


class RunVJP(_lib.NativeHandle):
  """The reverse-mode tables of one compiled graph (pgx_vjp)."""

  _destroy_symbol = "pgx_vjp_destroy"

  def __init__(self, context):
    """``context``: an InfererContext (``bp.context``) of a graph of Enum, OR,
    AND and Pool factors."""
    self._setup(batch_mod.FlatWiring.from_context(context), context.engine)

  @classmethod
  def from_wiring(cls, flat, eng):
    """From flat wiring arrays (batch.FlatWiring) and the Engine compiled from
    them, for graphs built without a FactorGraph."""
    obj = cls.__new__(cls)
    obj._setup(flat, eng)
    return obj

  def _setup(self, flat, eng):
    torch = _lib.require_gpu()
    del torch
    row_cfg, row_es = flat.enum_rows()
    i32 = lambda a: np.ascontiguousarray(np.asarray(a, np.int32))
    self._arrays = [i32(flat.var_states_for_edge_states), i32(flat.edge_indices_for_edge_states),
                    i32(row_cfg), i32(row_es)]
    self.engine = eng
    g = self.engine.graph
    h = ctypes.c_void_p()
    ptr = lambda a: a.ctypes.data if a.size else None
    _lib.check(_lib.load().pgx_vjp_create(
        g.handle, ptr(self._arrays[0]), ptr(self._arrays[1]), ptr(self._arrays[2]),
        ptr(self._arrays[3]), int(self._arrays[2].size), ctypes.byref(h)), "pgx_vjp_create")
    self._adopt(h)
    for code in (batch_mod.OR, batch_mod.AND):  # OR / AND wiring (logical.py:457-488)
      lw = flat.logical.get(code)
      if lw is None or np.asarray(lw[2]).size == 0:
        continue
      pfi, pmi, ces = (i32(a) for a in lw[:3])
      self._arrays += [pfi, pmi, ces]
      _lib.check(_lib.load().pgx_vjp_set_logical(
          self.handle, code, ptr(pfi), ptr(pmi), int(pfi.size), ptr(ces), int(ces.size),
          int(lw[3])), "pgx_vjp_set_logical")
    pw = flat.logical.get(batch_mod.POOL)  # Pool wiring (pool.py:275-326)
    if pw is not None and np.asarray(pw[2]).size:
      cfi, cmi, ies = (i32(a) for a in pw[:3])
      self._arrays += [cfi, cmi, ies]
      _lib.check(_lib.load().pgx_vjp_set_pool(
          self.handle, ptr(cfi), ptr(cmi), int(cfi.size), ptr(ies), int(ies.size)),
          "pgx_vjp_set_pool")

  def __call__(self, log_potentials, evidence, ftov_msgs, num_iters: int, damping: float,
               temperature: float, g_out):
    """(g_lp, g_ev, g_ftov) for the cotangent g_out of the final messages.
    Inputs may carry a leading batch axis; a gradient has its input's shape
    (summed over the batch for an input shared by the batch)."""
    torch = _lib.require_gpu()
    g = self.engine.graph
    lp, lp_lead = engine._as_batched(log_potentials, g.num_lp, "log_potentials")  # pylint: disable=protected-access
    ev, ev_lead = engine._as_batched(evidence, g.num_vs, "evidence")  # pylint: disable=protected-access
    ft, ft_lead = engine._as_batched(ftov_msgs, g.num_es, "ftov_msgs")  # pylint: disable=protected-access
    go, go_lead = engine._as_batched(g_out, g.num_es, "g_out")  # pylint: disable=protected-access
    leads = {l for l in (lp_lead, ev_lead, ft_lead, go_lead) if l is not None}
    if len(leads) > 1:
      raise ValueError(f"inconsistent batch shapes {sorted(leads)}")
    lead = leads.pop() if leads else None
    nb = int(np.prod(lead)) if lead is not None else 1
    if go_lead is None and nb > 1:
      go = go.expand(nb, g.num_es).contiguous()
    shape = lambda lead_, n: ((nb, n) if lead_ is not None else (n,))
    g_lp = torch.empty(shape(lp_lead, g.num_lp), device="cuda", dtype=torch.float32)
    g_ev = torch.empty(shape(ev_lead, g.num_vs), device="cuda", dtype=torch.float32)
    g_ft = torch.empty(shape(ft_lead, g.num_es), device="cuda", dtype=torch.float32)
    _lib.check(_lib.load().pgx_run_vjp(
        self.handle, _lib.ptr(lp), g.num_lp if lp_lead is not None else 0,
        _lib.ptr(ev), g.num_vs if ev_lead is not None else 0,
        _lib.ptr(ft), g.num_es if ft_lead is not None else 0, _lib.ptr(go),
        _lib.ptr(g_lp), _lib.ptr(g_ev), _lib.ptr(g_ft), nb, int(num_iters), float(damping),
        float(temperature), _lib.current_stream()), "pgx_run_vjp")

    def shaped(t, lead_, n):
      return t.reshape(lead_ + (n,)) if lead_ is not None else t
    return (shaped(g_lp, lp_lead, g.num_lp), shaped(g_ev, ev_lead, g.num_vs),
            shaped(g_ft, ft_lead, g.num_es))


def run_differentiable(bp, log_potentials, evidence, ftov_msgs, num_iters: int,
                       damping: float = 0.5, temperature=None):
  """Final messages of ``bp.run`` as a torch tensor that carries gradients
  to ``log_potentials``, ``evidence`` and ``ftov_msgs`` (device tensors;
  any of them may require grad)."""
  torch = _lib.require_gpu()
  ctx = bp.context
  t = ctx.temperature if temperature is None else float(temperature)
  vjp = getattr(ctx, "_pgx_vjp", None)
  if vjp is None:
    vjp = RunVJP(ctx)
    object.__setattr__(ctx, "_pgx_vjp", vjp)

  class _Run(torch.autograd.Function):  # pylint: disable=abstract-method

    @staticmethod
    def forward(fctx, lp, ev, ft):  # pylint: disable=arguments-differ
      fctx.save_for_backward(lp, ev, ft)
      out, _ = ctx.engine.run(lp.detach(), ft.detach(), ev.detach(), num_iters, damping, t)
      return out

    @staticmethod
    def backward(fctx, g_out):  # pylint: disable=arguments-differ
      lp, ev, ft = fctx.saved_tensors
      g_lp, g_ev, g_ft = vjp(lp.detach(), ev.detach(), ft.detach(), num_iters, damping, t,
                             g_out.contiguous())
      return g_lp, g_ev, g_ft

  as_t = lambda x: x if hasattr(x, "detach") else torch.as_tensor(
      np.asarray(x, np.float32), device="cuda")
  return _Run.apply(as_t(log_potentials), as_t(evidence), as_t(ftov_msgs))


def flat_beliefs(context, evidence, ftov_msgs):
  """evidence + the sum of incoming messages per variable state
  (pgmax/infer/inferer.py:218-222) as differentiable torch operations."""
  torch = _lib.require_gpu()
  vsi = getattr(context, "_pgx_vsi", None)
  if vsi is None:
    flat = batch_mod.FlatWiring.from_context(context)
    vsi = torch.as_tensor(np.asarray(flat.var_states_for_edge_states, np.int64), device="cuda")
    object.__setattr__(context, "_pgx_vsi", vsi)
  lead = tuple(ftov_msgs.shape[:-1])
  out = evidence.expand(lead + evidence.shape[-1:]).clone() if lead else evidence.clone()
  return out.index_add(out.dim() - 1, vsi, ftov_msgs)
