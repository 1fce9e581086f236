"""Python driver of the HIP engine (libpgx.so).

``Engine`` owns one compiled graph (``pgx_graph``) per ``InfererContext``
and runs BP, beliefs and MAP decoding on device-resident torch tensors.
Arrays may carry leading batch axes (the reference's ``jax.vmap`` over
BPArrays); arrays without them are shared across the batch.
"""

import collections
import ctypes
import hashlib
from typing import Any, Optional, Tuple

import numpy as np

from pgmax_amd import _arrays
from pgmax_amd import _lib
from pgmax_amd.factor import FAC_TO_VAR_UPDATES


def _i32(x) -> np.ndarray:
  return np.ascontiguousarray(np.asarray(x, dtype=np.int64).astype(np.int32))


def _ptr32(a: np.ndarray):
  return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)) if a.size else None


class Wiring:
  """A ``pgx_wiring`` struct plus the int32 arrays it points into."""

  def __init__(self, *, num_var_states, num_log_potentials, var_states_for_edge_states,
               edge_indices_for_edge_states, factor_indices_for_edge_states,
               evidence_to_vars, msgs_range, potentials_range, enum_args=None,
               logical_args=None):
    self.keep = []

    def keep(a):
      a = _i32(a)
      self.keep.append(a)
      return a

    w = _lib.Wiring()
    vs = keep(var_states_for_edge_states)
    w.num_edge_states = vs.shape[0]
    w.num_var_states = int(num_var_states)
    w.num_log_potentials = int(num_log_potentials)
    w.var_states_for_edge_states = _ptr32(vs)
    w.edge_indices_for_edge_states = _ptr32(keep(edge_indices_for_edge_states))
    w.factor_indices_for_edge_states = _ptr32(keep(factor_indices_for_edge_states))
    w.evidence_to_vars = _ptr32(keep(evidence_to_vars))
    for t in range(4):
      w.msgs_range[t][0], w.msgs_range[t][1] = (int(v) for v in msgs_range[t])
      w.potentials_range[t][0], w.potentials_range[t][1] = (int(v) for v in potentials_range[t])
    if isinstance(enum_args, str):
      if enum_args != "full":
        raise ValueError(f"enum_args must be (indices, edge_states) or 'full', got {enum_args!r}")
      w.enum_num_rows = -1  # PGX_ENUM_FULL_TABLES
    elif enum_args is not None:
      ci = keep(enum_args[0])
      w.enum_num_rows = ci.shape[0]
      w.enum_factor_configs_indices = _ptr32(ci)
      w.enum_factor_configs_edge_states = _ptr32(keep(enum_args[1]))
    for t, args in (logical_args or {}).items():
      pfi, pmi, ces, off = args
      lw = w.logical[t]
      pfi = keep(pfi)
      lw.num_parents = pfi.shape[0]
      ces = keep(ces)
      lw.num_factors = ces.shape[0]
      lw.parents_factor_indices = _ptr32(pfi)
      lw.parents_msg_indices = _ptr32(keep(pmi))
      lw.children_edge_states = _ptr32(ces)
      lw.edge_states_offset = int(off)
    self.struct = w


def wiring_from_context(ctx) -> Wiring:
  """The compiled wiring of an InfererContext (pgmax/infer/inferer.py:65-118)."""
  types = list(FAC_TO_VAR_UPDATES.keys())
  args = ctx.inference_arguments
  enum_t, or_t, and_t, pool_t = types
  logical = {}
  for code, t in ((_lib.PGX_OR, or_t), (_lib.PGX_AND, and_t)):
    a = args[t]
    logical[code] = (a["parents_factor_indices"], a["parents_msg_indices"],
                     a["children_edge_states"], a["edge_states_offset"])
  a = args[pool_t]
  logical[_lib.PGX_POOL] = (a["pool_choices_factor_indices"], a["pool_choices_msg_indices"],
                            a["pool_indicators_edge_states"], 1)
  fg_state = ctx.bp_state.fg_state
  return Wiring(
      num_var_states=fg_state.num_var_states,
      num_log_potentials=fg_state.log_potentials.shape[0],
      var_states_for_edge_states=ctx.var_states_for_edge_states,
      edge_indices_for_edge_states=ctx.edge_indices_for_edge_states,
      factor_indices_for_edge_states=ctx.factor_indices_for_edge_states,
      evidence_to_vars=ctx.evidence_to_vars,
      msgs_range=[ctx.factor_type_to_msgs_range[t] for t in types],
      potentials_range=[ctx.factor_type_to_potentials_range[t] for t in types],
      enum_args=(args[enum_t]["factor_configs_indices"],
                 args[enum_t]["factor_configs_edge_states"]),
      logical_args=logical,
  )


class Graph(_lib.NativeHandle):
  """An owned ``pgx_graph`` handle (``close()`` frees its device tables)."""

  _destroy_symbol = "pgx_graph_destroy"

  def __init__(self, wiring: Wiring):
    lib = _lib.load()
    _lib.require_gpu()
    handle = ctypes.c_void_p()
    _lib.check(lib.pgx_graph_create(ctypes.byref(wiring.struct), ctypes.byref(handle)),
               "pgx_graph_create")
    self._adopt(handle)
    info = (ctypes.c_int64 * 8)()
    _lib.check(lib.pgx_graph_info(handle, info), "pgx_graph_info")
    (self.num_es, self.num_vs, self.num_lp, self.num_edges, self.num_vars,
     self.num_factors, self.num_tasks, self.num_blocks) = (int(v) for v in info)


def tune_graph(g: "Graph", lp, lp_stride, ev, ev_stride, ft, ft_stride, batch, damping,
               temperature, converge, probe_iters):
  """pgx_autotune on device tensors; returns the chosen shape."""
  chosen = (ctypes.c_int64 * 7)()
  _lib.check(_lib.load().pgx_autotune(
      g.handle, _lib.ptr(lp), lp_stride, _lib.ptr(ev), ev_stride, _lib.ptr(ft), ft_stride,
      int(batch), float(damping), float(temperature), int(bool(converge)), int(probe_iters),
      _lib.current_stream(), chosen), "pgx_autotune")
  return {"mode": int(chosen[0]), "block": int(chosen[1]), "fuse": int(chosen[2]),
          "vec": int(chosen[3]), "tile": int(chosen[4]), "lanes": int(chosen[5]),
          "xcd": int(chosen[6])}


def _host_f32(x) -> np.ndarray:
  """A writable float32 NumPy copy or view of x (torch refuses to wrap the
  read-only arrays of frozen BPArrays / FactorGraphState without a warning)."""
  a = np.asarray(_arrays.to_numpy(x), dtype=np.float32)
  return a if a.flags.writeable else a.copy()


def _as_batched(x, n: int, name: str):
  """(device f32 contiguous tensor of shape (B, n) or (n,), lead shape or None)."""
  torch = _lib.require_gpu()
  t = x if _arrays.is_torch(x) else torch.as_tensor(_host_f32(x))
  t = t.to(device="cuda", dtype=torch.float32)
  if t.shape[-1] != n:
    raise ValueError(f"{name} has {t.shape[-1]} entries per element, expected {n}")
  if t.ndim == 1:
    return t.contiguous(), None
  lead = tuple(t.shape[:-1])
  return t.reshape(-1, n).contiguous(), lead


class Engine:
  """BP runs, beliefs and decoding for one compiled factor graph."""

  def __init__(self, ctx):
    self.ctx = ctx
    self._graph: Optional[Graph] = None

  @classmethod
  def from_wiring(cls, wiring: "Wiring") -> "Engine":
    """An engine over a wiring built without an InfererContext (e.g. batch.grid_wiring)."""
    eng = cls(None)
    eng._graph = Graph(wiring)
    return eng

  @property
  def graph(self) -> Graph:
    if self._graph is None:
      self._graph = Graph(wiring_from_context(self.ctx))
    return self._graph

  def run(self, log_potentials, ftov_msgs, evidence, num_iters: int, damping: float,
          temperature: float):
    """pgmax/infer/bp.py:63-155 on the GPU; returns (ftov_msgs, msgs_deltas)."""
    torch = _lib.require_gpu()
    g = self.graph
    lp, lp_lead = _as_batched(log_potentials, g.num_lp, "log_potentials")
    ft, ft_lead = _as_batched(ftov_msgs, g.num_es, "ftov_msgs")
    ev, ev_lead = _as_batched(evidence, g.num_vs, "evidence")
    leads = {l for l in (lp_lead, ft_lead, ev_lead) if l is not None}
    if len(leads) > 1:
      raise ValueError(f"inconsistent batch shapes {sorted(leads)}")
    lead = leads.pop() if leads else None
    batch = int(np.prod(lead)) if lead is not None else 1
    n_it = max(int(num_iters), 1)
    out = torch.empty((batch, g.num_es), device="cuda", dtype=torch.float32)
    deltas = torch.empty((batch, n_it), device="cuda", dtype=torch.float32)
    stride = lambda t, lead_, n: n if lead_ is not None else 0
    _lib.check(_lib.load().pgx_run(
        g.handle,
        _lib.ptr(lp), stride(lp, lp_lead, g.num_lp),
        _lib.ptr(ev), stride(ev, ev_lead, g.num_vs),
        _lib.ptr(ft), stride(ft, ft_lead, g.num_es),
        _lib.ptr(out), _lib.ptr(deltas), batch, n_it, float(damping),
        float(temperature), _lib.current_stream()), "pgx_run")
    if lead is None:
      return out[0], deltas[0]
    return out.reshape(lead + (g.num_es,)), deltas.reshape(lead + (n_it,))

  def autotune(self, log_potentials, ftov_msgs, evidence, damping: float = 0.5,
               temperature: float = 0.0, converge: bool = False, probe_iters: int = 10):
    """Times the graph's execution shapes on these inputs and keeps the fastest
    for later runs of the same batch size and temperature class. Returns
    {mode, block, fuse, vec, tile, lanes, xcd}."""
    g = self.graph
    lp, lp_lead = _as_batched(log_potentials, g.num_lp, "log_potentials")
    ft, ft_lead = _as_batched(ftov_msgs, g.num_es, "ftov_msgs")
    ev, ev_lead = _as_batched(evidence, g.num_vs, "evidence")
    leads = {l for l in (lp_lead, ft_lead, ev_lead) if l is not None}
    if len(leads) > 1:
      raise ValueError(f"inconsistent batch shapes {sorted(leads)}")
    lead = leads.pop() if leads else None
    batch = int(np.prod(lead)) if lead is not None else 1
    return tune_graph(g, lp, g.num_lp if lp_lead else 0, ev, g.num_vs if ev_lead else 0,
                      ft, g.num_es if ft_lead else 0, batch, damping, temperature, converge,
                      probe_iters)

  def run_converge(self, log_potentials, ftov_msgs, evidence, max_iters: int, tol: float,
                   damping: float, temperature: float):
    """BP until every (batch element, component) reaches max |delta| <= tol.

    Returns (ftov_msgs, deltas [batch, tasks, max_iters], iterations
    [batch, tasks]) with the batch dimensions of the inputs.
    """
    torch = _lib.require_gpu()
    g = self.graph
    lp, lp_lead = _as_batched(log_potentials, g.num_lp, "log_potentials")
    ft, ft_lead = _as_batched(ftov_msgs, g.num_es, "ftov_msgs")
    ev, ev_lead = _as_batched(evidence, g.num_vs, "evidence")
    leads = {l for l in (lp_lead, ft_lead, ev_lead) if l is not None}
    if len(leads) > 1:
      raise ValueError(f"inconsistent batch shapes {sorted(leads)}")
    lead = leads.pop() if leads else None
    batch = int(np.prod(lead)) if lead is not None else 1
    n_it = max(int(max_iters), 1)
    out = torch.empty((batch, g.num_es), device="cuda", dtype=torch.float32)
    deltas = torch.empty((batch, g.num_tasks, n_it), device="cuda", dtype=torch.float32)
    iters = torch.empty((batch, g.num_tasks), device="cuda", dtype=torch.int32)
    stride = lambda lead_, n: n if lead_ is not None else 0
    _lib.check(_lib.load().pgx_run_converge(
        g.handle,
        _lib.ptr(lp), stride(lp_lead, g.num_lp),
        _lib.ptr(ev), stride(ev_lead, g.num_vs),
        _lib.ptr(ft), stride(ft_lead, g.num_es),
        _lib.ptr(out), _lib.ptr(deltas), _lib.ptr(iters), batch, n_it, float(damping),
        float(temperature), float(tol), _lib.current_stream()), "pgx_run_converge")
    if lead is None:
      return out[0], deltas[0], iters[0]
    return (out.reshape(lead + (g.num_es,)), deltas.reshape(lead + deltas.shape[1:]),
            iters.reshape(lead + iters.shape[1:]))

  def beliefs(self, evidence, ftov_msgs):
    """evidence + per-state sum of ftov (pgmax/infer/inferer.py:218-222)."""
    torch = _lib.require_gpu()
    g = self.graph
    ev, ev_lead = _as_batched(evidence, g.num_vs, "evidence")
    ft, ft_lead = _as_batched(ftov_msgs, g.num_es, "ftov_msgs")
    leads = {l for l in (ev_lead, ft_lead) if l is not None}
    if len(leads) > 1:
      raise ValueError(f"inconsistent batch shapes {sorted(leads)}")
    lead = leads.pop() if leads else None
    batch = int(np.prod(lead)) if lead is not None else 1
    out = torch.empty((batch, g.num_vs), device="cuda", dtype=torch.float32)
    _lib.check(_lib.load().pgx_get_beliefs(
        g.handle, _lib.ptr(ev), g.num_vs if ev_lead is not None else 0,
        _lib.ptr(ft), g.num_es if ft_lead is not None else 0,
        _lib.ptr(out), batch, _lib.current_stream()), "pgx_get_beliefs")
    return out[0] if lead is None else out.reshape(lead + (g.num_vs,))

  def beliefs_decode(self, evidence, ftov_msgs, marginals: bool = True):
    """(flat beliefs, MAP state per variable, flat marginals or None) in one
    call (pgmax/infer/inferer.py:211-264, pgmax/infer/bp.py:263-288)."""
    torch = _lib.require_gpu()
    g = self.graph
    ev, ev_lead = _as_batched(evidence, g.num_vs, "evidence")
    ft, ft_lead = _as_batched(ftov_msgs, g.num_es, "ftov_msgs")
    leads = {l for l in (ev_lead, ft_lead) if l is not None}
    if len(leads) > 1:
      raise ValueError(f"inconsistent batch shapes {sorted(leads)}")
    lead = leads.pop() if leads else None
    batch = int(np.prod(lead)) if lead is not None else 1
    bel = torch.empty((batch, g.num_vs), device="cuda", dtype=torch.float32)
    dec = torch.empty((batch, g.num_vars), device="cuda", dtype=torch.int32)
    marg = torch.empty((batch, g.num_vs), device="cuda", dtype=torch.float32) if marginals else None
    _lib.check(_lib.load().pgx_beliefs_decode(
        g.handle, _lib.ptr(ev), g.num_vs if ev_lead is not None else 0,
        _lib.ptr(ft), g.num_es if ft_lead is not None else 0,
        _lib.ptr(bel), _lib.ptr(dec), _lib.ptr(marg) if marginals else None, batch,
        _lib.current_stream()), "pgx_beliefs_decode")
    if lead is None:
      return bel[0], dec[0], (marg[0] if marginals else None)
    shape = lambda t, n: t.reshape(lead + (n,))
    return (shape(bel, g.num_vs), shape(dec, g.num_vars),
            shape(marg, g.num_vs) if marginals else None)

  def decode(self, flat_beliefs):
    """First-max state of every variable from flat beliefs (pgmax/infer/inferer.py:251-264)."""
    torch = _lib.require_gpu()
    g = self.graph
    b, lead = _as_batched(flat_beliefs, g.num_vs, "beliefs")
    batch = 1 if lead is None else int(np.prod(lead))
    out = torch.empty((batch, g.num_vars), device="cuda", dtype=torch.int32)
    _lib.check(_lib.load().pgx_decode_map(g.handle, _lib.ptr(b), _lib.ptr(out), batch,
                                          _lib.current_stream()), "pgx_decode_map")
    return out[0] if lead is None else out.reshape(lead + (g.num_vars,))


# ---------------------------------------------------------------------------
# Plugin API: one factor type's update (FAC_TO_VAR_UPDATES entries)
# ---------------------------------------------------------------------------

class _OpCache:
  """Bounded LRU of the tables the plugin-API ops compile from their wiring
  arguments, keyed by the arguments' contents (dtype, shape and a digest of
  their bytes), so equal wirings share an entry and different ones never
  do. Evicted graphs release their device memory at once."""

  def __init__(self, max_entries: int = 16):
    self.max_entries = max_entries
    self._d: "collections.OrderedDict[Tuple, Any]" = collections.OrderedDict()

  def get(self, key):
    v = self._d.get(key)
    if v is not None:
      self._d.move_to_end(key)
    return v

  def put(self, key, value) -> None:
    self._d[key] = value
    self._d.move_to_end(key)
    while len(self._d) > self.max_entries:
      _, old = self._d.popitem(last=False)
      if isinstance(old, _lib.NativeHandle):
        old.close()

  def clear(self) -> None:
    while self._d:
      _, old = self._d.popitem(last=False)
      if isinstance(old, _lib.NativeHandle):
        old.close()

  def __len__(self) -> int:
    return len(self._d)


_op_cache = _OpCache()


def _content_key(*arrays) -> Tuple:
  """Digest of each integer wiring array's values (int64 view, so the
  same indices in another dtype share the key)."""
  key = []
  for a in arrays:
    v = np.ascontiguousarray(np.asarray(_arrays.to_numpy(a), dtype=np.int64))
    key.append((v.shape, hashlib.blake2b(v.tobytes(), digest_size=16).hexdigest()))
  return tuple(key)


def _dev(x):
  torch = _lib.require_gpu()
  return (x if _arrays.is_torch(x) else torch.as_tensor(_host_f32(x))).to(
      device="cuda", dtype=torch.float32).contiguous()


def op_enum_fac_to_var(vtof_msgs, factor_configs_indices, factor_configs_edge_states,
                       log_potentials, num_val_configs, temperature, normalize=True):
  """pgmax/factor/enum.py:397-475 via the row-level HIP kernels."""
  torch = _lib.require_gpu()
  ci = np.asarray(_arrays.to_numpy(factor_configs_indices), dtype=np.int64)
  ce = np.asarray(_arrays.to_numpy(factor_configs_edge_states), dtype=np.int64)
  vt = _dev(vtof_msgs)
  n_es = vt.shape[0]
  key = ("enum",) + _content_key(ci, ce) + (int(num_val_configs), n_es)
  tabs = _op_cache.get(key)
  if tabs is None:
    order = np.argsort(ci, kind="stable")
    cfg_ptr = np.searchsorted(ci[order], np.arange(num_val_configs + 1))
    row_es = ce[order]
    es_order = np.argsort(ce, kind="stable")
    es_ptr = np.searchsorted(ce[es_order], np.arange(n_es + 1))
    es_cfg = ci[es_order]
    tabs = tuple(torch.as_tensor(_i32(a), device="cuda")
                 for a in (cfg_ptr, row_es, es_ptr, es_cfg))
    _op_cache.put(key, tabs)
  lp = _dev(log_potentials)
  summary = torch.empty((max(int(num_val_configs), 1),), device="cuda", dtype=torch.float32)
  out = torch.empty_like(vt)
  _lib.check(_lib.load().pgx_enum_rows_fac_to_var(
      *(_lib.ptr(t) for t in tabs), int(num_val_configs), n_es, _lib.ptr(vt), _lib.ptr(lp),
      _lib.ptr(summary), _lib.ptr(out), float(temperature), _lib.current_stream()),
             "pgx_enum_rows_fac_to_var")
  return out


def _logical_graph(kind, pfi, pmi, ces, off, n_es):
  """A graph holding only logical (or pool) factors, built from their arguments.

  Every edge is its own binary variable; edges follow pgmax's layout:
  each factor's parents, then its child (pgmax/factor/logical.py:457-488).
  """
  key = (kind,) + _content_key(pfi, pmi, ces) + (int(off), int(n_es))
  g = _op_cache.get(key)
  if g is not None:
    return g
  pfi = np.asarray(_arrays.to_numpy(pfi), dtype=np.int64)
  pmi = np.asarray(_arrays.to_numpy(pmi), dtype=np.int64)
  ces = np.asarray(_arrays.to_numpy(ces), dtype=np.int64)
  rel = 0 if kind == _lib.PGX_POOL else (1 - off) // 2
  nf = ces.shape[0]
  edge_starts = np.concatenate([pmi - rel, ces - rel])
  edge_fac = np.concatenate([pfi, np.arange(nf)])
  order = np.argsort(edge_starts, kind="stable")
  edge_starts, edge_fac = edge_starts[order], edge_fac[order]
  es = np.arange(n_es)
  edge_of_es = np.repeat(np.arange(edge_starts.shape[0]), 2)
  if edge_of_es.shape[0] != n_es or not np.array_equal(edge_starts, es[::2]):
    raise ValueError("logical wiring does not tile its message range with binary edges")
  types = [(0, 0) if t < kind else ((0, n_es) if t == kind else (n_es, n_es))
           for t in range(4)]
  w = Wiring(
      num_var_states=n_es, num_log_potentials=0, var_states_for_edge_states=es,
      edge_indices_for_edge_states=edge_of_es,
      factor_indices_for_edge_states=np.repeat(edge_fac, 2),
      evidence_to_vars=es // 2, msgs_range=types, potentials_range=[(0, 0)] * 4,
      logical_args={kind: (pfi, pmi, ces, off)},
  )
  g = Graph(w)
  _op_cache.put(key, g)
  return g


def _run_logical_op(kind, vtof_msgs, pfi, pmi, ces, off, temperature, normalize):
  torch = _lib.require_gpu()
  vt = _dev(vtof_msgs)
  g = _logical_graph(kind, pfi, pmi, ces, off, vt.shape[0])
  out = torch.zeros_like(vt)
  _lib.check(_lib.load().pgx_fac_to_var(
      g.handle, kind, _lib.ptr(vt), 0, _lib.ptr(out), float(temperature), int(bool(normalize)),
      _lib.current_stream()), "pgx_fac_to_var")
  return out


def op_logical_fac_to_var(vtof_msgs, parents_factor_indices, parents_msg_indices,
                          children_edge_states, edge_states_offset, temperature, normalize):
  """pgmax/factor/logical.py:491-779 via the HIP logical-factor kernel."""
  kind = _lib.PGX_OR if edge_states_offset == 1 else _lib.PGX_AND
  return _run_logical_op(kind, vtof_msgs, parents_factor_indices, parents_msg_indices,
                         children_edge_states, edge_states_offset, temperature, normalize)


def op_pool_fac_to_var(vtof_msgs, pool_choices_factor_indices, pool_choices_msg_indices,
                       pool_indicators_edge_states, temperature, normalize):
  """pgmax/factor/pool.py:275-474 via the HIP pool-factor kernel."""
  return _run_logical_op(_lib.PGX_POOL, vtof_msgs, pool_choices_factor_indices,
                         pool_choices_msg_indices, pool_indicators_edge_states, 1,
                         temperature, normalize)
