"""ctypes driver of oracle/lbp_cpu.c -- TEST INFRASTRUCTURE ONLY.

The multi-threaded C restatement of the reference's Enum BP iteration (and of
max-product OR / AND BP), used as bench.py's timed CPU baseline and as a fast
bitwise oracle for full-size runs (tests/test_oracle.py checks it bit for bit
against the NumPy oracle). Built by ``build()`` (gcc -O3 -fopenmp) into
oracle/liblbp_cpu.so.
"""

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(_HERE, "lbp_cpu.c")
LIB = os.path.join(_HERE, "liblbp_cpu.so")
_lib = None


def build() -> str:
  # FMA instructions for fmaf; no contraction of any other multiply-add, so
  # every operation rounds as in the NumPy oracle
  cmd = ["gcc", "-O3", "-march=x86-64-v2", "-mfma", "-ffp-contract=off", "-fno-fast-math",
         "-fopenmp", "-fPIC", "-shared", "-o", LIB + ".tmp", SRC, "-lm"]
  subprocess.run(cmd, check=True)
  os.replace(LIB + ".tmp", LIB)
  return LIB


def _load():
  global _lib
  if _lib is None:
    if not os.path.exists(LIB):
      raise RuntimeError(f"{LIB} is missing: run oracle.cpu.build()")
    lib = ctypes.CDLL(LIB)
    p, i64, i32, f = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_float
    lib.lbp_cpu_threads.restype = ctypes.c_int
    lib.lbp_cpu_threads.argtypes = [ctypes.c_int]
    lib.lbp_cpu_run_enum.restype = ctypes.c_int
    lib.lbp_cpu_run_enum.argtypes = [i64, i64, i64, i64, i64, p, p, p, p, i64, p, i64, p, i64,
                                     p, p, p, i32, i32, f, f]
    lib.lbp_cpu_math.restype = None
    lib.lbp_cpu_math.argtypes = [ctypes.c_int, p, p, i64]
    lib.lbp_cpu_run_enum_graphs.restype = ctypes.c_int
    lib.lbp_cpu_run_enum_graphs.argtypes = [i32, p, p, p, p, p, p, p, p, p, p, i32, f, f]
    lib.lbp_cpu_run_logical_t0.restype = ctypes.c_int
    lib.lbp_cpu_run_logical_t0.argtypes = [i64, i64, i64, p, p, i32, p, p, p, p, p, p, p, i32,
                                           f]
    _lib = lib
  return _lib


def set_threads(n: int = 0) -> int:
  """Threads for the next runs (0: OpenMP default); returns the count in use."""
  return int(_load().lbp_cpu_threads(int(n)))


def math(op: str, x) -> np.ndarray:
  """The C restatement's exp or log, elementwise."""
  x = _f32(x).reshape(-1)
  y = np.empty_like(x)
  _load().lbp_cpu_math({"exp": 0, "log": 1}[op], x.ctypes.data, y.ctypes.data, x.size)
  return y


def _i32(a):
  return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _f32(a):
  return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def run_enum(g, log_potentials, ftov_in, evidence, num_iters, damping, temperature,
             threads=None):
  """BP on a graph of Enum factors only (oracle.lbp graph dict); inputs may
  carry a leading batch axis. Returns (ftov, deltas)."""
  lib = _load()
  set_threads(threads or 0)
  m0, m1 = g["msgs_range"]["enum"]
  num_es = int(g["var_states_for_edge_states"].shape[0])
  if (m0, m1) != (0, num_es):
    raise ValueError("run_enum handles graphs of Enum factors only")
  p0, p1 = g["potentials_range"]["enum"]
  args = g["args"]["enum"]
  cfg = _i32(args["factor_configs_indices"])
  es = _i32(np.asarray(args["factor_configs_edge_states"]) + m0)
  vs = _i32(g["var_states_for_edge_states"])
  ed = _i32(g["edge_indices_for_edge_states"])
  lp = _f32(log_potentials)
  ev = _f32(evidence)
  ft = _f32(ftov_in)
  batched = ev.ndim == 2
  batch = ev.shape[0] if batched else 1
  num_vs = ev.shape[-1]
  n_it = max(int(num_iters), 1)
  out = np.empty((batch, num_es), np.float32)
  deltas = np.empty((batch, n_it), np.float32)
  ft_b = np.broadcast_to(ft, (batch, num_es)).copy() if ft.ndim == 1 else ft
  rc = lib.lbp_cpu_run_enum(
      num_es, num_vs, int(g["num_edges"]), int(p1 - p0), int(cfg.shape[0]),
      vs.ctypes.data, ed.ctypes.data, cfg.ctypes.data, es.ctypes.data, int(p0),
      lp.ctypes.data, lp.shape[-1] if lp.ndim == 2 else 0,
      ev.ctypes.data, num_vs if batched else 0,
      ft_b.ctypes.data, out.ctypes.data, deltas.ctypes.data, batch, n_it,
      float(damping), float(temperature))
  if rc != 0:
    raise MemoryError("lbp_cpu_run_enum ran out of memory")
  return (out, deltas) if batched else (out[0], deltas[0])


def run_enum_graphs(graphs, log_potentials, evidence, ftov_in, num_iters, damping,
                    temperature, threads=None):
  """BP on independent Enum-only graphs, in parallel over the graphs.

  ``graphs`` are oracle.lbp-style dicts (explicit Enum rows); the three input
  lists hold one flat array per graph. Returns ([ftov per graph], deltas
  [graphs, iters]).
  """
  lib = _load()
  set_threads(threads or 0)
  n_it = max(int(num_iters), 1)
  sizes, vs, ed, cfg, es = [], [], [], [], []
  for g, ev_g in zip(graphs, evidence):
    num_es = int(g["var_states_for_edge_states"].shape[0])
    if tuple(g["msgs_range"]["enum"]) != (0, num_es):
      raise ValueError("run_enum_graphs handles graphs of Enum factors only")
    args = g["args"]["enum"]
    ci = np.asarray(args["factor_configs_indices"])
    num_cfg = int(ci.max()) + 1 if ci.size else 0
    sizes.append([num_es, int(np.asarray(ev_g).size), int(g["num_edges"]), num_cfg, ci.size, 0])
    vs.append(_i32(g["var_states_for_edge_states"]))
    ed.append(_i32(g["edge_indices_for_edge_states"]))
    cfg.append(_i32(ci))
    es.append(_i32(args["factor_configs_edge_states"]))
  cat = lambda xs, f: f(np.concatenate(xs)) if xs else f(np.zeros(0))
  sz = np.ascontiguousarray(np.asarray(sizes, np.int64))
  vs, ed, cfg, es = (cat(x, _i32) for x in (vs, ed, cfg, es))
  lp = cat([np.asarray(x).reshape(-1) for x in log_potentials], _f32)
  ev = cat([np.asarray(x).reshape(-1) for x in evidence], _f32)
  ft = cat([np.asarray(x).reshape(-1) for x in ftov_in], _f32)
  out = np.empty_like(ft)
  deltas = np.empty((len(graphs), n_it), np.float32)
  rc = lib.lbp_cpu_run_enum_graphs(len(graphs), sz.ctypes.data, vs.ctypes.data, ed.ctypes.data,
                                   cfg.ctypes.data, es.ctypes.data, lp.ctypes.data,
                                   ev.ctypes.data, ft.ctypes.data, out.ctypes.data,
                                   deltas.ctypes.data, n_it, float(damping), float(temperature))
  if rc != 0:
    raise MemoryError("lbp_cpu_run_enum_graphs ran out of memory")
  bounds = np.concatenate([[0], np.cumsum(sz[:, 0])])
  return [out[bounds[i]:bounds[i + 1]] for i in range(len(graphs))], deltas


def run_logical_t0(g, ftov_in, evidence, num_iters, damping, threads=None):
  """Max-product BP on a graph of OR and AND factors only (oracle.lbp graph
  dict, one graph). Returns (ftov, deltas)."""
  lib = _load()
  set_threads(threads or 0)
  num_es = int(g["var_states_for_edge_states"].shape[0])
  for t in ("enum", "pool"):
    m0, m1 = g["msgs_range"][t]
    if m1 > m0:
      raise ValueError("run_logical_t0 handles graphs of OR and AND factors only")
  sizes, arrays, offs, keep = [], [], [], []
  for t in ("or", "and"):
    m0, m1 = g["msgs_range"][t]
    if m1 == m0:
      continue
    a = g["args"][t]
    pfi, pmi, ces = (_i32(a[k]) for k in ("parents_factor_indices", "parents_msg_indices",
                                          "children_edge_states"))
    keep += [pfi, pmi, ces]
    sizes += [int(m0), int(pfi.size), int(ces.size)]
    arrays += [pfi.ctypes.data, pmi.ctypes.data, ces.ctypes.data]
    offs.append(int(a["edge_states_offset"]))
  vs = _i32(g["var_states_for_edge_states"])
  ed = _i32(g["edge_indices_for_edge_states"])
  ev = _f32(evidence).reshape(-1)
  ft = _f32(ftov_in).reshape(-1)
  n_it = max(int(num_iters), 1)
  out = np.empty(num_es, np.float32)
  deltas = np.empty(n_it, np.float32)
  sz = np.ascontiguousarray(np.asarray(sizes, np.int64))
  ar = (ctypes.c_void_p * max(len(arrays), 1))(*arrays)
  of = np.ascontiguousarray(np.asarray(offs, np.int32))
  rc = lib.lbp_cpu_run_logical_t0(
      num_es, ev.size, int(g["num_edges"]), vs.ctypes.data, ed.ctypes.data, len(offs),
      sz.ctypes.data, ctypes.cast(ar, ctypes.c_void_p), of.ctypes.data, ev.ctypes.data,
      ft.ctypes.data, out.ctypes.data, deltas.ctypes.data, n_it, float(damping))
  del keep
  if rc != 0:
    raise MemoryError("lbp_cpu_run_logical_t0 failed (memory, or an edge's states not consecutive)")
  return out, deltas
