"""pgx_run_noalloc (the allocation-free run) against pgx_run, through the C ABI.

Two calls in a row share the library's per-GPU scratch array; the second has
a smaller input. Each must equal pgx_run on the same inputs bit for bit
(messages and deltas). A run whose scratch would not fit is refused before
any launch.
"""

import ctypes

import numpy as np
import pytest

import graphs
from pgmax_amd import _lib
from pgmax_amd import batch
from pgmax_amd import infer
from pgmax_amd import workloads

pytestmark = pytest.mark.gpu


def _np(x):
  return x.detach().cpu().numpy()


def _run(fn, g, lp, ev, ft, batch_size, n_it, temperature, lp_stride, ev_stride, ft_stride):
  import torch  # pylint: disable=g-import-not-at-top
  out = torch.full((batch_size, g.num_es), np.nan, device="cuda")
  deltas = torch.full((batch_size, n_it), np.nan, device="cuda")
  rc = fn(g.handle, _lib.ptr(lp), lp_stride, _lib.ptr(ev), ev_stride, _lib.ptr(ft), ft_stride,
          _lib.ptr(out), _lib.ptr(deltas), batch_size, n_it, ctypes.c_float(0.5),
          ctypes.c_float(temperature), _lib.current_stream())
  _lib.check(rc, fn.__name__)
  torch.cuda.synchronize()
  return _np(out), _np(deltas)


def _compare(g, lp, ev, ft, batch_size, n_it, temperature, strides):
  lib = _lib.load()
  ref = _run(lib.pgx_run, g, lp, ev, ft, batch_size, n_it, temperature, *strides)
  got = _run(lib.pgx_run_noalloc, g, lp, ev, ft, batch_size, n_it, temperature, *strides)
  np.testing.assert_array_equal(got[0], ref[0])
  np.testing.assert_array_equal(got[1], ref[1])


@pytest.mark.parametrize("temperature", [0.0, 1.0])
def test_noalloc_equals_run_twice_smaller_second(gpu, temperature):
  """A batch of 3 spin glasses, then a batch of 1 (planar 2x2 layout), then a
  Potts union of 2..8 states and a smaller one: every call equals pgx_run."""
  import torch  # pylint: disable=g-import-not-at-top
  rng = np.random.default_rng(7)
  fg, _, _ = graphs.ising_torus(side=24, seed=3)
  bp = infer.BP(fg.bp_state, temperature=temperature)
  g = bp.context.engine.graph
  for b in (3, 1):  # second call: smaller input
    lp = torch.as_tensor(rng.normal(size=(b, g.num_lp)).astype(np.float32), device="cuda")
    ev = torch.as_tensor(rng.normal(size=(b, g.num_vs)).astype(np.float32), device="cuda")
    ft = torch.as_tensor(rng.normal(size=(b, g.num_es)).astype(np.float32), device="cuda")
    _compare(g, lp, ev, ft, b, 7, temperature, (g.num_lp, g.num_vs, g.num_es))
  for n in (6, 3):  # second call: a smaller graph
    pb = workloads.potts_batch(num_graphs=n, seed=17, side_range=(9, 25), state_range=(2, 8))
    bb = batch.BatchedBP(pb.wirings, temperature=temperature)
    lp = torch.as_tensor(rng.normal(size=pb.log_potentials.shape).astype(np.float32),
                         device="cuda")
    ev = torch.as_tensor(pb.evidence, device="cuda")
    ft = torch.as_tensor(rng.normal(size=bb.graph.num_es).astype(np.float32), device="cuda")
    _compare(bb.graph, lp, ev, ft, 1, 9, temperature, (0, 0, 0))
    bb.close()


def test_noalloc_refuses_oversize_runs(gpu):
  """A run needing more than the 64 MiB scratch is refused with its size."""
  import torch  # pylint: disable=g-import-not-at-top
  fg, _, _ = graphs.ising_torus(side=24, seed=3)
  bp = infer.BP(fg.bp_state, temperature=1.0)
  g = bp.context.engine.graph
  b = 4096  # 4096 x (2 x 4,608 message floats + ...) > 64 MiB
  lp = torch.zeros((g.num_lp,), device="cuda")
  ev = torch.zeros((g.num_vs,), device="cuda")
  ft = torch.zeros((g.num_es,), device="cuda")
  out = torch.zeros((b, g.num_es), device="cuda")
  deltas = torch.zeros((b, 2), device="cuda")
  rc = _lib.load().pgx_run_noalloc(g.handle, _lib.ptr(lp), 0, _lib.ptr(ev), 0, _lib.ptr(ft), 0,
                                   _lib.ptr(out), _lib.ptr(deltas), b, 2, ctypes.c_float(0.5),
                                   ctypes.c_float(1.0), _lib.current_stream())
  assert rc == _lib.PGX_EINVAL
  msg = _lib.load().pgx_last_error().decode()
  assert "scratch" in msg and "pgx_run" in msg, msg
