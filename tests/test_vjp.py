"""Reverse-mode BP (pgx_run_vjp, infer.vjp) against the FP64 reverse-mode
oracle, which tests/test_oracle.py pins by central finite differences.

The GPU product is FP32 along the engine's FP32 trajectory, the oracle FP64
along its own, so they agree to a relative tolerance, not bit for bit; the
bars are ten times the measured differences.
"""

import numpy as np
import pytest

from oracle import lbp
from oracle import lbp_vjp
from pgmax_amd import fgraph
from pgmax_amd import fgroup
from pgmax_amd import infer
from pgmax_amd import vgroup
from pgmax_amd.infer import vjp as vjp_mod

pytestmark = pytest.mark.gpu


def _np(x):
  return x.detach().cpu().numpy() if hasattr(x, "detach") else np.asarray(x)


def _graph(seed):
  rng = np.random.default_rng(seed)
  v = vgroup.NDVarArray(num_states=3, shape=(6,))
  fg = fgraph.FactorGraph(v)
  fg.add_factors(fgroup.PairwiseFactorGroup(
      [[v[0], v[1]], [v[1], v[2]], [v[2], v[3]], [v[3], v[0]], [v[0], v[2]], [v[4], v[5]]],
      rng.normal(size=(6, 3, 3))))
  cfg = np.array([[a, b, c] for a in range(3) for b in range(3) for c in range(3)
                  if (a + b + c) % 2 == 0])
  fg.add_factors(fgroup.EnumFactorGroup([[v[1], v[3], v[4]]], cfg,
                                        log_potentials=rng.normal(size=(1, cfg.shape[0]))))
  bp = infer.BP(fg.bp_state, temperature=1.0)
  g = lbp.graph_from_context(bp.context)
  args = g["args"]["enum"]
  eg = lbp_vjp.EnumGraph(g["var_states_for_edge_states"], g["edge_indices_for_edge_states"],
                         args["factor_configs_indices"], args["factor_configs_edge_states"],
                         bp.context.bp_state.fg_state.num_var_states,
                         int(np.asarray(args["factor_configs_indices"]).max()) + 1)
  lp = np.asarray(fg.bp_state.log_potentials.value, np.float32)
  return bp, eg, lp, rng


def _rel_err(got, want):
  return float(np.max(np.abs(got - want)) / max(np.max(np.abs(want)), 1e-12))


@pytest.mark.parametrize("temperature,damping", [(1.0, 0.5), (1.0, 0.0), (0.4, 0.3), (0.0, 0.5)])
def test_run_vjp_matches_oracle(gpu, temperature, damping):
  """Batched inputs (one set per element) and shared ones (stride 0, whose
  gradients sum over the batch)."""
  bp, eg, lp, rng = _graph(5)
  n, nb = 7, 3
  lpb = (lp[None] + rng.normal(0, 0.3, size=(nb, lp.size))).astype(np.float32)
  ev = rng.normal(size=(nb, eg.num_vs)).astype(np.float32)
  ft = rng.normal(size=(nb, eg.num_es)).astype(np.float32)
  w = rng.normal(size=(nb, eg.num_es)).astype(np.float32)
  run_vjp = vjp_mod.RunVJP(bp.context)
  g_lp, g_ev, g_ft = (_np(x) for x in run_vjp(lpb, ev, ft, n, damping, temperature, w))
  errs = []
  for b in range(nb):
    r_lp, r_ev, r_ft = lbp_vjp.vjp(eg, lpb[b], ev[b], ft[b], n, damping, temperature, w[b])
    errs += [_rel_err(g_lp[b], r_lp), _rel_err(g_ev[b], r_ev), _rel_err(g_ft[b], r_ft)]
  # shared log-potentials and initial messages, batched evidence
  s_lp, s_ev, s_ft = (_np(x) for x in run_vjp(lp, ev, ft[0], n, damping, temperature, w))
  assert s_lp.shape == lp.shape and s_ev.shape == ev.shape and s_ft.shape == ft[0].shape
  refs = [lbp_vjp.vjp(eg, lp, ev[b], ft[0], n, damping, temperature, w[b]) for b in range(nb)]
  errs += [_rel_err(s_lp, sum(r[0] for r in refs)), _rel_err(s_ft, sum(r[2] for r in refs))]
  errs += [_rel_err(s_ev[b], refs[b][1]) for b in range(nb)]
  print(f"T={temperature} d={damping}: max relative error {max(errs):.2e}")
  assert max(errs) < 3e-5  # measured <= 2.4e-6
  run_vjp.close()


def test_gradient_of_a_marginal_loss(gpu):
  """The reference's learning loop (examples/grid_mrf.ipynb cell 15): a
  cross-entropy of marginals after sum-product BP, differentiated with
  respect to the log-potentials and the evidence through torch autograd and
  run_differentiable; equal to the oracle's product with the loss's
  cotangent of the messages."""
  import torch  # pylint: disable=g-import-not-at-top
  bp, eg, lp, rng = _graph(6)
  n, damping = 15, 0.0
  ev = rng.normal(size=eg.num_vs).astype(np.float32)
  target = rng.integers(0, 3, size=6)
  lp_t = torch.tensor(lp, device="cuda", requires_grad=True)
  ev_t = torch.tensor(ev, device="cuda", requires_grad=True)
  ft_t = torch.zeros(eg.num_es, device="cuda")
  msgs = vjp_mod.run_differentiable(bp, lp_t, ev_t, ft_t, n, damping, 1.0)
  msgs.retain_grad()
  beliefs = vjp_mod.flat_beliefs(bp.context, ev_t, msgs).reshape(6, 3)
  marg = torch.softmax(beliefs, dim=-1)
  loss = -torch.log(marg[torch.arange(6), torch.as_tensor(target, device="cuda")]).mean()
  loss.backward()
  g_out = _np(msgs.grad).astype(np.float64)
  r_lp, r_ev, _ = lbp_vjp.vjp(eg, lp, ev, np.zeros(eg.num_es), n, damping, 1.0, g_out)
  # the beliefs' own dependence on the evidence (outside the run)
  b64 = ev.astype(np.float64).copy()
  np.add.at(b64, eg.vsi, lbp_vjp.forward(eg, lp, ev, np.zeros(eg.num_es), n, damping, 1.0)[0])
  p = np.exp(b64.reshape(6, 3) - b64.reshape(6, 3).max(-1, keepdims=True))
  p /= p.sum(-1, keepdims=True)
  onehot = np.eye(3)[target]
  direct = ((p - onehot) / 6).reshape(-1)
  e_lp, e_ev = _rel_err(_np(lp_t.grad), r_lp), _rel_err(_np(ev_t.grad), r_ev + direct)
  print(f"marginal loss: relative error lp {e_lp:.2e}, evidence {e_ev:.2e}")
  assert e_lp < 2e-6 and e_ev < 2e-6  # measured 1.2e-7


def test_reverse_mode_needs_enum_factors(gpu):
  v = vgroup.NDVarArray(num_states=2, shape=(3,))
  fg = fgraph.FactorGraph(v)
  fg.add_factors(fgroup.ORFactorGroup([[v[0], v[1], v[2]]]))
  bp = infer.BP(fg.bp_state, temperature=1.0)
  with pytest.raises(ValueError, match="Enum factors only"):
    vjp_mod.RunVJP(bp.context)
