"""Loopy belief propagation driver (pgmax/infer/bp.py).

``BP(bp_state, temperature)`` returns the same bundle of functions as the
reference (pgmax/infer/bp.py:36-188). ``run`` / ``run_with_diffs`` hand the
whole loop (clip, normalise, ``num_iters`` sweeps of var->factor,
factor->var, damping, normalise-and-clip, delta) to the native engine: one
call, stream-ordered, no host round trip per iteration.
"""

import dataclasses
from typing import Any, Callable, Dict, Hashable, Optional, Tuple
import warnings

import numpy as np

from pgmax_amd import _arrays
from pgmax_amd.infer.bp_state import BPArrays
from pgmax_amd.infer.bp_state import BPState
from pgmax_amd.infer.inferer import Inferer
from pgmax_amd.infer.inferer import InfererContext


@dataclasses.dataclass(frozen=True, eq=False)
class BeliefPropagation(Inferer):
  """pgmax/infer/bp.py:36-46."""

  run_bp: Callable[..., BPArrays]
  run_with_diffs: Callable[..., Tuple[BPArrays, Any]]


def BP(bp_state: BPState, temperature: Optional[float] = 0.0) -> BeliefPropagation:  # pylint: disable=invalid-name
  """Belief-propagation functions for a BPState (pgmax/infer/bp.py:50-188).

  Args:
    bp_state: Belief propagation state.
    temperature: Default temperature; 1.0 is sum-product, 0.0 max-product.
  """
  inferer_context = InfererContext(bp_state)

  def run_with_diffs(
      bp_arrays: BPArrays,
      num_iters: int,
      damping: float = 0.5,
      temperature: float = temperature,
  ):
    """Run ``num_iters`` BP iterations (pgmax/infer/bp.py:63-155).

    Returns the new BPArrays (log-potentials and evidence passed through
    unchanged, pgmax/infer/bp.py:148-152) and the per-iteration max-abs
    message change. ``num_iters <= 1`` runs exactly one iteration
    (pgmax/infer/bp.py:142-146).
    """
    ftov_msgs, msgs_deltas = inferer_context.engine.run(
        bp_arrays.log_potentials,
        bp_arrays.ftov_msgs,
        bp_arrays.evidence,
        num_iters=int(num_iters),
        damping=float(damping),
        temperature=float(temperature),
    )
    return (
        BPArrays(
            log_potentials=bp_arrays.log_potentials,
            ftov_msgs=ftov_msgs,
            evidence=bp_arrays.evidence,
        ),
        msgs_deltas,
    )

  def run(
      bp_arrays: BPArrays,
      num_iters: int,
      damping: float = 0.5,
      temperature: float = temperature,
  ) -> BPArrays:
    """pgmax/infer/bp.py:157-165."""
    bp_arrays, _ = run_with_diffs(bp_arrays, num_iters, damping, temperature)
    return bp_arrays

  def run_bp(bp_arrays: BPArrays, num_iters: int, damping: float = 0.5) -> BPArrays:
    """pgmax/infer/bp.py:167-176."""
    warnings.warn("BP.run_bp is deprecated. Please consider using BP.run instead.")
    return run(bp_arrays, num_iters, damping, temperature)

  bp = BeliefPropagation(
      init=inferer_context.init,
      update=inferer_context.update,
      to_bp_state=inferer_context.to_bp_state,
      get_beliefs=inferer_context.get_beliefs,
      run=run,
      run_bp=run_bp,
      run_with_diffs=run_with_diffs,
  )
  object.__setattr__(bp, "context", inferer_context)
  object.__setattr__(inferer_context, "temperature", temperature)
  return bp


def get_marginals(beliefs: Dict[Hashable, Any]) -> Dict[Hashable, Any]:
  """Softmax of beliefs over each variable's states (pgmax/infer/bp.py:263-288)."""

  def _softmax(x):
    if _arrays.is_torch(x):
      if x.numel() == 0:
        return x
      return (x - x.logsumexp(dim=-1, keepdim=True)).exp()
    x = np.asarray(x)
    if x.size == 0:
      return x
    m = np.max(x, axis=-1, keepdims=True)
    lse = m + np.log(np.sum(np.exp(x - m), axis=-1, keepdims=True))
    return np.exp(x - lse)

  out = {}
  for key, value in beliefs.items():
    if isinstance(value, dict):
      out[key] = {k: _softmax(v) for k, v in value.items()}
    else:
      out[key] = _softmax(value)
  return out
