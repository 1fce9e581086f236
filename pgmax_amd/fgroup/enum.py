"""Groups of enumeration factors (pgmax/fgroup/enum.py)."""

import dataclasses
from typing import Optional, Type

import numpy as np

from pgmax_amd import _arrays
from pgmax_amd.factor import enum
from pgmax_amd.factor import factor
from pgmax_amd.fgroup import fgroup


def _tile(data, reps):
  if _arrays.is_torch(data):
    return data.reshape(-1).repeat(reps)
  return np.tile(np.asarray(data).reshape(-1), reps)


@dataclasses.dataclass(frozen=True, eq=False)
class EnumFactorGroup(fgroup.FactorGroup):
  """EnumFactors sharing one config table (pgmax/fgroup/enum.py:29-206)."""

  factor_configs: np.ndarray
  log_potentials: Optional[np.ndarray] = None
  factor_type: Type[factor.Factor] = dataclasses.field(
      init=False, default=enum.EnumFactor
  )

  def __post_init__(self):
    super().__post_init__()
    num_val_configs = self.factor_configs.shape[0]
    if self.log_potentials is None:
      log_potentials = np.zeros((self.num_factors, num_val_configs), dtype=float)
    else:
      if self.log_potentials.shape != (num_val_configs,) and (
          self.log_potentials.shape != (self.num_factors, num_val_configs)
      ):
        raise ValueError(
            f"Expected log potentials shape: {(num_val_configs,)} or"
            f" {(self.num_factors, num_val_configs)}. Got"
            f" {self.log_potentials.shape}."
        )
      log_potentials = np.broadcast_to(
          np.asarray(self.log_potentials), (self.num_factors, num_val_configs)
      )
    if not np.issubdtype(log_potentials.dtype, np.floating):
      raise ValueError(f"Potentials should be floats. Got {log_potentials.dtype}.")
    object.__setattr__(self, "log_potentials", log_potentials)

  def _make_factor(self, f: int) -> enum.EnumFactor:
    return enum.EnumFactor(
        variables=list(self.given_factor(f)),
        factor_configs=self.factor_configs,
        log_potentials=np.asarray(self.log_potentials[f]),
    )

  @property
  def _edge_states_per_factor(self) -> int:
    e = self.factor_edges
    return int(e.num_states[e.offsets[0]:e.offsets[1]].sum())

  def flatten(self, data):
    """pgmax/fgroup/enum.py:127-162."""
    num_factors = self.num_factors
    es = self._edge_states_per_factor
    nc = self.factor_configs.shape[0]
    shape = tuple(data.shape)
    if shape not in ((num_factors, nc), (num_factors, es), (nc,)):
      raise ValueError(
          "data should be of shape"
          f" {(num_factors, nc)} or {(num_factors, es)} or {(nc,)}. Got"
          f" {shape}."
      )
    if shape == (nc,):
      return _tile(data, num_factors)
    return data.reshape(-1)

  def unflatten(self, flat_data):
    """pgmax/fgroup/enum.py:164-206."""
    if flat_data.ndim != 1:
      raise ValueError(
          f"Can only unflatten 1D array. Got a {flat_data.ndim}D array."
      )
    num_factors = self.num_factors
    es = self._edge_states_per_factor
    nc = self.factor_configs.shape[0]
    if flat_data.shape[0] == num_factors * nc:
      return flat_data.reshape((num_factors, nc))
    if flat_data.shape[0] == num_factors * es:
      return flat_data.reshape((num_factors, es))
    raise ValueError(
        "flat_data should be compatible with shape"
        f" {(num_factors, nc)} or {(num_factors, es)}. Got"
        f" {tuple(flat_data.shape)}."
    )


@dataclasses.dataclass(frozen=True, eq=False)
class PairwiseFactorGroup(fgroup.FactorGroup):
  """Pairwise EnumFactors with all S1 x S2 configs (pgmax/fgroup/enum.py:209-435).

  Configs are in row-major ``mgrid`` order: config ``a * S2 + b`` is
  ``(a, b)`` (pgmax/fgroup/enum.py:303-310).
  """

  log_potential_matrix: Optional[np.ndarray] = None
  factor_type: Type[factor.Factor] = dataclasses.field(
      init=False, default=enum.EnumFactor
  )

  def __post_init__(self):
    super().__post_init__()
    edges = self.factor_edges
    num_factors = edges.num_factors
    sizes = edges.sizes
    if self.log_potential_matrix is None:
      first = edges.num_states[edges.offsets[0]:edges.offsets[0] + 2]
      lpm = np.zeros(tuple(int(n) for n in first))
    else:
      lpm = np.asarray(self.log_potential_matrix)
    if not (lpm.ndim == 2 or lpm.ndim == 3):
      raise ValueError(
          "log_potential_matrix should be either a 2D array, specifying shared"
          " parameters for all pairwise factors, or 3D array, specifying"
          " parameters for individual pairwise factors. Got a"
          f" {lpm.ndim}D log_potential_matrix array."
      )
    if not np.issubdtype(lpm.dtype, np.floating):
      raise ValueError(f"Potential matrix should be floats. Got {lpm.dtype}.")
    if lpm.ndim == 3 and lpm.shape[0] != num_factors:
      raise ValueError(
          f"Expected log_potential_matrix for {num_factors} factors. Got"
          f" log_potential_matrix for {lpm.shape[0]} factors."
      )
    lp_shape = lpm.shape[-2:]
    # one pass over the edge table: the first factor that is not a pair, or
    # whose two state counts differ from the matrix, is reported
    not_pair = sizes != 2
    second = np.minimum(edges.offsets[:-1] + 1, max(edges.num_edges - 1, 0))
    first_ns = edges.num_states[np.minimum(edges.offsets[:-1], max(edges.num_edges - 1, 0))]
    mismatch = (first_ns != lp_shape[0]) | (edges.num_states[second] != lp_shape[1])
    bad = np.flatnonzero(not_pair | mismatch)
    if bad.size:
      f = int(bad[0])
      vs = self.given_factor(f)
      if not_pair[f]:
        raise ValueError(
            "All pairwise factors should connect to exactly 2 variables. Got a"
            f" factor connecting to {len(vs)} variables ({vs})."
        )
      raise ValueError(
          f"The specified pairwise factor {vs} (with"
          f" {(vs[0][1], vs[1][1])}configurations) does not match the"
          " specified log_potential_matrix (with"
          f" {lp_shape} configurations)."
      )
    object.__setattr__(self, "log_potential_matrix", lpm)
    # configuration a * S2 + b is (a, b): row-major over the two state counts
    rows, cols = np.divmod(np.arange(lp_shape[0] * lp_shape[1]), lp_shape[1])
    object.__setattr__(self, "factor_configs", np.stack([rows, cols], axis=1))
    full = np.broadcast_to(lpm, (num_factors,) + tuple(lp_shape))
    object.__setattr__(self, "log_potentials", full.reshape(num_factors, -1))

  def _make_factor(self, f: int) -> enum.EnumFactor:
    return enum.EnumFactor(
        variables=list(self.given_factor(f)),
        factor_configs=self.factor_configs,
        log_potentials=np.asarray(self.log_potentials[f]),
    )

  def flatten(self, data):
    """pgmax/fgroup/enum.py:353-386."""
    num_factors = self.num_factors
    s2 = tuple(self.log_potential_matrix.shape[-2:])
    shape = tuple(data.shape)
    if shape not in ((num_factors,) + s2, (num_factors, int(np.sum(s2))), s2):
      raise ValueError(
          "data should be of shape"
          f" {(num_factors,) + s2} or {(num_factors, int(np.sum(s2)))} or"
          f" {s2}. Got {shape}."
      )
    if shape == s2:
      return _tile(data, num_factors)
    return data.reshape(-1)

  def unflatten(self, flat_data):
    """pgmax/fgroup/enum.py:388-435."""
    if flat_data.ndim != 1:
      raise ValueError(
          f"Can only unflatten 1D array. Got a {flat_data.ndim}D array."
      )
    num_factors = self.num_factors
    s2 = tuple(self.log_potential_matrix.shape[-2:])
    if flat_data.shape[0] == num_factors * int(np.prod(s2)):
      return flat_data.reshape((num_factors,) + s2)
    if flat_data.shape[0] == num_factors * int(np.sum(s2)):
      return flat_data.reshape((num_factors, int(np.sum(s2))))
    raise ValueError(
        "flat_data should be compatible with shape"
        f" {(num_factors,) + s2} or {(num_factors, int(np.sum(s2)))}. Got"
        f" {tuple(flat_data.shape)}."
    )
