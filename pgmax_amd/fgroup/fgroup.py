"""Groups of factors of one type (the interface of pgmax/fgroup/fgroup.py:28-296).

A group's factors are given by ``variables_for_factors``, in any of three
forms:

* a sequence of variable lists, one per factor (the reference's form);
* an int array of shape (num_factors, num_variables, 2) of (hash,
  num_states) pairs, for groups whose factors all have the same arity;
* a ``RaggedVariables`` (flat (hash, num_states) pairs plus factor sizes),
  for groups of millions of factors of varying arity.

Whatever the form, the group keeps its factors as one ``FactorEdges`` table
-- the (hash, num_states) of every (factor, variable) edge in factor order
and each factor's first edge -- so that wirings compile with array
operations and ``Factor`` objects are only built when asked for.
"""

import dataclasses
import inspect
import itertools
from typing import Any, Dict, FrozenSet, List, Optional, Sequence, Tuple, Type

import numpy as np

from pgmax_amd.factor import factor
from pgmax_amd.utils import cached_property


@dataclasses.dataclass(frozen=True)
class RaggedVariables:
  """The variables of factors of varying arity as flat arrays.

  Attributes:
    pairs: int array (num_edges, 2): (hash, num_states) of every edge, the
      edges of factor 0 first, then those of factor 1, ...
    sizes: int array (num_factors,): edges per factor.
  """

  pairs: np.ndarray
  sizes: np.ndarray


class FactorEdges:
  """Every (factor, variable) edge of a group as flat arrays.

  ``hashes[e]`` / ``num_states[e]`` identify the variable of edge ``e``
  (edges in factor order, variables in each factor's order) and factor
  ``f`` owns edges ``[offsets[f], offsets[f + 1])``.
  """

  __slots__ = ("hashes", "num_states", "offsets", "arity")

  def __init__(self, hashes: np.ndarray, num_states: np.ndarray, offsets: np.ndarray,
               arity: Optional[int]):
    self.hashes = hashes
    self.num_states = num_states
    self.offsets = offsets
    self.arity = arity  # the common number of variables per factor, or None

  @classmethod
  def build(cls, variables_for_factors: Any) -> "FactorEdges":
    if isinstance(variables_for_factors, RaggedVariables):
      pairs = np.asarray(variables_for_factors.pairs, dtype=np.int64).reshape(-1, 2)
      sizes = np.asarray(variables_for_factors.sizes, dtype=np.int64).reshape(-1)
      if sizes.sum() != pairs.shape[0]:
        raise ValueError(
            f"RaggedVariables sizes add up to {int(sizes.sum())} edges, but"
            f" {pairs.shape[0]} (hash, num_states) pairs were given."
        )
      offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
      arity = int(sizes[0]) if sizes.size and np.all(sizes == sizes[0]) else None
      return cls(pairs[:, 0].copy(), pairs[:, 1].copy(), offsets, arity)
    if isinstance(variables_for_factors, np.ndarray) and variables_for_factors.ndim == 3:
      arr = np.asarray(variables_for_factors, dtype=np.int64)
      if arr.shape[-1] != 2:
        raise ValueError(
            "An array of variables_for_factors should have shape (num_factors,"
            f" num_variables, 2). Got {arr.shape}."
        )
      n, v = arr.shape[0], arr.shape[1]
      offsets = np.arange(0, n * v + 1, v, dtype=np.int64) if v else np.zeros(n + 1, np.int64)
      flat = arr.reshape(-1, 2)
      return cls(flat[:, 0].copy(), flat[:, 1].copy(), offsets, v)
    rows = list(variables_for_factors)
    sizes = np.fromiter((len(r) for r in rows), dtype=np.int64, count=len(rows))
    flat = np.fromiter(
        itertools.chain.from_iterable(
            itertools.chain.from_iterable(r) for r in rows),
        dtype=np.int64, count=2 * int(sizes.sum()))
    flat = flat.reshape(-1, 2)
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    arity = int(sizes[0]) if sizes.size and np.all(sizes == sizes[0]) else None
    return cls(flat[:, 0].copy(), flat[:, 1].copy(), offsets, arity)

  @property
  def num_factors(self) -> int:
    return int(self.offsets.shape[0]) - 1

  @property
  def num_edges(self) -> int:
    return int(self.hashes.shape[0])

  @property
  def sizes(self) -> np.ndarray:
    return np.diff(self.offsets)

  def factor(self, f: int) -> List[Tuple[int, int]]:
    """The variables of factor ``f`` as (hash, num_states) tuples."""
    a, z = int(self.offsets[f]), int(self.offsets[f + 1])
    return list(zip(self.hashes[a:z].tolist(), self.num_states[a:z].tolist()))

  def rows(self) -> List[List[Tuple[int, int]]]:
    return [self.factor(f) for f in range(self.num_factors)]


@dataclasses.dataclass(frozen=True, eq=False)
class FactorGroup:
  """A group of factors of one type.

  Subclasses set ``factor_type`` and the attributes its ``compile_wiring``
  takes (looked up on the group by parameter name, as in the reference's
  plugin API), and build one factor object on request (``_make_factor``).
  Groups compare and hash by identity.
  """

  variables_for_factors: Any
  factor_configs: np.ndarray = dataclasses.field(init=False)
  log_potentials: np.ndarray = dataclasses.field(
      init=False, default_factory=lambda: np.empty((0,))
  )
  factor_type: Type[Any] = dataclasses.field(init=False)

  def __post_init__(self):
    edges = FactorEdges.build(self.variables_for_factors)
    if edges.num_factors == 0:
      raise ValueError("Cannot create a FactorGroup with no Factor.")
    object.__setattr__(self, "factor_edges", edges)
    object.__setattr__(self, "_serial", next(_SERIALS))

  def __lt__(self, other: "FactorGroup") -> bool:
    """Groups order by creation."""
    return self._serial < other._serial

  def given_factor(self, f: int) -> Any:
    """The variables of factor ``f`` as the caller gave them (a list of
    (hash, num_states) tuples when they came as an array)."""
    if isinstance(self.variables_for_factors, (list, tuple)):
      return self.variables_for_factors[f]
    return self.factor_edges.factor(f)

  # ---- sizes ------------------------------------------------------------
  @property
  def num_factors(self) -> int:
    return self.factor_edges.num_factors

  @property
  def factor_sizes(self) -> np.ndarray:
    return self.factor_edges.sizes

  @property
  def factor_edges_num_states(self) -> np.ndarray:
    """States of every (factor, variable) edge, in edge order."""
    return self.factor_edges.num_states

  @cached_property
  def factor_group_log_potentials(self) -> np.ndarray:
    return np.asarray(self.log_potentials).reshape(-1)

  # ---- factor objects (built on request) --------------------------------
  def _make_factor(self, f: int) -> Any:
    raise NotImplementedError(
        "Please subclass the FactorGroup class and override this method"
    )

  @cached_property
  def _factor_of_variables(self) -> Dict[FrozenSet[Any], int]:
    out = {}
    for f, row in enumerate(self.factor_edges.rows()):
      out.setdefault(frozenset(row), f)
    return out

  def __getitem__(self, variables: Sequence[Tuple[int, int]]) -> Any:
    """The factor over a set of variables."""
    key = frozenset(variables)
    f = self._factor_of_variables.get(frozenset((int(h), int(n)) for h, n in key))
    if f is None:
      raise ValueError(
          f"The queried factor connected to the set of variables {key} is"
          " not present in the factor group."
      )
    return self.factors[f]

  @cached_property
  def factors(self) -> Tuple[Any, ...]:
    return tuple(self._make_factor(f) for f in range(self.num_factors))

  def flatten(self, data):
    raise NotImplementedError(
        "Please subclass the FactorGroup class and override this method"
    )

  def unflatten(self, flat_data):
    raise NotImplementedError(
        "Please subclass the FactorGroup class and override this method"
    )

  # ---- wiring -------------------------------------------------------------
  def _wiring_attr(self, name: str) -> Any:
    return getattr(self, name)

  def compile_wiring(self, vars_to_starts) -> Any:
    """The factor type's ``compile_wiring`` on this group.

    Every parameter of the factor type's ``compile_wiring`` other than
    ``vars_to_starts`` is taken from the group's attribute of that name. The
    package's own factor types take ``variables_for_factors`` as the group's
    ``FactorEdges`` table; other factor types get variable lists.
    """
    kwargs = {}
    own = getattr(self.factor_type, "accepts_factor_edges", False)
    for name in _wiring_params(self.factor_type):
      if name == "variables_for_factors":
        kwargs[name] = self.factor_edges if own else self.factor_edges.rows()
      else:
        kwargs[name] = self._wiring_attr(name)
    return self.factor_type.compile_wiring(vars_to_starts=vars_to_starts, **kwargs)


_SERIALS = itertools.count()
_PARAMS: Dict[Any, Tuple[str, ...]] = {}


def _wiring_params(factor_type) -> Tuple[str, ...]:
  """Parameter names of ``factor_type.compile_wiring`` except vars_to_starts."""
  names = _PARAMS.get(factor_type)
  if names is None:
    sig = inspect.signature(factor_type.compile_wiring)
    names = tuple(n for n in sig.parameters if n != "vars_to_starts")
    _PARAMS[factor_type] = names
  return names


@dataclasses.dataclass(frozen=True, eq=False)
class SingleFactorGroup(FactorGroup):
  """The group of one factor, as ``FactorGraph.add_factors(factor)`` makes it."""

  single_factor: factor.Factor

  def __post_init__(self):
    super().__post_init__()
    if self.num_factors != 1:
      raise ValueError(
          f"SingleFactorGroup should only contain one factor. Got {self.num_factors}"
      )
    object.__setattr__(self, "factor_type", type(self.single_factor))
    lp = np.asarray(self.single_factor.log_potentials)
    object.__setattr__(self, "log_potentials", lp[None] if lp.shape[0] > 0 else lp)

  def _wiring_attr(self, name: str) -> Any:
    # the factor's own attributes (factor_configs, edge_states_offset, ...)
    # unless the group derives them (sizes, edge states, number of factors)
    if name in ("factor_edges_num_states", "factor_sizes", "num_factors"):
      return getattr(self, name)
    return getattr(self.single_factor, name)

  def _make_factor(self, f: int) -> factor.Factor:
    return self.single_factor

  def flatten(self, data):
    raise NotImplementedError(
        "SingleFactorGroup does not support vectorized factor operations."
    )

  def unflatten(self, flat_data):
    raise NotImplementedError(
        "SingleFactorGroup does not support vectorized factor operations."
    )
