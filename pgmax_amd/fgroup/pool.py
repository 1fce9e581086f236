"""Groups of pool factors (pgmax/fgroup/pool.py:27-61)."""

import collections
import dataclasses
from typing import Optional, Type

import numpy as np

from pgmax_amd.factor import factor
from pgmax_amd.factor import pool
from pgmax_amd.fgroup import fgroup


@dataclasses.dataclass(frozen=True, eq=False)
class PoolFactorGroup(fgroup.FactorGroup):
  """pgmax/fgroup/pool.py:27-61."""

  factor_configs: Optional[np.ndarray] = dataclasses.field(init=False, default=None)
  factor_type: Type[factor.Factor] = dataclasses.field(
      init=False, default=pool.PoolFactor
  )

  def _get_variables_to_factors(self):
    return collections.OrderedDict([
        (frozenset(vs), pool.PoolFactor(variables=vs))
        for vs in self.variables_for_factors
    ])
