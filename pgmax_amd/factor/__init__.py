"""Factor types and the factor-to-variable update registry.

``FAC_TO_VAR_UPDATES`` mirrors pgmax/factor/__init__.py:37-44: the factor
types in the order Enum, OR, AND, Pool, each mapped to its update function.
Here every update function launches a hand-written HIP kernel.
"""

import collections
from typing import Callable, OrderedDict, Type

from pgmax_amd.factor import enum
from pgmax_amd.factor import logical
from pgmax_amd.factor import pool
from pgmax_amd.factor.enum import EnumFactor
from pgmax_amd.factor.enum import EnumWiring
from pgmax_amd.factor.factor import concatenate_var_states_for_edges
from pgmax_amd.factor.factor import Factor
from pgmax_amd.factor.factor import Wiring
from pgmax_amd.factor.logical import ANDFactor
from pgmax_amd.factor.logical import LogicalFactor
from pgmax_amd.factor.logical import LogicalWiring
from pgmax_amd.factor.logical import ORFactor
from pgmax_amd.factor.pool import PoolFactor
from pgmax_amd.factor.pool import PoolWiring

FAC_TO_VAR_UPDATES: OrderedDict[Type[Factor], Callable] = collections.OrderedDict([
    (EnumFactor, enum.pass_enum_fac_to_var_messages),
    (ORFactor, logical.pass_logical_fac_to_var_messages),
    (ANDFactor, logical.pass_logical_fac_to_var_messages),
    (PoolFactor, pool.pass_pool_fac_to_var_messages),
])
