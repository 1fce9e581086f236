"""Factor groups (mirror of pgmax/fgroup/__init__.py)."""

from pgmax_amd.fgroup.enum import EnumFactorGroup
from pgmax_amd.fgroup.enum import PairwiseFactorGroup
from pgmax_amd.fgroup.fgroup import FactorEdges
from pgmax_amd.fgroup.fgroup import FactorGroup
from pgmax_amd.fgroup.fgroup import RaggedVariables
from pgmax_amd.fgroup.fgroup import SingleFactorGroup
from pgmax_amd.fgroup.logical import ANDFactorGroup
from pgmax_amd.fgroup.logical import LogicalFactorGroup
from pgmax_amd.fgroup.logical import ORFactorGroup
from pgmax_amd.fgroup.pool import PoolFactorGroup
