"""Variable groups (mirror of pgmax/vgroup/__init__.py)."""

from pgmax_amd.vgroup.varray import NDVarArray
from pgmax_amd.vgroup.vdict import VarDict
from pgmax_amd.vgroup.vgroup import VarGroup
