"""Factor graph (mirror of pgmax/fgraph/__init__.py)."""

from pgmax_amd.fgraph.fgraph import FactorGraph
from pgmax_amd.fgraph.fgraph import FactorGraphState
