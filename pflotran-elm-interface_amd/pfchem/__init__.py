"""Host-side package of the MI355X chemistry step (C ABI in include/pfchem.h)."""
from . import abi  # noqa: F401
from .network import ReactionNetwork, SolverOptions, CellState  # noqa: F401
