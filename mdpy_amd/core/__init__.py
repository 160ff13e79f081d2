from .particle import Particle
from .topology import Topology
from .cell_list import CellList
from .state import State

__all__ = ["Particle", "Topology", "CellList", "State"]
