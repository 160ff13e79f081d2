from .particle import Particle
from .topology import Topology
from .cell_list import CellList
from .state import State
from .trajectory import Trajectory

__all__ = ["Particle", "Topology", "CellList", "State", "Trajectory"]
