"""Force-term plugins (mdpy/constraint/__init__.py)."""
import numpy as np

from .constraint import Constraint
from .nonbonded import CharmmNonbondedConstraint, ElectrostaticConstraint, ElectrostaticPMEConstraint
from .bonded import (CharmmAngleConstraint, CharmmBondConstraint, CharmmDihedralConstraint,
                     CharmmImproperConstraint)

NUM_NEIGHBOR_CELLS = 27
# 27 offsets over {-1, 0, 1}^3, i-major then j then k (mdpy/constraint/__init__.py:12-19)
NEIGHBOR_CELL_TEMPLATE = np.array([[i, j, k] for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)],
                                  dtype=np.int64)
LONG_RANGE_SOLVER = ["PME"]

__all__ = ["Constraint", "CharmmNonbondedConstraint", "ElectrostaticConstraint", "ElectrostaticPMEConstraint",
           "CharmmBondConstraint", "CharmmAngleConstraint", "CharmmDihedralConstraint", "CharmmImproperConstraint",
           "NEIGHBOR_CELL_TEMPLATE", "NUM_NEIGHBOR_CELLS"]
