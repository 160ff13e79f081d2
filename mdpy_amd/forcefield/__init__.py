"""Force fields (mdpy/forcefield/__init__.py)."""
from .forcefield import Forcefield
from .charmm_forcefield import CharmmForcefield

__all__ = ["Forcefield", "CharmmForcefield"]
