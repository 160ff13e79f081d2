"""mdpy_amd — the MI355X (gfx950) force-and-integrate step of mdpy, behind mdpy's own API.

Host side in Python (this package, mirroring mdpy's Particle / Topology / State / CellList /
Ensemble / Constraint / Integrator / Simulation); compute in hand-written HIP kernels of
libmdx.so, called through a ctypes C-ABI (include/mdx.h). There is no CPU compute path.
"""
SPATIAL_DIM = 3

from .environment import env  # noqa: E402
from . import unit, error, systems  # noqa: E402,F401


def __getattr__(name):
    # lazy: the device-facing modules import torch only when used
    import importlib
    lazy = {"core": ".core", "constraint": ".constraint", "integrator": ".integrator",
            "engine": ".engine", "io": ".io", "forcefield": ".forcefield",
            "minimizer": ".minimizer", "dumper": ".dumper", "analyser": ".analyser", "utils": ".utils"}
    if name in lazy:
        return importlib.import_module(lazy[name], __name__)
    if name == "Ensemble":
        from .ensemble import Ensemble
        return Ensemble
    if name in ("Simulation", "LogDumper", "Dumper", "PDBDumper"):
        from . import simulation
        return getattr(simulation, name)
    if name == "HDF5Dumper":
        from .dumper import HDF5Dumper
        return HDF5Dumper
    raise AttributeError(name)
