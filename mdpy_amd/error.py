"""Exception types of the force-and-integrate path.

Names and meaning follow mdpy/error.py (the reference raises these from the same
call sites; each class below cites the reference class it stands in for).
"""


class EnvironmentVariableError(Exception):
    """mdpy/error.py:10 — unsupported precision/platform request."""


class UnitDimensionDismatchedError(Exception):
    """mdpy/error.py:19 — a Quantity with the wrong dimension was passed."""


class GeomtryDimError(Exception):
    """mdpy/error.py:28 — bond/angle/dihedral with the wrong number of atoms."""


class ArrayDimError(Exception):
    """mdpy/error.py:37 — an array of the wrong shape was passed."""


class ParticleConflictError(Exception):
    """mdpy/error.py:52 — duplicate or self reference in a topology connection."""


class ConstraintConflictError(Exception):
    """mdpy/error.py:69 — the same constraint was added twice to an ensemble."""


class ModifyJoinedTopologyError(Exception):
    """mdpy/error.py:78 — a joined topology was modified."""


class NonBoundedError(Exception):
    """mdpy/error.py:87 — a constraint was used before bind_ensemble."""


class PBCPoorDefinedError(Exception):
    """mdpy/error.py:113 — singular periodic box matrix."""


class CellListPoorDefinedError(Exception):
    """mdpy/error.py:123 — cutoff too large for the box, or a cell overflowed."""


class DumperPoorDefinedError(Exception):
    """mdpy/error.py:171 — Simulation.integrate without a dumper."""


class ParticleLossError(Exception):
    """mdpy/error.py:203 — an atom moved two or more box images in one update."""


class DeviceError(RuntimeError):
    """A HIP call or kernel launch inside libmdx failed (no reference counterpart)."""


class NumericalInstabilityError(RuntimeError):
    """Non-finite positions or forces were detected on the device."""


class FileFormatError(Exception):
    """mdpy/error.py:96 — a file whose name or content is not of the expected format."""


class TrajectoryPoorDefinedError(Exception):
    """mdpy/error.py:133 — inconsistent trajectory frames."""


class ParameterPoorDefinedError(Exception):
    """mdpy/error.py:142 — a force-field parameter the topology needs is missing."""


class HDF5FilePoorDefinedError(Exception):
    """mdpy/error.py:151 — an HDF5 file that does not follow the mdpy layout."""


class ParserPoorDefinedError(Exception):
    """mdpy/error.py:160 — a parser asked for data it was told not to read."""


class AnalyserPoorDefinedError(Exception):
    """mdpy/error.py:193 — an analyser given inconsistent selections or options."""


class SelectionConditionPoorDefinedError(Exception):
    """mdpy/error.py:184 — a selection condition with an unknown keyword or shape."""
