"""State (mdpy/core/state.py:17-108): positions, velocities and box, resident on the GPU.

Positions and velocities live in float64 device tensors owned by the context's Engine.
set_positions wraps them with the reference formula (pbc.py:28-36, ParticleLossError past two
images) on the device; the reference cell list is re-binned lazily, when something reads it.
Host copies are returned in the environment precision, as the reference stores them.
"""
import numpy as np

from ..environment import env
from ..error import ArrayDimError
from ..unit import KB, check_quantity_value
from .cell_list import CellList, check_pbc_matrix

DEFAULT_CELL_CUTOFF = 12.0  # cell_list.py:20


def initial_cell_cutoff(box_diag):
    """The reference's default cell-list cutoff is 12 A, which it rejects for boxes shorter
    than 24 A (cell_list.py:56-69, State creates it unconditionally at state.py:28). Such boxes
    start with half the shortest edge instead; the grid rule n = max(3, floor((L + 2)/rc)) gives
    the same n = 3 grid either way."""
    if np.all(np.floor(np.asarray(box_diag) / DEFAULT_CELL_CUTOFF) >= 2):
        return DEFAULT_CELL_CUTOFF
    return float(np.min(box_diag)) / 2.0


class State:
    def __init__(self, topology, pbc_matrix, device=0) -> None:
        from ..engine import Engine
        self._particles = topology.particles
        self._masses = topology.masses
        self._num_particles = len(self._particles)
        self._matrix_shape = [self._num_particles, 3]
        self.set_pbc_matrix(pbc_matrix, _init=True)
        self._engine = Engine.bare(self._num_particles, np.diag(self._pbc_matrix), device=device)
        self._engine.set_masses(np.asarray(topology.masses, dtype=np.float64).reshape(-1))
        self._engine.set_charges(np.asarray(topology.charges, dtype=np.float64).reshape(-1))
        self._cell_list = CellList(self._pbc_matrix.copy(), initial_cell_cutoff(np.diag(self._pbc_matrix)),
                                   engine=self._engine)
        self._host_pos = None
        self._host_vel = None

    def __repr__(self) -> str:
        return "<mdpy_amd.core.State object with %d particles at %x>" % (self._num_particles, id(self))

    __str__ = __repr__

    def _check_matrix_shape(self, matrix):
        if not isinstance(matrix, np.ndarray):
            raise TypeError("Matrix should be numpy.ndarray, instead of %s" % type(matrix))
        if list(matrix.shape) != self._matrix_shape:
            raise ArrayDimError("The dimension of array should be [%d, %d], while array %s is provided"
                                % (self._matrix_shape[0], self._matrix_shape[1], list(matrix.shape)))

    def set_pbc_matrix(self, pbc_matrix, _init=False):
        pbc = check_pbc_matrix(check_quantity_value(pbc_matrix))
        self._pbc_matrix = np.ascontiguousarray(pbc, dtype=env.NUMPY_FLOAT)
        self._pbc_inv = np.ascontiguousarray(np.linalg.inv(self._pbc_matrix), dtype=env.NUMPY_FLOAT)
        if not _init:
            self._engine.set_box(np.diag(self._pbc_matrix).astype(np.float64))
            self.invalidate()

    def set_positions(self, positions):
        """state.py:56-61: wrap into the primary image and rebuild the cell list."""
        self._check_matrix_shape(positions)
        self._engine.set_positions(np.asarray(positions, dtype=env.NUMPY_FLOAT).astype(np.float64), wrap=True)
        self.invalidate()

    def set_velocities(self, velocities):
        self._check_matrix_shape(velocities)
        self._engine.set_velocities(np.asarray(velocities, dtype=env.NUMPY_FLOAT).astype(np.float64))
        self._host_vel = None

    def set_velocities_to_temperature(self, temperature):
        """state.py:67-80: uniform velocities of width sqrt(3 kB T / m) per component
        (numpy's global RNG, as the reference)."""
        temperature = float(check_quantity_value(temperature))
        factor = 3 * KB * temperature
        m = np.asarray(self._masses, dtype=np.float64).reshape(-1)
        v = np.random.rand(self._num_particles, 3)
        width = np.sqrt(factor / m)[:, None]
        self.set_velocities((v * 2 * width - width).astype(env.NUMPY_FLOAT))

    def invalidate(self):
        """Device positions changed: drop host copies and the materialised cell table."""
        self._host_pos = None
        self._host_vel = None
        self._cell_list.mark_stale()

    @property
    def engine(self):
        return self._engine

    @property
    def positions(self):
        if self._host_pos is None:
            self._host_pos = self._engine.positions().astype(env.NUMPY_FLOAT)
        return self._host_pos

    @property
    def velocities(self):
        if self._host_vel is None:
            self._host_vel = self._engine.velocities().astype(env.NUMPY_FLOAT)
        return self._host_vel

    matrix_shape = property(lambda self: self._matrix_shape)
    pbc_matrix = property(lambda self: self._pbc_matrix)
    pbc_inv = property(lambda self: self._pbc_inv)
    pbc_info = property(lambda self: (self._pbc_matrix, self._pbc_inv))
    cell_list = property(lambda self: self._cell_list)
