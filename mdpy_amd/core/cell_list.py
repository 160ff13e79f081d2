"""CellList (mdpy/core/cell_list.py:19-148), binned on the GPU.

Same grid rule as the reference: n = max(3, floor((L + 2 A)/rc)) cells per axis of width
(L + 2)/n, cell index floor((x - min x)/w) in float64, dense [nx, ny, nz, maxpc] table with
atoms ascending inside a cell and -1 padding. The kernels (libmdx mdx_celllist_reference /
mdx_celllist_table) reproduce the reference's binning bit for bit. The force kernels do not
use this table (they use their own sorted grid and neighbour lists); it is materialised on
demand for callers that read it.
"""
import numpy as np

from ..error import CellListPoorDefinedError, PBCPoorDefinedError, ArrayDimError
from ..unit import check_quantity_value

CELL_LIST_SKIN = 2.0  # cell_list.py:17


def check_pbc_matrix(pbc_matrix):
    """mdpy/utils/pbc.py:15-26 plus the orthorhombic restriction of this build."""
    pbc_matrix = np.asarray(pbc_matrix, dtype=np.float64)
    if pbc_matrix.shape != (3, 3):
        raise ArrayDimError("The pbc matrix should have shape [3, 3], while matrix %s is provided"
                            % (list(pbc_matrix.shape),))
    if np.linalg.det(pbc_matrix) == 0:
        raise PBCPoorDefinedError("PBC is poor defined. Two or more column vectors are linear corellated")
    if np.any(pbc_matrix - np.diag(np.diag(pbc_matrix)) != 0):
        raise PBCPoorDefinedError("only orthorhombic (diagonal) boxes are supported on the device path")
    return pbc_matrix


class CellList:
    def __init__(self, pbc_matrix, cutoff_radius=12, engine=None) -> None:
        self._engine = engine
        self.set_pbc_matrix(pbc_matrix, is_update=False)
        self.set_cutoff_radius(cutoff_radius, is_update=False)
        self._cell_list_skin = CELL_LIST_SKIN
        self._particle_cell_index = None
        self._cell_list = None
        self._num_particles_per_cell = 0
        self._stale = False
        self._update_attributes()

    def __repr__(self) -> str:
        x, y, z = self._num_cells_vec
        return "<mdpy_amd.core.CellList object with %d x %d x %d cells at %x>" % (x, y, z, id(self))

    __str__ = __repr__

    def __getitem__(self, matrix_id):
        """cell_list.py:45-47 (an int gives the atom's cell row; see the reference for tuples)."""
        self._materialise()
        x, y, z = self._particle_cell_index[matrix_id, :]
        return self._cell_list[x, y, z, :]

    def set_pbc_matrix(self, pbc_matrix, is_update=True):
        self._pbc_matrix = check_pbc_matrix(check_quantity_value(pbc_matrix))
        self._pbc_diag = self._pbc_matrix.diagonal().copy()
        if is_update:
            self._update_attributes()

    def set_cutoff_radius(self, cutoff_radius, is_update=True):
        """cell_list.py:56-72: zero cutoff or floor(L/rc) < 2 raises CellListPoorDefinedError."""
        cutoff_radius = float(check_quantity_value(cutoff_radius))
        if cutoff_radius == 0:
            raise CellListPoorDefinedError("Cutoff radius is poor defined, current value %.3f" % cutoff_radius)
        if np.any(np.floor(self._pbc_diag / cutoff_radius) < 2):
            raise CellListPoorDefinedError("The cutoff_radius is too large to create cell list")
        self._cutoff_radius = cutoff_radius
        if is_update:
            self._update_attributes()

    def _update_attributes(self):
        n = np.floor((self._pbc_diag + CELL_LIST_SKIN) / (np.ones(3) * self._cutoff_radius)).astype(np.int64)
        n[n < 3] = 3
        self._num_cells_vec = n
        self._num_cells = int(np.prod(n))
        self._cell_matrix = np.diag((self._pbc_diag + CELL_LIST_SKIN) / n)
        self._cell_inv = np.linalg.inv(self._cell_matrix)

    def bind_engine(self, engine):
        self._engine = engine

    def update(self, positions=None):
        """Bin `positions` (or, when bound to a State, its current device positions)."""
        if positions is not None:
            positions = np.asarray(positions, dtype=np.float64)
            if self._engine is None or self._engine.n != positions.shape[0]:
                from ..engine import Engine
                self._engine = Engine.bare(positions.shape[0], self._pbc_diag)
            self._engine.set_positions(positions, wrap=False)
        self._stale = True
        self._materialise()

    def mark_stale(self):
        self._stale = True
        self._particle_cell_index = None
        self._cell_list = None

    def _materialise(self):
        if not self._stale:
            return
        if self._engine is None:
            raise CellListPoorDefinedError("cell list has no positions yet")
        idx, table, dims = self._engine.celllist_reference(self._cutoff_radius)
        if not np.array_equal(dims, self._num_cells_vec):
            raise CellListPoorDefinedError("device grid %s differs from %s" % (dims, self._num_cells_vec))
        self._particle_cell_index, self._cell_list = idx, table
        self._num_particles_per_cell = table.shape[3]
        self._stale = False

    cutoff_radius = property(lambda self: self._cutoff_radius)
    pbc_matrix = property(lambda self: self._pbc_matrix)
    cell_matrix = property(lambda self: self._cell_matrix)
    cell_inv = property(lambda self: self._cell_inv)
    num_cell_vec = property(lambda self: self._num_cells_vec)

    @property
    def particle_cell_index(self):
        self._materialise()
        return self._particle_cell_index

    @property
    def cell_list(self):
        self._materialise()
        return self._cell_list

    @property
    def num_particles_per_cell(self):
        self._materialise()
        return self._num_particles_per_cell
