"""Pair constraints: CHARMM LJ, smooth-PME electrostatics and plain Coulomb.

CharmmNonbondedConstraint  mdpy/constraint/charmm_nonbonded_constraint.py:22-62 (API and
                           per-atom [eps, sigma, eps14, sigma14] table), evaluated by the
                           libmdx nonbonded kernel.
ElectrostaticPMEConstraint new (mdpy/constraint/__init__.py:22 names PME; nothing implements it):
                           erfc direct space over the same pair set as the LJ kernel, smooth PME
                           reciprocal space, self / exclusion / background corrections.
ElectrostaticConstraint    mdpy/constraint/electrostatic_constraint.py:23-174: all pairs,
                           minimum image, no cutoff (float64 kernel).
"""
import math

import numpy as np

from .. import _lib as L
from ..unit import check_quantity_value
from .constraint import Constraint


def pme_grid_for(box, spacing=1.0):
    """Smallest 2^a 3^b 5^c >= L/spacing per axis (about one grid point per A)."""
    from ..systems import fft_friendly
    return tuple(fft_friendly(int(math.ceil(L / spacing))) for L in box)


class CharmmNonbondedConstraint(Constraint):
    TERMS = L.TERM_LJ
    ENERGY_NAMES = ("lj",)

    def __init__(self, parameters, cutoff_radius=12, force_id: int = 0, force_group: int = 0) -> None:
        super().__init__(parameters, force_id=force_id, force_group=force_group)
        self._cutoff_radius = float(check_quantity_value(cutoff_radius))
        self._parameters_list = None
        self._num_nonbonded_pairs = 0

    def bind_ensemble(self, ensemble):
        """charmm_nonbonded_constraint.py:48-62: 2-parameter entries repeat (eps, sigma) for 1-4."""
        self._bind_common(ensemble)
        rows = []
        for p in ensemble.topology.particles:
            prm = self._parameters[p.particle_type]
            if len(prm) == 2:
                rows.append([prm[0], prm[1], prm[0], prm[1]])
            else:
                rows.append(list(prm[:4]))
        self._parameters_list = np.asarray(rows, dtype=np.float64)
        ensemble.state.engine.set_lj(self._parameters_list)
        ensemble._configure_pairs()

    @property
    def num_nonbonded_pairs(self):
        return self._num_nonbonded_pairs


class ElectrostaticPMEConstraint(Constraint):
    TERMS = L.TERM_COUL_DIRECT | L.TERM_COUL_RECIP
    ENERGY_NAMES = ("coulomb_direct", "coulomb_reciprocal")

    def __init__(self, cutoff_radius=12, ewald_error=1e-5, grid=None, order=4, force_id: int = 0,
                 force_group: int = 0) -> None:
        super().__init__(None, force_id=force_id, force_group=force_group)
        self._cutoff_radius = float(check_quantity_value(cutoff_radius))
        self._ewald_error = float(ewald_error)
        self._grid = None if grid is None else tuple(int(k) for k in grid)
        self._order = int(order)
        # beta = sqrt(-ln(2 tol))/rc (OpenMM's rule); erfc(beta rc) ~ 1.3e-5 for tol 1e-5
        self._beta = math.sqrt(-math.log(2.0 * self._ewald_error)) / self._cutoff_radius

    def bind_ensemble(self, ensemble):
        self._bind_common(ensemble)
        box = np.diag(ensemble.state.pbc_matrix).astype(np.float64)
        grid = self._grid or pme_grid_for(box)
        self._grid = grid
        ensemble.state.engine.configure_pme(grid, self._beta, self._order)
        ensemble._configure_pairs()

    beta = property(lambda self: self._beta)
    grid = property(lambda self: self._grid)
    order = property(lambda self: self._order)
    ewald_error = property(lambda self: self._ewald_error)


class ElectrostaticConstraint(Constraint):
    TERMS = L.TERM_COUL_PLAIN
    ENERGY_NAMES = ("coulomb_plain",)

    def __init__(self, parameters=None, force_id: int = 0, force_group: int = 0) -> None:
        super().__init__(parameters, force_id=force_id, force_group=force_group)

    def bind_ensemble(self, ensemble):
        self._bind_common(ensemble)
