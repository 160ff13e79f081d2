"""CHARMM force field (mdpy/forcefield/charmm_forcefield.py:20-111): parameters from CHARMM
files, checked against a topology, turned into an Ensemble with one constraint per term kind.

Electrostatics: the reference adds ElectrostaticConstraint (all pairs, no cutoff) and leaves its
`long_range_solver` / `ewald_error` arguments unused. Here long_range_solver 'PME' or 'PPPM'
(the default, a mesh solver as the reference names it) gives ElectrostaticPMEConstraint: erfc
direct space over the LJ pair set within the cutoff plus smooth-PME reciprocal space at the
requested Ewald error; long_range_solver None (or 'NONE' / 'PLAIN') gives the reference's
all-pairs ElectrostaticConstraint. Either way the ensemble holds six constraints for a topology
with every term kind.
"""
import numpy as np

from ..constraint import (CharmmAngleConstraint, CharmmBondConstraint, CharmmDihedralConstraint,
                          CharmmImproperConstraint, CharmmNonbondedConstraint, ElectrostaticConstraint,
                          ElectrostaticPMEConstraint)
from ..core import Topology
from ..core.cell_list import check_pbc_matrix
from ..error import ParameterPoorDefinedError
from ..io import CharmmTopparParser
from ..unit import check_quantity_value
from .forcefield import Forcefield

MESH_SOLVERS = ("PME", "PPPM")
PLAIN_SOLVERS = (None, "NONE", "PLAIN")


class CharmmForcefield(Forcefield):
    def __init__(self, topology: Topology, pbc_matrix, cutoff_radius=12, long_range_solver="PPPM",
                 ewald_error=1e-6, is_SHAKE: bool = True) -> None:
        super().__init__(topology)
        self._pbc_matrix = check_pbc_matrix(check_quantity_value(pbc_matrix))
        self._cutoff_radius = float(check_quantity_value(cutoff_radius))
        solver = long_range_solver.upper() if isinstance(long_range_solver, str) else long_range_solver
        if solver not in MESH_SOLVERS + PLAIN_SOLVERS:
            raise ValueError("long_range_solver must be one of PME, PPPM or None (plain Coulomb), got %r"
                             % (long_range_solver,))
        self._long_range_solver = solver
        self._ewald_error = float(ewald_error)
        self._is_SHAKE = is_SHAKE  # accepted for the reference's signature; no constraint solver

    def set_param_files(self, *file_paths) -> None:
        self._parameters = CharmmTopparParser(*file_paths).parameters

    @staticmethod
    def _name(topology, ids):
        return "-".join(topology.particles[i].particle_type for i in ids)

    def check_parameters(self):
        """Every particle type and every bonded term's type key must have parameters
        (charmm_forcefield.py:35-94)."""
        t, prm = self._topology, self._parameters
        for p in t.particles:
            if p.particle_type not in prm["nonbonded"]:
                raise ParameterPoorDefinedError("no nonbonded parameters for particle %d (type %s)"
                                                % (p.matrix_id, p.particle_type))
        for kind, terms in (("bond", t.bonds), ("angle", t.angles), ("dihedral", t.dihedrals),
                            ("improper", t.impropers)):
            for ids in terms:
                name = self._name(t, ids)
                if name not in prm[kind]:
                    raise ParameterPoorDefinedError("no %s parameters for %s (types %s)"
                                                    % (kind, "-".join(str(i) for i in ids), name))

    def _electrostatics(self):
        if self._long_range_solver in MESH_SOLVERS:
            return ElectrostaticPMEConstraint(self._cutoff_radius, ewald_error=self._ewald_error)
        return ElectrostaticConstraint()

    def create_ensemble(self):
        """charmm_forcefield.py:96-111, with the electrostatics chosen by long_range_solver."""
        from ..ensemble import Ensemble
        self.check_parameters()
        ensemble = Ensemble(self._topology, self._pbc_matrix)
        t = self._topology
        constraints = []
        if t.num_particles != 0:
            constraints.append(self._electrostatics())
            constraints.append(CharmmNonbondedConstraint(self._parameters["nonbonded"], self._cutoff_radius))
        if t.num_bonds != 0:
            constraints.append(CharmmBondConstraint(self._parameters["bond"]))
        if t.num_angles != 0:
            constraints.append(CharmmAngleConstraint(self._parameters["angle"]))
        if t.num_dihedrals != 0:
            constraints.append(CharmmDihedralConstraint(self._parameters["dihedral"]))
        if t.num_impropers != 0:
            constraints.append(CharmmImproperConstraint(self._parameters["improper"]))
        ensemble.add_constraints(*constraints)
        return ensemble

    cutoff_radius = property(lambda self: self._cutoff_radius)
    long_range_solver = property(lambda self: self._long_range_solver)
    ewald_error = property(lambda self: self._ewald_error)
    pbc_matrix = property(lambda self: self._pbc_matrix.copy())
