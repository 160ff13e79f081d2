"""CHARMM bonded constraints (float64 device kernels).

Parameter lookup follows each reference bind_ensemble: the key is the '-'-joined particle
types of the term (charmm_bond_constraint.py:32-47, charmm_angle_constraint.py:32-50,
charmm_dihedral_constraint.py:33-56 — one row per cosine term, charmm_improper_constraint.py:32-50).
"""
import numpy as np

from .. import _lib as L
from .constraint import Constraint


def _key(topology, ids):
    return "-".join(topology.particles[i].particle_type for i in ids)


class _Bonded(Constraint):
    WIDTH = 2
    NPAR = 2

    def __init__(self, parameters, force_id: int = 0, force_group: int = 0) -> None:
        super().__init__(parameters, force_id=force_id, force_group=force_group)
        self._int_parameters = np.zeros((0, self.WIDTH), dtype=np.int64)
        self._float_parameters = np.zeros((0, self.NPAR))
        self._count = 0

    def _terms(self, topology):
        raise NotImplementedError

    def _collect(self, topology):
        ints, floats = [], []
        for ids in self._terms(topology):
            ints.append([topology.particles[i].matrix_id for i in ids])
            floats.append(self._parameters[_key(topology, ids)])
        self._count = len(ints)
        self._int_parameters = np.array(ints, dtype=np.int64).reshape(-1, self.WIDTH)
        self._float_parameters = np.array(floats, dtype=np.float64).reshape(-1, self.NPAR)


class CharmmBondConstraint(_Bonded):
    TERMS = L.TERM_BOND
    ENERGY_NAMES = ("bond",)

    def _terms(self, topology):
        return topology.bonds

    def bind_ensemble(self, ensemble):
        self._bind_common(ensemble)
        self._collect(ensemble.topology)
        ensemble.state.engine.set_bonds(self._int_parameters, self._float_parameters)

    num_bonds = property(lambda self: self._count)


class CharmmAngleConstraint(_Bonded):
    TERMS = L.TERM_ANGLE
    ENERGY_NAMES = ("angle",)
    WIDTH, NPAR = 3, 4

    def _terms(self, topology):
        return topology.angles

    def bind_ensemble(self, ensemble):
        self._bind_common(ensemble)
        self._collect(ensemble.topology)
        ensemble.state.engine.set_angles(self._int_parameters, self._float_parameters)

    num_angles = property(lambda self: self._count)


class CharmmDihedralConstraint(_Bonded):
    """refcompat=True keeps the reference's force_val = -k(1 - n sin(n phi - delta))
    (charmm_dihedral_constraint.py:81); the default uses the energy gradient."""
    TERMS = L.TERM_DIHEDRAL
    ENERGY_NAMES = ("dihedral",)
    WIDTH, NPAR = 4, 3

    def __init__(self, parameters, force_id: int = 0, force_group: int = 0, refcompat: bool = False) -> None:
        super().__init__(parameters, force_id=force_id, force_group=force_group)
        self._refcompat = bool(refcompat)
        self._num_dihedrals = 0

    def bind_ensemble(self, ensemble):
        self._bind_common(ensemble)
        top = ensemble.topology
        ints, floats = [], []
        for ids in top.dihedrals:
            for row in self._parameters[_key(top, ids)]:
                ints.append([top.particles[i].matrix_id for i in ids])
                floats.append(row)
        self._num_dihedrals = len(top.dihedrals)
        self._int_parameters = np.array(ints, dtype=np.int64).reshape(-1, 4)
        self._float_parameters = np.array(floats, dtype=np.float64).reshape(-1, 3)
        ensemble.state.engine.set_dihedrals(self._int_parameters, self._float_parameters, self._refcompat)

    num_dihedrals = property(lambda self: self._num_dihedrals)


class CharmmImproperConstraint(_Bonded):
    TERMS = L.TERM_IMPROPER
    ENERGY_NAMES = ("improper",)
    WIDTH, NPAR = 4, 2

    def _terms(self, topology):
        return topology.impropers

    def bind_ensemble(self, ensemble):
        self._bind_common(ensemble)
        self._collect(ensemble.topology)
        ensemble.state.engine.set_impropers(self._int_parameters, self._float_parameters)

    num_impropers = property(lambda self: self._count)
