"""Ensemble (mdpy/ensemble.py:15-103): topology, state and constraints; sums forces/energies.

update() evaluates every constraint's terms in one fused device call (one neighbour-list
pass shared by LJ and PME direct space); per-constraint forces are produced when a caller
asks for them (Constraint.update), at the same positions. Kinetic energy is computed from the
state's velocities, as the reference does (ensemble.py:63-71).
"""
import numpy as np

from . import _lib as L
from .core import State, Topology
from .error import ConstraintConflictError
from .environment import env


class Ensemble:
    def __init__(self, topology: Topology, pbc_matrix, device=0, pair_set="exact", skin=1.5) -> None:
        if not topology.is_joined:
            topology.join()
        self._topology = topology
        self._state = State(topology, pbc_matrix, device=device)
        self._matrix_shape = self._state.matrix_shape
        self._forces = np.zeros(self._matrix_shape)
        self._total_energy = 0
        self._potential_energy = 0
        self._kinetic_energy = 0
        self._constraints = []
        self._pair_set = pair_set
        self._skin = float(skin)
        self._last_energies = None
        ep, ei = topology.exclusion_csr()
        sp, si = topology.scaling_csr()
        eng = self._state.engine
        eng.set_exclusions(ep, ei)
        eng.set_scaling(sp, si)

    def __repr__(self) -> str:
        return "<mdpy_amd.Ensemble object: %d constraints at %x>" % (len(self._constraints), id(self))

    __str__ = __repr__

    def add_constraints(self, *constraints):
        """ensemble.py:40-51: duplicates raise; the cell-list cutoff grows to the largest one."""
        for constraint in constraints:
            if constraint in self._constraints:
                raise ConstraintConflictError("%s has added twice to %s" % (constraint, self))
            self._constraints.append(constraint)
            constraint.bind_ensemble(self)
            if constraint.cutoff_radius > self.state.cell_list.cutoff_radius:
                self.state.cell_list.set_cutoff_radius(constraint.cutoff_radius)
        self._configure_pairs()

    def _configure_pairs(self):
        """LJ and PME direct space share one neighbour list, so their cutoffs must agree."""
        pair = [c for c in self._constraints if c.TERMS & (L.TERM_LJ | L.TERM_COUL_DIRECT)]
        if not pair:
            return
        cuts = {float(c.cutoff_radius) for c in pair}
        if len(cuts) != 1:
            raise ValueError("LJ and PME direct-space cutoffs must be equal on the device path (got %s)"
                             % sorted(cuts))
        self._state.engine.configure_nonbonded(cuts.pop(), skin=self._skin, pair_mode=self._pair_set,
                                               cell_cutoff=self._state.cell_list.cutoff_radius)

    @property
    def term_bits(self):
        bits = 0
        for c in self._constraints:
            bits |= c.TERMS
        return bits

    def _evaluate(self, terms):
        """Forces (N,3, environment precision) and per-term energies of `terms` (device call)."""
        F, en = self._state.engine.compute(terms)
        return F.cpu().numpy().astype(env.NUMPY_FLOAT), en

    def update(self):
        """ensemble.py:53-61 (fused: one device call for every constraint's terms)."""
        F, en = self._state.engine.compute(self.term_bits)
        self._forces = F.cpu().numpy().astype(np.float64)
        self._last_energies = en
        self._potential_energy = 0
        for c in self._constraints:
            c._potential_energy = env.NUMPY_FLOAT(sum(en[k] for k in c.ENERGY_NAMES))
            c._forces = None  # per-term forces: Constraint.update()
            self._potential_energy += c._potential_energy
        self._kinetic_energy = en["kinetic"]
        self._total_energy = self._potential_energy + self._kinetic_energy

    def _update_kinetic_energy(self):
        v = np.asarray(self._state.velocities, dtype=np.float64)
        m = np.asarray(self._topology.masses, dtype=np.float64).reshape(-1, 1)
        self._kinetic_energy = float(((v ** 2).sum(1).reshape(-1, 1) * m).sum() / 2)

    topology = property(lambda self: self._topology)
    state = property(lambda self: self._state)
    forces = property(lambda self: self._forces)
    total_energy = property(lambda self: self._total_energy)
    potential_energy = property(lambda self: self._potential_energy)
    kinetic_energy = property(lambda self: self._kinetic_energy)
    constraints = property(lambda self: self._constraints)
    num_constraints = property(lambda self: len(self._constraints))
