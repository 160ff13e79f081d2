"""Topology (mdpy/core/topology.py:16-301): particles, bonded terms and the exclusion tables.

add_bond excludes 1-2 partners, add_angle excludes the (end, end) 1-3 pair and add_dihedral
records (end, end) 1-4 partners (topology.py:122-200). join() builds the per-atom arrays the
reference exposes (masses/charges (N,1), -1 padded bonded/scaling tables) and, for the device
path, CSR tables. The reference sizes the 1-4 table by the largest bonded count
(topology.py:74-76); here it is sized by the largest 1-4 count.
"""
import numpy as np

from ..environment import env
from ..error import GeomtryDimError, ModifyJoinedTopologyError, ParticleConflictError
from .particle import Particle


class Topology:
    def __init__(self) -> None:
        self._particles = []
        self._bonds, self._angles, self._dihedrals, self._impropers = [], [], [], []
        self._is_joined = False
        self._masses = self._charges = None
        self._bonded_particles = self._scaling_particles = None

    def __repr__(self) -> str:
        return "<mdpy_amd.core.Topology object: %d particles at %x>" % (self.num_particles, id(self))

    def __str__(self) -> str:
        return "Topology: %d particles, %d bonds, %d angles, %d dihedrals, %d impropers" % (
            self.num_particles, self.num_bonds, self.num_angles, self.num_dihedrals, self.num_impropers)

    def _check_matrix_ids(self, *ids):
        for k, i in enumerate(ids):
            if i >= self.num_particles:
                raise ParticleConflictError("atom index %d is outside this topology's %d particles" % (i, self.num_particles))
            if i in ids[k + 1:]:
                raise ParticleConflictError("the same atom is listed more than once in one bonded term")

    def _check_joined(self):
        if self._is_joined:
            raise ModifyJoinedTopologyError("the topology is already joined and can no longer be modified (%s)" % self)

    def add_particles(self, particles):
        self._check_joined()
        for p in particles:
            if not isinstance(p, Particle):
                raise TypeError("expected a mdpy_amd.core.Particle, got %s" % type(p))
            p.change_matrix_id(len(self._particles))
            self._particles.append(p)

    def add_bond(self, bond):
        self._check_joined()
        if len(bond) != 2:
            raise GeomtryDimError("a bond joins 2 atoms; %d indices were given" % len(bond))
        p1, p2 = bond
        self._check_matrix_ids(p1, p2)
        self._bonds.append(list(bond))
        self._particles[p1].add_bonded_particle(p2)
        self._particles[p2].add_bonded_particle(p1)

    def add_angle(self, angle):
        self._check_joined()
        if len(angle) != 3:
            raise GeomtryDimError("an angle spans 3 atoms; %d indices were given" % len(angle))
        p1, p2, p3 = angle
        self._check_matrix_ids(p1, p2, p3)
        self._angles.append(list(angle))
        self._particles[p1].add_bonded_particle(p3)
        self._particles[p3].add_bonded_particle(p1)

    def add_dihedral(self, dihedral, scaling_factor=1):
        self._check_joined()
        if len(dihedral) != 4:
            raise GeomtryDimError("a dihedral spans 4 atoms; %d indices were given" % len(dihedral))
        p1, p2, p3, p4 = dihedral
        self._check_matrix_ids(p1, p2, p3, p4)
        self._dihedrals.append(list(dihedral))
        self._particles[p1].add_scaling_particle(p4, scaling_factor)
        self._particles[p4].add_scaling_particle(p1, scaling_factor)

    def del_particles(self, particles):
        """Remove the given particles (matched by identity) and every bond, angle, dihedral and
        improper that names one of them (topology.py:101-120). The remaining particles keep
        their matrix ids, as the remaining terms refer to them."""
        self._check_joined()
        gone = [p for p in particles if any(p is q for q in self._particles)]
        ids = {p.matrix_id for p in gone}
        self._particles = [q for q in self._particles if not any(q is p for p in gone)]
        self._bonds = [t for t in self._bonds if not ids.intersection(t)]
        self._angles = [t for t in self._angles if not ids.intersection(t)]
        self._dihedrals = [t for t in self._dihedrals if not ids.intersection(t)]
        self._impropers = [t for t in self._impropers if not ids.intersection(t)]

    def _remove_term(self, terms, term, reversible=True):
        """Remove `term` (or, when reversible, its reversed order) from `terms`; True if found."""
        for candidate in ([list(term), list(term)[::-1]] if reversible else [list(term)]):
            if candidate in terms:
                terms.remove(candidate)
                return True
        return False

    def del_bond(self, bond):
        """topology.py:136-153: the bond in either order; its ends leave each other's lists."""
        self._check_joined()
        if len(bond) != 2:
            raise GeomtryDimError("a bond joins 2 atoms; %d indices were given" % len(bond))
        p1, p2 = bond
        self._check_matrix_ids(p1, p2)
        if self._remove_term(self._bonds, bond):
            self._particles[p1].del_bonded_particle(p2)
            self._particles[p2].del_bonded_particle(p1)

    def del_angle(self, angle):
        """topology.py:169-186: the angle as given or reversed."""
        self._check_joined()
        if len(angle) != 3:
            raise GeomtryDimError("an angle spans 3 atoms; %d indices were given" % len(angle))
        p1, p2, p3 = angle
        self._check_matrix_ids(p1, p2, p3)
        if self._remove_term(self._angles, angle):
            self._particles[p1].del_bonded_particle(p3)
            self._particles[p3].del_bonded_particle(p1)

    def del_dihedral(self, dihedral):
        """topology.py:202-219: the dihedral as given or reversed; its ends lose their 1-4 link."""
        self._check_joined()
        if len(dihedral) != 4:
            raise GeomtryDimError("a dihedral spans 4 atoms; %d indices were given" % len(dihedral))
        p1, p2, p3, p4 = dihedral
        self._check_matrix_ids(p1, p2, p3, p4)
        if self._remove_term(self._dihedrals, dihedral):
            self._particles[p1].del_scaling_particle(p4)
            self._particles[p4].del_scaling_particle(p1)

    def del_improper(self, improper):
        """topology.py:232-241: the improper exactly as given."""
        self._check_joined()
        if len(improper) != 4:
            raise GeomtryDimError("an improper spans 4 atoms; %d indices were given" % len(improper))
        self._check_matrix_ids(*improper)
        self._remove_term(self._impropers, improper, reversible=False)

    def add_improper(self, improper):
        self._check_joined()
        if len(improper) != 4:
            raise GeomtryDimError("an improper spans 4 atoms; %d indices were given" % len(improper))
        self._check_matrix_ids(*improper)
        self._impropers.append(list(improper))

    def join(self):
        n = self.num_particles
        self._masses = np.array([[p.mass if p.mass is not None else 0.0] for p in self._particles],
                                dtype=env.NUMPY_FLOAT).reshape(n, 1)
        self._charges = np.array([[p.charge if p.charge is not None else 0.0] for p in self._particles],
                                 dtype=env.NUMPY_FLOAT).reshape(n, 1)
        wb = max([p.num_bonded_particles for p in self._particles] + [0])
        ws = max([p.num_scaling_particles for p in self._particles] + [0])
        self._bonded_particles = -np.ones((n, wb), dtype=env.NUMPY_INT)
        self._scaling_particles = -np.ones((n, ws), dtype=env.NUMPY_INT)
        for i, p in enumerate(self._particles):
            self._bonded_particles[i, :p.num_bonded_particles] = p.bonded_particles
            self._scaling_particles[i, :p.num_scaling_particles] = p.scaling_particles
        self._is_joined = True

    def split(self):
        self._masses = self._charges = self._bonded_particles = self._scaling_particles = None
        self._is_joined = False

    @staticmethod
    def _csr(rows):
        ptr = np.zeros(len(rows) + 1, dtype=np.int64)
        for i, r in enumerate(rows):
            ptr[i + 1] = ptr[i] + len(r)
        idx = np.array([j for r in rows for j in sorted(r)], dtype=np.int64)
        return ptr, idx

    def exclusion_csr(self):
        """1-2 + 1-3 partners as CSR (ptr, idx), rows sorted."""
        return self._csr([p.bonded_particles for p in self._particles])

    def scaling_csr(self):
        """1-4 partners as CSR (ptr, idx), rows sorted."""
        return self._csr([p.scaling_particles for p in self._particles])

    @classmethod
    def from_system(cls, system, name_prefix=""):
        """Build a Topology from a flat System with the reference's add_* rules."""
        t = cls()
        t.add_particles([Particle(particle_id=i, particle_type=system.types[i], particle_name=system.types[i],
                                  mass=float(system.masses[i]), charge=float(system.charges[i]))
                         for i in range(system.num_atoms)])
        for b in system.bonds:
            t.add_bond([int(b[0]), int(b[1])])
        for a in system.angles:
            t.add_angle([int(a[0]), int(a[1]), int(a[2])])
        seen = set()
        for d in system.dihedrals:  # one row per cosine term in the System; one dihedral here
            key = tuple(int(v) for v in d)
            if key not in seen:
                seen.add(key)
                t.add_dihedral(list(key))
        for d in system.impropers:
            t.add_improper([int(v) for v in d])
        return t

    masses = property(lambda self: self._masses)
    charges = property(lambda self: self._charges)
    bonded_particles = property(lambda self: self._bonded_particles)
    scaling_particles = property(lambda self: self._scaling_particles)
    particles = property(lambda self: self._particles)
    num_particles = property(lambda self: len(self._particles))
    bonds = property(lambda self: self._bonds)
    num_bonds = property(lambda self: len(self._bonds))
    angles = property(lambda self: self._angles)
    num_angles = property(lambda self: len(self._angles))
    dihedrals = property(lambda self: self._dihedrals)
    num_dihedrals = property(lambda self: len(self._dihedrals))
    impropers = property(lambda self: self._impropers)
    num_impropers = property(lambda self: len(self._impropers))
    is_joined = property(lambda self: self._is_joined)
