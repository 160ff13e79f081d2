"""Particle (mdpy/core/particle.py:14-143): identity, mass, charge, bonded and 1-4 partners."""
from ..error import ParticleConflictError
from ..unit import check_quantity_value


class Particle:
    def __init__(self, particle_id=None, particle_type=None, particle_name=None, matrix_id=None,
                 molecule_id=None, molecule_type=None, chain_id=None, mass=None, charge=None) -> None:
        self._particle_id = particle_id
        self._particle_type = particle_type
        self._particle_name = particle_name
        self._matrix_id = matrix_id
        self._molecule_id = molecule_id
        self._molecule_type = molecule_type
        self._chain_id = chain_id
        self._mass = check_quantity_value(mass)
        self._charge = check_quantity_value(charge)
        self._bonded_particles = []
        self._scaling_particles = []
        self._scaling_factors = []

    def __repr__(self) -> str:
        return "<mdpy_amd.core.Particle object: %s-%s at %x>" % (self._particle_name, self._particle_id, id(self))

    __str__ = __repr__

    def __eq__(self, o) -> bool:
        return id(self) == id(o)

    __hash__ = object.__hash__

    def change_matrix_id(self, matrix_id: int):
        self._matrix_id = matrix_id

    def add_bonded_particle(self, particle_matrix_id):
        """particle.py:53-64: duplicates and self references are errors."""
        if particle_matrix_id in self._bonded_particles:
            raise ParticleConflictError("bonded partner %d is already listed for atom %d"
                                        % (particle_matrix_id, self._matrix_id))
        if particle_matrix_id == self._matrix_id:
            raise ParticleConflictError("an atom cannot be bonded to itself")
        self._bonded_particles.append(particle_matrix_id)

    def del_bonded_particle(self, particle_matrix_id):
        if particle_matrix_id in self._bonded_particles:
            self._bonded_particles.remove(particle_matrix_id)

    def add_scaling_particle(self, particle_matrix_id: int, factor):
        """particle.py:71-80: the same 1-4 partner twice is ignored (benzene-like rings)."""
        if particle_matrix_id == self._matrix_id:
            raise ParticleConflictError("an atom cannot be its own 1-4 (scaled) partner")
        if particle_matrix_id not in self._scaling_particles:
            self._scaling_particles.append(particle_matrix_id)
            self._scaling_factors.append(factor)

    def del_scaling_particle(self, particle_matrix_id):
        if particle_matrix_id in self._scaling_particles:
            k = self._scaling_particles.index(particle_matrix_id)
            del self._scaling_particles[k]
            del self._scaling_factors[k]

    particle_id = property(lambda self: self._particle_id)
    particle_type = property(lambda self: self._particle_type)
    particle_name = property(lambda self: self._particle_name)
    matrix_id = property(lambda self: self._matrix_id)
    molecule_id = property(lambda self: self._molecule_id)
    molecule_type = property(lambda self: self._molecule_type)
    chain_id = property(lambda self: self._chain_id)
    mass = property(lambda self: self._mass)
    charge = property(lambda self: self._charge)
    bonded_particles = property(lambda self: self._bonded_particles)
    num_bonded_particles = property(lambda self: len(self._bonded_particles))
    scaling_particles = property(lambda self: self._scaling_particles)
    scaling_factors = property(lambda self: self._scaling_factors)
    num_scaling_particles = property(lambda self: len(self._scaling_particles))
