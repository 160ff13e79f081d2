"""Constraint plugin base (mdpy/constraint/constraint.py:16-73).

A constraint contributes one or more force terms (bits of the libmdx term mask) and the
parameters those terms need; bind_ensemble uploads the parameters to the ensemble's device
context. update() evaluates only this constraint's terms on the GPU and leaves the result in
``forces`` ((N,3) in the environment precision) and ``potential_energy``, like the reference.
"""
from ..environment import env
from ..error import NonBoundedError
from ..unit import check_quantity_value


class Constraint:
    TERMS = 0            # libmdx term bits this constraint evaluates
    ENERGY_NAMES = ()    # keys of Engine.energies() that make up its potential energy

    def __init__(self, parameters, force_id: int = 0, force_group: int = 0) -> None:
        self._parameters = parameters
        self._force_id = force_id
        self._force_group = force_group
        self._parent_ensemble = None
        self._forces = None
        self._potential_energy = None
        self._cutoff_radius = env.NUMPY_FLOAT(0)

    def __repr__(self) -> str:
        return "<mdpy_amd.constraint.%s object>" % type(self).__name__

    __str__ = __repr__

    def __eq__(self, o) -> bool:
        return id(self) == id(o)

    __hash__ = object.__hash__

    def bind_ensemble(self, ensemble):
        raise NotImplementedError("%s must implement bind_ensemble()" % type(self).__name__)

    def _bind_common(self, ensemble):
        self._parent_ensemble = ensemble
        self._force_id = ensemble.constraints.index(self)

    def _check_bound_state(self):
        if self._parent_ensemble is None:
            raise NonBoundedError("%s is not bound to an Ensemble yet (add it with Ensemble.add_constraints)" % self)

    def update(self):
        """Evaluate this constraint's terms at the ensemble's current positions (on the GPU)."""
        self._check_bound_state()
        self._forces, energies = self._parent_ensemble._evaluate(self.TERMS)
        self._potential_energy = env.NUMPY_FLOAT(sum(energies[k] for k in self.ENERGY_NAMES))

    @property
    def force_id(self):
        return self._force_id

    @property
    def force_group(self):
        return self._force_group

    @property
    def parent_ensemble(self):
        return self._parent_ensemble

    @property
    def forces(self):
        # after a fused Ensemble.update() the per-term forces are evaluated on first access
        if self._forces is None and self._parent_ensemble is not None and self._potential_energy is not None:
            self.update()
        return self._forces

    @property
    def potential_energy(self):
        return self._potential_energy

    def set_cutoff_radius(self, val):
        self._cutoff_radius = check_quantity_value(val)

    @property
    def cutoff_radius(self):
        return self._cutoff_radius
