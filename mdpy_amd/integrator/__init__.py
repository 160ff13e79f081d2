"""Integrators (mdpy/integrator/): every step runs on the GPU as one hipGraph-captured launch
sequence (forces + update), with no host synchronisation inside integrate().

VerletIntegrator(time_step)                         velocity Verlet (NVE);
VerletIntegrator(time_step, refcompat=True)         verlet_integrator.py:20-50 as written
                                                    (position Verlet, +a dt^2 start-up,
                                                    v = mi(x_n - x_{n-1})/(2 dt) on exit);
LangevinIntegrator(dt, T, gamma)                    G-JF Langevin with the reference's a, b, sigma;
LangevinIntegrator(dt, T, gamma, refcompat=True)    langevin_integrator.py:37-72 as written
                                                    (restarts from x_{n-1}, v_{n-1}, subtracts forces).
"""
from .. import _lib as L
from ..unit import check_quantity_value


class Integrator:
    KIND = None

    def __init__(self, time_step) -> None:
        self._time_step = float(check_quantity_value(time_step))
        self._cur_positions = None
        self._pre_positions = None
        self._cached_engine = None
        self.use_graph = True

    def erase_cache(self):
        """integrator.py:17-22: forget history (the next call starts from the state)."""
        self._cur_positions = None
        self._pre_positions = None
        if self._cached_engine is not None:
            self._cached_engine.reset_integrator()
        self._cached_engine = None

    def _configure(self, engine):
        raise NotImplementedError

    def integrate(self, ensemble, num_steps: int = 1):
        self._advance(ensemble, num_steps)
        self._finish(ensemble)

    def _advance(self, ensemble, num_steps, energy_every=None):
        """Enqueue num_steps steps on the device and return without waiting for them; potential
        energies every energy_every steps (default: every step)."""
        engine = ensemble.state.engine
        if self._cached_engine is not engine:
            engine.reset_integrator()
        self._configure(engine)
        every = int(energy_every) if energy_every else 1
        engine.integrate(ensemble.term_bits, int(num_steps), use_graph=self.use_graph, energy_every=every)
        self._cached_engine = engine

    def _finish(self, ensemble):
        """Wait for the device, raise on flags, refresh the ensemble's forces and energies."""
        engine = ensemble.state.engine
        status = engine.status()
        if status & 2:
            from ..error import ParticleLossError
            raise ParticleLossError("Atom(s) moved beyond 2 PBC image.")
        if status & 1:
            from ..error import CellListPoorDefinedError
            raise CellListPoorDefinedError("neighbour-list or cell capacity exceeded during integration")
        self._cached_engine = engine
        self._cur_positions = True
        ensemble.state.invalidate()
        ensemble.update()

    @property
    def time_step(self):
        return self._time_step

    @time_step.setter
    def time_step(self, time_step):
        self._time_step = float(check_quantity_value(time_step))

    @property
    def cur_positions(self):
        return self._cur_positions

    @property
    def pre_positions(self):
        return self._pre_positions

    @property
    def is_cached(self):
        if self._cur_positions is None:
            self.erase_cache()
            return False
        return True


class VerletIntegrator(Integrator):
    def __init__(self, time_step, refcompat: bool = False) -> None:
        super().__init__(time_step)
        self.refcompat = bool(refcompat)

    def _configure(self, engine):
        kind = L.INT_VERLET_REF if self.refcompat else L.INT_VELOCITY_VERLET
        engine.configure_integrator(kind, self._time_step)


class LangevinIntegrator(Integrator):
    def __init__(self, time_step, temperature, friction_rate, seed: int = 0, refcompat: bool = False) -> None:
        super().__init__(time_step)
        self._temperature = float(check_quantity_value(temperature))
        self._gamma = float(check_quantity_value(friction_rate))
        self.seed = int(seed)
        self.refcompat = bool(refcompat)

    def _configure(self, engine):
        kind = L.INT_LANGEVIN_REF if self.refcompat else L.INT_LANGEVIN
        engine.configure_integrator(kind, self._time_step, temperature=self._temperature, friction=self._gamma,
                                    seed=self.seed)

    temperature = property(lambda self: self._temperature)
    friction_rate = property(lambda self: self._gamma)


__all__ = ["Integrator", "VerletIntegrator", "LangevinIntegrator"]
