"""Simulation driver and a minimal dumper (mdpy/simulation.py:20-90, mdpy/dumper/).

Simulation.integrate runs the integrator in chunks of gcd(dump frequencies) and calls each
dumper whose frequency divides the current step, like the reference. Dumping reads energies
and positions back from the device, so it is the only host synchronisation in a run.
"""
import math
import time

import numpy as np

from .error import DumperPoorDefinedError
from .unit import KB, ns_per_day


class Dumper:
    """mdpy/dumper/dumper.py:13-46: dump(simulation) every dump_frequency steps."""

    def __init__(self, file_path=None, dump_frequency: int = 1) -> None:
        self._file_path = file_path
        self._dump_frequency = int(dump_frequency)

    def dump(self, simulation):
        raise NotImplementedError

    dump_frequency = property(lambda self: self._dump_frequency)
    file_path = property(lambda self: self._file_path)


class LogDumper(Dumper):
    """Energy/temperature/ns-per-day log (the columns of mdpy/dumper/log_dumper.py:75-208),
    kept in memory (``records``) and appended to ``file_path`` when given."""

    def __init__(self, file_path=None, dump_frequency: int = 1) -> None:
        super().__init__(file_path, dump_frequency)
        self.records = []
        self._t0 = None
        self._step0 = 0

    def dump(self, simulation):
        ens = simulation.ensemble
        now = time.perf_counter()
        if self._t0 is None:
            self._t0, self._step0 = now, simulation.cur_step
            speed = float("nan")
        else:
            steps = simulation.cur_step - self._step0
            speed = ns_per_day(simulation.integrator.time_step, (now - self._t0) / steps) if steps else float("nan")
        n = ens.topology.num_particles
        rec = {"step": simulation.cur_step, "potential_energy": float(ens.potential_energy),
               "kinetic_energy": float(ens.kinetic_energy), "total_energy": float(ens.total_energy),
               "temperature": 2 * float(ens.kinetic_energy) / (3 * n * KB), "ns_per_day": speed}
        self.records.append(rec)
        if self._file_path:
            with open(self._file_path, "a") as f:
                f.write(" ".join("%s=%.8g" % (k, v) for k, v in rec.items()) + "\n")


class Simulation:
    def __init__(self, ensemble, integrator, minimizer=None) -> None:
        self._ensemble = ensemble
        self._integrator = integrator
        self._minimizer = minimizer
        self._cur_step = 0
        self._dumpers = []
        self._minimum_dump_frequency = 0

    def add_dumpers(self, *dumpers):
        for d in dumpers:
            self._dumpers.append(d)
            self._minimum_dump_frequency = (d.dump_frequency if self._minimum_dump_frequency == 0
                                            else math.gcd(self._minimum_dump_frequency, d.dump_frequency))

    def _dump(self):
        for d in self._dumpers:
            if self._cur_step % d.dump_frequency == 0:
                d.dump(self)

    def dump_initial_state(self):
        self._ensemble.update()
        for d in self._dumpers:
            d.dump(self)

    def reset_simulation_step(self):
        self._cur_step = 0

    def integrate(self, num_steps: int):
        if not self._dumpers or self._minimum_dump_frequency == 0:
            raise DumperPoorDefinedError("No dumper has been added to Simulation yet.")
        target = self._cur_step + num_steps
        while self._cur_step < target:
            self._integrator.integrate(self._ensemble, self._minimum_dump_frequency)
            self._cur_step += self._minimum_dump_frequency
            self._dump()

    ensemble = property(lambda self: self._ensemble)
    integrator = property(lambda self: self._integrator)
    dumpers = property(lambda self: self._dumpers)
    num_dumpers = property(lambda self: len(self._dumpers))
    cur_step = property(lambda self: self._cur_step)
