"""Simulation driver (mdpy/simulation.py:19-90).

Simulation.integrate advances the ensemble in chunks of gcd(dump frequencies) steps, like the
reference, and hands each dumper whose frequency divides the current step a frame of that step.
The chunks are enqueued on the GPU back to back: a frame's positions and energies are copied to
pinned host buffers on the stream right after the steps that produced them, and a writer thread
formats and writes the frame once its copy has landed, so no host synchronisation happens
between chunks (only a bounded number of frames may be in flight). The device state is checked
(particle loss, list overflow) and the ensemble refreshed once, at the end of the call.
"""
import math
import queue
import threading
import time

import numpy as np

from .dumper import Dumper, Frame, HDF5Dumper, LogDumper, PDBDumper  # noqa: F401  (re-exported)
from .error import DumperPoorDefinedError

FRAMES_IN_FLIGHT = 3


class _FrameWriter:
    """Pinned buffers for frames in flight and the thread that writes them in order."""

    def __init__(self, simulation, need_positions):
        import torch
        self.sim = simulation
        ens = simulation.ensemble
        eng = ens.state.engine
        self.torch = torch
        self.eng = eng
        n = ens.topology.num_particles
        from . import _lib as L
        self.nterms = L.NTERMS
        self.pos = [torch.empty((n, 3), dtype=torch.float64, pin_memory=True) if need_positions else None
                    for _ in range(FRAMES_IN_FLIGHT)]
        self.ehist = [torch.empty((2, L.NTERMS + 1), dtype=torch.float64, pin_memory=True)
                      for _ in range(FRAMES_IN_FLIGHT)]
        self.free = queue.Queue()
        for k in range(FRAMES_IN_FLIGHT):
            self.free.put(k)
        self.todo = queue.Queue()
        self.error = None
        self.pbc = np.asarray(ens.state.pbc_matrix)
        self.masses = np.asarray(ens.topology.masses)
        self.thread = threading.Thread(target=self._run, daemon=True)
        self.thread.start()

    def submit(self, step, dumpers):
        slot = self.free.get()  # waits only when FRAMES_IN_FLIGHT frames are still being written
        rows = self.eng.energy_history_async(self.ehist[slot])
        if self.pos[slot] is not None:
            self.pos[slot].copy_(self.eng.pos, non_blocking=True)
        ev = self.torch.cuda.Event()
        ev.record()
        self.todo.put((slot, ev, rows, step, dumpers))

    def _run(self):
        while True:
            item = self.todo.get()
            if item is None:
                return
            slot, ev, rows, step, dumpers = item
            try:
                ev.synchronize()
                wall = time.perf_counter()
                e = self.ehist[slot].numpy()[rows - 1]
                x = self.pos[slot].numpy().copy() if self.pos[slot] is not None else None
                frame = Frame(step, x, float(e[:self.nterms].sum()), float(e[self.nterms]), self.pbc, self.masses,
                              self.sim.integrator.time_step, wall)
                for d in dumpers:
                    d.dump_frame(frame, self.sim)
            except Exception as ex:  # reported by close()
                self.error = ex
            finally:
                self.free.put(slot)

    def close(self):
        self.todo.put(None)
        self.thread.join()
        if self.error is not None:
            raise self.error


class Simulation:
    def __init__(self, ensemble, integrator, minimizer=None) -> None:
        from .minimizer import SteepestDescentMinimizer
        self._ensemble = ensemble
        self._integrator = integrator
        # simulation.py:17: steepest descent with alpha 0.01 unless another minimiser is given
        self._minimizer = minimizer if minimizer is not None else SteepestDescentMinimizer(0.01)
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

    def set_minimizer(self, minimizer):
        self._minimizer = minimizer

    def minimize(self, energy_tolerance, max_iterations):
        """simulation.py:56-58: minimise the ensemble's potential energy, then dump."""
        self._minimizer.minimize(self._ensemble, energy_tolerance, max_iterations)
        self._dump()

    def integrate(self, num_steps: int):
        """simulation.py:60-70, with asynchronous frame output (see the module docstring)."""
        if not self._dumpers or self._minimum_dump_frequency == 0:
            raise DumperPoorDefinedError("No dumper has been added to Simulation yet.")
        chunk = self._minimum_dump_frequency
        need_pos = any(getattr(d, "NEEDS_POSITIONS", False) for d in self._dumpers)
        writer = _FrameWriter(self, need_pos)
        target = self._cur_step + num_steps
        try:
            while self._cur_step < target:
                # energies on the chunk's last step: the frame's row of the energy history
                self._integrator._advance(self._ensemble, chunk, energy_every=chunk)
                self._cur_step += chunk
                due = [d for d in self._dumpers if self._cur_step % d.dump_frequency == 0]
                if due:
                    writer.submit(self._cur_step, due)
        finally:
            writer.close()
        self._integrator._finish(self._ensemble)

    ensemble = property(lambda self: self._ensemble)
    integrator = property(lambda self: self._integrator)
    minimizer = property(lambda self: self._minimizer)
    dumpers = property(lambda self: self._dumpers)
    num_dumpers = property(lambda self: len(self._dumpers))
    cur_step = property(lambda self: self._cur_step)
