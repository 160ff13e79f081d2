"""Dumpers (mdpy/dumper/*.py): LogDumper, PDBDumper and HDF5Dumper.

A dumper formats a Frame: the step, the positions, the per-term potential energies and the
kinetic energy at that step, the box and the wall-clock time the frame was reached. During
Simulation.integrate frames are copied from the device asynchronously (pinned host buffers,
ordered on the stream after the steps that produced them) and written by a background thread,
so the GPU keeps stepping while a frame is formatted; dump(simulation) formats the ensemble's
current state synchronously, as the reference does.
"""
import math
import time
from datetime import timedelta

import numpy as np

from ..error import DumperPoorDefinedError, FileFormatError
from ..unit import KB, KJ_PER_MOL


class Frame:
    def __init__(self, step, positions, potential_energy, kinetic_energy, pbc_matrix, masses, time_step, wall):
        self.step = step
        self.positions = positions
        self.potential_energy = potential_energy
        self.kinetic_energy = kinetic_energy
        self.pbc_matrix = pbc_matrix
        self.masses = masses
        self.time_step = time_step
        self.wall = wall

    @classmethod
    def of(cls, simulation, positions=True):
        ens = simulation.ensemble
        return cls(simulation.cur_step, np.asarray(ens.state.positions, dtype=np.float64) if positions else None,
                   float(ens.potential_energy), float(ens.kinetic_energy), np.asarray(ens.state.pbc_matrix),
                   np.asarray(ens.topology.masses), simulation.integrator.time_step, time.perf_counter())


class Dumper:
    """mdpy/dumper/dumper.py:13-46: a file of the right suffix, a positive dump frequency."""
    SUFFIX = None
    NEEDS_POSITIONS = False

    def __init__(self, file_path: str, dump_frequency: int, suffix: str = None) -> None:
        suffix = suffix or self.SUFFIX
        if file_path is not None:
            if suffix and not file_path.endswith(suffix):
                raise FileFormatError("a %s dumper writes a .%s file, got %s" % (type(self).__name__, suffix, file_path))
            open(file_path, "w").close()
        if int(dump_frequency) <= 0:
            raise DumperPoorDefinedError("dump_frequency must be a positive integer")
        self._file_path = file_path
        self._dump_frequency = int(dump_frequency)
        self._num_dumpped_frames = 0

    def dump(self, simulation):
        self.dump_frame(Frame.of(simulation, positions=self.NEEDS_POSITIONS), simulation)

    def dump_frame(self, frame, simulation):
        raise NotImplementedError("%s must implement dump_frame()" % type(self).__name__)

    def _append(self, text):
        if self._file_path is not None and text:
            with open(self._file_path, "a") as f:
                f.write(text)

    file_path = property(lambda self: self._file_path)
    num_dumpped_frames = property(lambda self: self._num_dumpped_frames)

    @property
    def dump_frequency(self):
        return self._dump_frequency

    @dump_frequency.setter
    def dump_frequency(self, val):
        self._dump_frequency = int(val)


class LogDumper(Dumper):
    """mdpy/dumper/log_dumper.py:15-208: the selected columns, tab separated, energies in kJ/mol.
    With no column selected (or no file) the values are still kept in ``records``."""
    SUFFIX = "log"

    def __init__(self, file_path: str = None, dump_frequency: int = 1, step: bool = False, sim_time: bool = False,
                 sim_speed: bool = False, potential_energy: bool = False, kinetic_energy: bool = False,
                 total_energy: bool = False, temperature: bool = False, pressure: bool = False, volume: bool = False,
                 density: bool = False, cpu_time: bool = False, rest_time: bool = False, total_step: int = 0,
                 seperator: str = "\t") -> None:
        super().__init__(file_path, dump_frequency)
        self._cols = {"step": step, "sim_time": sim_time, "potential_energy": potential_energy,
                      "kinetic_energy": kinetic_energy, "total_energy": total_energy, "temperature": temperature,
                      "volume": volume, "pressure": pressure, "density": density, "cpu_time": cpu_time,
                      "sim_speed": sim_speed, "rest_time": rest_time}
        if total_step == 0 and rest_time:
            raise DumperPoorDefinedError("rest_time needs total_step")
        self._total_step = total_step
        self._sep = seperator
        self.records = []
        self._t_start = None     # wall time of the first frame after step 0
        self._step_start = 0

    _HEADERS = {"step": "Step", "sim_time": "Time (ns)", "potential_energy": "E_p (kj/mol)",
                "kinetic_energy": "E_k (kj/mol)", "total_energy": "E_t (kj/mol)", "temperature": "Temperature (K)",
                "volume": "Volume (nm^3)", "pressure": "Pressure (atm)", "density": "Density (1e3*kg/m^3)",
                "cpu_time": "CPU Time", "sim_speed": "Speed (ns/day)", "rest_time": "Rest Time"}

    def dump_frame(self, frame, simulation):
        n = len(frame.masses)
        if self._t_start is None:
            self._t_start, self._step_start = frame.wall, frame.step
        elapsed = frame.wall - self._t_start
        steps = frame.step - self._step_start
        speed = 0.0864 * frame.time_step * steps / elapsed if elapsed > 0 and steps > 0 else float("nan")
        p = frame.pbc_matrix
        volume = abs(float(np.dot(np.cross(p[0], p[1]), p[2])))
        rec = {"step": frame.step, "potential_energy": frame.potential_energy, "kinetic_energy": frame.kinetic_energy,
               "total_energy": frame.potential_energy + frame.kinetic_energy,
               "temperature": 2 * frame.kinetic_energy / (3 * n * KB), "ns_per_day": speed,
               "volume_nm3": volume * 1e-3, "wall_s": elapsed}
        self.records.append(rec)
        text = ""
        if self._num_dumpped_frames == 0 and any(self._cols.values()):
            text += "".join(self._HEADERS[k] + self._sep for k, on in self._cols.items() if on) + "\n"
        cells = []
        for k, on in self._cols.items():
            if not on:
                continue
            if k == "step":
                cells.append("%d" % frame.step)
            elif k == "sim_time":
                cells.append("%.2e" % (frame.step * frame.time_step * 1e-6))
            elif k in ("potential_energy", "kinetic_energy", "total_energy"):
                cells.append("%.3f" % (rec[k] / KJ_PER_MOL))
            elif k == "temperature":
                cells.append("%.2f" % rec["temperature"])
            elif k == "volume":
                cells.append("%.3f" % (volume * 1e-3))
            elif k == "pressure":
                continue  # the reference leaves the pressure column empty (log_dumper.py:120-121)
            elif k == "density":
                # Da/A^3 -> kg/m^3 is 1660.539; the column is in units of 1e3 kg/m^3
                cells.append("%.3f" % (float(np.sum(frame.masses)) / volume * 1660.53904 / 1000))
            elif k == "cpu_time":
                cells.append("%s" % timedelta(seconds=int(elapsed)))
            elif k == "sim_speed":
                cells.append("--:--:--" if not math.isfinite(speed) else "%.4f" % speed)
            elif k == "rest_time":
                if frame.step == 0 or steps <= 0:
                    cells.append("--:--:--")
                else:
                    left = max(self._total_step - frame.step, 0)
                    cells.append("%s" % timedelta(seconds=int(elapsed / steps * left)))
        if cells:
            text += "".join(c + self._sep for c in cells) + "\n"
        self._append(text)
        self._num_dumpped_frames += 1


class PDBDumper(Dumper):
    """mdpy/dumper/pdb_dumper.py:14-24: one MODEL per dumped frame."""
    SUFFIX = "pdb"
    NEEDS_POSITIONS = True

    def __init__(self, file_path: str, dump_frequency: int) -> None:
        from ..io import PDBWriter
        super().__init__(file_path, dump_frequency)
        self._writer = PDBWriter(file_path, "w")

    def dump_frame(self, frame, simulation):
        if self._num_dumpped_frames == 0:
            self._writer.topology = simulation.ensemble.topology
            self._writer.pbc_matrix = frame.pbc_matrix
        self._writer.write(frame.positions)
        self._num_dumpped_frames += 1


class HDF5Dumper(Dumper):
    """mdpy/dumper/hdf5_dumper.py:14-23: the topology and box at the first dumped frame, then the
    positions of every dumped frame (mdpy's HDF5 layout, mdpy_amd.io.HDF5Writer)."""
    SUFFIX = "hdf5"
    NEEDS_POSITIONS = True

    def __init__(self, file_path: str, dump_frequency: int) -> None:
        from ..io import HDF5Writer
        super().__init__(file_path, dump_frequency)
        self._writer = HDF5Writer(file_path, "w")

    def dump_frame(self, frame, simulation):
        if self._num_dumpped_frames == 0:
            self._writer.topology = simulation.ensemble.topology
            self._writer.pbc_matrix = frame.pbc_matrix
        self._writer.write(frame.positions)
        self._num_dumpped_frames += 1


__all__ = ["Dumper", "LogDumper", "PDBDumper", "HDF5Dumper", "Frame"]
