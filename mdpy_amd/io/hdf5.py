"""HDF5 trajectory files in mdpy's layout (mdpy/io/__init__.py:21-72): HDF5Writer
(mdpy/io/hdf5_writer.py:20-153) and HDF5Parser (mdpy/io/hdf5_parser.py:35-206), on the small
HDF5 container reader/writer of hdf5_format.py instead of h5py.

Layout: /topology/particles/{particle_id, particle_type, particle_name, matrix_id, molecule_id,
molecule_type, chain_id, mass, charge}, /topology/{num_particles, bonds, num_bonds, angles,
num_angles, dihedrals, num_dihedrals, impropers, num_impropers}, /positions/{frame-0 ..,
num_frames}, /pbc_matrix; strings as variable-length UTF-8, a missing particle attribute as the
value 10000 of the column's type.

The writer keeps the file's content in memory and rewrites the file on every change (a topology
or box set, a frame written), so the file on disk is complete after each call.
"""
from __future__ import annotations

import copy

import numpy as np

from ..core import Particle, Topology, Trajectory
from ..core.cell_list import check_pbc_matrix
from ..environment import env
from ..error import ArrayDimError, FileFormatError, HDF5FilePoorDefinedError, ParserPoorDefinedError
from ..unit import check_quantity_value
from .hdf5_format import H5Reader, H5Writer

NONE_LABEL = 10000
SPATIAL_DIM = 3
ROOT_KEYS = ["topology", "positions", "pbc_matrix"]
TOPOLOGY_KEYS = ["particles", "num_particles", "bonds", "num_bonds", "angles", "num_angles", "dihedrals",
                 "num_dihedrals", "impropers", "num_impropers"]
PARTICLES_KEYS = ["particle_id", "particle_type", "particle_name", "molecule_id", "molecule_type", "chain_id",
                  "matrix_id", "mass", "charge"]


class HDF5Writer:
    def __init__(self, file_path: str, mode: str = "w", topology: Topology = None,
                 pbc_matrix=np.diag([1] * 3).astype(env.NUMPY_FLOAT)) -> None:
        if not file_path.endswith(".hdf5"):
            raise FileFormatError("an HDF5 file name must end in .hdf5: %s" % file_path)
        self._file_path = file_path
        self._mode = mode
        topology = Topology() if topology is None else topology
        if not isinstance(topology, Topology):
            raise TypeError("topology must be a mdpy_amd.core.Topology")
        self._topology = topology
        self._pbc_matrix = check_pbc_matrix(check_quantity_value(pbc_matrix))
        self._frames = []
        self._flush()

    # ------------------------------------------------------------------ content
    @staticmethod
    def _value(val, target_type):
        return target_type(NONE_LABEL) if val is None else target_type(val)

    def _topology_tree(self):
        t = self._topology
        parts = t.particles
        n = len(parts)

        def ints(name):
            return np.array([self._value(getattr(p, name), env.NUMPY_INT) for p in parts], dtype=env.NUMPY_INT)

        def floats(name):
            return np.array([self._value(getattr(p, name), env.NUMPY_FLOAT) for p in parts], dtype=env.NUMPY_FLOAT)

        def strs(name):
            return [self._value(getattr(p, name), str) for p in parts]

        def terms(lst, width):
            return np.array(lst, dtype=env.NUMPY_INT).reshape(-1, width) if len(lst) else np.zeros((0,), dtype=env.NUMPY_INT)

        particles = {"particle_id": ints("particle_id"), "particle_type": strs("particle_type"),
                     "particle_name": strs("particle_name"), "matrix_id": ints("matrix_id"),
                     "molecule_id": ints("molecule_id"), "molecule_type": strs("molecule_type"),
                     "chain_id": strs("chain_id"), "mass": floats("mass"), "charge": floats("charge")}
        return {"particles": particles, "num_particles": env.NUMPY_INT(n),
                "bonds": terms(t.bonds, 2), "num_bonds": env.NUMPY_INT(t.num_bonds),
                "angles": terms(t.angles, 3), "num_angles": env.NUMPY_INT(t.num_angles),
                "dihedrals": terms(t.dihedrals, 4), "num_dihedrals": env.NUMPY_INT(t.num_dihedrals),
                "impropers": terms(t.impropers, 4), "num_impropers": env.NUMPY_INT(t.num_impropers)}

    def _flush(self):
        positions = {"frame-%d" % k: f for k, f in enumerate(self._frames)}
        positions["num_frames"] = np.int64(len(self._frames))
        H5Writer().write(self._file_path, {"topology": self._topology_tree(), "positions": positions,
                                           "pbc_matrix": np.asarray(self._pbc_matrix, dtype=env.NUMPY_FLOAT)})

    # ------------------------------------------------------------------ interface
    def write(self, positions=None):
        positions = np.asarray(positions)
        shape, n = positions.shape, self._topology.num_particles
        ok = (positions.ndim == 2 and shape[0] == n) or (positions.ndim == 3 and shape[1] == n)
        if not ok:
            raise ArrayDimError("the topology holds %d particles, positions have shape %s" % (n, list(shape)))
        frames = positions[None] if positions.ndim == 2 else positions
        for f in frames:
            self._frames.append(np.ascontiguousarray(f, dtype=env.NUMPY_FLOAT))
        self._flush()

    @property
    def file_path(self):
        return self._file_path

    @property
    def mode(self):
        return self._mode

    @property
    def topology(self):
        return self._topology

    @topology.setter
    def topology(self, topology: Topology):
        if not isinstance(topology, Topology):
            raise TypeError("topology must be a mdpy_amd.core.Topology")
        self._topology = topology
        self._flush()

    @property
    def pbc_matrix(self):
        return self._pbc_matrix

    @pbc_matrix.setter
    def pbc_matrix(self, pbc_matrix):
        self._pbc_matrix = check_pbc_matrix(check_quantity_value(pbc_matrix))
        self._flush()


class HDF5Parser:
    def __init__(self, file_path: str, is_parse_all=True) -> None:
        if not file_path.endswith(".hdf5"):
            raise FileFormatError("an HDF5 file name must end in .hdf5: %s" % file_path)
        self._file_path = file_path
        self._is_parse_all = is_parse_all
        f = H5Reader(file_path)
        self._check(f)
        self._topology = self._parse_topology(f)
        self._pbc_matrix = np.asarray(f.read("pbc_matrix")).astype(env.NUMPY_FLOAT)
        self._num_frames = int(f.read("positions/num_frames"))
        self._num_particles = self._topology.num_particles
        if self._is_parse_all:
            self._positions = self._parse_positions(f)
            self._trajectory = Trajectory(self._topology)
            if self._positions is not None:
                self._trajectory.append(positions=self._positions)
            self._trajectory.set_pbc_matrix(self._pbc_matrix)

    @staticmethod
    def _check(f):
        ok = all(k in f.keys("/") for k in ROOT_KEYS)
        ok = ok and all(k in f.keys("topology") for k in TOPOLOGY_KEYS)
        ok = ok and all(k in f.keys("topology/particles") for k in PARTICLES_KEYS)
        ok = ok and "num_frames" in f.keys("positions")
        if not ok:
            raise HDF5FilePoorDefinedError("%s does not follow the mdpy HDF5 layout (mdpy_amd.io.HDF5_FILE_HIERARCHY)"
                                           % f.path)

    @staticmethod
    def _none(val):
        return None if val == type(val)(NONE_LABEL) else val

    def _parse_topology(self, f):
        t = Topology()
        p = "topology/particles/"
        pid = f.read(p + "particle_id").astype(env.NUMPY_INT)
        ptype, pname = f.read(p + "particle_type"), f.read(p + "particle_name")
        mid, molid = f.read(p + "matrix_id").astype(env.NUMPY_INT), f.read(p + "molecule_id").astype(env.NUMPY_INT)
        moltype, chain = f.read(p + "molecule_type"), f.read(p + "chain_id")
        mass, charge = f.read(p + "mass").astype(env.NUMPY_FLOAT), f.read(p + "charge").astype(env.NUMPY_FLOAT)
        particles = []
        for k in range(int(f.read("topology/num_particles"))):
            particles.append(Particle(
                particle_id=self._none(pid[k]), particle_type=self._none(bytes.decode(ptype[k])),
                particle_name=self._none(bytes.decode(pname[k])), matrix_id=self._none(mid[k]),
                molecule_id=self._none(molid[k]), molecule_type=self._none(bytes.decode(moltype[k])),
                chain_id=self._none(bytes.decode(chain[k])), mass=self._none(mass[k]), charge=self._none(charge[k])))
        t.add_particles(particles)
        for name, add in (("bonds", t.add_bond), ("angles", t.add_angle), ("dihedrals", t.add_dihedral),
                          ("impropers", t.add_improper)):
            arr = f.read("topology/" + name)
            for k in range(int(f.read("topology/num_" + name))):
                add(list(arr[k, :]))
        t.join()
        return t

    def _parse_positions(self, f):
        if self._num_frames == 0:
            return None
        out = np.empty([self._num_frames, self._num_particles, SPATIAL_DIM], dtype=env.NUMPY_FLOAT)
        for k in range(self._num_frames):
            out[k] = f.read("positions/frame-%d" % k)
        return out if self._num_frames != 1 else out[0]

    def get_positions(self, *frames):
        for frame in frames:
            if frame >= self._num_frames:
                raise ArrayDimError("frame %d is beyond the %d frames of %s" % (frame, self._num_frames,
                                                                              self._file_path))
        f = H5Reader(self._file_path)
        out = np.stack([np.asarray(f.read("positions/frame-%d" % k), dtype=env.NUMPY_FLOAT) for k in frames])
        return out[0] if len(frames) == 1 else out

    @property
    def num_frames(self):
        return self._num_frames

    @property
    def num_particles(self):
        return self._num_particles

    @property
    def topology(self) -> Topology:
        return copy.copy(self._topology)

    @property
    def positions(self) -> np.ndarray:
        if not self._is_parse_all:
            raise ParserPoorDefinedError("positions are not parsed with is_parse_all=False; "
                                         "read frames with get_positions")
        return None if self._positions is None else self._positions.copy()

    @property
    def pbc_matrix(self) -> np.ndarray:
        return self._pbc_matrix.copy()

    @property
    def trajectory(self):
        if not self._is_parse_all:
            raise ParserPoorDefinedError("the trajectory is not parsed with is_parse_all=False")
        return copy.copy(self._trajectory)
