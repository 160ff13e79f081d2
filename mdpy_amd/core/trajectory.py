"""Trajectory (mdpy/core/trajectory.py:19-200): frames of positions (optionally velocities and
forces) of one topology, with its box."""
import numpy as np

from ..environment import env
from ..error import ArrayDimError, TrajectoryPoorDefinedError
from ..unit import check_quantity_value
from .cell_list import check_pbc_matrix

SPATIAL_DIM = 3


class Trajectory:
    def __init__(self, topology, contain_positions=True, contain_velocities=False, contain_forces=False) -> None:
        self._topology = topology
        self._contain = {"positions": contain_positions, "velocities": contain_velocities, "forces": contain_forces}
        self._num_frames = 0
        shape = [0, topology.num_particles, SPATIAL_DIM]
        self._data = {k: np.zeros(shape, dtype=env.NUMPY_FLOAT) for k in self._contain}
        self._unwrapped_positions = np.zeros(shape, dtype=env.NUMPY_FLOAT)
        self._pbc_matrix = np.zeros([SPATIAL_DIM, SPATIAL_DIM], dtype=env.NUMPY_FLOAT)
        self._pbc_inv = np.zeros([SPATIAL_DIM, SPATIAL_DIM], dtype=env.NUMPY_FLOAT)
        self._is_pbc_specified = False
        self._time_step = None

    def __repr__(self) -> str:
        return "<mdpy_amd.core.Trajectory object of %d particles with %d frames at %x>" % (
            self._topology.num_particles, self._num_frames, id(self))

    __str__ = __repr__

    def set_time_step(self, time_step):
        self._time_step = check_quantity_value(time_step)

    def set_pbc_matrix(self, pbc_matrix):
        pbc = check_pbc_matrix(check_quantity_value(pbc_matrix))
        self._pbc_matrix = np.ascontiguousarray(pbc, dtype=env.NUMPY_FLOAT)
        self._pbc_inv = np.ascontiguousarray(np.linalg.inv(self._pbc_matrix), dtype=env.NUMPY_FLOAT)
        self._is_pbc_specified = True

    def _check_array(self, array):
        if not isinstance(array, np.ndarray):
            raise TypeError("expected a numpy.ndarray, got %s" % type(array))
        n = self._topology.num_particles
        if array.ndim == 2:
            array = array.reshape(1, *array.shape)
        if array.ndim != 3 or array.shape[1:] != (n, SPATIAL_DIM):
            raise ArrayDimError("frames must have shape [%d, %d] (or [x, %d, %d]), got %s"
                                % (n, SPATIAL_DIM, n, SPATIAL_DIM, list(array.shape)))
        return array.astype(env.NUMPY_FLOAT)

    def append(self, positions=None, velocities=None, forces=None):
        given = {"positions": positions, "velocities": velocities, "forces": forces}
        frames = []
        for k, v in given.items():
            if self._contain[k] and v is None:
                raise TrajectoryPoorDefinedError("this trajectory holds %s: pass them" % k)
            if not self._contain[k] and v is not None:
                raise TrajectoryPoorDefinedError("this trajectory does not hold %s: pass None" % k)
            if v is not None:
                given[k] = self._check_array(v)
                frames.append(given[k].shape[0])
        if not frames:
            return
        if len(set(frames)) != 1:
            raise ArrayDimError("the arrays hold different numbers of frames: %s" % frames)
        for k, v in given.items():
            if v is not None:
                self._data[k] = np.vstack([self._data[k], v])
        self._num_frames += frames[0]

    def unwrap_positions(self):
        """Continuous positions: each frame adds the minimum-image step from the previous one."""
        check_pbc_matrix(self._pbc_matrix)
        s = np.dot(self._data["positions"], self._pbc_inv)
        ds = s[1:] - s[:-1]
        ds -= np.round(ds)
        diff = np.dot(ds, self._pbc_matrix)
        out = np.zeros_like(self._data["positions"], dtype=env.NUMPY_FLOAT)
        if self._num_frames:
            out[0] = self._data["positions"][0]
        for f in range(1, self._num_frames):
            out[f] = out[f - 1] + diff[f - 1]
        self._unwrapped_positions = out

    topology = property(lambda self: self._topology)
    num_frames = property(lambda self: self._num_frames)
    positions = property(lambda self: self._data["positions"])
    unwrapped_positions = property(lambda self: self._unwrapped_positions)
    velocities = property(lambda self: self._data["velocities"])
    forces = property(lambda self: self._data["forces"])
    pbc_matrix = property(lambda self: self._pbc_matrix)
    pbc_inv = property(lambda self: self._pbc_inv)
    time_step = property(lambda self: self._time_step)
    contain_positions = property(lambda self: self._contain["positions"])
    contain_velocities = property(lambda self: self._contain["velocities"])
    contain_forces = property(lambda self: self._contain["forces"])
