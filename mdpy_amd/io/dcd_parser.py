"""DCD trajectory reader with the interface of mdpy's DCDParser (mdpy/io/dcd_parser.py:14-70),
reading the file itself instead of through MDAnalysis.

The CHARMM/NAMD/OpenMM DCD layout is a sequence of Fortran unformatted records (each framed by
its byte length before and after):
  1. 84 bytes: "CORD" and 20 int32 control words — [0] frames, [8] fixed atoms, [9] time step
     (float32), [10] unit cell present, [11] fourth dimension present, [19] CHARMM version;
  2. titles: int32 count, then 80-byte lines;
  3. int32 number of atoms;
then per frame: the unit cell (6 float64: a, gamma, b, beta, alpha, c — angles as cosines when
all lie in [-1, 1], else in degrees) when present, and the x, y and z coordinates (N float32
each, in that order). Either byte order is accepted (the first record length tells).
"""
from __future__ import annotations

import numpy as np

from ..environment import env
from ..error import ArrayDimError, FileFormatError, ParserPoorDefinedError

SPATIAL_DIM = 3


def _triclinic(a, b, c, alpha, beta, gamma):
    """Box vectors as matrix rows from edge lengths and angles in degrees (a along x, b in the
    xy plane)."""
    al, be, ga = np.radians([alpha, beta, gamma])
    cg, sg = np.cos(ga), np.sin(ga)
    v3x = c * np.cos(be)
    v3y = c * (np.cos(al) - np.cos(be) * cg) / sg
    v3z = np.sqrt(max(c * c - v3x * v3x - v3y * v3y, 0.0))
    m = np.array([[a, 0.0, 0.0], [b * cg, b * sg, 0.0], [v3x, v3y, v3z]])
    m[np.abs(m) < 1e-9 * max(a, b, c)] = 0.0
    return m


class DCDParser:
    def __init__(self, file_path: str, is_parse_all=True) -> None:
        if not file_path.endswith(".dcd"):
            raise FileFormatError("a DCD file name must end in .dcd: %s" % file_path)
        self._file_path = file_path
        self._is_parse_all = is_parse_all
        raw = np.memmap(file_path, dtype=np.uint8, mode="r")
        self._read_header(raw)
        if self._is_parse_all:
            pos = self._frames(raw, range(self._num_frames))
            self._positions = pos[0] if self._num_frames == 1 else pos
        self._pbc_matrix = self._cell(raw, 0)

    # ------------------------------------------------------------------ layout
    def _record(self, raw, off, length=None):
        """(payload start, end of record) of the Fortran record at `off`."""
        if off + 4 > len(raw):
            raise FileFormatError("%s: truncated DCD file" % self._file_path)
        n = int(np.frombuffer(raw[off:off + 4].tobytes(), dtype=self._i4)[0])
        if (length is not None and n != length) or off + 8 + n > len(raw):
            raise FileFormatError("%s: unexpected DCD record length %d at byte %d" % (self._file_path, n, off))
        tail = int(np.frombuffer(raw[off + 4 + n:off + 8 + n].tobytes(), dtype=self._i4)[0])
        if tail != n:
            raise FileFormatError("%s: DCD record markers differ (%d, %d) at byte %d" % (self._file_path, n, tail, off))
        return off + 4, off + 8 + n

    def _read_header(self, raw):
        first = raw[:4].tobytes()
        if np.frombuffer(first, dtype="<i4")[0] == 84:
            self._i4, self._f4, self._f8 = "<i4", "<f4", "<f8"
        elif np.frombuffer(first, dtype=">i4")[0] == 84:
            self._i4, self._f4, self._f8 = ">i4", ">f4", ">f8"
        else:
            raise FileFormatError("%s is not a DCD file (first record is not 84 bytes)" % self._file_path)
        p, off = self._record(raw, 0, 84)
        if raw[p:p + 4].tobytes() != b"CORD":
            raise FileFormatError("%s is not a coordinate DCD file" % self._file_path)
        ic = np.frombuffer(raw[p + 4:p + 84].tobytes(), dtype=self._i4)
        if ic[8] != 0:
            raise FileFormatError("%s: DCD files with fixed atoms are not supported" % self._file_path)
        self._has_cell = ic[10] != 0
        self._has_4d = ic[11] != 0
        self._time_step = float(np.frombuffer(raw[p + 40:p + 44].tobytes(), dtype=self._f4)[0])
        p, off = self._record(raw, off)  # titles
        p, off = self._record(raw, off, 4)
        self._num_particles = int(np.frombuffer(raw[p:p + 4].tobytes(), dtype=self._i4)[0])
        self._first = off
        n4 = 4 * self._num_particles + 8
        self._frame_bytes = (56 if self._has_cell else 0) + 3 * n4 + (n4 if self._has_4d else 0)
        # frames the file holds (the header count can lag behind a file still being written)
        self._num_frames = (len(raw) - self._first) // self._frame_bytes
        if self._num_frames < 1:
            raise FileFormatError("%s holds no complete frame" % self._file_path)

    def _frame_start(self, frame):
        return self._first + frame * self._frame_bytes

    def _cell(self, raw, frame):
        if not self._has_cell:
            return np.zeros((SPATIAL_DIM, SPATIAL_DIM), dtype=env.NUMPY_FLOAT)
        p, _ = self._record(raw, self._frame_start(frame), 48)
        a, g, b, be, al, c = np.frombuffer(raw[p:p + 48].tobytes(), dtype=self._f8).astype(np.float64)
        ang = np.array([al, be, g])
        if np.all(np.abs(ang) <= 1.0):  # cosines
            ang = np.degrees(np.arccos(ang))
        return _triclinic(a, b, c, *ang).astype(env.NUMPY_FLOAT)

    def _frames(self, raw, frames):
        n = self._num_particles
        out = np.empty((len(frames), n, SPATIAL_DIM), dtype=env.NUMPY_FLOAT)
        for k, f in enumerate(frames):
            off = self._frame_start(f) + (56 if self._has_cell else 0)
            for d in range(SPATIAL_DIM):
                p, off = self._record(raw, off, 4 * n)
                out[k, :, d] = np.frombuffer(raw[p:p + 4 * n].tobytes(), dtype=self._f4)
        return out

    # ------------------------------------------------------------------ interface
    def get_positions(self, *frames):
        for frame in frames:
            if frame >= self._num_frames:
                raise ArrayDimError("frame %d is beyond the %d frames of %s" % (frame, self._num_frames,
                                                                              self._file_path))
        raw = np.memmap(self._file_path, dtype=np.uint8, mode="r")
        pos = self._frames(raw, list(frames))
        return pos[0] if len(frames) == 1 else pos

    @property
    def num_frames(self):
        return self._num_frames

    @property
    def num_particles(self):
        return self._num_particles

    @property
    def time_step(self):
        """The header's time step (float32, in the writer's units; AKMA for CHARMM/OpenMM)."""
        return self._time_step

    @property
    def positions(self) -> np.ndarray:
        if not self._is_parse_all:
            raise ParserPoorDefinedError("positions are not parsed with is_parse_all=False; "
                                         "read frames with get_positions")
        return self._positions.copy()

    @property
    def pbc_matrix(self) -> np.ndarray:
        return self._pbc_matrix.copy()

