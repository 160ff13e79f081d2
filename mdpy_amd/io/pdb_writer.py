"""PDB trajectory writer (mdpy/io/pdb_writer.py:60-280).

Record layout (fixed columns, as the reference writes them): a HEADER and a CRYST1 line (edge
lengths and angles of the box) once, then per frame MODEL n, one ATOM line per particle (serial,
name, residue name, chain, residue number, x y z with 3 decimals, occupancy and B-factor 0,
element = the particle type), a TER line where the chain changes and after the last particle,
ENDMDL; END when asked."""
import datetime

import numpy as np

from ..core import Topology
from ..error import ArrayDimError, FileFormatError
from ..unit import check_quantity_value


def _atom_line(serial, name, resname, chain, resid, x, y, z, element):
    s = "ATOM  " + "%5s" % serial + " " * (7 - len("%5d" % serial))
    s += "%-4s%-4s%1s" % (name[:4], resname[:4], chain[:1])
    s += "%4d" % resid + " " * (8 - len("%4d" % resid))
    s += "%8.3f%8.3f%8.3f%6.2f%6.2f" % (x, y, z, 0, 0)
    return s + " " * 10 + "%-2s\n" % element[:2]


def _ter_line(serial, resname, chain, resid):
    s = "TER   " + "%5s" % serial + " " * (11 - len("%5d" % serial))
    s += "%-4s%1s" % (resname[:4], chain[:1])
    return s + "%4d" % resid + " " * (8 - len("%4d" % resid)) + "\n"


def _angle_deg(a, b):
    c = float(np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b)))
    return float(np.degrees(np.arccos(np.clip(c, -1.0, 1.0))))


class PDBWriter:
    def __init__(self, file_path: str, mode: str = "w", topology: Topology = None, pbc_matrix=None) -> None:
        if not file_path.endswith(".pdb"):
            raise FileFormatError("expected a .pdb file, got %s" % file_path)
        self._file_path = file_path
        self._mode = mode
        self._topology = topology if topology is not None else Topology()
        if not isinstance(self._topology, Topology):
            raise TypeError("topology must be a mdpy_amd.core.Topology")
        self._pbc_matrix = np.zeros((3, 3)) if pbc_matrix is None else np.asarray(check_quantity_value(pbc_matrix))
        open(file_path, mode).close()
        self._is_header = True
        self._cur_model = 0

    def _text_header(self):
        head = "HEADER" + " " * 4 + "%-40s%-11s\n" % ("PDB FILE CREATED WITH MDPY_AMD",
                                                     datetime.date.today().strftime("%d-%b-%Y").upper())
        p = self._pbc_matrix
        lens = np.linalg.norm(p, axis=0)
        if np.all(lens > 0):
            ang = (_angle_deg(p[0], p[1]), _angle_deg(p[0], p[2]), _angle_deg(p[1], p[2]))
        else:
            ang = (0.0, 0.0, 0.0)
        return head + "CRYST1" + "%9.3f%9.3f%9.3f%7.2f%7.2f%7.2f\n" % (lens[0], lens[1], lens[2], *ang)

    def _text_model(self, x):
        out = ["MODEL " + " " * 4 + "%4d\n" % self._cur_model]
        parts = self._topology.particles
        chain, serial = (parts[0].chain_id or "")[:1], 1
        for k, p in enumerate(parts):
            cid = (p.chain_id or "")[:1]
            if cid != chain:
                chain = cid
                q = parts[k - 1]
                out.append(_ter_line(serial, q.molecule_type or "", q.chain_id or "", q.molecule_id or 0))
                serial += 1
            out.append(_atom_line(serial, p.particle_name or "", p.molecule_type or "", p.chain_id or "",
                                  p.molecule_id or 0, x[k, 0], x[k, 1], x[k, 2], p.particle_type or ""))
            serial += 1
        q = parts[-1]
        out.append(_ter_line(serial, q.molecule_type or "", q.chain_id or "", q.molecule_id or 0))
        out.append("ENDMDL\n")
        self._cur_model += 1
        return "".join(out)

    def frame_text(self, positions: np.ndarray, is_end=False):
        """The text write() appends for these positions (one frame (N,3) or several (F,N,3))."""
        n = self._topology.num_particles
        x = np.asarray(positions)
        if x.ndim == 2:
            x = x[None]
        if x.ndim != 3 or x.shape[1:] != (n, 3):
            raise ArrayDimError("the topology has %d particles; positions of shape %s were given"
                                % (n, list(np.asarray(positions).shape)))
        text = ""
        if self._is_header and "a" not in self._mode:
            text += self._text_header()
            self._is_header = False
        text += "".join(self._text_model(f) for f in x)
        if is_end:
            text += "END"
        return text

    def write(self, positions: np.ndarray, is_end=False):
        text = self.frame_text(positions, is_end)
        with open(self._file_path, "a") as f:
            f.write(text)

    def _write_end(self):
        with open(self._file_path, "a") as f:
            f.write("END")

    @property
    def topology(self):
        return self._topology

    @topology.setter
    def topology(self, topology):
        if not isinstance(topology, Topology):
            raise TypeError("topology must be a mdpy_amd.core.Topology")
        self._topology = topology

    @property
    def pbc_matrix(self):
        return self._pbc_matrix

    @pbc_matrix.setter
    def pbc_matrix(self, pbc_matrix):
        self._pbc_matrix = np.asarray(check_quantity_value(pbc_matrix))
