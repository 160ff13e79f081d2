"""PDB coordinate files (mdpy/io/pdb_parser.py:17-127), read directly.

The reference reads the file through MDAnalysis 2.0.0 (PDBReader / PDBParser). This reader
takes the same fields from the fixed-column records: ATOM / HETATM (serial 7-11, name 13-16,
residue name 18-21, chain 22, residue number 23-26, x y z 31-54), MODEL / ENDMDL frames and
CRYST1 (a b c alpha beta gamma -> the 3 x 3 box matrix with edge vectors as rows, as
MDAnalysis' triclinic_dimensions; None when the file has no CRYST1). Particle types are guessed
from the atom names by MDAnalysis' rule (guess_atom_element), restated below.
"""
import math
import re

import numpy as np

from ..environment import env
from ..error import ArrayDimError, FileFormatError, ParserPoorDefinedError

SPATIAL_DIM = 3

# MDAnalysis 2.0.0 topology/tables.py: names that translate directly (atomelements) and the
# element symbols the name-based guess accepts (elements)
_NAME_TO_ELEMENT = {
    "CAL": "CA", "C0": "CA", "CA2+": "CA", "CES": "CS", "CLA": "CL", "CLAL": "CL", "CL": "CL", "CL-": "CL",
    "FE": "FE", "FE2": "FE", "LIT": "LI", "LI": "LI", "LI+": "LI", "QL": "LI", "MG": "MG", "MG2+": "MG",
    "NE": "NE", "HE": "HE", "K": "K", "POT": "K", "K+": "K", "QK": "K", "SOD": "NA", "NA": "NA", "NA+": "NA",
    "QN": "NA", "ZN": "ZN", "CU": "CU", "CS": "CS", "CS+": "CS", "QC": "CE", "RB": "RB", "QR": "RB",
    "BC": "C", "AC": "C", "MW": "DUMMY",
}
_ELEMENTS = ("H", "LI", "BE", "B", "C", "N", "O", "F", "NA", "MG", "AL", "P", "SI", "S", "CL", "K")


def guess_atom_type(name):
    """Element guessed from an atom name (MDAnalysis' guess_atom_element): a direct translation
    if the name is a known ion/special name, else the leading letters with symbols and digits
    removed, shortened until an element symbol is found."""
    if name == "":
        return ""
    if name.upper() in _NAME_TO_ELEMENT:
        return _NAME_TO_ELEMENT[name.upper()]
    no_symbols = re.sub(r"[*+\-']", "", name)
    parts = [p for p in re.split(r"[0-9]", no_symbols) if p]
    cand = parts[0].upper() if parts else ""
    if cand in _NAME_TO_ELEMENT:
        return _NAME_TO_ELEMENT[cand]
    while cand:
        if cand in _ELEMENTS:
            return cand
        if cand[:-1] in _ELEMENTS:
            return cand[:-1]
        if cand[1:] in _ELEMENTS:
            return cand[1:]
        if len(cand) <= 2:
            return cand[0]
        cand = cand[:-1]
    return no_symbols


def _box_matrix(a, b, c, alpha, beta, gamma):
    """Edge vectors (rows) from lengths and angles in degrees, a along x, b in the xy plane."""
    if abs(alpha - 90) < 1e-9 and abs(beta - 90) < 1e-9 and abs(gamma - 90) < 1e-9:
        return np.diag([a, b, c]).astype(np.float32)
    al, be, ga = (math.radians(v) for v in (alpha, beta, gamma))
    bx, by = b * math.cos(ga), b * math.sin(ga)
    cx = c * math.cos(be)
    cy = c * (math.cos(al) - math.cos(be) * math.cos(ga)) / math.sin(ga)
    cz = math.sqrt(max(c * c - cx * cx - cy * cy, 0.0))
    return np.array([[a, 0, 0], [bx, by, 0], [cx, cy, cz]], dtype=np.float32)


class PDBParser:
    def __init__(self, file_path, is_parse_all=True) -> None:
        if not file_path.endswith(".pdb"):
            raise FileFormatError("expected a .pdb file, got %s" % file_path)
        self._file_path = file_path
        self._is_parse_all = is_parse_all
        frames, cur, atoms, box = [], [], [], None
        with open(file_path) as f:
            for line in f:
                rec = line[:6]
                if rec in ("ATOM  ", "HETATM"):
                    if not frames:  # the first frame defines the atoms
                        serial = line[6:11].strip()
                        resnum = line[22:26].strip()
                        atoms.append((int(serial) if serial.lstrip("-").isdigit() else serial, line[12:16].strip(),
                                      line[17:21].strip(), line[21:22].strip(),
                                      int(resnum) if resnum.lstrip("-").isdigit() else resnum))
                    cur.append([float(line[30:38]), float(line[38:46]), float(line[46:54])])
                elif rec == "ENDMDL":
                    frames.append(cur)
                    cur = []
                elif rec == "CRYST1" and box is None:
                    box = _box_matrix(*[float(v) for v in (line[6:15], line[15:24], line[24:33], line[33:40],
                                                           line[40:47], line[47:54])])
        if cur:
            frames.append(cur)
        self._num_particles = len(atoms)
        self._particle_ids = [a[0] for a in atoms]
        self._particle_names = [a[1] for a in atoms]
        self._particle_types = [guess_atom_type(a[1]) for a in atoms]
        self._molecule_types = [a[2] for a in atoms]
        self._chain_ids = [a[3] for a in atoms]
        self._molecule_ids = [a[4] for a in atoms]
        self._matrix_ids = list(np.arange(self._num_particles, dtype=env.NUMPY_INT))
        self._frames = [np.asarray(fr, dtype=np.float64) for fr in frames]
        self._num_frames = len(self._frames)
        if is_parse_all:
            if self._num_frames == 1:
                self._positions = self._frames[0].astype(env.NUMPY_FLOAT)
            else:
                self._positions = np.stack(self._frames).astype(env.NUMPY_FLOAT)
        self._pbc_matrix = box

    def get_matrix_id(self, particle_id):
        return self._particle_ids.index(particle_id)

    def get_particle_info(self, particle_id):
        k = self.get_matrix_id(particle_id)
        return {"particle_id": self._particle_ids[k], "particle_type": self._particle_types[k],
                "particle_name": self._particle_names[k], "molecule_id": self._molecule_ids[k],
                "molecule_type": self._molecule_types[k], "chain_id": self._chain_ids[k], "matrix_id": k,
                "position": self._frames[0][k, :].astype(env.NUMPY_FLOAT)}

    def get_positions(self, *frames):
        for fr in frames:
            if fr >= self._num_frames:
                raise ArrayDimError("frame %d requested; the file holds %d" % (fr, self._num_frames))
        if len(frames) == 1:
            return self._frames[frames[0]].astype(env.NUMPY_FLOAT)
        return np.stack([self._frames[fr] for fr in frames]).astype(env.NUMPY_FLOAT)

    particle_ids = property(lambda self: self._particle_ids)
    particle_types = property(lambda self: self._particle_types)
    particle_names = property(lambda self: self._particle_names)
    molecule_ids = property(lambda self: self._molecule_ids)
    molecule_types = property(lambda self: self._molecule_types)
    chain_ids = property(lambda self: self._chain_ids)
    num_frames = property(lambda self: self._num_frames)
    num_particles = property(lambda self: self._num_particles)

    @property
    def positions(self) -> np.ndarray:
        if not self._is_parse_all:
            raise ParserPoorDefinedError("positions are not kept with is_parse_all=False; use get_positions()")
        return self._positions.copy()

    @property
    def pbc_matrix(self):
        return None if self._pbc_matrix is None else self._pbc_matrix.copy()
