"""CHARMM / X-PLOR PSF topology files (mdpy/io/psf_parser.py:19-140), read directly.

The reference reads the file through MDAnalysis 2.0.0's PSFParser; this reader takes the
sections that parser uses from the file itself: the !NATOM table (id, segment, residue id,
residue name, atom name, atom type, charge, mass) and the !NBOND, !NTHETA, !NPHI and !NIMPHI
index lists (1-based in the file, 0-based here). As in the reference, a Particle's type is the
PSF type column (e.g. NH3) and its name the PSF name column (e.g. N); masses stay in Da and
charges in e (the internal units).
"""
from copy import copy

import numpy as np

from ..core import Particle, Topology
from ..environment import env
from ..error import FileFormatError

# section tag -> number of atom indices per entry
_INDEX_SECTIONS = {"!NBOND": 2, "!NTHETA": 3, "!NPHI": 4, "!NIMPHI": 4}


def _read_psf(path):
    with open(path) as f:
        lines = f.read().split("\n")
    if not lines or not lines[0].startswith("PSF"):
        raise FileFormatError("%s does not start with a PSF header" % path)
    atoms, sections = None, {}
    i = 0
    while i < len(lines):
        fields = lines[i].split()
        if len(fields) >= 2 and fields[1].startswith("!N"):
            tag, count = fields[1].rstrip(":"), int(fields[0])
            if tag == "!NATOM":
                atoms = [lines[i + 1 + k].split() for k in range(count)]
                i += 1 + count
                continue
            if tag in _INDEX_SECTIONS:
                width = _INDEX_SECTIONS[tag]
                need, vals = count * width, []
                i += 1
                while len(vals) < need:
                    vals.extend(int(v) for v in lines[i].split())
                    i += 1
                sections[tag] = (np.array(vals[:need], dtype=np.int64).reshape(count, width) - 1).tolist()
                continue
        i += 1
    if atoms is None:
        raise FileFormatError("%s has no !NATOM section" % path)
    return atoms, sections


class PSFParser:
    def __init__(self, file_path: str) -> None:
        if not file_path.endswith(".psf"):
            raise FileFormatError("expected a .psf file, got %s" % file_path)
        atoms, sections = _read_psf(file_path)
        self._num_particles = len(atoms)
        # atom line: id segid resid resname name type charge mass [imove ...]
        self._particle_ids = [int(a[0]) for a in atoms]
        self._chain_ids = [a[1] for a in atoms]
        self._molecule_ids = [int(a[2]) if a[2].lstrip("-").isdigit() else a[2] for a in atoms]
        self._molecule_types = [a[3] for a in atoms]
        self._particle_names = [a[4] for a in atoms]
        self._particle_types = [a[5] for a in atoms]
        self._charges = [float(a[6]) for a in atoms]
        self._masses = [float(a[7]) for a in atoms]
        self._matrix_ids = list(np.arange(self._num_particles, dtype=env.NUMPY_INT))
        self._bonds = sections.get("!NBOND", [])
        self._angles = sections.get("!NTHETA", [])
        self._dihedrals = sections.get("!NPHI", [])
        self._impropers = sections.get("!NIMPHI", [])
        self._num_bonds, self._num_angles = len(self._bonds), len(self._angles)
        self._num_dihedrals, self._num_impropers = len(self._dihedrals), len(self._impropers)
        self._topology = self._create_topology()

    def _create_topology(self):
        topology = Topology()
        topology.add_particles([
            Particle(particle_id=self._particle_ids[i], particle_type=self._particle_types[i],
                     particle_name=self._particle_names[i], matrix_id=self._matrix_ids[i],
                     molecule_id=self._molecule_ids[i], molecule_type=self._molecule_types[i],
                     chain_id=self._chain_ids[i], mass=self._masses[i], charge=self._charges[i])
            for i in range(self._num_particles)])
        for b in self._bonds:
            topology.add_bond(b)
        for a in self._angles:
            topology.add_angle(a)
        for d in self._dihedrals:
            topology.add_dihedral(d)
        for m in self._impropers:
            topology.add_improper(m)
        topology.join()
        return topology

    def get_matrix_id(self, particle_id):
        return self._particle_ids.index(particle_id)

    def get_particle_info(self, particle_id):
        k = self.get_matrix_id(particle_id)
        return {"particle_id": self._particle_ids[k], "particle_type": self._particle_types[k],
                "particle_name": self._particle_names[k], "molecule_id": self._molecule_ids[k],
                "molecule_type": self._molecule_types[k], "chain_id": self._chain_ids[k], "matrix_id": k}

    num_particles = property(lambda self: self._num_particles)
    num_bonds = property(lambda self: self._num_bonds)
    num_angles = property(lambda self: self._num_angles)
    num_dihedrals = property(lambda self: self._num_dihedrals)
    num_impropers = property(lambda self: self._num_impropers)
    particle_ids = property(lambda self: self._particle_ids)
    particle_types = property(lambda self: self._particle_types)
    particle_names = property(lambda self: self._particle_names)
    molecule_ids = property(lambda self: self._molecule_ids)
    molecule_types = property(lambda self: self._molecule_types)
    chain_ids = property(lambda self: self._chain_ids)

    @property
    def topology(self) -> Topology:
        return copy(self._topology)
