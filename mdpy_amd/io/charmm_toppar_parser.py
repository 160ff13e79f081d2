"""CHARMM parameter (.prm), topology (.rtf) and combined (.str) files -> the parameter dictionary
the constraints bind (mdpy/io/charmm_toppar_parser.py:21-277).

The dictionary is the reference's interface, so its layout and lookup rules are kept:
    'atom'      atom types in file order (the X wildcard expands over those known so far)
    'mass'      type -> Da
    'charge'    'RESI-NAME' -> e (an ion whose atom name equals its residue: 'NAME')
    'bond'      'A-B' and 'B-A' -> [k, r0]                         (k (r - r0)^2)
    'angle'     'A-B-C' and 'C-B-A' -> [k, theta0, k_ub, u0]        (theta0 in rad)
    'dihedral'  'A-B-C-D' and 'D-C-B-A' -> [[k, n, delta], ...]     (one row per cosine term)
    'improper'  every ordering of the 4 types -> [k, psi0]
    'nonbonded' type -> [eps, sigma] or [eps, sigma, eps14, sigma14], eps = -epsilon_CHARMM
                and sigma = 2 Rmin/2 2^(-1/6) (charmm_toppar_parser.py:16,209-222)
Energies are converted to internal units (kcal/mol -> mdpy_amd.unit.KCAL_PER_MOL), lengths stay
in A, masses in Da, charges in e, angles go to radians. Earlier entries win for bonds, angles,
impropers and wildcard dihedrals; explicit dihedral terms accumulate (a palindromic key such as
CT2-CT2-CT2-CT2 receives each of its terms twice, once per order, as the reference's does).
"""
import itertools
import math

from ..error import FileFormatError
from ..unit import KCAL_PER_MOL, RMIN_TO_SIGMA_FACTOR

USED_BLOCK_LABELS = ["ATOMS", "BONDS", "ANGLES", "DIHEDRALS", "IMPROPER", "NONBONDED"]
UNUSED_BLOCK_LABELS = ["CMAP", "NBFIX", "HBOND", "END"]
BLOCK_LABELS = USED_BLOCK_LABELS + UNUSED_BLOCK_LABELS


def _tokens(line):
    """Fields of a data line, the '!' comment removed."""
    return line.strip().split("!")[0].split()


class CharmmTopparParser:
    def __init__(self, *file_path_list) -> None:
        self._file_path_list = file_path_list
        self._parameters = {"atom": [], "mass": {}, "charge": {}, "bond": {}, "angle": {},
                            "nonbonded": {}, "dihedral": {}, "improper": {}}
        for path in file_path_list:
            # dispatch on the path string as the reference does (charmm_toppar_parser.py:34-45)
            if path.startswith("toppar") or path.endswith("str"):
                self.parse_toppar_file(path)
            elif path.startswith("par") or path.endswith("prm"):
                self.parse_par_file(path)
            elif path.startswith("top") or path.endswith("rtf"):
                self.parse_top_file(path)
            else:
                raise FileFormatError("%s is neither a CHARMM parameter (par*, .prm), topology (top*, .rtf) "
                                      "nor stream (toppar*, .str) file" % path.split("/")[-1])

    @property
    def parameters(self) -> dict:
        return self._parameters.copy()

    # ------------------------------------------------------------------ blocks of a .prm
    @staticmethod
    def _split_blocks(lines):
        """Data rows of each used block: a block runs from a line starting with a block label to
        the next; comment-only and empty lines are dropped, as is the label line itself."""
        blocks, cur = {}, None
        for line in lines:
            if any(line.startswith(lb) for lb in BLOCK_LABELS):
                head = line.lstrip()
                cur = next((lb for lb in USED_BLOCK_LABELS if head.startswith(lb)), None)
                if cur is not None:
                    blocks[cur] = []  # a later block of the same label replaces an earlier one
                continue
            if cur is None or line.lstrip().startswith("!"):
                continue
            toks = _tokens(line)
            if toks:
                blocks[cur].append(toks)
        return blocks

    def parse_par_file(self, file_path):
        with open(file_path) as f:
            lines = f.read().split("\n")
        self._parse_par_blocks(self._split_blocks(lines))

    def _parse_par_blocks(self, blocks):
        self._parse_par_mass_block(blocks.get("ATOMS", []))
        self._parse_par_bond_block(blocks.get("BONDS", []))
        self._parse_par_angle_block(blocks.get("ANGLES", []))
        self._parse_par_dihedral_block(blocks.get("DIHEDRALS", []))
        self._parse_par_improper_block(blocks.get("IMPROPER", []))
        self._parse_par_nonbonded_block(blocks.get("NONBONDED", []))

    def _embed_x_element(self, pair):
        """'A-X-B' -> every key with X replaced by each known atom type (itertools.product order)."""
        if "X" not in pair:
            return [pair]
        atoms = self._parameters["atom"]
        choices = [atoms if t == "X" else [t] for t in pair.split("-")]
        return ["-".join(c) for c in itertools.product(*choices)]

    def _first_wins(self, table, keys, value):
        for k in keys:
            if k not in table:
                table[k] = value

    def _parse_par_mass_block(self, rows):
        for r in rows:  # MASS  -1  TYPE  mass
            self._parameters["atom"].append(r[2])
            self._parameters["mass"][r[2]] = float(r[3])

    def _parse_par_bond_block(self, rows):
        for r in rows:  # A B Kb b0
            value = [float(r[2]) * KCAL_PER_MOL, float(r[3])]
            keys = self._embed_x_element("%s-%s" % (r[0], r[1])) + self._embed_x_element("%s-%s" % (r[1], r[0]))
            self._first_wins(self._parameters["bond"], keys, value)

    def _parse_par_angle_block(self, rows):
        for r in rows:  # A B C Ktheta theta0 [Kub S0]
            if len(r) not in (5, 7):
                continue
            ub = (float(r[5]) * KCAL_PER_MOL, float(r[6])) if len(r) == 7 else (0.0, 0.0)
            value = [float(r[3]) * KCAL_PER_MOL, math.radians(float(r[4])), ub[0], ub[1]]
            keys = (self._embed_x_element("%s-%s-%s" % (r[0], r[1], r[2])) +
                    self._embed_x_element("%s-%s-%s" % (r[2], r[1], r[0])))
            self._first_wins(self._parameters["angle"], keys, value)

    def _parse_par_dihedral_block(self, rows):
        table = self._parameters["dihedral"]
        wild = []
        for r in rows:  # A B C D Kchi n delta
            if "X" in r[:4]:
                wild.append(r)
                continue
            value = [float(r[4]) * KCAL_PER_MOL, float(r[5]), math.radians(float(r[6]))]
            for k in ("%s-%s-%s-%s" % tuple(r[:4]), "%s-%s-%s-%s" % tuple(r[3::-1])):
                table.setdefault(k, []).append(value)
        for r in wild:  # after the explicit terms of this file; only for keys still missing
            value = [float(r[4]) * KCAL_PER_MOL, float(r[5]), math.radians(float(r[6]))]
            keys = (self._embed_x_element("%s-%s-%s-%s" % tuple(r[:4])) +
                    self._embed_x_element("%s-%s-%s-%s" % tuple(r[3::-1])))
            for k in keys:
                if k not in table:
                    table[k] = [value]

    def _parse_par_improper_block(self, rows):
        for r in rows:  # A B C D Kpsi 0 psi0
            value = [float(r[4]) * KCAL_PER_MOL, math.radians(float(r[6]))]
            keys = []
            for perm in itertools.permutations(r[:4]):
                keys.extend(self._embed_x_element("%s-%s-%s-%s" % perm))
            self._first_wins(self._parameters["improper"], keys, value)

    def _parse_par_nonbonded_block(self, rows):
        # the first row after the NONBONDED label continues its option list (cutnb ... e14fac ...)
        for r in rows[1:]:  # TYPE 0 -eps Rmin/2 [0 -eps14 Rmin14/2]
            entry = [-float(r[2]) * KCAL_PER_MOL, float(r[3]) * 2 * RMIN_TO_SIGMA_FACTOR]
            if len(r) != 4:
                entry += [-float(r[5]) * KCAL_PER_MOL, float(r[6]) * 2 * RMIN_TO_SIGMA_FACTOR]
            self._parameters["nonbonded"][r[0]] = entry

    # ------------------------------------------------------------------ residue topology (.rtf)
    @staticmethod
    def _split_residues(lines):
        """ATOM rows of every RESI / PRES entry, keyed by residue name."""
        out, cur = {}, None
        for line in lines:
            if line.startswith("RESI") or line.startswith("PRES"):
                cur = line.split()[1]
                out[cur] = []
            elif cur is not None and line.startswith("ATOM"):
                out[cur].append(_tokens(line))
        return out

    def _parse_top_charge_block(self, residues):
        for res, rows in residues.items():
            for r in rows:  # ATOM name type charge: keyed by residue and atom TYPE
                key = res if r[2] == res else "%s-%s" % (res, r[2])
                self._parameters["charge"][key] = float(r[3])

    def parse_top_file(self, file_path):
        with open(file_path) as f:
            self._parse_top_charge_block(self._split_residues(f.read().split("\n")))

    # ------------------------------------------------------------------ stream files (.str)
    def parse_toppar_file(self, file_path):
        """Topology part up to the first line starting with END, parameters from there on."""
        with open(file_path) as f:
            lines = f.read().split("\n")
        end = next(i for i, line in enumerate(lines) if line.startswith("END"))
        self._parse_top_charge_block(self._split_residues(lines[:end + 1]))
        self._parse_par_blocks(self._split_blocks(lines[end:]))
