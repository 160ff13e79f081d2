"""Atom selection by conditions (mdpy/utils/select.py:19-336), topological keywords.

A condition list [{keyword: [values...], ...}, ...] selects the union over its dicts of the
intersection over each dict's keywords; a keyword containing 'not' selects the complement.
Trajectories give one list per frame; the steric keywords 'nearby' and 'in sphere' are
evaluated per frame with minimum-image distances (select.py:160-190), 'in cubic' and
'in cylinder' select nothing, as in the reference.
"""
import numpy as np

from ..core import Topology
from ..error import SelectionConditionPoorDefinedError

SELECTION_SUPPORTED_TOPOLOGICAL_KEYWORDS = [
    "all", "particle type", "particle name", "particle id", "molecule type", "molecule id", "chain id",
    "water", "ion", "protein", "backbone", "nucleic acid"]
SELECTION_SUPPORTED_STERIC_KEYWORDS = ["nearby", "in sphere", "in cubic", "in cylinder"]
SELECTION_SUPPORTED_KEYWORDS = SELECTION_SUPPORTED_TOPOLOGICAL_KEYWORDS + SELECTION_SUPPORTED_STERIC_KEYWORDS

# residue names the reference counts as protein (select.py:108-131) and its backbone atom types
PROTEIN_MOLECULE_TYPE_LIST = [
    "ALA", "CYS", "ASP", "GLU", "PHE", "GLY", "HIS", "ILE", "LYS", "LEU", "MET", "ASN", "PRO", "GLN", "ARG",
    "SER", "THR", "VAL", "TRP", "TYR", "PYL", "SEC",
    "DAL", "DSN", "DCY", "DPR", "DVA", "DTH", "DLE", "DIL", "DSG", "DAS", "MED", "DGN", "DGL", "DLY", "DHI",
    "DPN", "DAR", "DTY", "DTR",
    "HSD", "HSE", "HSP", "NTER", "GLYP", "PROP", "ACE", "ACED", "ACP", "ACPD", "NNEU", "CTER", "CNEU", "CTP",
    "CT1", "CT2", "CT3", "ASPP", "GLUP", "LSN", "LINK", "DISU", "HS2", "LIG1", "LIG2", "LIG3"]
PROTEIN_BACKBONE_PARTICLE_TYPE = ["N", "C", "CA", "O"]


def _topology(target):
    from ..core.trajectory import Trajectory
    if isinstance(target, Topology):
        return target
    if isinstance(target, Trajectory):
        return target.topology
    raise TypeError("selection targets a Topology or a Trajectory, not %s" % type(target))


def _where(target, pred):
    return [p.matrix_id for p in _topology(target).particles if pred(p)]


def select_all(target):
    return list(range(_topology(target).num_particles))


def select_particle_type(target, particle_types):
    return _where(target, lambda p: p.particle_type in particle_types)


def select_particle_name(target, particle_names):
    return _where(target, lambda p: p.particle_name in particle_names)


def select_particle_id(target, particle_ids):
    return _where(target, lambda p: p.particle_id in particle_ids)


def select_molecule_type(target, molecule_types):
    return _where(target, lambda p: p.molecule_type in molecule_types)


def select_molecule_id(target, molecule_ids):
    return _where(target, lambda p: p.molecule_id in molecule_ids)


def select_chain_id(target, chain_ids):
    return _where(target, lambda p: p.chain_id in chain_ids)


def select_water(target):
    return []  # the reference's select_water returns no atoms (select.py:97-100)


def select_ion(target=None):
    return []


def select_protein(target):
    return _where(target, lambda p: p.molecule_type in PROTEIN_MOLECULE_TYPE_LIST)


def select_backbone(target):
    return _where(target, lambda p: p.molecule_type in PROTEIN_MOLECULE_TYPE_LIST and
                  p.particle_type in PROTEIN_BACKBONE_PARTICLE_TYPE)


def select_nucleic_acid(target=None):
    return []


def _min_image(d, pbc):
    box = np.diag(np.asarray(pbc, dtype=np.float64))
    s = d / box
    return (s - np.rint(s)) * box


def select_nearby(target, frame, matrix_ids, radius):
    """Atoms closer than radius to any of matrix_ids (each excluding itself), with repeats."""
    x = np.asarray(target.positions[frame], dtype=np.float64)
    out = []
    for m in matrix_ids:
        d = _min_image(x[m] - x, target.pbc_matrix)
        near = [int(k) for k in np.where(np.sqrt((d ** 2).sum(1)) < radius)[0]]
        near.remove(m)
        out.extend(near)
    return out


def select_in_sphere(target, frame, center, radius):
    x = np.asarray(target.positions[frame], dtype=np.float64)
    d = _min_image(np.asarray(center, dtype=np.float64).reshape(3) - x, target.pbc_matrix)
    return [int(k) for k in np.where(np.sqrt((d ** 2).sum(1)) < radius)[0]]


def select_in_cubic(target=None, *args):
    return []


def select_in_cylinder(target=None, *args):
    return []


STERIC_METHODS = {"nearby": select_nearby, "in sphere": select_in_sphere, "in cubic": select_in_cubic,
                  "in cylinder": select_in_cylinder}

METHOD_DICT = {
    "all": select_all, "particle type": select_particle_type, "particle name": select_particle_name,
    "particle id": select_particle_id, "molecule type": select_molecule_type, "molecule id": select_molecule_id,
    "chain id": select_chain_id, "water": select_water, "ion": select_ion, "protein": select_protein,
    "backbone": select_backbone, "nucleic acid": select_nucleic_acid,
}


def _check(conditions, allowed, what):
    if not isinstance(conditions, list) or not conditions or not isinstance(conditions[0], dict):
        raise SelectionConditionPoorDefinedError("a selection condition is a list of dicts")
    for condition in conditions:
        for key in condition:
            if key.split("not")[-1].strip() not in allowed:
                raise SelectionConditionPoorDefinedError("%r is not a supported %s keyword" % (key, what))


def check_selection_condition(conditions):
    _check(conditions, SELECTION_SUPPORTED_KEYWORDS, "selection")


def check_topological_selection_condition(conditions):
    _check(conditions, SELECTION_SUPPORTED_TOPOLOGICAL_KEYWORDS, "topological selection")


def parse_selection_condition(conditions):
    """'key: v1, v2 and key: v or ...' (select.py:260-271)."""
    out = []
    for condition in conditions:
        parts = ["%s: %s" % (key, ", ".join("%s" % v for v in value)) for key, value in condition.items()]
        out.append(" and ".join(parts))
    return " or ".join(out)


def _select_topological(target, conditions):
    n = _topology(target).num_particles
    everyone = set(range(n))
    chosen = set()
    for condition in conditions:
        cur = set(everyone)
        for key, value in condition.items():
            inverse = "not" in key
            key = key.split("not")[-1].strip()
            if key not in SELECTION_SUPPORTED_TOPOLOGICAL_KEYWORDS:
                raise SelectionConditionPoorDefinedError("%r is not a topological selection keyword" % key)
            res = set(METHOD_DICT[key](target, *value))
            cur &= (everyone ^ res) if inverse else res
        chosen |= cur
    return sorted(chosen)


def _select_frame(target, conditions, frame, cache):
    n = target.topology.num_particles
    everyone = set(range(n))
    chosen = set()
    for index, condition in enumerate(conditions):
        cur = set(everyone)
        for key, value in condition.items():
            inverse = "not" in key
            key = key.split("not")[-1].strip()
            if key in SELECTION_SUPPORTED_TOPOLOGICAL_KEYWORDS:
                tag = (index, key, repr(value))
                if tag not in cache:
                    cache[tag] = set(METHOD_DICT[key](target, *value))
                res = cache[tag]
            elif key in SELECTION_SUPPORTED_STERIC_KEYWORDS:
                res = set(STERIC_METHODS[key](target, frame, *value))
            else:
                raise SelectionConditionPoorDefinedError("%r is not a selection keyword" % key)
            cur &= (everyone ^ res) if inverse else res
        chosen |= cur
    return sorted(chosen)


def select(target, conditions):
    """Matrix ids of a Topology (one list) or of each frame of a Trajectory (list of lists)."""
    from ..core.trajectory import Trajectory
    if isinstance(target, Topology):
        return _select_topological(target, conditions)
    if isinstance(target, Trajectory):
        cache = {}
        return [_select_frame(target, conditions, f, cache) for f in range(target.num_frames)]
    raise TypeError("selection targets a Topology or a Trajectory, not %s" % type(target))
