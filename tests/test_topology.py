"""Topology (mdpy/core/topology.py): the reference's own topology tests restated
(mdpy/test/test_topology.py), and the exclusion / 1-4 tables a joined Topology hands to the
device checked against the oracle's restatement of the add_* rules on the water boxes, the
solvated chain and 6PO6 — and against the tables the GPU parity tests use."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

from mdpy_amd.core import Particle, Topology
from mdpy_amd.error import GeomtryDimError, ModifyJoinedTopologyError, ParticleConflictError


def four(t):
    ps = [Particle(0, "CA", molecule_id=0, molecule_type="ASN"), Particle(1, "CB", molecule_id=0, molecule_type="ASN"),
          Particle(2, "N", molecule_id=0, molecule_type="ASN"), Particle(3, "H", molecule_id=1, molecule_type="HOH")]
    t.add_particles(ps)
    return ps


def test_add_particles_and_join():
    # test_topology.py:36-63
    t = Topology()
    t.add_particles([Particle(0, "CA", mass=1, charge=2), Particle(0, "CB", mass=2), Particle(0, "N")])
    assert t.num_particles == 3
    assert t.particles[1].particle_id == 0 and t.particles[1].matrix_id == 1
    t.join()
    assert t.charges[0] == 2 and t.charges[-1] == 0 and t.masses[1] == 2
    assert t.masses.shape == (3, 1) and t.charges.shape == (3, 1)
    with pytest.raises(ModifyJoinedTopologyError):
        t.add_bond([0, 1])
    t.split()
    t.add_bond([0, 1])


def test_del_particles():
    # test_topology.py:67-92
    t = Topology()
    p1, p2, p3, p4 = four(t)
    t.del_particles([p1, p2, p3])
    assert t.num_particles == 1
    t.del_particles([p4])
    assert t.num_particles == 0
    t.add_particles([p1, p2, p3, p4])
    t.add_bond([0, 1])
    t.add_angle([0, 1, 2])
    t.add_dihedral([0, 1, 2, 3])
    assert (t.num_particles, t.num_bonds, t.num_angles, t.num_dihedrals) == (4, 1, 1, 1)
    t.del_particles([p4])
    assert (t.num_particles, t.num_bonds, t.num_angles, t.num_dihedrals) == (3, 1, 1, 0)


def test_add_del_bond():
    # test_topology.py:93-143
    t = Topology()
    p1, p2, _, _ = four(t)
    t.add_bond([0, 1])
    assert p1.num_bonded_particles == 1 and t.num_bonds == 1
    t.add_bond([1, 2])
    assert p2.num_bonded_particles == 2 and t.num_bonds == 2
    with pytest.raises(GeomtryDimError):
        t.add_bond([0, 1, 2])
    with pytest.raises(ParticleConflictError):
        t.add_bond([1, 1])
    with pytest.raises(ParticleConflictError):
        t.add_bond([1, 4])
    t.del_bond([1, 0])
    assert t.num_bonds == 1 and p1.num_bonded_particles == 0
    t.del_bond([0, 1])  # absent: no change
    assert t.num_bonds == 1
    with pytest.raises(GeomtryDimError):
        t.del_bond([1, 2, 3])
    with pytest.raises(ParticleConflictError):
        t.del_bond([1, 1])
    with pytest.raises(ParticleConflictError):
        t.del_bond([1, 9])


def test_add_del_angle():
    # test_topology.py:145-194
    t = Topology()
    four(t)
    t.add_angle([0, 1, 2])
    t.add_angle([2, 0, 1])
    assert t.num_angles == 2
    with pytest.raises(GeomtryDimError):
        t.add_angle([1, 1, 2, 3])
    with pytest.raises(ParticleConflictError):
        t.add_angle([2, 1, 1])
    with pytest.raises(ParticleConflictError):
        t.add_angle([2, 1, 9])
    t.del_angle([2, 1, 0])  # reversed
    assert t.num_angles == 1
    t.del_angle([0, 2, 1])  # not the remaining one ([2, 0, 1]) in either order
    assert t.num_angles == 1
    with pytest.raises(GeomtryDimError):
        t.del_angle([1, 2, 3, 4])
    with pytest.raises(ParticleConflictError):
        t.del_angle([1, 2, 1])


def test_add_del_dihedral():
    # test_topology.py:196-252
    t = Topology()
    ps = four(t)
    t.add_dihedral([0, 1, 2, 3])
    assert ps[0].scaling_particles[0] == 3 and ps[3].scaling_factors[0] == 1 and t.num_dihedrals == 1
    t.add_dihedral([3, 2, 1, 0])
    assert ps[0].num_scaling_particles == 1 and t.num_dihedrals == 2
    t.add_dihedral([0, 2, 1, 3])
    assert t.num_dihedrals == 3
    with pytest.raises(GeomtryDimError):
        t.add_dihedral([0, 1])
    with pytest.raises(ParticleConflictError):
        t.add_dihedral([1, 1, 2, 0])
    with pytest.raises(ParticleConflictError):
        t.add_dihedral([1, 1, 2, 9])
    t2 = Topology()
    ps2 = four(t2)
    t2.add_dihedral([0, 1, 2, 3])
    t2.del_dihedral([3, 2, 1, 0])
    assert t2.num_dihedrals == 0 and ps2[0].num_scaling_particles == 0
    t2.add_dihedral([0, 1, 2, 3])
    t2.del_dihedral([1, 2, 3, 0])
    assert t2.num_dihedrals == 1
    with pytest.raises(ParticleConflictError):
        t2.del_dihedral([0, 0, 1, 2])


def test_add_del_improper():
    # test_topology.py:254-290
    t = Topology()
    four(t)
    t.add_improper([0, 1, 2, 3])
    t.add_improper([3, 2, 1, 0])
    assert t.num_impropers == 2
    with pytest.raises(GeomtryDimError):
        t.add_improper([1, 2, 3])
    with pytest.raises(ParticleConflictError):
        t.add_improper([2, 1, 1, 3])
    t.del_improper([0, 1, 2, 3])
    assert t.num_impropers == 1
    t.del_improper([0, 1, 2, 3])  # only the exact order is removed
    assert t.num_impropers == 1


def oracle_csr(n, bonds, angles, dihedrals):
    from oracle import mdref as R
    bonded, scaling = R.topology_tables(n, bonds, angles, dihedrals)
    return [sorted(r) for r in bonded], [sorted(r) for r in scaling]


def rows(ptr, idx):
    return [sorted(int(v) for v in idx[ptr[i]:ptr[i + 1]]) for i in range(len(ptr) - 1)]


@pytest.mark.parametrize("name", ["water648", "chain"])
def test_topology_tables_match_oracle_and_parity_inputs(name):
    """The joined Topology's exclusion / 1-4 tables equal the oracle's restatement of the add_*
    rules and the tables the GPU parity tests feed the device (System.exclusions /
    scaling_pairs)."""
    from mdpy_amd import systems
    s = systems.water_box_small() if name == "water648" else systems.solvated_chain()
    t = Topology.from_system(s)
    t.join()
    dihedral_quads = sorted({tuple(int(v) for v in d) for d in s.dihedrals})
    ob, osc = oracle_csr(s.num_atoms, s.bonds, s.angles, dihedral_quads)
    assert rows(*t.exclusion_csr()) == ob
    assert rows(*t.scaling_csr()) == osc
    assert rows(*s.exclusions()) == ob
    assert rows(*s.scaling_pairs()) == osc
    # the joined dense tables hold the same partners, -1 padded
    for i in (0, 1, s.num_atoms // 2, s.num_atoms - 1):
        assert sorted(v for v in t.bonded_particles[i] if v >= 0) == ob[i]
        assert sorted(v for v in t.scaling_particles[i] if v >= 0) == osc[i]


def test_topology_tables_6po6():
    """6PO6 from its PSF: the Topology the parser builds (bonds, angles, dihedrals from the file)
    against the oracle's add_* rules on the same term lists."""
    from mdpy_amd.io import PSFParser
    t = PSFParser(os.path.join(GOLDEN, "data", "6PO6.psf")).topology
    ob, osc = oracle_csr(t.num_particles, t.bonds, t.angles, t.dihedrals)
    assert rows(*t.exclusion_csr()) == ob
    assert rows(*t.scaling_csr()) == osc
    assert sum(len(r) for r in ob) == 2 * (t.num_bonds + t.num_angles)
