"""The mdpy-shaped API on the GPU: the reference's own unit tests, restated against this
package (same topologies, parameters and expected values), plus a Simulation run."""
import json
import math
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu
KAT = json.load(open(os.path.join(GOLDEN, "kat.json")))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def four_particles(types, charges=(0, 0, 0, 0), masses=(12, 14, 1, 12)):
    from mdpy_amd.core import Particle, Topology
    t = Topology()
    t.add_particles([Particle(particle_id=i, particle_type=types[i], particle_name=types[i], mass=masses[i],
                              charge=charges[i]) for i in range(4)])
    return t


def kcal(v):
    from mdpy_amd.unit import KCAL_PER_MOL
    return float(v) * KCAL_PER_MOL


def test_nonbonded_constraint_reference_kat():
    # mdpy/test/test_charmm_nonbonded_constraint.py:93-127
    from mdpy_amd import Ensemble
    from mdpy_amd.constraint import CharmmNonbondedConstraint
    from mdpy_amd.unit import RMIN_TO_SIGMA_FACTOR
    say("case: API nonbonded KAT")
    prm = {}
    for name in ("CA", "NY", "CPT"):
        row = KAT["prm"]["nonbonded_" + name]
        prm[name] = [-kcal(row[2]), float(row[3]) * 2 * RMIN_TO_SIGMA_FACTOR]
    t = four_particles(["CA", "NY", "CPT", "CA"])
    ens = Ensemble(t, np.eye(3) * 30)
    p = np.array([[0, 0, 0], [0, 10, 0], [0, 21, 0], [0, 11, 0]], dtype=np.float32)
    ens.state.set_positions(p)
    c = CharmmNonbondedConstraint(prm)
    c.set_cutoff_radius(9.1)
    ens.add_constraints(c)
    c.update()
    F = c.forces
    assert F.sum() == pytest.approx(0, abs=1e-8)
    eps = math.sqrt(kcal(0.07) * kcal(0.099))
    sr = (1.9924 + 1.86) * RMIN_TO_SIGMA_FACTOR / 9
    fval = -24 * eps / 9 * (2 * sr ** 12 - sr ** 6)
    assert F[0] == pytest.approx(fval * np.array([0, -1.0, 0]), abs=1e-8)
    eps = math.sqrt(kcal(0.07) * kcal(0.2))
    sr = (1.9924 + 1.85) * RMIN_TO_SIGMA_FACTOR
    assert c.potential_energy == pytest.approx(4 * eps * (sr ** 12 - sr ** 6), abs=1e-3)


def test_electrostatic_constraint_reference_kat():
    # mdpy/test/test_electrostatic_constraint.py:197-221
    from mdpy_amd import Ensemble
    from mdpy_amd.constraint import ElectrostaticConstraint
    from mdpy_amd.unit import COULOMB_K
    say("case: API plain Coulomb KAT")
    t = four_particles(["C", "N", "CA", "C"], charges=(1, 2, 0, 0))
    ens = Ensemble(t, np.eye(3) * 100)
    ens.state.set_positions(np.array([[0, 0, 0], [0, 10, 0], [0, 21, 0], [0, 11, 0]], dtype=np.float32))
    c = ElectrostaticConstraint()
    ens.add_constraints(c)
    c.update()
    F = c.forces
    assert F[2, 0] == 0 and F[3, 1] == 0
    f1 = -1 * 2 / COULOMB_K / 100 * np.array([0, 1.0, 0])
    assert F[0] == pytest.approx(f1) and F[1] == pytest.approx(-f1)
    assert c.potential_energy == pytest.approx(2 / COULOMB_K / 10)


def test_bonded_constraints_reference_kat():
    # test_charmm_bond_constraint.py:93-112, test_charmm_angle_constraint.py:93-125,
    # test_charmm_dihedral_constraint.py:95-125, test_charmm_improper_constraint.py:96-134
    from mdpy_amd import Ensemble
    from mdpy_amd.constraint import (CharmmAngleConstraint, CharmmBondConstraint, CharmmDihedralConstraint,
                                     CharmmImproperConstraint)
    say("case: API bonded KATs")
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.float32)
    # bond CA-CA
    row = KAT["prm"]["bond_CA_CA"]
    k, r0 = kcal(row[2]), float(row[3])
    t = four_particles(["CA", "N", "HA1", "CA"])
    t.add_bond([0, 3])
    ens = Ensemble(t, np.eye(3) * 30)
    ens.state.set_positions(pos)
    c = CharmmBondConstraint({"CA-CA": [k, r0]})
    ens.add_constraints(c)
    c.update()
    fv = -2 * k * (1 - r0)
    assert c.num_bonds == 1
    assert c.forces[0, 2] == pytest.approx(-fv) and c.forces[3, 2] == pytest.approx(fv)
    assert c.potential_energy == pytest.approx(k * (1 - r0) ** 2, rel=1e-6)
    # angle CA-CA-CA with Urey-Bradley
    row = KAT["prm"]["angle_CA_CA_CA"]
    ka, th0, ku, u0 = kcal(row[3]), math.radians(float(row[4])), kcal(row[5]), float(row[6])
    t = four_particles(["CA", "CA", "HA1", "CA"])
    t.add_angle([0, 1, 3])
    ens = Ensemble(t, np.eye(3) * 30)
    ens.state.set_positions(pos)
    c = CharmmAngleConstraint({"CA-CA-CA": [ka, th0, ku, u0]})
    ens.add_constraints(c)
    c.update()
    theta = np.pi / 4
    assert c.potential_energy == pytest.approx(ka * (theta - th0) ** 2 + ku * (1 - u0) ** 2, rel=1e-6)
    assert c.forces.sum() == pytest.approx(0, abs=1e-8)
    # dihedral CA-NY-CPT-CA
    row = KAT["prm"]["dihedral_CA_NY_CPT_CA"]
    kd, n, d = kcal(row[4]), float(row[5]), math.radians(float(row[6]))
    t = four_particles(["CA", "NY", "CPT", "CA"])
    t.add_dihedral([0, 1, 2, 3])
    ens = Ensemble(t, np.eye(3) * 30)
    ens.state.set_positions(pos)
    c = CharmmDihedralConstraint({"CA-NY-CPT-CA": [[kd, n, d]]})
    ens.add_constraints(c)
    c.update()
    assert c.num_dihedrals == 1
    assert c.forces.sum() == pytest.approx(0, abs=1e-8)
    # improper HE2-HE2-CE2-CE2
    row = KAT["prm"]["improper_HE2_HE2_CE2_CE2"]
    ki, psi0 = kcal(row[4]), math.radians(float(row[6]))
    t = four_particles(["HE2", "HE2", "CE2", "CE2"])
    t.add_improper([0, 1, 2, 3])
    ens = Ensemble(t, np.eye(3) * 30)
    ens.state.set_positions(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0.5, 0.5, 1]], dtype=np.float32))
    c = CharmmImproperConstraint({"HE2-HE2-CE2-CE2": [ki, psi0]})
    ens.add_constraints(c)
    c.update()
    assert c.potential_energy == pytest.approx(ki * (np.deg2rad(90) - psi0) ** 2, rel=1e-6)
    assert c.forces.sum() == pytest.approx(0, abs=1e-8)


def test_cell_list_api_6po6():
    # mdpy/test/test_cell_list.py:41-58
    from mdpy_amd.core import CellList
    from mdpy_amd.error import CellListPoorDefinedError
    say("case: API CellList 6PO6")
    cl = CellList(np.diag([30, 30, 30]), 9)
    assert cl.cutoff_radius == 9 and cl.pbc_matrix[0, 0] == 30
    with pytest.raises(CellListPoorDefinedError):
        cl.set_cutoff_radius(0)
    with pytest.raises(CellListPoorDefinedError):
        cl.set_cutoff_radius(24)
    pos = np.array(KAT["6PO6_positions"])
    cl.update(pos)
    cell_id = np.floor(np.dot(pos - pos.min(0), cl.cell_inv))
    atoms = [a for a in range(len(pos)) if (cell_id[a] == 0).all()]
    assert atoms
    for k, a in enumerate(atoms):
        assert a == cl.cell_list[0, 0, 0, k]


def test_water_simulation_through_api():
    from mdpy_amd import Ensemble, LogDumper, Simulation, systems
    from mdpy_amd.constraint import (CharmmAngleConstraint, CharmmBondConstraint, CharmmNonbondedConstraint,
                                     ElectrostaticPMEConstraint)
    from mdpy_amd.core import Topology
    from mdpy_amd.integrator import LangevinIntegrator
    say("case: API water648 Simulation, 200 Langevin steps")
    s = systems.water_box_small()
    t = Topology.from_system(s)
    ens = Ensemble(t, s.pbc_matrix)
    ens.state.set_positions(s.positions)
    nb = {"OT": list(s.lj[0, :2]), "HT": list(s.lj[1, :2])}
    ens.add_constraints(CharmmNonbondedConstraint(nb, cutoff_radius=9.0),
                        ElectrostaticPMEConstraint(cutoff_radius=9.0, grid=(24, 24, 24)),
                        CharmmBondConstraint({"OT-HT": list(s.bond_params[0]), "HT-OT": list(s.bond_params[0])}),
                        CharmmAngleConstraint({"HT-OT-HT": list(s.angle_params[0])}))
    ens.state.set_velocities_to_temperature(300)
    ens.update()
    e0 = ens.potential_energy
    assert np.isfinite(e0)
    sim = Simulation(ens, LangevinIntegrator(1.0, 300, 0.01))
    log = LogDumper(dump_frequency=50)
    sim.add_dumpers(log)
    sim.integrate(200)
    assert sim.cur_step == 200 and len(log.records) == 4
    T = [r["temperature"] for r in log.records]
    say("  temperatures", ["%.0f" % x for x in T])
    # the ideal lattice melts in the first ~50 fs and its potential energy heats it to ~600 K
    # (+-20 K at 648 atoms); the thermostat (1/gamma = 100 fs) then pulls it back down
    assert all(100 < x < 1000 for x in T)
    assert T[-1] < T[0]
    x = ens.state.positions
    assert np.all(np.abs(x) <= 18.6 / 2 + 1e-6)  # wrapped to [-L/2, L/2]
