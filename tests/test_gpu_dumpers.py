"""GPU test: Simulation.integrate with a LogDumper, a PDBDumper and an HDF5Dumper writes its frames from
asynchronous device copies; every frame is the state at its step (the last one equals the
ensemble after the run) and the log's energies are the device energies of those steps."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_simulation_async_frames(tmp_path):
    from mdpy_amd import Ensemble, HDF5Dumper, LogDumper, PDBDumper, Simulation, systems
    from mdpy_amd.constraint import (CharmmAngleConstraint, CharmmBondConstraint, CharmmNonbondedConstraint,
                                     ElectrostaticPMEConstraint)
    from mdpy_amd.core import Topology
    from mdpy_amd.integrator import LangevinIntegrator
    from mdpy_amd.io import HDF5Parser, PDBParser
    from mdpy_amd.unit import KJ_PER_MOL
    print("case: Simulation with asynchronous log and PDB frames, water648, 200 steps", flush=True)
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
    sim = Simulation(ens, LangevinIntegrator(1.0, 300, 0.01))
    log_path, pdb_path = str(tmp_path / "run.log"), str(tmp_path / "run.pdb")
    log = LogDumper(log_path, 20, step=True, potential_energy=True, kinetic_energy=True, total_energy=True,
                    temperature=True, sim_speed=True)
    pdb = PDBDumper(pdb_path, 50)
    h5_path = str(tmp_path / "run.hdf5")
    sim.add_dumpers(log, pdb, HDF5Dumper(h5_path, 100))
    sim.integrate(200)
    assert sim.cur_step == 200
    rows = [r.rstrip("\t").split("\t") for r in open(log_path).read().strip().split("\n")]
    assert rows[0][0] == "Step" and len(rows) == 11 and rows[-1][0] == "200"
    assert [int(r[0]) for r in rows[1:]] == list(range(20, 201, 20))
    # the last frame's energies are those of the final positions (evaluated again at the end)
    assert float(rows[-1][1]) == pytest.approx(ens.potential_energy / KJ_PER_MOL, abs=2e-3)
    assert log.records[-1]["potential_energy"] == pytest.approx(ens.potential_energy, rel=1e-6)
    assert log.records[-1]["kinetic_energy"] == pytest.approx(ens.kinetic_energy, rel=1e-9)
    p = PDBParser(pdb_path)
    assert p.num_frames == 4
    last = p.get_positions(3).astype(np.float64)
    assert np.abs(last - np.asarray(ens.state.positions, np.float64)).max() < 1e-3
    assert np.abs(p.get_positions(0) - last).max() > 0.01  # frames differ: the run moved
    h = HDF5Parser(h5_path)
    assert h.num_frames == 2 and h.topology.num_particles == t.num_particles
    assert np.abs(h.get_positions(1) - np.asarray(ens.state.positions, np.float64)).max() < 1e-4
    np.testing.assert_allclose(h.get_positions(0), p.get_positions(1), atol=1e-3)  # step 100 in both files
