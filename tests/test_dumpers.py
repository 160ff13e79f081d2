"""CPU tests of the trajectory/log output: the PDB writer's records read back by the PDB parser
(mdpy/test/test_pdb_writer.py, test_pdb_dumper.py restated) and the LogDumper's columns
(mdpy/dumper/log_dumper.py:75-208) on synthetic frames."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from mdpy_amd.dumper import Frame, LogDumper, PDBDumper
from mdpy_amd.error import ArrayDimError, DumperPoorDefinedError, FileFormatError
from mdpy_amd.io import PDBParser, PDBWriter, PSFParser
from mdpy_amd.unit import KB, KJ_PER_MOL

DATA = os.path.join(GOLDEN, "data")


def test_pdb_writer_round_trip(tmp_path):
    t = PSFParser(os.path.join(DATA, "6PO6.psf")).topology
    x = PDBParser(os.path.join(DATA, "6PO6.pdb")).positions.astype(np.float64)
    path = str(tmp_path / "w.pdb")
    w = PDBWriter(path, topology=t, pbc_matrix=np.diag([30.0, 31.0, 32.0]))
    w.write(x)
    w.write(np.stack([x + 1.0, x - 2.5]), is_end=True)
    p = PDBParser(path)
    assert p.num_frames == 3 and p.num_particles == 49
    assert np.allclose(p.get_positions(0), x, atol=5e-4)
    assert np.allclose(p.get_positions(2), x - 2.5, atol=5e-4)
    assert np.allclose(np.diag(p.pbc_matrix), [30, 31, 32])
    # the reference's ATOM layout starts names one column late (pdb_writer.py:28-31), so a parser
    # reading columns 13-16 keeps the first three characters of a four-character name
    assert p.particle_names == [q.particle_name[:3] for q in t.particles]
    assert p.molecule_types[20] == t.particles[20].molecule_type
    lines = open(path).read().split("\n")
    atom = next(line for line in lines if line.startswith("ATOM"))
    assert atom[12:16].strip() == "N" and atom[17:20] == "VAL" and atom[21] == "A"
    assert sum(line.startswith("TER") for line in lines) == 3 and lines[-1] == "END"
    with pytest.raises(ArrayDimError):
        w.write(x[:10])
    with pytest.raises(FileFormatError):
        PDBWriter(str(tmp_path / "w.pd"))


def test_dumper_arguments(tmp_path):
    with pytest.raises(FileFormatError):
        LogDumper(str(tmp_path / "a.txt"), 10)
    with pytest.raises(DumperPoorDefinedError):
        LogDumper(str(tmp_path / "a.log"), 0)
    with pytest.raises(DumperPoorDefinedError):
        LogDumper(str(tmp_path / "a.log"), 10, rest_time=True)
    with pytest.raises(FileFormatError):
        PDBDumper(str(tmp_path / "a.log"), 10)


def test_log_dumper_columns(tmp_path):
    class Sim:
        pass
    path = str(tmp_path / "run.log")
    log = LogDumper(path, 100, step=True, sim_time=True, potential_energy=True, kinetic_energy=True,
                    total_energy=True, temperature=True, volume=True, density=True, sim_speed=True, rest_time=True,
                    total_step=300)
    n = 30
    masses = np.full((n, 1), 18.0)
    pbc = np.diag([10.0, 10.0, 10.0])
    for k, (ep, ek, wall) in enumerate(((-5.0, 1.0, 0.0), (-6.0, 1.5, 2.0), (-6.5, 1.2, 4.0))):
        log.dump_frame(Frame(100 * k, None, ep * KJ_PER_MOL, ek * KJ_PER_MOL, pbc, masses, 2.0, wall), Sim())
    rows = [r.rstrip("\t").split("\t") for r in open(path).read().strip().split("\n")]
    assert rows[0][0] == "Step" and rows[0][2] == "E_p (kj/mol)" and rows[0][-1] == "Rest Time"
    assert rows[2][0] == "100" and rows[2][1] == "%.2e" % (100 * 2.0 * 1e-6)
    assert rows[2][2] == "-6.000" and rows[2][3] == "1.500" and rows[2][4] == "-4.500"
    assert float(rows[2][5]) == pytest.approx(2 * 1.5 * KJ_PER_MOL / (3 * n * KB), abs=0.01)
    assert rows[2][6] == "1.000"  # 1000 A^3 = 1 nm^3
    assert float(rows[2][7]) == pytest.approx(30 * 18 / 1000 * 1.66053904, abs=1e-3)
    # 100 steps of 2 fs in 2 s of wall time: 0.1 ps/s = 8.64 ns/day; 200 steps left take 4 s
    assert rows[2][8] == "8.6400" and rows[2][9] == "0:00:04"
    assert log.records[-1]["total_energy"] == pytest.approx(-5.3 * KJ_PER_MOL)


def test_hdf5_dumper_frames(tmp_path):
    """mdpy/dumper/hdf5_dumper.py: the first dumped frame writes the ensemble's topology and box,
    every frame its positions; read back with the HDF5 parser."""
    from mdpy_amd.dumper import HDF5Dumper
    from mdpy_amd.io import HDF5Parser

    class Ens:
        topology = PSFParser(os.path.join(DATA, "6PO6.psf")).topology

    class Sim:
        ensemble = Ens()
    with pytest.raises(FileFormatError):
        HDF5Dumper(str(tmp_path / "a.pdb"), 10)
    x = PDBParser(os.path.join(DATA, "6PO6.pdb")).positions.astype(np.float64)
    path = str(tmp_path / "run.hdf5")
    d = HDF5Dumper(path, 10)
    pbc = np.diag([30.0, 31.0, 32.0])
    for k in range(3):
        d.dump_frame(Frame(10 * k, x + k, 0.0, 0.0, pbc, None, 1.0, 0.0), Sim())
    p = HDF5Parser(path)
    assert p.num_frames == 3 and p.topology.num_particles == 49 and p.topology.num_bonds == 49
    np.testing.assert_allclose(p.positions[2], x + 2, rtol=1e-6)
    np.testing.assert_allclose(p.pbc_matrix, pbc)
