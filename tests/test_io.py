"""CPU tests of the file readers and the CHARMM force-field wiring: the reference's own
known-answer tests (mdpy/test/test_charmm_toppar_parser.py:29-140, test_psf_parser.py,
test_pdb_parser.py, test_charmm_forcefield.py:27-41) restated on the data files they use,
which are kept under tests/golden/data."""
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN
from mdpy_amd.error import (ArrayDimError, FileFormatError, ParameterPoorDefinedError, ParserPoorDefinedError,
                            PBCPoorDefinedError)
from mdpy_amd.io import CharmmTopparParser, PDBParser, PSFParser
from mdpy_amd.unit import KCAL_PER_MOL, RMIN_TO_SIGMA_FACTOR

DATA = os.path.join(GOLDEN, "data")
STR, PRM, RTF_PROT, RTF_NA = (os.path.join(DATA, f) for f in
                              ("toppar_water_ions_namd.str", "par_all36_prot.prm", "top_all36_prot.rtf",
                               "top_all36_na.rtf"))


def kcal(v):
    return v * KCAL_PER_MOL


def test_toppar_parser_rejects_unknown_files():
    with pytest.raises(FileFormatError):
        CharmmTopparParser("x.p")
    with pytest.raises(FileFormatError):
        CharmmTopparParser("to.p")


def test_embed_x_element():
    c = CharmmTopparParser()
    c._parameters["atom"] = "C CA N H S".split()
    assert c._embed_x_element("SS-CS") == ["SS-CS"]
    r = c._embed_x_element("X-A")
    assert len(r) == 5 and r[0] == "C-A" and r[-1] == "S-A"
    r = c._embed_x_element("X-A-X")
    assert len(r) == 25 and r[0] == "C-A-C" and r[5] == "CA-A-C" and r[-1] == "S-A-S"
    r = c._embed_x_element("X-A-X-X")
    assert len(r) == 125 and r[2] == "C-A-C-N"


def test_parse_par_file():
    c = CharmmTopparParser()
    c.parse_par_file(PRM)
    p = c.parameters
    assert p["mass"]["H"] == 1.008
    assert p["bond"]["CE1-CE1"][0] == kcal(440)
    assert p["bond"]["CC-CP1"][1] == 1.49  # listed as CP1 CC: the opposite order is a key too
    assert p["angle"]["NH2-CT1-HB1"][0] == kcal(38)
    assert p["angle"]["CPH1-CPH1-CT2"][1] == pytest.approx(np.deg2rad(130), rel=1e-15)
    assert p["angle"]["NH2-CT1-HB1"][2] == kcal(50)
    assert p["angle"]["CAI-CA-HP"][3] == 2.1525
    assert p["dihedral"]["H-NH2-CT1-HB1"][0][0] == kcal(0.11)
    assert p["dihedral"]["CT2-CT2-CT2-CT2"][0][1] == 2
    assert p["dihedral"]["C-CT1-NH1-H"][0][2] == 0
    assert p["improper"]["HR1-NR1-NR2-CPH2"][0] == kcal(0.5)
    assert p["improper"]["O-NH2-CT1-CC"][1] == 0
    assert p["nonbonded"]["CA"][0] == kcal(0.07)
    assert p["nonbonded"]["SM"][1] == pytest.approx(1.975 * RMIN_TO_SIGMA_FACTOR * 2)
    assert p["nonbonded"]["CP1"][2] == kcal(0.01)
    assert p["nonbonded"]["CP2"][3] == pytest.approx(1.9 * RMIN_TO_SIGMA_FACTOR * 2)


def test_dihedral_terms_accumulate_per_order():
    # explicit multi-term dihedrals keep every term, once per order; a palindromic key therefore
    # holds each term twice (charmm_toppar_parser.py:158-176 appends for both orders)
    c = CharmmTopparParser(PRM)
    p = c.parameters
    rows = [r for r in open(PRM).read().split("\n") if r.split()[:4] == ["CT2", "CT2", "CT2", "CT2"]]
    assert len(rows) >= 2
    assert len(p["dihedral"]["CT2-CT2-CT2-CT2"]) == 2 * len(rows)
    # a wildcard entry fills only keys no explicit entry set
    x_rows = [r.split() for r in open(PRM).read().split("\n") if r.split()[:1] == ["X"] and len(r.split()) >= 7]
    assert x_rows


def test_parse_top_file():
    c = CharmmTopparParser()
    c.parse_top_file(RTF_NA)
    p = c.parameters
    assert p["charge"]["GUA-NN2B"] == -0.02
    assert p["charge"]["3POM-ON3"] == -0.78
    assert p["charge"]["URA-CN7B"] == 0.14


def test_parse_toppar_file():
    c = CharmmTopparParser()
    c.parse_toppar_file(STR)
    p = c.parameters
    assert p["charge"]["LIT"] == 1.0
    assert p["mass"]["RUB"] == 85.4678
    assert p["bond"]["HT-HT"][0] == 0
    assert p["bond"]["OX-HX"][1] == 0.97
    assert p["angle"]["HT-OT-HT"][0] == kcal(55)
    assert p["angle"]["HT-OT-HT"][1] == pytest.approx(np.deg2rad(104.52), rel=1e-15)
    assert p["nonbonded"]["OX"][0] == kcal(0.12)
    assert p["nonbonded"]["BAR"][1] == pytest.approx(1.890 * 2 * RMIN_TO_SIGMA_FACTOR)


def test_parse_multiple_files():
    p = CharmmTopparParser(STR, PRM, RTF_NA).parameters
    assert p["charge"]["GUA-NN2B"] == -0.02 and p["charge"]["LIT"] == 1.0
    assert p["mass"]["RUB"] == 85.4678 and p["mass"]["H"] == 1.008
    assert p["bond"]["HT-HT"][0] == 0 and p["bond"]["CE1-CE1"][0] == kcal(440)


def _psf_atom_rows(path):
    lines = open(path).read().split("\n")
    k = next(i for i, line in enumerate(lines) if "!NATOM" in line)
    n = int(lines[k].split()[0])
    return [lines[k + 1 + i].split() for i in range(n)]


def test_psf_parser_6po6():
    path = os.path.join(DATA, "6PO6.psf")
    psf = PSFParser(path)
    assert (psf.num_particles, psf.num_bonds, psf.num_angles, psf.num_dihedrals, psf.num_impropers) == \
        (49, 49, 87, 122, 5)
    rows = _psf_atom_rows(path)
    assert psf.particle_ids[1] == 2 and psf.particle_names[3] == "HT3" and psf.particle_types[0] == "NH3"
    assert psf.molecule_types[21] == rows[21][3] and psf.molecule_ids[30] == int(rows[30][2])
    assert psf.chain_ids[0] == "A"
    assert psf._charges[2] == pytest.approx(0.33) and psf._masses[5] == pytest.approx(1.008)
    t = psf.topology
    assert (t.num_particles, t.num_bonds, t.num_angles, t.num_dihedrals, t.num_impropers) == \
        (psf.num_particles, psf.num_bonds, psf.num_angles, psf.num_dihedrals, psf.num_impropers)
    # the first bond of the file, 1-based there
    lines = open(path).read().split("\n")
    k = next(i for i, line in enumerate(lines) if "!NBOND" in line)
    a, b = (int(v) - 1 for v in lines[k + 1].split()[:2])
    assert t.bonds[0] == [a, b]
    with pytest.raises(FileFormatError):
        PSFParser("x.ps")


def test_pdb_parser_6po6():
    path = os.path.join(DATA, "6PO6.pdb")
    pdb = PDBParser(path)
    assert pdb.num_particles == 49 and pdb.num_frames == 1
    # ATOM      1  N   VAL A   1       2.347  -0.970   3.962  1.00  0.00      A    N
    assert pdb.positions[0] == pytest.approx([2.347, -0.970, 3.962])
    assert pdb.particle_ids[2] == 3 and pdb.particle_names[0] == "N" and pdb.particle_types[0] == "N"
    assert pdb.particle_types[4] == "C"  # CA: carbon by the name-based guess, not calcium
    info = pdb.get_particle_info(1)
    assert info["molecule_type"] == "VAL" and info["chain_id"] == "A" and info["molecule_id"] == 1
    assert info["position"][1] == pytest.approx(-0.970)
    assert pdb.pbc_matrix is None  # no CRYST1 record
    with pytest.raises(FileFormatError):
        PDBParser("test.pd")
    with pytest.raises(ParserPoorDefinedError):
        PDBParser(path, is_parse_all=False).positions
    with pytest.raises(ArrayDimError):
        PDBParser(path, is_parse_all=False).get_positions(100)
    with pytest.raises(ArrayDimError):
        PDBParser(path, is_parse_all=False).get_positions(0, 101)


def test_charmm_forcefield_parameters_6po6():
    from mdpy_amd.forcefield import CharmmForcefield
    topology = PSFParser(os.path.join(DATA, "6PO6.psf")).topology
    with pytest.raises(PBCPoorDefinedError):
        CharmmForcefield(topology, np.diag([0, 30, 30]))
    ff = CharmmForcefield(topology, np.diag([30, 30, 30]))
    ff.set_param_files(STR, PRM, RTF_PROT)
    ff.check_parameters()
    # a type with no nonbonded entry is reported
    ff._parameters["nonbonded"] = {k: v for k, v in ff._parameters["nonbonded"].items() if k != "NH3"}
    with pytest.raises(ParameterPoorDefinedError):
        ff.check_parameters()


# ------------------------------------------------------------------ DCD (mdpy/test/test_dcd_parser.py)
def test_dcd_parser_attributes():
    # test_dcd_parser.py:25-31, on the reference's own file (an OpenMM trajectory)
    from mdpy_amd.io import DCDParser
    p = DCDParser(os.path.join(DATA, "test_dcd_parser.dcd"))
    assert p.positions.shape[0] == 20
    assert p.positions.shape[1] == 14984
    assert p.pbc_matrix[0, 0] == 100
    # the coordinates lie in the file's 100 x 40 x 40 A box (x, y, z records in that order)
    x = p.positions
    for d, edge in enumerate((100.0, 40.0, 40.0)):
        assert -2.0 < x[..., d].min() < 2.0 and edge - 2.0 < x[..., d].max() < edge + 2.0
    np.testing.assert_array_equal(np.diag(p.pbc_matrix), [100.0, 40.0, 40.0])
    assert np.count_nonzero(p.pbc_matrix - np.diag(np.diag(p.pbc_matrix))) == 0


def test_dcd_parser_exceptions():
    # test_dcd_parser.py:33-46
    from mdpy_amd.io import DCDParser
    path = os.path.join(DATA, "test_dcd_parser.dcd")
    with pytest.raises(FileFormatError):
        DCDParser("te.dc")
    with pytest.raises(ParserPoorDefinedError):
        DCDParser(path, is_parse_all=False).positions
    with pytest.raises(ArrayDimError):
        DCDParser(path, is_parse_all=False).get_positions(100)
    with pytest.raises(ArrayDimError):
        DCDParser(path, is_parse_all=False).get_positions(100, 101)
    lazy = DCDParser(path, is_parse_all=False)
    full = DCDParser(path)
    np.testing.assert_array_equal(lazy.get_positions(7), full.positions[7])
    np.testing.assert_array_equal(lazy.get_positions(0, 19), full.positions[[0, 19]])


def _dcd_bytes(x, cell, order):
    """A DCD file written record by record (frames x atoms x 3 float32, a 6-double unit cell)."""
    import struct
    i4, f4, f8 = order + "i", order + "f", order + "d"

    def rec(payload):
        return struct.pack(i4, len(payload)) + payload + struct.pack(i4, len(payload))
    ic = [len(x), 0, 1, len(x), 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 24]
    head = b"CORD" + b"".join(struct.pack(i4, v) for v in ic[:9]) + struct.pack(f4, 0.5) + \
        b"".join(struct.pack(i4, v) for v in ic[10:])
    out = rec(head) + rec(struct.pack(i4, 1) + b"synthetic".ljust(80, b" ")) + rec(struct.pack(i4, x.shape[1]))
    for frame in x:
        out += rec(b"".join(struct.pack(f8, v) for v in cell))
        for d in range(3):
            out += rec(frame[:, d].astype(order + "f4").tobytes())
    return out


@pytest.mark.parametrize("order", ["<", ">"])
def test_dcd_parser_byte_order_and_triclinic_cell(tmp_path, order):
    """Both byte orders; a triclinic cell given in degrees (a, gamma, b, beta, alpha, c)."""
    from mdpy_amd.io import DCDParser
    rng = np.random.default_rng(0)
    x = rng.uniform(0, 20, size=(3, 7, 3)).astype(np.float32)
    cell = (20.0, 80.0, 22.0, 90.0, 90.0, 25.0)  # gamma = 80 degrees
    path = str(tmp_path / "t.dcd")
    with open(path, "wb") as f:
        f.write(_dcd_bytes(x, cell, order))
    p = DCDParser(path)
    np.testing.assert_array_equal(p.positions, x)
    m = p.pbc_matrix
    g = math.radians(80.0)
    np.testing.assert_allclose(m[0], [20.0, 0, 0], atol=1e-5)
    np.testing.assert_allclose(m[1], [22.0 * math.cos(g), 22.0 * math.sin(g), 0], atol=1e-5)
    np.testing.assert_allclose(m[2], [0, 0, 25.0], atol=1e-5)
    assert abs(p.time_step - 0.5) < 1e-7


# ------------------------------------------------------------------ HDF5 (mdpy/test/test_hdf5_*.py)
def test_hdf5_parser_attributes():
    # test_hdf5_parser.py:28-40, on the reference's own file (written by h5py)
    from mdpy_amd.io import HDF5Parser
    p = HDF5Parser(os.path.join(DATA, "test_hdf5_parser.hdf5"))
    assert p.topology.num_particles == 49
    assert p.topology.num_bonds == 49
    assert p.positions.shape[0] == 2
    assert p.positions.shape[1] == 49
    assert p.pbc_matrix[0, 0] == 10
    assert p.trajectory.num_frames == 2
    # the same molecule as the 6PO6 files: masses, charges and the bond graph per particle (the
    # file was written by an older mdpy, which stored the PSF atom name as particle_type)
    t = PSFParser(os.path.join(DATA, "6PO6.psf")).topology
    assert [q.particle_type for q in p.topology.particles] == [q.particle_name for q in t.particles]
    np.testing.assert_allclose([q.mass for q in p.topology.particles], [q.mass for q in t.particles], rtol=1e-6)
    np.testing.assert_allclose([q.charge for q in p.topology.particles], [q.charge for q in t.particles], rtol=1e-6)
    assert sorted(tuple(sorted(b)) for b in p.topology.bonds) == sorted(tuple(sorted(b)) for b in t.bonds)


def test_hdf5_parser_exceptions():
    # test_hdf5_parser.py:43-58
    from mdpy_amd.io import HDF5Parser
    path = os.path.join(DATA, "test_hdf5_parser.hdf5")
    with pytest.raises(FileFormatError):
        HDF5Parser("te.hd")
    with pytest.raises(ParserPoorDefinedError):
        HDF5Parser(path, is_parse_all=False).positions
    with pytest.raises(ParserPoorDefinedError):
        HDF5Parser(path, is_parse_all=False).trajectory
    with pytest.raises(ArrayDimError):
        HDF5Parser(path, is_parse_all=False).get_positions(100)
    with pytest.raises(ArrayDimError):
        HDF5Parser(path, is_parse_all=False).get_positions(100, 101)
    np.testing.assert_array_equal(HDF5Parser(path, is_parse_all=False).get_positions(1), HDF5Parser(path).positions[1])


def test_hdf5_writer_round_trip(tmp_path):
    # test_hdf5_writer.py:36-45 (6PO6 topology and positions written twice), read back
    from mdpy_amd.io import HDF5Parser, HDF5Writer
    with pytest.raises(FileFormatError):
        HDF5Writer("te.hd")
    t = PSFParser(os.path.join(DATA, "6PO6.psf")).topology
    x = PDBParser(os.path.join(DATA, "6PO6.pdb")).positions
    path = str(tmp_path / "w.hdf5")
    # the reference passes Quantity(np.diag([1] * 3), nanometer); raw numbers are in Angstrom here
    w = HDF5Writer(path, topology=t, pbc_matrix=np.diag([10.0] * 3))
    w.write(x)
    w.write(x + 1)
    p = HDF5Parser(path)
    assert p.num_frames == 2 and p.topology.num_particles == 49
    np.testing.assert_array_equal(p.positions[1], (x + 1).astype(p.positions.dtype))
    np.testing.assert_allclose(p.pbc_matrix, np.diag([10.0] * 3))
    for a, b in zip(p.topology.particles, t.particles):
        assert (a.particle_type, a.particle_name, a.molecule_type, a.chain_id, a.molecule_id) == \
            (b.particle_type, b.particle_name, b.molecule_type, b.chain_id, b.molecule_id)
    assert p.topology.bonds == t.bonds and p.topology.angles == t.angles
    assert p.topology.dihedrals == t.dihedrals and p.topology.impropers == t.impropers


def test_hdf5_container_round_trip(tmp_path):
    """The HDF5 container writer read back by the reader: nested groups (one of 20 members, more
    than a default symbol node holds), scalars, int/float arrays of several ranks and variable-
    length UTF-8 strings."""
    from mdpy_amd.io.hdf5_format import H5Reader, H5Writer
    rng = np.random.default_rng(0)
    big = {"d%02d" % k: rng.standard_normal(k + 1) for k in range(20)}
    tree = {"g": {"h": big, "s": ["N", "CA", "Ω-x", ""], "i": np.arange(12, dtype=np.int32).reshape(3, 4)},
            "x": np.float64(2.5), "n": np.int64(-7), "f": rng.standard_normal((2, 3, 4)).astype(np.float32)}
    path = str(tmp_path / "c.hdf5")
    H5Writer().write(path, tree)
    r = H5Reader(path)
    assert sorted(r.keys("/")) == ["f", "g", "n", "x"] and sorted(r.keys("g/h")) == sorted(big)
    for k, v in big.items():
        np.testing.assert_array_equal(r.read("g/h/" + k), v)
    assert [s.decode() for s in r.read("g/s")] == ["N", "CA", "Ω-x", ""]
    np.testing.assert_array_equal(r.read("g/i"), tree["g"]["i"])
    assert r.read("x") == 2.5 and r.read("n") == -7
    np.testing.assert_array_equal(r.read("f"), tree["f"])
    assert "g/h/d05" in r and "g/zz" not in r
