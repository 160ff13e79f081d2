"""GPU tests of the analysers: the device pair-distance histograms equal numpy.histogram's
counts bit for bit, and RDFAnalyser / CoordinationNumberAnalyser results equal the float64
restatement of the reference's loops (oracle/analyse.py) on frames of a water run."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def water_trajectory(frames=4, every=25):
    from mdpy_amd import _lib as L, systems
    from mdpy_amd.core import Topology, Trajectory
    from mdpy_amd.engine import Engine
    s = systems.water_box_small()
    e = Engine(s, skin=1.0)
    e.set_positions(s.positions)
    e.configure_integrator(L.INT_LANGEVIN, dt=1.0, temperature=300.0, friction=0.01, seed=3)
    t = Trajectory(Topology.from_system(s))
    t.set_pbc_matrix(s.pbc_matrix)
    terms = L.TERM_LJ | L.TERM_PME | L.TERM_BOND | L.TERM_ANGLE
    for _ in range(frames):
        e.integrate(terms, every, True, every)
        t.append(e.positions().astype(np.float64))
    e.close()
    return s, t


def test_pair_histogram_bit_exact():
    from mdpy_amd.analyser.pair_histogram import pair_histograms
    from oracle import analyse as A
    say("case: pair histograms, water648, 4 frames, all atoms x oxygens")
    s, t = water_trajectory()
    ids1 = list(range(0, s.num_atoms, 7))
    ids2 = list(range(0, s.num_atoms, 3))
    got, edges = pair_histograms(t, ids1, ids2, 9.0, 90)
    ref, ref_edges = A.pair_counts(np.asarray(t.positions, np.float64), s.box, ids1, ids2, 9.0, 90)
    assert np.array_equal(edges, ref_edges)
    assert np.array_equal(got, ref)
    assert got.sum() > 0


def test_rdf_and_coordination_number_match_oracle():
    from mdpy_amd.analyser import CoordinationNumberAnalyser, RDFAnalyser
    from oracle import analyse as A
    say("case: RDF and coordination number O-O and O-H, water648")
    s, t = water_trajectory()
    x = np.asarray(t.positions, np.float64)
    o = [i for i in range(s.num_atoms) if s.types[i] == s.types[0]]
    h = [i for i in range(s.num_atoms) if s.types[i] != s.types[0]]
    c_o = [{"particle type": [[s.types[0]]]}]
    c_h = [{"not particle type": [[s.types[0]]]}]
    for c1, c2, i1, i2 in ((c_o, c_o, o, o), (c_o, c_h, o, h)):
        r = RDFAnalyser(c1, c2, 9.0, 60).analysis(t)
        m, sd, edge = A.rdf(x, s.box, i1, i2, 9.0, 60)
        assert np.array_equal(r.data["mean"], m) and np.array_equal(r.data["std"], sd)
        assert np.array_equal(r.data["bin_edge"], edge)
        cn = CoordinationNumberAnalyser(c1, c2, 9.0, 60).analysis(t)
        m, sd, _ = A.coordination_number(x, s.box, i1, i2, 9.0, 60)
        assert np.array_equal(cn.data["mean"], m) and np.array_equal(cn.data["std"], sd)
    # O-O repulsion: no pair closer than 2.1 A (the LJ energy there is ~70 kcal/mol above its
    # minimum; closest approaches in water at 300 K stay above ~2.2 A)
    cn = CoordinationNumberAnalyser(c_o, c_o, 9.0, 60).analysis(t)
    first = int(round(2.1 / 0.15))
    assert np.all(cn.data["mean"][1:first] == cn.data["mean"][1])
