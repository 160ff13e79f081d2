"""GPU tests of the energy minimisers on the device force path against the float64 oracle
restatement of the reference's minimisers (oracle/minimize.py) on the 648-atom water box, plus
the Simulation.minimize path on 6PO6 built from its CHARMM files."""
import ctypes
import math
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def water_setup():
    from mdpy_amd import _lib as L, systems
    from mdpy_amd.engine import Engine
    from mdpy_amd.unit import ONE_OVER_4PI_EPS0
    from oracle import ewald as EW
    from oracle import mdref as R
    s = systems.water_box_small()
    e = Engine(s, skin=1.0)
    e.set_positions(s.positions)
    terms = L.TERM_LJ | L.TERM_PME | L.TERM_BOND | L.TERM_ANGLE
    pbc = s.pbc_matrix
    inv = np.linalg.inv(pbc)
    ep, ei = s.exclusions()
    ex = np.array(sorted({(a, int(b)) for a in range(s.num_atoms) for b in ei[ep[a]:ep[a + 1]] if b > a}))

    def force_energy(x):
        xw = R.wrap_positions(x, pbc, inv)
        pairs = R.exact_pairs(xw, s.box, s.cutoff, ep, ei)
        F1, E1 = R.lj_pair_terms(xw, s.lj, pairs, s.box)
        F2, E2, _ = EW.pme_total(xw, s.charges, s.box, pairs, ex, s.pme_grid, 4, e.beta, ONE_OVER_4PI_EPS0)
        F3, E3 = R.bond_terms(xw, s.bonds, s.bond_params, s.box)
        F4, E4 = R.angle_terms(xw, s.angles, s.angle_params, s.box)
        return F1 + F2 + F3 + F4, E1 + E2 + E3 + E4

    wrap = lambda x: R.wrap_positions(x, pbc, inv)
    return s, e, terms, force_energy, wrap


def test_steepest_descent_matches_oracle():
    from mdpy_amd import _lib as L
    from oracle import minimize as M
    say("case: steepest descent, water648, 8 iterations")
    s, e, terms, fe, wrap = water_setup()
    it, en = ctypes.c_int(), ctypes.c_double()
    L.call("mdx_minimize_sd", e.ctx, terms, 0.01, 1e-15, 8, ctypes.byref(it), ctypes.byref(en))
    x_gpu = e.positions()
    x_ref, it_ref, en_ref = M.steepest_descent(wrap(s.positions), fe, wrap, 0.01, 1e-15, 8)
    assert it.value == it_ref == 8
    assert en_ref[-1] < en_ref[0]
    assert en.value == pytest.approx(en_ref[-1], rel=1e-5)
    # every atom moved 8 x 0.01 A along float32-accurate force directions
    assert np.abs(R_unwrap(x_gpu - x_ref, s.box)).max() < 1e-5
    e.close()


def R_unwrap(d, box):
    return d - box * np.round(d / box)


def test_conjugate_gradient_matches_oracle():
    """Finite-difference probes (theta = 1e-3) divide float32-rounded force differences by theta,
    so the device and float64 runs agree to ~1e-4 in the step length, not to float32 rounding."""
    from mdpy_amd import _lib as L
    from oracle import minimize as M
    say("case: conjugate gradient, water648, 2 outer iterations of 4 probes")
    s, e, terms, fe, wrap = water_setup()
    it, en = ctypes.c_int(), ctypes.c_double()
    L.call("mdx_minimize_cg", e.ctx, terms, 1e-3, 0.01, 4, 1e-15, 2, ctypes.byref(it), ctypes.byref(en))
    x_gpu = e.positions()
    x_ref, it_ref, en_ref = M.conjugate_gradient(wrap(s.positions), fe, wrap, 1e-3, 0.01, 4, 1e-15, 2)
    say("  energies device %.8e oracle %.8e (start %.8e)" % (en.value, en_ref[-1], en_ref[0]))
    assert it.value == it_ref == 2
    assert en.value == pytest.approx(en_ref[-1], rel=1e-4)
    assert np.abs(R_unwrap(x_gpu - x_ref, s.box)).max() < 1e-3
    e.close()


def test_simulation_minimize_6po6():
    from mdpy_amd import LogDumper, Simulation
    from mdpy_amd.forcefield import CharmmForcefield
    from mdpy_amd.integrator import VerletIntegrator
    from mdpy_amd.io import PDBParser, PSFParser
    from mdpy_amd.minimizer import ConjugateGradientMinimizer, SteepestDescentMinimizer
    say("case: Simulation.minimize on 6PO6 (steepest descent, then conjugate gradient)")
    data = os.path.join(GOLDEN, "data")
    pdb = PDBParser(os.path.join(data, "6PO6.pdb"))
    topology = PSFParser(os.path.join(data, "6PO6.psf")).topology
    ff = CharmmForcefield(topology, np.diag([30.0] * 3))
    ff.set_param_files(os.path.join(data, "par_all36_prot.prm"))
    ens = ff.create_ensemble()
    ens.state.set_positions(pdb.positions)
    ens.update()
    e0 = ens.potential_energy
    sim = Simulation(ens, VerletIntegrator(1.0))
    sim.add_dumpers(LogDumper(dump_frequency=10))
    assert isinstance(sim.minimizer, SteepestDescentMinimizer)
    sim.minimize(1e-6, 50)
    e1 = ens.potential_energy
    assert e1 < e0 and 0 < sim.minimizer.iterations <= 50
    sim.set_minimizer(ConjugateGradientMinimizer())
    sim.minimize(1e-6, 5)
    assert np.isfinite(ens.potential_energy) and sim.minimizer.iterations >= 1
