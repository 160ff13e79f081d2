"""GPU tests of the CHARMM force-field path from files: the reference's 6PO6 force-field and
integrator tests (mdpy/test/test_charmm_forcefield.py:43-53, test_verlet_integrator.py:33-48,
test_langevin_integrator.py:34-49) restated, and the 6PO6 forces and energies of the created
ensemble against the float64 oracle evaluated with the same parameters."""
import math
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu
DATA = os.path.join(GOLDEN, "data")
FILES = [os.path.join(DATA, f) for f in ("toppar_water_ions_namd.str", "par_all36_prot.prm", "top_all36_prot.rtf")]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ensemble_6po6(box=30.0, files=FILES):
    from mdpy_amd.forcefield import CharmmForcefield
    from mdpy_amd.io import PDBParser, PSFParser
    pdb = PDBParser(os.path.join(DATA, "6PO6.pdb"))
    topology = PSFParser(os.path.join(DATA, "6PO6.psf")).topology
    ff = CharmmForcefield(topology, np.diag(np.ones(3) * box))
    ff.set_param_files(*files)
    ens = ff.create_ensemble()
    return ens, pdb


def test_create_ensemble_6po6():
    say("case: CharmmForcefield 6PO6 create_ensemble")
    ens, _ = ensemble_6po6()
    assert ens.num_constraints == 6


@pytest.mark.parametrize("kind", ["verlet", "langevin"])
def test_integrator_one_step_6po6(kind):
    from mdpy_amd.integrator import LangevinIntegrator, VerletIntegrator
    say("case: 6PO6 one %s step in a 100 A box" % kind)
    ens, pdb = ensemble_6po6(box=100.0, files=[FILES[1]])
    ens.state.cell_list.set_cutoff_radius(12)
    ens.state.set_positions(pdb.positions)
    ens.state.set_velocities_to_temperature(300)
    integ = VerletIntegrator(1) if kind == "verlet" else LangevinIntegrator(1, 300, 0.001)
    integ.integrate(ens, 1)
    # ATOM      1  N   VAL A   1       2.347  -0.970   3.962
    assert ens.state.positions[0, 1] == pytest.approx(-0.970, abs=0.01)
    assert ens.state.positions[0, 0] == pytest.approx(2.347, abs=0.01)


def test_forces_6po6_match_oracle():
    """Every term of the created ensemble (LJ with 1-4 pairs, PME, bonds, angles + Urey-Bradley,
    dihedrals, impropers) against the oracle fed the same parameter arrays."""
    from mdpy_amd.constraint import (CharmmAngleConstraint, CharmmBondConstraint, CharmmDihedralConstraint,
                                     CharmmImproperConstraint, CharmmNonbondedConstraint,
                                     ElectrostaticPMEConstraint)
    from mdpy_amd.unit import ONE_OVER_4PI_EPS0
    from oracle import ewald as EW
    from oracle import mdref as R
    say("case: 6PO6 ensemble forces vs float64 oracle")
    ens, pdb = ensemble_6po6()
    ens.state.set_positions(pdb.positions)
    ens.update()
    F = np.asarray(ens.forces, dtype=np.float64)
    by = {type(c): c for c in ens.constraints}
    box = np.diag(ens.state.pbc_matrix).astype(np.float64)
    pbc = np.diag(box)
    x = R.wrap_positions(np.asarray(pdb.positions, dtype=np.float64), pbc, np.linalg.inv(pbc))
    t = ens.topology
    ep, ei = t.exclusion_csr()
    sp, si = t.scaling_csr()
    q = np.asarray(t.charges, dtype=np.float64).reshape(-1)
    nb = by[CharmmNonbondedConstraint]
    pairs = R.exact_pairs(x, box, nb.cutoff_radius, ep, ei)
    Fl, El = R.lj_pair_terms(x, nb._parameters_list, pairs, box, sp, si)
    pme = by[ElectrostaticPMEConstraint]
    ex = np.array(sorted({(a, int(b)) for a in range(len(q)) for b in ei[ep[a]:ep[a + 1]] if b > a}))
    Fp, Ep, _ = EW.pme_total(x, q, box, pairs, ex, pme.grid, pme.order, pme.beta, ONE_OVER_4PI_EPS0)
    b, a, d, m = (by[k] for k in (CharmmBondConstraint, CharmmAngleConstraint, CharmmDihedralConstraint,
                                  CharmmImproperConstraint))
    Fb, Eb = R.bond_terms(x, b._int_parameters, b._float_parameters, box)
    Fa, Ea = R.angle_terms(x, a._int_parameters, a._float_parameters, box)
    Fd, Ed = R.dihedral_terms(x, d._int_parameters, d._float_parameters, box)
    Fm, Em = R.improper_terms(x, m._int_parameters, m._float_parameters, box)
    Fref = Fl + Fp + Fb + Fa + Fd + Fm
    err = math.sqrt(((F - Fref) ** 2).sum() / (Fref ** 2).sum())
    say("  relative force RMS %.3e" % err)
    assert err < 1e-5
    Eref = El + Ep + Eb + Ea + Ed + Em
    scale = abs(El) + abs(Ep) + abs(Eb) + abs(Ea) + abs(Ed) + abs(Em)
    assert abs(ens.potential_energy - Eref) < 1e-5 * scale
    # per term through Constraint.update (the reference's per-constraint forces)
    for c, Fo in ((b, Fb), (a, Fa), (d, Fd), (m, Fm), (nb, Fl)):
        c.update()
        assert np.sqrt(((np.asarray(c.forces, np.float64) - Fo) ** 2).sum() / (Fo ** 2).sum()) < 1e-5, type(c)
