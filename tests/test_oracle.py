"""CPU tests that pin the float64 oracle against the reference's own known-answer tests,
finite differences, closed forms and independent algorithms."""
import json
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN
from mdpy_amd.unit import COULOMB_K, KCAL_PER_MOL, ONE_OVER_4PI_EPS0, RMIN_TO_SIGMA_FACTOR
from oracle import ewald as EW
from oracle import mdref as R

KAT = json.load(open(os.path.join(GOLDEN, "kat.json")))


def lj_param(name):
    row = KAT["prm"]["nonbonded_" + name]   # ATOM 0.0 -eps Rmin/2  (par_all36_prot.prm)
    return [-float(row[2]) * KCAL_PER_MOL, float(row[3]) * 2 * RMIN_TO_SIGMA_FACTOR]


def fd_forces(fun, x, h=1e-6):
    F = np.zeros_like(x)
    for i in range(x.shape[0]):
        for k in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, k] += h
            xm[i, k] -= h
            F[i, k] = -(fun(xp)[1] - fun(xm)[1]) / (2 * h)
    return F


# ----------------------------------------------------------------------------- units
def test_constants_match_textbook():
    # 1/(4 pi eps0) = 332.0637 kcal A/(mol e^2) (CODATA); the reference's eps0/N_A inputs
    # (mdpy/unit/__init__.py:56, unit_definition.py:117) land within 3e-5 of it.
    assert ONE_OVER_4PI_EPS0 / KCAL_PER_MOL == pytest.approx(332.0637, rel=3e-5)
    assert KCAL_PER_MOL == pytest.approx(4.18403e-4, rel=1e-5)


# ----------------------------------------------------------------------------- PBC
def test_wrap_and_unwrap_vectors():
    # mdpy/test/test_pbc.py:16-17,26-81, box 10. The reference test runs in the default SINGLE
    # precision, so its box and inverse are float32 (0.1f > 0.1: 5 * 0.1f rounds up to one image).
    pbc = np.diag(np.ones(3) * 10.0).astype(np.float32)
    inv = np.linalg.inv(pbc).astype(np.float32)
    p = np.array([[0, 0, 0], [4, 5, 1], [-4, -1, -5], [6, 8, 9], [8, 0, 1], [-7, -8, 9],
                  [11, 12, 3], [-3, -12., -14]])
    w = R.wrap_positions(p, pbc, inv)
    assert w[0, 0] == 0 and w[3, 0] == -4 and w[1, 1] == -5 and w[2, 2] == 5
    assert w[-1, 1] == -2 and w[-2, 0] == 1 and w[-2, 2] == 3
    with pytest.raises(R.ParticleLossError):
        R.wrap_positions(np.array([[16.0, 0, 1]]), pbc, inv)
    u = R.unwrap_vec(np.array([0, 6, 1.0], dtype=np.float32), pbc, inv)
    assert u[0] == 0 and u[1] == pytest.approx(-4) and u[2] == 1
    u = R.unwrap_vec(np.array([-5, -6, 9.0], dtype=np.float32), pbc, inv)
    assert u[0] == -5 and u[1] == pytest.approx(4) and u[2] == pytest.approx(-1)
    # diagonal shortcut is bit-identical to the matrix form
    rng = np.random.default_rng(0)
    v = rng.uniform(-30, 30, (1000, 3))
    box = np.array([10.0, 11.3, 9.7])
    assert np.array_equal(R.unwrap_diag(v, box), R.unwrap_vec(v, np.diag(box), np.linalg.inv(np.diag(box))))


def test_pbc_dihedral_sign():
    # mdpy/test/test_geometry.py:97-171 (a subset)
    box = np.array([100.0] * 3)
    cases = [([0, 1, 1], [0, 0, 0], [1, 0, 0], [1, 1, 0], -np.pi / 4),
             ([0, 1, 1], [0, 0, 0], [1, 0, 0], [1, 0, -1], -3 * np.pi / 4),
             ([0, 1, 1], [0, 0, 0], [1, 0, 0], [1, 0, 1], np.pi / 4),
             ([0, 0, 0], [1, 0, 0], [0, 1, 0], [0.5, 0.5, 1], np.pi / 2)]
    for a, b, c, d, ref in cases:
        v = R.get_pbc_dihedral(*[np.array([x], float) for x in (a, b, c, d)], box)
        assert v[0] == pytest.approx(ref)


# ----------------------------------------------------------------------------- cell list
def test_cell_list_6po6():
    # mdpy/test/test_cell_list.py:41-58: CellList(diag 30, 9 A) on 6PO6.pdb
    pos = np.array(KAT["6PO6_positions"])
    assert pos.shape == (49, 3)
    R.check_cell_cutoff(np.array([30.0] * 3), 9.0)
    n, w, inv = R.cell_grid(np.array([30.0] * 3), 9.0)
    assert list(n) == [3, 3, 3]
    idx, table = R.cell_list_build(pos, inv, n)
    atoms = [a for a in range(49) if (idx[a] == 0).all()]
    assert len(atoms) != 0
    for k, a in enumerate(atoms):
        assert table[0, 0, 0, k] == a
    assert 23 in R.cell_list_getitem(idx, table, (-1, -1, 0))
    assert 38 in R.cell_list_getitem(idx, table, (-1, -1, -1))
    assert 0 in R.cell_list_getitem(idx, table, 0)
    with pytest.raises(R.CellListPoorDefinedError):
        R.check_cell_cutoff(np.array([30.0] * 3), 24)
    with pytest.raises(R.CellListPoorDefinedError):
        R.check_cell_cutoff(np.array([30.0] * 3), 0)


def test_cell_list_binning_bruteforce():
    # every atom sits in the cell whose [k w, (k+1) w) interval holds its shifted position
    rng = np.random.default_rng(3)
    box = np.array([61.9] * 3)
    x = rng.uniform(-31, 31, (5000, 3))
    n, w, inv = R.cell_grid(box, 12.0)
    assert list(n) == [5, 5, 5]
    idx, table = R.cell_list_build(x, inv, n)
    s = x - x.min(0)
    assert np.all(idx >= 0) and np.all(idx < n)
    assert np.all(s >= idx * w - 1e-9) and np.all(s < (idx + 1) * w + 1e-9)
    for c in np.ndindex(*n):
        members = table[c][table[c] >= 0]
        assert np.all(np.diff(members) > 0)
        assert np.all((idx[members] == np.array(c)).all(1))
    assert (table >= 0).sum() == len(x)


def test_reference_pairs_literal_matches_vectorised():
    d = np.load(os.path.join(GOLDEN, "water648.npz"))
    from mdpy_amd import systems
    s = systems.water_box_small()
    ep, ei = s.exclusions()
    xw = d["wrapped"]
    lit = R.reference_pairs_literal(xw, s.box, s.cutoff, ep, ei, d["cell_index"], d["cell_table"], d["cell_n"])
    vec = {tuple(p) for p in d["ref_pairs"].tolist()}
    assert lit == vec
    assert len(lit) == 95830


def test_reference_stencil_misses_pairs_at_61_9():
    # n = 5, w = 12.78 > L/n: the grid spans L + 2 A but wraps at n, so some in-cutoff
    # pairs are never visited (cell_list.py:76-80 with charmm_nonbonded_constraint.py:78-80).
    rng = np.random.default_rng(0)
    box = np.array([61.9] * 3)
    x = rng.uniform(-30.95, 30.95, (6000, 3))
    ep = np.zeros(6001, dtype=np.int64)
    ei = np.zeros(0, dtype=np.int64)
    n, w, inv = R.cell_grid(box, 12.0)
    idx, _ = R.cell_list_build(x, inv, n)
    exact = R.exact_pairs(x, box, 12.0, ep, ei)
    ref = R.reference_pairs(x, box, 12.0, ep, ei, idx, n)
    miss = 1 - len(ref) / len(exact)
    assert 0 < miss < 2e-3


def test_exact_pairs_bruteforce():
    rng = np.random.default_rng(5)
    box = np.array([20.0, 22.0, 25.0])
    x = rng.uniform(0, 1, (300, 3)) * box
    p = R.exact_pairs(x, box, 9.0)
    i, j = np.triu_indices(300, 1)
    dd = R.unwrap_diag(x[j] - x[i], box)
    r = np.sqrt((dd * dd).sum(1))
    bf = np.stack([i[r <= 9.0], j[r <= 9.0]], 1)
    assert np.array_equal(p, bf)


# ----------------------------------------------------------------------------- force KATs
def test_lj_kat():
    # mdpy/test/test_charmm_nonbonded_constraint.py:93-127
    CA, NY, CPT = lj_param("CA"), lj_param("NY"), lj_param("CPT")
    lj = np.array([CA + CA, NY + NY, CPT + CPT, CA + CA])
    x = np.array([[0, 0, 0], [0, 10, 0], [0, 21, 0], [0, 11, 0.0]])
    box = np.array([30.0] * 3)
    pairs = R.exact_pairs(x, box, 9.1)
    F, E = R.lj_pair_terms(x, lj, pairs, box)
    assert F.sum() == pytest.approx(0, abs=1e-8)
    eps = math.sqrt(0.07 * KCAL_PER_MOL * 0.099 * KCAL_PER_MOL)
    sigma = (1.9924 + 1.86) * RMIN_TO_SIGMA_FACTOR
    sr = sigma / 9
    fval = -24 * eps / 9 * (2 * sr ** 12 - sr ** 6)
    assert F[0] == pytest.approx(fval * np.array([0, -1.0, 0]), abs=1e-8)
    eps = math.sqrt(0.07 * KCAL_PER_MOL * 0.2 * KCAL_PER_MOL)
    sr = (1.9924 + 1.85) * RMIN_TO_SIGMA_FACTOR
    assert E == pytest.approx(4 * eps * (sr ** 12 - sr ** 6), abs=1e-3)


def test_coulomb_kat():
    # mdpy/test/test_electrostatic_constraint.py:197-221
    x = np.array([[0, 0, 0], [0, 10, 0], [0, 21, 0], [0, 11, 0.0]])
    q = np.array([1.0, 2.0, 0.0, 0.0])
    F, E = R.coulomb_allpairs(x, q, np.zeros(5, dtype=np.int64), np.zeros(0, dtype=np.int64),
                              np.array([100.0] * 3), COULOMB_K)
    assert F[2, 0] == 0 and F[3, 1] == 0
    f1 = -1 * 2 / COULOMB_K / 100 * np.array([0, 1.0, 0])
    assert F[0] == pytest.approx(f1) and F[1] == pytest.approx(-f1)
    assert E == pytest.approx(2 / COULOMB_K / 10)


def test_bond_kat():
    # mdpy/test/test_charmm_bond_constraint.py:93-112, CA-CA (par_all36_prot.prm:112)
    row = KAT["prm"]["bond_CA_CA"]
    k, r0 = float(row[2]) * KCAL_PER_MOL, float(row[3])
    x = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    F, E = R.bond_terms(x, np.array([[0, 3]]), np.array([[k, r0]]), np.array([30.0] * 3))
    fv = -2 * k * (1 - r0)
    assert F[1, 0] == 0 and F[2, 1] == 0
    assert tuple(F[0]) == (0, 0, -fv) and tuple(F[3]) == (0, 0, fv)
    assert F.sum() == 0
    assert E == pytest.approx(k * (1 - r0) ** 2)


def test_angle_kat():
    # mdpy/test/test_charmm_angle_constraint.py:93-125, CA-CA-CA (+UB) (prm:397)
    row = KAT["prm"]["angle_CA_CA_CA"]
    k, th0 = float(row[3]) * KCAL_PER_MOL, math.radians(float(row[4]))
    ku, u0 = float(row[5]) * KCAL_PER_MOL, float(row[6])
    x = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    F, E = R.angle_terms(x, np.array([[0, 1, 3]]), np.array([[k, th0, ku, u0]]), np.array([30.0] * 3))
    theta = np.pi / 4
    fval = 2 * k * (np.deg2rad(45) - th0) / np.abs(np.sin(theta))
    v0, v1 = np.array([-1, 0, 0.0]), np.array([-1, 0, 1.0]) / math.sqrt(2)
    f0 = fval * (v1 - v0 * np.cos(theta))
    f2 = fval * (v0 - v1 * np.cos(theta)) / math.sqrt(2)
    fu = 2 * ku * (1 - u0)
    f0 = f0 + fu * np.array([0, 0, 1.0])
    f2 = f2 - fu * np.array([0, 0, 1.0])
    assert F[0] == pytest.approx(f0, abs=1e-8)
    assert F[3] == pytest.approx(f2, abs=1e-8)
    assert F.sum() == pytest.approx(0, abs=1e-8)
    torque = np.cross(v0, F[0]) + np.cross(np.array([-1, 0, 1.0]), F[3])
    assert torque == pytest.approx(np.zeros(3), abs=1e-8)
    assert E == pytest.approx(k * (theta - th0) ** 2 + ku * (1 - u0) ** 2)


def _torque_four(x, F):
    vbc = x[2] - x[1]
    oc, ob = vbc / 2, -vbc / 2
    oa, od = ob + x[0] - x[1], oc + x[3] - x[2]
    return np.cross(oa, F[0]) + np.cross(ob, F[1]) + np.cross(oc, F[2]) + np.cross(od, F[3])


@pytest.mark.parametrize("refcompat", [True, False])
def test_dihedral_kat(refcompat):
    # mdpy/test/test_charmm_dihedral_constraint.py:95-125, CA-NY-CPT-CA (prm:1025)
    row = KAT["prm"]["dihedral_CA_NY_CPT_CA"]
    k, n, d = float(row[4]) * KCAL_PER_MOL, float(row[5]), math.radians(float(row[6]))
    x = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    F, E = R.dihedral_terms(x, np.array([[0, 1, 2, 3]]), np.array([[k, n, d]]), np.array([30.0] * 3),
                            refcompat=refcompat)
    assert F.sum() == pytest.approx(0, abs=1e-8)
    assert _torque_four(x, F) == pytest.approx(np.zeros(3), abs=1e-8)
    phi = R.get_pbc_dihedral(*[x[i:i + 1] for i in range(4)], np.array([30.0] * 3))[0]
    assert E == pytest.approx(k * (1 + np.cos(n * phi - d)))


def test_improper_kat():
    # mdpy/test/test_charmm_improper_constraint.py:96-134, HE2-HE2-CE2-CE2 (prm:2111)
    row = KAT["prm"]["improper_HE2_HE2_CE2_CE2"]
    k, psi0 = float(row[4]) * KCAL_PER_MOL, math.radians(float(row[6]))
    x = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0.5, 0.5, 1]])
    F, E = R.improper_terms(x, np.array([[0, 1, 2, 3]]), np.array([[k, psi0]]), np.array([30.0] * 3))
    assert F.sum() == pytest.approx(0, abs=1e-8)
    assert _torque_four(x, F) == pytest.approx(np.zeros(3), abs=1e-8)
    fval = -2 * k * (np.deg2rad(90) - psi0)
    vab, vbc = x[1] - x[0], x[2] - x[1]
    th_abc = math.acos(np.dot(x[0] - x[1], x[2] - x[1]) / (np.linalg.norm(x[0] - x[1]) * np.linalg.norm(x[2] - x[1])))
    c = np.cross(-vab, vbc)
    fa = fval / (np.linalg.norm(vab) * math.sin(th_abc)) * c / np.linalg.norm(c)
    assert F[0] == pytest.approx(fa, rel=1e-12, abs=1e-15)
    assert E == pytest.approx(k * (np.deg2rad(90) - psi0) ** 2)


# ----------------------------------------------------------------------------- finite differences
X4 = np.array([[0, 0, 0], [1.2, 0.1, 0], [1.5, 1.3, 0.2], [2.6, 1.5, 1.1]]) + 5
BOX = np.array([30.0] * 3)


@pytest.mark.parametrize("name", ["bond", "angle", "dihedral", "improper", "lj", "coulomb", "erfc", "erf_excl"])
def test_forces_are_minus_energy_gradient(name):
    q4 = np.array([[0, 1, 2, 3]])
    pairs = np.array([[0, 2], [0, 3], [1, 3]])
    ch = np.array([1.0, -0.5, 0.3, -0.8])
    funs = {
        "bond": lambda y: R.bond_terms(y, np.array([[0, 2]]), np.array([[0.3, 1.1]]), BOX),
        "angle": lambda y: R.angle_terms(y, np.array([[0, 1, 2]]), np.array([[0.3, 1.9, 0.2, 2.0]]), BOX),
        "dihedral": lambda y: R.dihedral_terms(y, q4, np.array([[0.3, 2.0, 0.5]]), BOX),
        "improper": lambda y: R.improper_terms(y, q4, np.array([[0.3, 0.2]]), BOX),
        "lj": lambda y: R.lj_pair_terms(y, np.array([[1e-3, 3.0, 1e-3, 3.0]] * 4), pairs, BOX),
        "coulomb": lambda y: R.coulomb_allpairs(y, ch, np.array([0, 1, 2, 2, 2]), np.array([1, 0]), BOX, COULOMB_K),
        "erfc": lambda y: EW.direct_terms(y, ch, pairs, BOX, 0.3, 1.0),
        "erf_excl": lambda y: EW.exclusion_terms(y, ch, pairs, BOX, 0.3, 1.0),
    }
    f = funs[name]
    F, _ = f(X4)
    Ff = fd_forces(f, X4)
    assert np.abs(F - Ff).max() <= 1e-6 * max(1.0, np.abs(Ff).max())


def test_reference_dihedral_force_is_not_the_gradient():
    # the reference's force_val carries an extra -k (charmm_dihedral_constraint.py:81)
    q4 = np.array([[0, 1, 2, 3]])
    f = lambda y: R.dihedral_terms(y, q4, np.array([[0.3, 2.0, 0.5]]), BOX, refcompat=True)
    F, _ = f(X4)
    assert np.abs(F - fd_forces(f, X4)).max() > 1e-2


# ----------------------------------------------------------------------------- Ewald / PME
def test_spme_converges_to_bruteforce_ewald():
    rng = np.random.default_rng(1)
    L = np.array([10.0, 11.0, 12.0])
    x = rng.uniform(0, 1, (8, 3)) * L
    q = rng.uniform(-1, 1, 8)
    q -= q.mean()
    Fb, Eb = EW.ewald_reciprocal_bruteforce(x, q, L, 0.35, 1.0, kmax=12)
    F, E = EW.spme_reciprocal(x, q, L, (64, 64, 64), 4, 0.35, 1.0)
    assert abs(E - Eb) / abs(Eb) < 2e-6
    assert np.sqrt(((F - Fb) ** 2).sum() / (Fb ** 2).sum()) < 5e-5
    F6, E6 = EW.spme_reciprocal(x, q, L, (32, 32, 32), 6, 0.35, 1.0)
    assert abs(E6 - Eb) / abs(Eb) < 1e-7


def test_bruteforce_reciprocal_forces_fd():
    rng = np.random.default_rng(2)
    L = np.array([9.0, 9.0, 9.0])
    x = rng.uniform(0, 9, (5, 3))
    q = np.array([1, -1, 0.5, -0.5, 0.0])
    f = lambda y: EW.ewald_reciprocal_bruteforce(y, q, L, 0.4, 1.0, kmax=8)
    F, _ = f(x)
    assert np.abs(F - fd_forces(f, x, 1e-5)).max() < 1e-7


def test_spme_forces_fd():
    rng = np.random.default_rng(7)
    L = np.array([12.0, 12.0, 12.0])
    x = rng.uniform(0, 12, (6, 3))
    q = np.array([1, -1, 0.5, -0.5, 0.3, -0.3])
    f = lambda y: EW.spme_reciprocal(y, q, L, (16, 16, 16), 4, 0.3, 1.0)
    F, _ = f(x)
    assert np.abs(F - fd_forces(f, x, 1e-5)).max() < 1e-7 * max(1, np.abs(F).max())


def test_nacl_madelung_constant_and_beta_independence():
    pos, qs = [], []
    for a in range(2):
        for b in range(2):
            for c in range(2):
                pos.append([a, b, c])
                qs.append(1.0 if (a + b + c) % 2 == 0 else -1.0)
    pos = np.array(pos, float) + 0.25
    qs = np.array(qs)
    Lb = np.array([2.0, 2.0, 2.0])
    for beta in (1.5, 2.5):
        Er = EW.ewald_real_bruteforce(pos, qs, Lb, beta, 1.0, nimg=4)
        _, Ek = EW.ewald_reciprocal_bruteforce(pos, qs, Lb, beta, 1.0, kmax=10)
        Es = EW.self_energy(qs, beta, 1.0)
        # 8 ions, nearest-neighbour distance 1: E = -8/2 * M
        assert -(Er + Ek + Es) / 4 == pytest.approx(1.747565, abs=2e-6)


def test_pme_total_matches_bruteforce_ewald_with_exclusions():
    # direct(r<=rc, non-excluded) + SPME + excl + self == full Ewald of the non-excluded pairs
    rng = np.random.default_rng(11)
    L = np.array([14.0, 14.0, 14.0])
    x = rng.uniform(0, 14, (10, 3))
    q = rng.uniform(-1, 1, 10)
    q -= q.mean()
    ex = np.array([[0, 1], [2, 3]])
    ep = np.zeros(11, dtype=np.int64)
    ei = []
    rows = {0: [1], 1: [0], 2: [3], 3: [2]}
    for a in range(10):
        ep[a + 1] = ep[a] + len(rows.get(a, []))
        ei += rows.get(a, [])
    ei = np.array(ei, dtype=np.int64)
    beta = 0.45
    pairs = R.exact_pairs(x, L, 6.9, ep, ei)
    F, E, parts = EW.pme_total(x, q, L, pairs, ex, (48, 48, 48), 6, beta, 1.0)
    Er = EW.ewald_real_bruteforce(x, q, L, beta, 1.0, nimg=2, excl_keys={0 * 10 + 1, 1 * 10 + 0, 2 * 10 + 3, 3 * 10 + 2})
    _, Ek = EW.ewald_reciprocal_bruteforce(x, q, L, beta, 1.0, kmax=10)
    Eself = EW.self_energy(q, beta, 1.0)
    # true energy with the excluded pairs removed from the home image only:
    # real (excluded home pairs omitted) + reciprocal + self - erf part of the excluded pairs;
    # the PME form differs only by the direct-space cutoff truncation (erfc(beta rc) ~ 1e-5)
    _, Ex = EW.exclusion_terms(x, q, ex, L, beta, 1.0)
    assert parts["background"] == pytest.approx(0.0, abs=1e-12)
    assert E == pytest.approx(Er + Ek + Eself + Ex, rel=1e-4)


# ----------------------------------------------------------------------------- integrators
def test_velocity_verlet_harmonic_oscillator():
    k, m, dt = 0.5, 2.0, 0.01
    force = lambda x: -k * x
    x0, v0 = np.array([[1.0, 0, 0]]), np.zeros((1, 3))
    x, v, _ = R.velocity_verlet(x0, v0, force(x0), [m], force, dt, 1000)
    w = math.sqrt(k / m)
    assert x[0, 0] == pytest.approx(math.cos(w * 10.0), abs=1e-4)
    assert v[0, 0] == pytest.approx(-w * math.sin(w * 10.0), abs=1e-4)


def test_langevin_reference_free_particle_closed_form():
    # zero force, zero noise: the reference's update restarts from x_{n-1}, v_{n-1}
    # (langevin_integrator.py:56-70): two interleaved sequences x_{n+1} = x_{n-1} + b dt v_{n-1}
    dt, gamma = 1.0, 0.01
    integ = R.LangevinReference(dt, 0.0, gamma)
    x0, v0 = np.zeros((1, 3)), np.array([[1.0, 0, 0]])
    pbc = np.diag([1e6] * 3)
    xs, vs, _ = integ.integrate(x0, v0, np.zeros((1, 3)), [1.0], lambda y: np.zeros((1, 3)),
                                lambda s: np.zeros(s), pbc, np.linalg.inv(pbc), 4)
    a = (1 - gamma * dt / 2) / (1 + gamma * dt / 2)
    b = 1 / (1 + gamma * dt / 2)
    # hand recursion of the pair (cur, pre)
    cx, px, cv, pv = 0.0, 0.0, 1.0, 1.0
    for _ in range(4):
        cx, px = px + b * dt * pv, cx
        cv, pv = a * pv, cv
    assert xs[0, 0] == pytest.approx(cx) and vs[0, 0] == pytest.approx(cv)


def _verlet_ref_run(x0, v0, acc, dt, calls):
    """VerletReference on one atom of unit mass under a constant acceleration, in a box too large
    for the wrap to act; `calls` = steps per integrate call (the cache carries x_{n-1} across)."""
    integ = R.VerletReference(dt)
    pbc = np.diag([1e9] * 3)
    inv = np.linalg.inv(pbc)
    x, v = x0.copy(), v0.copy()
    for n in calls:
        x, v = integ.integrate(x, v, [1.0], lambda y: np.array([acc]), pbc, inv, n)
    return x, v


def test_verlet_reference_free_particle_ends_at_half_velocity():
    # verlet_integrator.py:28-50 with F = 0: x_{-1} = x0 - v0 dt, then x_n = x0 + n v0 dt and
    # the output velocity mi(x_n - x_{n-1}) / (2 dt) is v0 / 2 (the reference's central
    # difference over one step divided by 2 dt)
    x0, v0, dt = np.array([[1.0, -2.0, 0.5]]), np.array([[0.3, -0.1, 0.02]]), 0.5
    x, v = _verlet_ref_run(x0, v0, [0.0, 0.0, 0.0], dt, [7])
    assert np.allclose(x, x0 + 7 * v0 * dt, rtol=0, atol=1e-12)
    assert np.allclose(v, v0 / 2, rtol=0, atol=1e-12)


def test_verlet_reference_constant_force_closed_form():
    # x_{-1} = x0 - v0 dt + a dt^2 (start-up with +a dt^2, verlet_integrator.py:31-34) and
    # x_{n+1} = 2 x_n - x_{n-1} + a dt^2 solve to x_n = x0 + (v0 dt - a dt^2/2) n + a dt^2 n^2 / 2,
    # so the output velocity (x_n - x_{n-1}) / (2 dt) is (v0 + a dt (n - 1)) / 2
    a = np.array([0.2, -0.05, 0.01])
    x0, v0, dt = np.array([[0.0, 1.0, -1.0]]), np.array([[0.1, 0.0, -0.3]]), 0.25
    for n in (1, 2, 9):
        x, v = _verlet_ref_run(x0, v0, a, dt, [n])
        xn = x0 + (v0 * dt - a * dt ** 2 / 2) * n + a * dt ** 2 * n ** 2 / 2
        assert np.allclose(x, xn, rtol=0, atol=1e-12)
        assert np.allclose(v, (v0 + a * dt * (n - 1)) / 2, rtol=0, atol=1e-12)
    # the cache (integrator.py:17-22) makes 3 + 6 steps in two calls the same trajectory as 9 in one
    x2, v2 = _verlet_ref_run(x0, v0, a, dt, [3, 6])
    x1, v1 = _verlet_ref_run(x0, v0, a, dt, [9])
    assert np.allclose(x2, x1, rtol=0, atol=1e-12) and np.allclose(v2, v1, rtol=0, atol=1e-12)


def test_gjf_langevin_temperature():
    # harmonic wells, 300 K: <m v^2> -> kT per component (G-JF gives exact configurational T)
    from mdpy_amd.unit import KB
    rng = np.random.default_rng(0)
    n, m, k = 2000, 10.0, 1e-3
    kbt = KB * 300
    force = lambda x: -k * x
    x, v = np.zeros((n, 3)), np.zeros((n, 3))
    x, v, f = R.langevin_gjf(x, v, force(x), np.full(n, m), force, 2.0, kbt, 0.05,
                             lambda s: rng.standard_normal(s), 400)
    assert np.mean(k * x * x) / kbt == pytest.approx(1.0, rel=0.06)


def test_golden_water648_consistency():
    d = np.load(os.path.join(GOLDEN, "water648.npz"))
    from mdpy_amd import systems
    s = systems.water_box_small()
    assert str(d["sha256"]) == s.sha256()
    # exact and reference neighbour sets coincide at n = 3 (every cell is a stencil neighbour)
    assert np.array_equal(d["ref_pairs"], d["exact_pairs"])
    assert d["F_lj"].sum(0) == pytest.approx(np.zeros(3), abs=1e-10)
    # SPME does not conserve momentum exactly; its net force is a small fraction of the total
    assert np.abs(d["F_pme"].sum(0)).max() < 1e-4 * np.abs(d["F_pme"]).sum()


# ----------------------------------------------------------------------------- minimisers
def _harmonic(k, centre):
    def fe(x):
        d = x - centre
        return -k * d, 0.5 * k * float((d * d).sum()) + 1.0  # +1: the relative-change test divides by E
    return fe


def test_steepest_descent_closed_form():
    # steepest_descent_minimizer.py:37-40: every atom steps alpha along its force direction, so an
    # atom 0.35 from the minimum of an isotropic well is 0.35 - k alpha away after k steps
    from oracle import minimize as M
    centre = np.array([[1.0, 2.0, 3.0]])
    x0 = centre + np.array([[0.35, 0.0, 0.0]])
    box = 1e6
    wrap = lambda y: R.wrap_positions(y, np.diag([box] * 3), np.diag([1 / box] * 3))
    x, it, en = M.steepest_descent(x0, _harmonic(2.0, centre), wrap, 0.1, 1e-12, 3)
    assert it == 3 and x[0, 0] == pytest.approx(1.05, abs=1e-12)
    assert en[-1] == pytest.approx(0.5 * 2.0 * 0.05 ** 2 + 1.0, abs=1e-12)


def test_conjugate_gradient_as_written_closed_form():
    # conjugate_gradient_minimizer.py:46-69 on E = k/2 |x - c|^2. Each sub-iteration probes from the
    # already-probed positions: t_s = F(x_s) = k (1 - k theta)^s (c - x0); the probe gives
    # omega = k t_s, so alpha = 1/k; the update cur_f + alpha omega = 2 t_s never meets force_tol,
    # and only the last sub-iteration's d = t_s / k is applied:
    # x = x0 + (1 - k theta)^(S - 1) (c - x0) after S sub-iterations
    from oracle import minimize as M
    rng = np.random.default_rng(1)
    k, theta, S = 3.0, 1e-3, 10
    centre = rng.uniform(-1, 1, (5, 3))
    x0 = centre + rng.uniform(-0.3, 0.3, (5, 3))
    box = 1e6
    wrap = lambda y: R.wrap_positions(y, np.diag([box] * 3), np.diag([1 / box] * 3))
    x, it, en = M.conjugate_gradient(x0, _harmonic(k, centre), wrap, theta, 0.01, S, 1e-9, 1)
    assert it == 1
    assert np.allclose(x, x0 + (1 - k * theta) ** (S - 1) * (centre - x0), rtol=0, atol=1e-9)


# ----------------------------------------------------------------------------- analysers
def test_coordination_number_simple_cubic_closed_form():
    # a simple cubic lattice of spacing a: 6 neighbours at a, 12 at a sqrt 2, 8 at a sqrt 3,
    # 6 at 2a; 45 bins over 2.1a put every shell at least a tenth of a bin from an edge
    from oracle import analyse as A
    a, m = 3.0, 6
    g = np.arange(m) * a
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(1, -1, 3)
    box = np.array([m * a] * 3)
    ids = list(range(x.shape[1]))
    rc, nb = 2.1 * a, 45
    counts, edges = A.pair_counts(x, box, ids[:20], ids, rc, nb)
    w = rc / nb
    expect = {0: 1, int(a / w): 6, int(a * 2 ** 0.5 / w): 12, int(a * 3 ** 0.5 / w): 8, int(2 * a / w): 6}
    for i in range(20):
        assert {k: c for k, c in enumerate(counts[i]) if c} == expect
    mean, std, _ = A.coordination_number(x, box, ids, ids, rc, nb)
    assert mean[-1] == 1 + 6 + 12 + 8 + 6 and np.all(std == 0)
    # rdf of an ideal lattice: zero between shells
    rm, _, _ = A.rdf(x, box, ids, ids, rc, nb)
    assert rm[10] == 0 and rm[int(a / w)] > 0


def test_topology_tables_pins():
    """oracle.mdref.topology_tables on hand-worked cases: a 4-atom chain (1-2 and 1-3 partners
    in insertion order, one 1-4 pair), a repeated dihedral end pair (benzene-like, kept once),
    a three-ring whose angle closes an existing bond (the reference raises), a self bond."""
    from oracle import mdref as R
    b, s = R.topology_tables(4, [(0, 1), (1, 2), (2, 3)], [(0, 1, 2), (1, 2, 3)], [(0, 1, 2, 3)])
    assert b == [[1, 2], [0, 2, 3], [1, 3, 0], [2, 1]]
    assert s == [[3], [], [], [0]]
    _, s2 = R.topology_tables(4, [], [], [(0, 1, 2, 3), (0, 2, 1, 3)])
    assert s2 == [[3], [], [], [0]]
    with pytest.raises(R.TopologyConflict):
        R.topology_tables(3, [(0, 1), (1, 2), (2, 0)], [(0, 1, 2)], [])
    with pytest.raises(R.TopologyConflict):
        R.topology_tables(2, [(1, 1)], [], [])
