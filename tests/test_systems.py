"""The synthetic chain system: geometry and parameter rules, and energy conservation of its
bonded + nonbonded physics under velocity Verlet, integrated on the CPU with the oracle's terms."""
import math

import numpy as np

from oracle import mdref as R


def _chain_angles(s, n_chain):
    x, L = s.positions[:n_chain], s.box
    a = s.angles[: n_chain - 2]
    v1 = R.unwrap_diag(x[a[:, 0]] - x[a[:, 1]], L)
    v2 = R.unwrap_diag(x[a[:, 2]] - x[a[:, 1]], L)
    c = (v1 * v2).sum(1) / np.linalg.norm(v1, axis=1) / np.linalg.norm(v2, axis=1)
    return np.degrees(np.arccos(c))


def test_chain_geometry_and_parameter_rules():
    from mdpy_amd import systems
    s = systems.solvated_chain()
    nc = s.meta["n_chain"]
    th = _chain_angles(s, nc)
    # bond angles 105-130 deg: no angle near 180 deg, where the dihedral normals vanish
    assert th.min() > 104.9 and th.max() < 130.1
    # impropers are harmonic in (psi - psi0) without periodic wrap
    # (mdpy/constraint/charmm_improper_constraint.py:70,90): psi0 stays well inside (-180, 180)
    assert np.abs(s.improper_params[:, 1]).max() < 0.5 * math.pi + 0.1
    psi = R.get_pbc_dihedral(*[s.positions[s.impropers[:, c]] for c in range(4)], s.box)
    assert np.abs(psi - s.improper_params[:, 1]).max() <= 0.1 + 1e-12
    # U(-0.6, 0.6) charges shifted by their mean to a neutral chain
    assert abs(s.charges[:nc].sum()) < 1e-9 and np.abs(s.charges[:nc]).max() <= 0.65
    assert np.abs(s.charges[:nc]).max() > 0.55
    assert s.num_atoms == 31200


def test_chain_energy_conserved_under_velocity_verlet():
    """120-atom chain in vacuum, all bonded terms + LJ (1-4 scaled) + plain Coulomb, 1,000 VV
    steps of 0.5 fs from 300 K: the total energy stays within 0.1% of the kinetic energy."""
    from mdpy_amd import systems
    from mdpy_amd.unit import COULOMB_K, KB
    s = systems.chain_in_vacuum(n_chain=120)
    ep, ei = s.exclusions()
    sp, si = s.scaling_pairs()

    def forces(x):
        pairs = R.exact_pairs(x, s.box, s.cutoff, ep, ei)
        F, E = np.zeros_like(x), 0.0
        for fn, args in ((R.bond_terms, (s.bonds, s.bond_params)), (R.angle_terms, (s.angles, s.angle_params)),
                         (R.dihedral_terms, (s.dihedrals, s.dihedral_params)),
                         (R.improper_terms, (s.impropers, s.improper_params))):
            Fi, Ei = fn(x, *args, s.box)
            F += Fi
            E += Ei
        Fl, El = R.lj_pair_terms(x, s.lj, pairs, s.box, sp, si)
        Fc, Ec = R.coulomb_allpairs(x, s.charges, ep, ei, s.box, COULOMB_K)
        return F + Fl + Fc, E + El + Ec

    rng = np.random.default_rng(0)
    m = s.masses[:, None]
    v = rng.standard_normal(s.positions.shape) * np.sqrt(KB * 300.0 / m)
    x = s.positions.copy()
    F, U = forces(x)
    e0 = U + 0.5 * (m * v * v).sum()
    dev, dt = [], 0.5
    for k in range(1000):
        v += 0.5 * dt * F / m
        x += dt * v
        F, U = forces(x)
        v += 0.5 * dt * F / m
        ke = 0.5 * (m * v * v).sum()
        dev.append(U + ke - e0)
    assert np.all(np.isfinite(dev))
    assert np.abs(dev).max() < 1e-3 * ke
