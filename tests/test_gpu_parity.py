"""GPU parity: the HIP path (through the C-ABI) against the float64 oracle on the same seeded
inputs. Integer/index results must be bit-exact; float forces within the stated tolerances."""
import math
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

FORCE_RTOL = 1e-5      # relative force RMS, float32 pair/PME arithmetic vs float64 oracle, per term
                       # and in total (measured per term: LJ 3.4e-6 / 8.2e-6 / 6.0e-6, direct
                       # Coulomb 0.9e-6 / 2.0e-6 / 2.7e-6 at water648 / water23328 / chain)
ENERGY_RTOL = 1e-5     # relative energy error per term (large terms)
# The direct-space Coulomb energy is a sum of float32 pair energies that cancel: |E_direct| is
# 4.6e-3 (water648) and 1.2e-4 (water23328, chain) of the sum of |pair energies|. Measured errors:
# 0.9e-6, 2.0e-6 and 1.55e-5 |E_direct| (on the chain, with its +-0.6 e charges, 1.9e-9 of the
# sum of |pair energies|); the chain's bound is therefore 2e-5, the others 1e-5.
DIRECT_ENERGY_RTOL = {"water648": 1e-5, "water23328": 1e-5, "chain": 2e-5, "replica_small": 1e-5,
                      "replica_large": 1e-5}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_rms(F, Fref):
    return math.sqrt(float(((F - Fref) ** 2).sum()) / float((Fref ** 2).sum()))


@pytest.fixture(scope="module")
def water648():
    from mdpy_amd import systems
    return systems.water_box_small()


@pytest.fixture(scope="module")
def water23328():
    from mdpy_amd import systems
    return systems.water_box_large()


@pytest.fixture(scope="module")
def chain():
    from mdpy_amd import systems
    return systems.solvated_chain()


def _replica(which):
    """The smallest or largest of the 64 batched water replicas (r_c 9 A, PME ~1 point/A)."""
    from mdpy_amd import systems
    reps = systems.water_replicas(64, seed=0)
    sizes = [s.num_atoms for s in reps]
    return reps[int(np.argmin(sizes) if which == "small" else np.argmax(sizes))]


@pytest.fixture(scope="module")
def replica_small():
    return _replica("small")


@pytest.fixture(scope="module")
def replica_large():
    return _replica("large")


def wrapped(s):
    from oracle import mdref as R
    return R.wrap_positions(s.positions, s.pbc_matrix, np.linalg.inv(s.pbc_matrix))


# ----------------------------------------------------------------------------- cell list
# cell-list cutoff: the reference raises for floor(L/rc) < 2 (cell_list.py:56-69), so the 18.6 A box
# uses its constraint cutoff (9 A; same n = 3 grid as 12 A would give)
CELL_CUTOFF = {"water648": 9.0, "water23328": 12.0}


@pytest.mark.parametrize("name", ["water648", "water23328"])
def test_reference_celllist_bit_exact(name, request):
    from mdpy_amd.engine import Engine
    from oracle import mdref as R
    s = request.getfixturevalue(name)
    say("case: reference cell list", name)
    e = Engine(s, cell_cutoff=CELL_CUTOFF[name])
    e.set_positions(s.positions)  # wraps on the device
    xw = wrapped(s)
    assert np.array_equal(e.positions(), xw)
    n, w, inv = R.cell_grid(s.box, CELL_CUTOFF[name])
    cidx, table = R.cell_list_build(xw, inv, n)
    gidx, gtable, gn = e.celllist_reference(CELL_CUTOFF[name])
    assert np.array_equal(gn, n)
    assert np.array_equal(gidx, cidx)
    assert np.array_equal(gtable, table)
    e.close()


def test_reference_celllist_6po6_kat():
    # mdpy/test/test_cell_list.py:41-58 on the 6PO6 coordinates (30 A box, 9 A)
    import json
    from mdpy_amd.engine import Engine
    from mdpy_amd.systems import System
    from oracle import mdref as R
    say("case: 6PO6 cell list")
    pos = np.array(json.load(open(os.path.join(GOLDEN, "kat.json")))["6PO6_positions"])
    n = len(pos)
    s = System(name="6po6", positions=pos, box=np.array([30.0] * 3), masses=np.ones(n), charges=np.zeros(n),
               lj=np.zeros((n, 4)), types=["X"] * n, bonds=np.zeros((0, 2), np.int64), bond_params=np.zeros((0, 2)),
               angles=np.zeros((0, 3), np.int64), angle_params=np.zeros((0, 4)), dihedrals=np.zeros((0, 4), np.int64),
               dihedral_params=np.zeros((0, 3)), impropers=np.zeros((0, 4), np.int64), improper_params=np.zeros((0, 2)),
               cutoff=9.0, pme_grid=(32, 32, 32))
    e = Engine(s)
    e.set_positions(pos, wrap=False)  # the reference test bins the raw PDB coordinates
    idx, table, dims = e.celllist_reference(9.0)
    assert list(dims) == [3, 3, 3]
    atoms = [a for a in range(n) if (idx[a] == 0).all()]
    for k, a in enumerate(atoms):
        assert table[0, 0, 0, k] == a
    assert 23 in R.cell_list_getitem(idx, table, (-1, -1, 0))
    assert 38 in R.cell_list_getitem(idx, table, (-1, -1, -1))
    e.close()


# ----------------------------------------------------------------------------- neighbour sets
@pytest.mark.parametrize("name", ["water648", "water23328"])
def test_reference_neighbor_set_bit_exact(name, request):
    from mdpy_amd.engine import Engine
    from oracle import mdref as R
    s = request.getfixturevalue(name)
    say("case: reference neighbour set", name)
    e = Engine(s, pair_mode="reference", cell_cutoff=CELL_CUTOFF[name])
    e.set_positions(s.positions)
    xw = wrapped(s)
    ep, ei = s.exclusions()
    n, w, inv = R.cell_grid(s.box, CELL_CUTOFF[name])
    cidx, _ = R.cell_list_build(xw, inv, n)
    ref = R.reference_pairs(xw, s.box, s.cutoff, ep, ei, cidx, n)
    got = e.neighbor_pairs()
    assert got.shape == ref.shape
    assert np.array_equal(got, ref)
    e.close()


@pytest.mark.parametrize("name", ["chain", "water23328"])
def test_exact_neighbor_set(name, request):
    """The skinned list (wave-per-atom build, far candidate groups skipped) cut at r_c in float64
    holds exactly the oracle's pair set."""
    from mdpy_amd.engine import Engine
    from oracle import mdref as R
    s = request.getfixturevalue(name)
    say("case: exact neighbour set", name)
    e = Engine(s, skin=1.5)
    e.set_positions(s.positions)
    xw = wrapped(s)
    ep, ei = s.exclusions()
    ref = R.exact_pairs(xw, s.box, s.cutoff, ep, ei)
    got = e.neighbor_pairs()
    assert got.shape == ref.shape
    assert np.array_equal(got, ref)
    e.close()


# ----------------------------------------------------------------------------- forces
def oracle_terms(s, xw, pairs, beta):
    from mdpy_amd.unit import ONE_OVER_4PI_EPS0
    from oracle import ewald as EW, mdref as R
    ep, ei = s.exclusions()
    sp, si = s.scaling_pairs()
    out = {}
    out["lj"] = R.lj_pair_terms(xw, s.lj, pairs, s.box, sp, si)
    ex = np.array(sorted({(a, int(b)) for a in range(s.num_atoms) for b in ei[ep[a]:ep[a + 1]] if b > a}),
                  dtype=np.int64).reshape(-1, 2)
    Fd, Ed = EW.direct_terms(xw, s.charges, pairs, s.box, beta, ONE_OVER_4PI_EPS0)
    Fr, Er = EW.spme_reciprocal(xw, s.charges, s.box, s.pme_grid, 4, beta, ONE_OVER_4PI_EPS0)
    Fx, Ex = EW.exclusion_terms(xw, s.charges, ex, s.box, beta, ONE_OVER_4PI_EPS0)
    Es = EW.self_energy(s.charges, beta, ONE_OVER_4PI_EPS0)
    Eb = EW.background_energy(s.charges, s.box, beta, ONE_OVER_4PI_EPS0)
    out["coulomb_direct"] = (Fd, Ed)
    out["coulomb_reciprocal"] = (Fr + Fx, Er + Ex + Es + Eb)
    # sum of |pair energy| (the same sum over |q|): float32 pair arithmetic errs per pair, so the
    # error of a total that cancels to a small value scales with this, not with the total
    _, Ed_abs = EW.direct_terms(xw, np.abs(s.charges), pairs, s.box, beta, ONE_OVER_4PI_EPS0)
    out["_scale"] = {"coulomb_reciprocal": abs(Er) + abs(Ex) + abs(Es) + abs(Eb)}
    out["_direct_abs"] = Ed_abs
    out["bond"] = R.bond_terms(xw, s.bonds, s.bond_params, s.box)
    out["angle"] = R.angle_terms(xw, s.angles, s.angle_params, s.box)
    out["dihedral"] = R.dihedral_terms(xw, s.dihedrals, s.dihedral_params, s.box)
    out["improper"] = R.improper_terms(xw, s.impropers, s.improper_params, s.box)
    return out


TERM_BITS = {"lj": 1, "coulomb_direct": 2, "coulomb_reciprocal": 4, "bond": 8, "angle": 16, "dihedral": 32,
             "improper": 64}


@pytest.mark.parametrize("name", ["water648", "water23328", "chain", "replica_small", "replica_large"])
def test_forces_match_oracle(name, request):
    from mdpy_amd.engine import Engine
    from oracle import mdref as R
    s = request.getfixturevalue(name)
    say("case: forces vs oracle", name)
    e = Engine(s, skin=1.5)
    e.set_positions(s.positions)
    xw = wrapped(s)
    ep, ei = s.exclusions()
    pairs = R.exact_pairs(xw, s.box, s.cutoff, ep, ei)
    ref = oracle_terms(s, xw, pairs, e.beta)
    ref_scale = ref.pop("_scale")
    direct_abs = ref.pop("_direct_abs")
    Ftot_ref = np.zeros_like(xw)
    Etot_ref = 0.0
    for term, bit in TERM_BITS.items():
        Fref, Eref = ref[term]
        Ftot_ref += Fref
        Etot_ref += Eref
        if not np.any(Fref):
            continue
        F, en = e.compute(bit)
        F = F.cpu().numpy()
        rms_ref = math.sqrt(float((Fref ** 2).sum()) / len(Fref))
        if rms_ref < 1e-9:  # e.g. water bonds/angles at their exact reference geometry
            say("  %-20s |F| max %.3e (reference RMS %.1e)" % (term, np.abs(F).max(), rms_ref))
            assert np.abs(F - Fref).max() < 1e-10
            continue
        err = rel_rms(F, Fref)
        say("  %-20s rel force RMS %.3e  E %.10e ref %.10e" % (term, err, en[term], Eref))
        # float64 bonded kernels: only the int64 fixed-point quantum (2^-40) separates them
        tol = 1e-8 if term in ("bond", "angle", "dihedral", "improper") else FORCE_RTOL
        assert err < tol, (term, err)
        # energies: relative to |E| of the term, except the reciprocal term, which is grid energy
        # + self + exclusion correction, each much larger than their sum
        scale = ref_scale[term] if term in ref_scale else abs(Eref)
        rtol = DIRECT_ENERGY_RTOL[name] if term == "coulomb_direct" else ENERGY_RTOL
        if term == "coulomb_direct":
            say("  direct energy: |E| / sum |pair E| = %.4f" % (abs(Eref) / direct_abs))
        assert abs(en[term] - Eref) <= max(rtol * scale, 1e-9), (term, en[term], Eref, scale)
    F, en = e.compute(sum(TERM_BITS.values()))
    err = rel_rms(F.cpu().numpy(), Ftot_ref)
    say("  total rel force RMS %.3e" % err)
    assert err < FORCE_RTOL
    e.close()


def test_forces_deterministic(water23328):
    """Repeated evaluations are bitwise equal, and so are the resident-workgroup nonbonded
    launches (workgroups looping over the work) and one workgroup per unit of work."""
    from mdpy_amd.engine import Engine
    say("case: determinism water23328")
    e = Engine(water23328, skin=1.5)
    e.set_positions(water23328.positions)
    F1, E1 = e.compute(127)
    F1 = F1.cpu().numpy().copy()
    F2, E2 = e.compute(127)
    assert np.array_equal(F1, F2.cpu().numpy())
    assert E1 == E2
    Fo = e.compute_forces_only(127).cpu().numpy().copy()
    for per_cu in (0, 1, 3):
        e.set_nonbonded_persist(per_cu)
        F3, E3 = e.compute(127)
        assert np.array_equal(F1, F3.cpu().numpy()) and E1 == E3, per_cu
        assert np.array_equal(Fo, e.compute_forces_only(127).cpu().numpy()), per_cu
    assert e.status() == 0
    e.close()


@pytest.mark.parametrize("name", ["water648", "water23328", "chain"])
def test_nonbonded_variants_match_oracle(name, request):
    """Every nonbonded kernel variant (16/32/64 lanes per atom) against the oracle's LJ and
    direct-space Coulomb, forces-only launches (tabulated Ewald correction) and energy launches
    (erfc form) alike, with the tolerances of the full force test; each bitwise repeatable."""
    from mdpy_amd import _lib as L
    from mdpy_amd.engine import Engine
    from mdpy_amd.unit import ONE_OVER_4PI_EPS0
    from oracle import ewald as EW, mdref as R
    s = request.getfixturevalue(name)
    e = Engine(s, skin=1.5)
    e.set_positions(s.positions)
    xw = wrapped(s)
    ep, ei = s.exclusions()
    sp, si = s.scaling_pairs()
    pairs = R.exact_pairs(xw, s.box, s.cutoff, ep, ei)
    Flj, Elj = R.lj_pair_terms(xw, s.lj, pairs, s.box, sp, si)
    Fd, Ed = EW.direct_terms(xw, s.charges, pairs, s.box, e.beta, ONE_OVER_4PI_EPS0)
    for lanes in (16, 32, 64):
        say("case: nonbonded variant", lanes, name)
        assert e.tune_nonbonded(L.TERM_LJ | L.TERM_COUL_DIRECT, lanes) == lanes
        for bit, key, Fref, Eref, rtol in ((L.TERM_LJ, "lj", Flj, Elj, ENERGY_RTOL),
                                          (L.TERM_COUL_DIRECT, "coulomb_direct", Fd, Ed, DIRECT_ENERGY_RTOL[name])):
            F1, en1 = e.compute(bit)
            F1 = F1.cpu().numpy().copy()
            F2, en2 = e.compute(bit)
            assert np.array_equal(F1, F2.cpu().numpy()) and en1[key] == en2[key], (lanes, key)
            err = rel_rms(F1, Fref)
            F3 = e.compute_forces_only(bit).cpu().numpy().copy()
            assert np.array_equal(F3, e.compute_forces_only(bit).cpu().numpy()), (lanes, key)
            err3 = rel_rms(F3, Fref)
            say("  %-16s rel force RMS %.3e (forces-only launch %.3e)  E %.10e ref %.10e"
                % (key, err, err3, en1[key], Eref))
            assert err < FORCE_RTOL and err3 < FORCE_RTOL, (lanes, key, err, err3)
            assert abs(en1[key] - Eref) <= max(rtol * abs(Eref), 1e-9), (lanes, key, en1[key], Eref)
        assert e.status() == 0
    e.close()


@pytest.mark.parametrize("name", ["water648", "water23328", "chain"])
def test_force_phases_equal_two_stream_path(name, request):
    """Phase launches (the PME kernels carrying the nonbonded and bonded workgroups on one stream)
    against the two-stream path: each atom's nonbonded partial sums run in the same lanes and
    order, the PME spread is fixed-point and the gather per plane slot, so forces are bitwise
    equal; energies differ only in the per-workgroup rounding of the block sums."""
    from mdpy_amd import _lib as L
    from mdpy_amd.engine import Engine
    s = request.getfixturevalue(name)
    say("case: force phases vs two streams,", name)
    terms = sum(TERM_BITS.values())
    e = Engine(s, skin=1.5)
    e.set_positions(s.positions)
    for lanes in (16, 64):
        e.tune_nonbonded(L.TERM_LJ | L.TERM_COUL_DIRECT, lanes)
        out = {}
        for on in (False, True, True):
            e.set_force_phases(on, 0.3, 0.1)
            F, en = e.compute(terms)
            Fo = e.compute_forces_only(terms).cpu().numpy().copy()
            out.setdefault(on, []).append((F.cpu().numpy().copy(), dict(en), Fo))
            assert e.status() == 0, (lanes, on)
        (Fa, Ea, Foa), (Fb, Eb, Fob) = out[False][0], out[True][0]
        assert np.array_equal(Fa, Fb) and np.array_equal(Foa, Fob), lanes
        assert np.array_equal(out[True][0][0], out[True][1][0])  # repeatable
        for k in ("lj", "coulomb_direct", "coulomb_reciprocal", "bond", "angle", "dihedral", "improper"):
            assert abs(Ea[k] - Eb[k]) <= 4096 * 2.0 ** -40 + 1e-12 * abs(Ea[k]), (lanes, k, Ea[k], Eb[k])
    # inside the integrator's graph: same trajectory as the two-stream path
    e.tune_nonbonded(L.TERM_LJ | L.TERM_COUL_DIRECT, 16)
    xs = []
    for on in (False, True):
        e.set_positions(s.positions)
        e.set_velocities(np.zeros((e.n, 3)))
        e.set_force_phases(on, 0.3, 0.1)
        e.configure_integrator(L.INT_VELOCITY_VERLET, dt=0.5)
        e.integrate(terms, 60, True, 20)
        xs.append(e.positions().copy())
        assert e.status() == 0, on
    assert np.array_equal(xs[0], xs[1])
    e.close()


@pytest.mark.parametrize("name", ["water648", "water23328", "chain"])
def test_pme_one_launch_equals_three_launches(name, request):
    """The one-launch reciprocal kernel (grid barrier between its phases) and the three-kernel
    path run the same arithmetic per plane and per pencil: forces bitwise equal, energies equal
    up to the order of the per-block energy sums; no status flag (coverage, barrier timeout)."""
    from mdpy_amd import _lib as L
    from mdpy_amd.engine import Engine
    s = request.getfixturevalue(name)
    say("case: PME one launch vs three,", name)
    e = Engine(s, skin=1.5)
    e.set_positions(s.positions)
    out = {}
    for fused in (False, True, False, True):
        e.set_pme_fused(fused)
        F, en = e.compute(L.TERM_COUL_RECIP)
        out.setdefault(fused, []).append((F.cpu().numpy().copy(), en["coulomb_reciprocal"]))
        assert e.status() == 0, fused
    (Fa, Ea), (Fb, Eb) = out[False][0], out[True][0]
    assert np.array_equal(Fa, Fb)
    # each block's energy sum is rounded once to the 2^-40 fixed-point quantum: at most one
    # quantum per block of either launch shape (256 x-pass blocks at 64^3)
    assert abs(Ea - Eb) <= 256 * 2.0 ** -40
    for k in (False, True):  # repeated launches: the barrier re-arms itself
        assert np.array_equal(out[k][0][0], out[k][1][0]) and out[k][0][1] == out[k][1][1]
    # inside the integrator's graph, many replays
    e.set_pme_fused(True)
    e.set_velocities(np.zeros((e.n, 3)))
    e.configure_integrator(L.INT_LANGEVIN, dt=1.0, temperature=300.0, friction=1e-2, seed=1)
    e.integrate(L.TERM_LJ | L.TERM_PME | L.TERM_BOND | L.TERM_ANGLE, 200, True, 50)
    assert e.status() == 0
    e.close()


def test_plain_coulomb_matches_reference_kat_and_oracle(water648):
    # mdpy ElectrostaticConstraint semantics: all pairs, minimum image, no cutoff
    from mdpy_amd.engine import Engine
    from mdpy_amd.unit import COULOMB_K
    from oracle import mdref as R
    say("case: plain coulomb water648")
    e = Engine(water648)
    e.set_positions(water648.positions)
    xw = wrapped(water648)
    ep, ei = water648.exclusions()
    Fref, Eref = R.coulomb_allpairs(xw, water648.charges, ep, ei, water648.box, COULOMB_K)
    F, en = e.compute(128)
    assert rel_rms(F.cpu().numpy(), Fref) < 1e-10
    assert en["coulomb_plain"] == pytest.approx(Eref, rel=1e-10)
    e.close()


def test_dihedral_refcompat(chain):
    from mdpy_amd.engine import Engine
    from oracle import mdref as R
    say("case: dihedral refcompat chain")
    e = Engine(chain, dihedral_refcompat=True)
    e.set_positions(chain.positions)
    xw = wrapped(chain)
    Fref, Eref = R.dihedral_terms(xw, chain.dihedrals, chain.dihedral_params, chain.box, refcompat=True)
    F, en = e.compute(32)
    assert rel_rms(F.cpu().numpy(), Fref) < 1e-9   # int64 fixed-point quantum
    assert en["dihedral"] == pytest.approx(Eref, rel=1e-9)
    e.close()


@pytest.mark.parametrize("name", ["water23328", "chain"])
def test_ensemble_through_topology_matches_oracle(name, request):
    """The mdpy-shaped path end to end: Topology.from_system -> Ensemble with the CHARMM and PME
    constraints (type-keyed parameter tables, as a force field would give them) -> update(),
    against the float64 oracle whose exclusion and 1-4 tables come from its own restatement of
    the topology rules (oracle.mdref.topology_tables), not from the System."""
    from mdpy_amd import Ensemble
    from mdpy_amd.constraint import (CharmmAngleConstraint, CharmmBondConstraint, CharmmDihedralConstraint,
                                     CharmmImproperConstraint, CharmmNonbondedConstraint,
                                     ElectrostaticPMEConstraint)
    from mdpy_amd.core import Topology
    from mdpy_amd.unit import ONE_OVER_4PI_EPS0
    from oracle import ewald as EW, mdref as R
    s = request.getfixturevalue(name)
    say("case: Ensemble through Topology vs oracle", name)
    ty = list(s.types)
    key = lambda ids: "-".join(ty[int(i)] for i in ids)
    nb = {}
    for i, t in enumerate(ty):
        nb.setdefault(t, list(s.lj[i]))
    bonds = {key(b): list(p) for b, p in zip(s.bonds, s.bond_params)}
    bonds.update({key(b[::-1]): list(p) for b, p in zip(s.bonds, s.bond_params)})
    angles = {key(a): list(p) for a, p in zip(s.angles, s.angle_params)}
    dihedrals = {}
    for d, p in zip(s.dihedrals, s.dihedral_params):
        dihedrals.setdefault(key(d), []).append(list(p))
    impropers = {key(m): list(p) for m, p in zip(s.impropers, s.improper_params)}
    top = Topology.from_system(s)
    ens = Ensemble(top, s.pbc_matrix)
    ens.state.set_positions(s.positions)
    cons = [CharmmNonbondedConstraint(nb, cutoff_radius=s.cutoff),
            ElectrostaticPMEConstraint(cutoff_radius=s.cutoff, grid=s.pme_grid),
            CharmmBondConstraint(bonds), CharmmAngleConstraint(angles)]
    if len(s.dihedrals):
        cons.append(CharmmDihedralConstraint(dihedrals))
    if len(s.impropers):
        cons.append(CharmmImproperConstraint(impropers))
    ens.add_constraints(*cons)
    ens.update()
    F = np.asarray(ens.forces, dtype=np.float64)
    pme = cons[1]
    # oracle, with its own topology tables
    quads = sorted({tuple(int(v) for v in d) for d in s.dihedrals})
    bonded, scaling = R.topology_tables(s.num_atoms, s.bonds, s.angles, quads)
    csr = lambda rows: (np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int64),
                        np.array([j for r in rows for j in sorted(r)], dtype=np.int64))
    ep, ei = csr(bonded)
    sp, si = csr(scaling)
    # the mdpy-shaped State and Topology hold positions and charges at the environment's precision
    # (float32 by default, as the reference's env): the oracle takes the same rounded values
    x32 = np.asarray(s.positions, dtype=np.float32).astype(np.float64)
    xw = R.wrap_positions(x32, s.pbc_matrix, np.linalg.inv(s.pbc_matrix))
    q = np.asarray(top.charges, dtype=np.float64).reshape(-1)
    pairs = R.exact_pairs(xw, s.box, s.cutoff, ep, ei)
    Fl, El = R.lj_pair_terms(xw, s.lj, pairs, s.box, sp, si)
    ex = np.array(sorted({(a, int(b)) for a in range(s.num_atoms) for b in bonded[a] if b > a}), dtype=np.int64)
    Fp, Ep, _ = EW.pme_total(xw, q, s.box, pairs, ex, s.pme_grid, 4, pme.beta, ONE_OVER_4PI_EPS0)
    Fref = Fl + Fp + R.bond_terms(xw, s.bonds, s.bond_params, s.box)[0] + \
        R.angle_terms(xw, s.angles, s.angle_params, s.box)[0]
    if len(s.dihedrals):
        Fref += R.dihedral_terms(xw, s.dihedrals, s.dihedral_params, s.box)[0]
    if len(s.impropers):
        Fref += R.improper_terms(xw, s.impropers, s.improper_params, s.box)[0]
    err = rel_rms(F, Fref)
    say("  total rel force RMS %.3e" % err)
    assert err < FORCE_RTOL
