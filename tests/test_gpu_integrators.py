"""GPU integrator parity: reference-compatible Verlet/Langevin against the float64 oracle
(same seeded noise), hipGraph replay against direct launches, NVE conservation and the
Langevin temperature."""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ref_force_fn(s):
    """The reference force field: LJ (r <= rc) + plain all-pairs Coulomb + bond + angle."""
    from mdpy_amd.unit import COULOMB_K
    from oracle import mdref as R
    ep, ei = s.exclusions()

    def f(x):
        pairs = R.exact_pairs(x, s.box, s.cutoff, ep, ei)
        F = R.lj_pair_terms(x, s.lj, pairs, s.box)[0]
        F += R.coulomb_allpairs(x, s.charges, ep, ei, s.box, COULOMB_K)[0]
        F += R.bond_terms(x, s.bonds, s.bond_params, s.box)[0]
        F += R.angle_terms(x, s.angles, s.angle_params, s.box)[0]
        return F
    return f


REF_TERMS = 1 | 128 | 8 | 16   # LJ | plain Coulomb | bond | angle


def thermal_v(s, T, seed):
    from mdpy_amd.unit import KB
    rng = np.random.default_rng(seed)
    return rng.standard_normal((s.num_atoms, 3)) * np.sqrt(KB * T / s.masses)[:, None]


def test_langevin_refcompat_matches_oracle():
    from mdpy_amd import _lib as L, systems
    from mdpy_amd.engine import Engine
    from mdpy_amd.unit import KB
    from oracle import mdref as R
    say("case: Langevin refcompat vs oracle, water648, 4 steps")
    s = systems.water_box_small()
    nsteps = 4
    v0 = thermal_v(s, 300.0, 3)
    noise = np.random.RandomState(0).randn(nsteps, s.num_atoms, 3)
    e = Engine(s, skin=1.0)
    e.set_positions(s.positions)
    e.set_velocities(v0)
    e.configure_integrator(L.INT_LANGEVIN_REF, dt=1.0, temperature=300.0, friction=1e-3)
    e.set_external_noise(noise)
    e.integrate(REF_TERMS, nsteps, use_graph=False)
    xg, vg = e.positions(), e.velocities()
    pbc = s.pbc_matrix
    x0 = R.wrap_positions(s.positions, pbc, np.linalg.inv(pbc))
    f = ref_force_fn(s)
    it = R.LangevinReference(1.0, KB * 300.0, 1e-3)
    rows = iter(noise)
    xo, vo, _ = it.integrate(x0, v0, f(x0), s.masses, f, lambda shape: next(rows), pbc, np.linalg.inv(pbc), nsteps)
    dx = np.abs(xg - xo).max()
    dv = np.abs(vg - vo).max() / np.abs(vo).max()
    say("  max |dx| %.3e A, max rel dv %.3e" % (dx, dv))
    assert dx < 1e-7 and dv < 1e-5
    e.close()


def test_verlet_refcompat_matches_oracle():
    from mdpy_amd import _lib as L, systems
    from mdpy_amd.engine import Engine
    from oracle import mdref as R
    say("case: Verlet refcompat vs oracle, water648, 4 steps")
    s = systems.water_box_small()
    v0 = thermal_v(s, 300.0, 4)
    e = Engine(s, skin=1.0)
    e.set_positions(s.positions)
    e.set_velocities(v0)
    e.configure_integrator(L.INT_VERLET_REF, dt=1.0)
    e.integrate(REF_TERMS, 4, use_graph=False)
    xg, vg = e.positions(), e.velocities()
    pbc = s.pbc_matrix
    x0 = R.wrap_positions(s.positions, pbc, np.linalg.inv(pbc))
    it = R.VerletReference(1.0)
    xo, vo = it.integrate(x0, v0, s.masses, ref_force_fn(s), pbc, np.linalg.inv(pbc), 4)
    dx = np.abs(xg - xo).max()
    dv = np.abs(vg - vo).max() / np.abs(vo).max()
    say("  max |dx| %.3e A, max rel dv %.3e" % (dx, dv))
    assert dx < 1e-7 and dv < 1e-5
    e.close()


def test_verlet_refcompat_two_calls_match_oracle():
    """Two integrate calls (3 then 4 steps, graph replays): the second continues from the
    position history the first left (verlet_integrator.py:28-33 initialises only when nothing is
    cached), on the device as in the oracle."""
    from mdpy_amd import _lib as L, systems
    from mdpy_amd.engine import Engine
    from oracle import mdref as R
    say("case: Verlet refcompat two calls vs oracle, water648, 3 + 4 steps")
    s = systems.water_box_small()
    v0 = thermal_v(s, 300.0, 6)
    e = Engine(s, skin=1.0)
    e.set_positions(s.positions)
    e.set_velocities(v0)
    e.configure_integrator(L.INT_VERLET_REF, dt=1.0)
    pbc = s.pbc_matrix
    inv = np.linalg.inv(pbc)
    it = R.VerletReference(1.0)
    xo, vo = R.wrap_positions(s.positions, pbc, inv), v0
    for n in (3, 4):
        e.integrate(REF_TERMS, n, use_graph=True)
        xo, vo = it.integrate(xo, vo, s.masses, ref_force_fn(s), pbc, inv, n)
        xg, vg = e.positions(), e.velocities()
        dx = np.abs(xg - xo).max()
        dv = np.abs(vg - vo).max() / np.abs(vo).max()
        say("  after %d steps: max |dx| %.3e A, max rel dv %.3e" % (n, dx, dv))
        assert dx < 1e-7 and dv < 1e-5
    e.close()


def test_graph_replay_bitwise_equals_direct_launches():
    """Calls of 25, 7, 2, 40 and 3 G-JF steps: the first primes its forces (direct first step,
    step graphs, direct closing kernels); the later ones start from the last call's forces and
    run as head graph + step graphs + tail graph. Every call's state and energy history equal
    those of direct launches bit for bit."""
    from mdpy_amd import _lib as L, systems
    from mdpy_amd.engine import Engine
    say("case: hipGraph vs direct, water648, G-JF calls of 25, 7, 2, 40, 3 steps")
    s = systems.water_box_small()
    calls = (25, 7, 2, 40, 3)
    out = []
    for graph in (False, True):
        e = Engine(s, skin=1.0)
        e.set_positions(s.positions)
        e.set_velocities(thermal_v(s, 300.0, 5))
        e.configure_integrator(L.INT_LANGEVIN, dt=1.0, temperature=300.0, friction=1e-3, seed=7)
        terms = L.TERM_LJ | L.TERM_PME | L.TERM_BOND | L.TERM_ANGLE
        run = []
        for n in calls:
            e.integrate(terms, n, use_graph=graph, energy_every=10)
            run.append((e.positions(), e.velocities(), e.energy_history()))
        out.append(run)
        e.close()
    for k, n in enumerate(calls):
        for a, b in zip(out[0][k], out[1][k]):
            assert np.array_equal(a, b), "call %d (%d steps)" % (k, n)


def test_graph_replay_on_fresh_context_reads_reciprocal_forces():
    """A fresh context's first call captures its step graphs before it has evaluated anything:
    the 32-step graph's first prologue must still read the reciprocal forces of the evaluation
    that precedes it on replay. One 60-step call (energies every 50: step graphs of 32 + 16, an
    energy step, 8 + 1, the closing step) equals direct launches bit for bit."""
    from mdpy_amd import _lib as L, systems
    from mdpy_amd.engine import Engine
    say("case: hipGraph vs direct on a fresh context, water648, one G-JF call of 60 steps")
    s = systems.water_box_small()
    out = []
    for graph in (False, True):
        e = Engine(s, skin=1.0)
        e.set_positions(s.positions)
        e.set_velocities(thermal_v(s, 300.0, 11))
        e.configure_integrator(L.INT_LANGEVIN, dt=1.0, temperature=300.0, friction=1e-3, seed=3)
        e.integrate(L.TERM_LJ | L.TERM_PME | L.TERM_BOND | L.TERM_ANGLE, 60, use_graph=graph, energy_every=50)
        out.append((e.positions(), e.velocities(), e.energy_history()))
        e.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_scheduled_list_rebuilds_match_displacement_rebuilds():
    """Lists also rebuilt every 4 steps (mdx_nonbonded_rebuild_every) against rebuilds on the
    displacement test only: no device flag, more rebuilds, and the same trajectory up to the
    float32 summation order of the PME spreading (the cell order changes at other steps)."""
    from mdpy_amd import _lib as L, systems
    from mdpy_amd.engine import Engine
    say("case: scheduled list rebuilds vs displacement rebuilds, water648, G-JF 40 steps")
    s = systems.water_box_small()
    out = []
    for every in (0, 4):
        e = Engine(s, skin=1.0)
        e.set_positions(s.positions)
        e.set_velocities(thermal_v(s, 300.0, 2))
        e.configure_integrator(L.INT_LANGEVIN, dt=1.0, temperature=300.0, friction=1e-3, seed=5)
        e.set_rebuild_every(every)
        e.integrate(L.TERM_LJ | L.TERM_PME | L.TERM_BOND | L.TERM_ANGLE, 40, use_graph=True, energy_every=10)
        out.append((e.positions(), e.energy_history(), e.stats()["rebuilds"], e.status()))
        e.close()
    (p0, h0, r0, f0), (p1, h1, r1, f1) = out
    assert f0 == 0 and f1 == 0
    assert r1 >= 10 and r1 > r0  # every 4th of 40 steps, the displacement rebuilds pre-empted
    assert np.abs(p0 - p1).max() < 1e-6
    assert np.allclose(h0, h1, rtol=1e-6, atol=1e-6)


def test_list_grid_cell_sizes_give_the_same_forces_and_trajectory():
    """One, two or three list-grid cells per list radius (mdx_nonbonded_cells): the lists hold the
    same pairs (the fast shifted-candidate kernel on some grids, the general cell kernel on the
    others) in another row order. Each row's float32 partial sums follow the row order and the
    PME spreading the atoms' cell order, so forces, energies and a 30-step trajectory agree to
    float32 rounding (measured 6.5e-9 kJ/mol/A on forces of ~1e-2 between 2 and 3 cells)."""
    from mdpy_amd import _lib as L, systems
    from mdpy_amd.engine import Engine
    reps = systems.water_replicas(64, seed=0)
    mid = sorted(reps, key=lambda s: s.num_atoms)[40]
    for s in (mid, systems.water_box_large(seed=0)):
        say("case: list cells per radius 2 / 3 / 1, %d atoms, box %.1f A" % (s.num_atoms, s.pbc_matrix[0, 0]))
        direct = L.TERM_LJ | L.TERM_COUL_DIRECT | L.TERM_BOND | L.TERM_ANGLE
        full = direct | L.TERM_COUL_RECIP
        out = []
        for k in (2, 3, 1):
            e = Engine(s, skin=1.5)
            e.set_positions(s.positions)
            e.set_velocities(thermal_v(s, 300.0, 4))
            e.configure_integrator(L.INT_LANGEVIN, dt=1.0, temperature=300.0, friction=1e-3, seed=9)
            e.set_list_cells(k)
            f0, en0 = e.compute(direct)
            f0 = f0.cpu().numpy().copy()
            f1, en1 = e.compute(full)
            f1 = f1.cpu().numpy().copy()
            e.integrate(full, 30, use_graph=True, energy_every=10)
            out.append((f0, en0, f1, e.positions(), e.energy_history(), e.status()))
            e.close()
        base = out[0]
        for k, o in zip((3, 1), out[1:]):
            assert o[5] == 0 and base[5] == 0
            d0 = np.sqrt(np.mean((o[0] - base[0]) ** 2)) / np.sqrt(np.mean(base[0] ** 2))
            assert d0 < 1e-6, (k, d0)
            for t in ("lj", "coulomb_direct", "bond", "angle"):
                assert o[1][t] == pytest.approx(base[1][t], rel=1e-6, abs=1e-6), (k, t)
            scale = np.sqrt(np.mean(base[2] ** 2))
            assert np.sqrt(np.mean((o[2] - base[2]) ** 2)) < 1e-6 * scale, k
            assert np.abs(o[3] - base[3]).max() < 1e-5, k
            assert np.allclose(o[4], base[4], rtol=1e-6, atol=1e-4), k


def test_kernel_timing_leaves_the_trajectory_unchanged():
    """Timing the step prologue (state saved and restored) and the three PME plane launches
    between two integrate calls changes nothing: the next call starts from a fresh evaluation at
    the same positions and follows the untimed trajectory bit for bit. Timing the list build
    (which forces a full rebuild on the next evaluation) leaves a valid, finite run."""
    from mdpy_amd import _lib as L, systems
    from mdpy_amd.engine import Engine
    say("case: kernel timing modes between integrate calls, water648, G-JF")
    s = systems.water_box_small()
    terms = L.TERM_LJ | L.TERM_PME | L.TERM_BOND | L.TERM_ANGLE
    out = []
    for timed in (False, True):
        e = Engine(s, skin=1.0)
        e.set_positions(s.positions)
        e.set_velocities(thermal_v(s, 300.0, 11))
        e.configure_integrator(L.INT_LANGEVIN, dt=1.0, temperature=300.0, friction=1e-3, seed=3)
        e.integrate(terms, 12, use_graph=True, energy_every=4)
        if timed:
            for which in (4, 6, 7, 8):
                assert e.time_kernel(which, terms, reps=5) > 0
        e.integrate(terms, 12, use_graph=True, energy_every=4)
        out.append((e.positions(), e.velocities(), e.energy_history()))
        if timed:
            assert e.time_kernel(5, terms, reps=3) > 0
            e.integrate(terms, 12, use_graph=True, energy_every=4)
            assert e.status() == 0 and np.all(np.isfinite(e.positions()))
        e.close()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


def test_nve_energy_conservation_pme_water():
    from mdpy_amd import _lib as L, systems
    from mdpy_amd.engine import Engine
    say("case: NVE drift, water648, 400 VV steps at 0.5 fs")
    s = systems.water_box_small()
    e = Engine(s, skin=1.5)
    e.set_positions(s.positions)
    e.set_velocities(thermal_v(s, 300.0, 6))
    e.configure_integrator(L.INT_VELOCITY_VERLET, dt=0.5)
    terms = L.TERM_LJ | L.TERM_PME | L.TERM_BOND | L.TERM_ANGLE
    e.integrate(terms, 400, use_graph=True)
    h = e.energy_history()
    assert h.shape[0] == 401
    etot = h[:, :L.NTERMS].sum(1) + h[:, L.E_KINETIC]
    ke = h[:, L.E_KINETIC].mean()
    drift = np.polyfit(np.arange(len(etot)) * 0.5, etot, 1)[0]  # per fs
    say("  E0 %.6e  std %.3e  KE %.3e  drift %.3e /fs" % (etot[0], etot.std(), ke, drift))
    assert etot.std() < 2e-3 * ke
    e.close()


def test_gjf_langevin_thermalises():
    from mdpy_amd import _lib as L, systems
    from mdpy_amd.engine import Engine
    from mdpy_amd.unit import KB
    say("case: G-JF temperature, water648, 3000 steps, gamma 0.05/fs")
    s = systems.water_box_small()
    e = Engine(s, skin=1.5)
    e.set_positions(s.positions)
    e.configure_integrator(L.INT_LANGEVIN, dt=1.0, temperature=300.0, friction=0.05, seed=11)
    terms = L.TERM_LJ | L.TERM_PME | L.TERM_BOND | L.TERM_ANGLE
    e.integrate(terms, 3000, use_graph=True)
    h = e.energy_history()
    T = 2 * h[1000:, L.E_KINETIC] / (3 * s.num_atoms * KB)
    say("  <T> = %.1f K" % T.mean())
    assert 270 < T.mean() < 330
    e.close()
