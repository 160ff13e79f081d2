"""GPU tests: non-finite and runaway positions raise device flags and every kernel of a force
evaluation and of integrator steps still completes; the context recovers once valid positions
are set again (same forces as a fresh context)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FLAG_OVERFLOW, FLAG_NONFINITE = 1, 4


def _terms(L):
    return L.TERM_LJ | L.TERM_PME | L.TERM_BOND | L.TERM_ANGLE


def test_nonfinite_positions_flag_and_recover():
    import torch
    from mdpy_amd import _lib as L, systems
    from mdpy_amd.engine import Engine
    s = systems.water_box_small()
    terms = _terms(L)
    e = Engine(s, skin=1.0)
    e.set_positions(s.positions)
    F0 = e.compute(terms)[0].cpu().numpy().copy()
    assert e.status() == 0
    bad = s.positions.copy()
    bad[5] = np.nan
    bad[17, 1] = np.inf
    bad[30] = [1.0e7, -3.0e6, 2.5e8]   # runaway but finite
    print("non-finite case: water648, atoms 5 (NaN), 17 (Inf), 30 (1e7 A)", flush=True)
    e.set_positions(bad, wrap=False)
    e.compute(terms)
    torch.cuda.synchronize()
    assert e.status() & FLAG_NONFINITE
    e.configure_integrator(L.INT_LANGEVIN, dt=1.0, temperature=300.0, friction=1e-3, seed=0)
    e.integrate(terms, 12, use_graph=True, energy_every=4)
    torch.cuda.synchronize()
    assert e.status() & FLAG_NONFINITE
    # valid positions again: the next evaluation rebuilds and matches a fresh context
    e.set_positions(s.positions)
    F1 = e.compute(terms)[0].cpu().numpy()
    assert e.status() == 0
    assert np.allclose(F1, F0, rtol=0, atol=1e-12 * np.abs(F0).max())
    e.close()


def test_all_atoms_nonfinite_oversize_cell():
    """Every atom NaN: all of them bin to one cell, far beyond the per-cell sort capacity (1,024);
    the cell is filled unsorted, its atoms get empty neighbour lists, and the flags say so."""
    import torch
    from mdpy_amd import _lib as L, systems
    from mdpy_amd.engine import Engine
    s = systems.water_box_large()
    terms = _terms(L)
    e = Engine(s, skin=1.5)
    e.set_positions(s.positions)
    F0 = e.compute(terms)[0].cpu().numpy().copy()
    print("non-finite case: water23328, every atom NaN, list rebuilt", flush=True)
    # through the wrap: it requests a list rebuild, so the NaN atoms are binned (all into one cell)
    e.set_positions(np.full_like(s.positions, np.nan), wrap=True)
    e.compute(terms)
    torch.cuda.synchronize()
    f = e.status()
    assert f & FLAG_NONFINITE and f & FLAG_OVERFLOW
    e.set_positions(s.positions)
    F1 = e.compute(terms)[0].cpu().numpy()
    assert e.status() == 0
    assert np.allclose(F1, F0, rtol=0, atol=1e-12 * np.abs(F0).max())
    e.close()
