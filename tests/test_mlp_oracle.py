"""Pins the MLP-surrogate oracle (oracle/mlp.py): angular moment invariants against the explicit
double sum, forces against finite differences of the energy, the descriptor JVP against finite
differences of the descriptors, and the force-matching weight gradient against finite
differences of the loss."""
import numpy as np
import pytest

from oracle import mlp as M


@pytest.fixture(scope="module")
def small():
    from mdpy_amd import systems
    s = systems.water_box(4, 3.1, seed=3, cutoff=6.0, pme_grid=(12, 12, 12), name="w64")  # 192 atoms, 12.4 A
    elem = M.elements_of(s.types)
    return s.positions.copy(), np.asarray(s.box, dtype=np.float64), elem


def _surrogate(hidden, seed=0):
    rng = np.random.default_rng(seed)
    nets = [M.init_net(rng, hidden) for _ in range(M.N_ELEM)]
    for net in nets:
        for k in ("b1", "b2", "b3"):
            net[k] = rng.standard_normal(net[k].shape) * 0.1
    return nets


def _normaliser(pos, box, elem):
    G, _ = M.descriptors(pos, box, elem)
    return G.mean(0), G.std(0) + 1e-3


def test_angular_moments_equal_pair_double_sum(small):
    pos, box, elem = small
    sel = slice(0, 30)
    G, _ = M.descriptors(pos, box, elem)
    P = M.angular_bruteforce(pos, box, elem)
    ang = G.reshape(-1, M.N_ELEM, M.PER_ELEM)[:, :, M.N_RAD:].reshape(-1, M.N_ELEM, M.N_ANG_CH, M.N_L)
    np.testing.assert_allclose(ang[sel], P[sel], rtol=1e-12, atol=1e-12)
    # l = 0 is (sum g)^2 and l = 1 of a single neighbour is g^2: spot the closed forms
    assert np.all(ang[..., 0] >= ang[..., 1] - 1e-12)


def test_descriptor_jvp_matches_finite_difference(small):
    pos, box, elem = small
    rng = np.random.default_rng(1)
    w = rng.standard_normal(pos.shape)
    _, T = M.descriptors(pos, box, elem)
    jvp = M.descriptor_jvp(pos, box, elem, T, w)
    h = 1e-5
    gp, _ = M.descriptors(pos + h * w, box, elem)
    gm, _ = M.descriptors(pos - h * w, box, elem)
    fd = (gp - gm) / (2 * h)
    scale = np.abs(fd).max()
    np.testing.assert_allclose(jvp, fd, atol=1e-6 * scale)


def test_forces_are_minus_energy_gradient(small):
    pos, box, elem = small
    nets = _surrogate(16)
    mean, std = _normaliser(pos, box, elem)
    sur = M.Surrogate(nets, mean, std)
    F, E = sur.forces(pos, box, elem)
    assert abs(F.sum(0)).max() < 1e-9 * abs(F).max()  # translation invariance
    h = 1e-5
    for a, d in [(0, 0), (1, 2), (5, 1), (100, 0), (150, 2)]:
        x = pos.copy()
        x[a, d] += h
        ep = sur.energy(x, box, elem)
        x[a, d] -= 2 * h
        em = sur.energy(x, box, elem)
        assert F[a, d] == pytest.approx(-(ep - em) / (2 * h), rel=1e-5, abs=1e-7 * abs(F).max())


def test_force_matching_weight_gradient_matches_finite_difference(small):
    pos, box, elem = small
    nets = _surrogate(8, seed=2)
    mean, std = _normaliser(pos, box, elem)
    sur = M.Surrogate(nets, mean, std)
    rng = np.random.default_rng(4)
    f_ref = rng.standard_normal(pos.shape) * 0.5
    L0, grads, _ = sur.loss_and_grad(pos, box, elem, f_ref)
    h = 1e-6
    checks = [(0, "W1", (3, 2)), (1, "W1", (50, 5)), (0, "W2", (1, 7)), (1, "W3", (4, 4)), (0, "w4", (6,)),
              (1, "b1", (2,)), (0, "b2", (5,)), (1, "b3", (0,))]
    for e, k, idx in checks:
        orig = nets[e][k][idx]
        nets[e][k][idx] = orig + h
        lp = sur.loss_and_grad(pos, box, elem, f_ref)[0]
        nets[e][k][idx] = orig - h
        lm = sur.loss_and_grad(pos, box, elem, f_ref)[0]
        nets[e][k][idx] = orig
        fd = (lp - lm) / (2 * h)
        assert grads[e][k][idx] == pytest.approx(fd, rel=2e-5, abs=1e-9), (e, k, idx)


def test_tree_neighbour_search_equals_dense():
    """The periodic k-d tree search (used above 4,096 atoms) returns the dense search's pairs in
    the same order with bitwise the same vectors and distances: 192- and 648-atom boxes, and the
    648-atom box with atoms displaced out of the primary cell."""
    from mdpy_amd import systems
    cases = []
    for s in (systems.water_box(4, 3.1, seed=3, cutoff=6.0, pme_grid=(12, 12, 12), name="w64"),
              systems.water_box_small()):
        cases.append((np.asarray(s.positions, dtype=np.float64), np.asarray(s.box, dtype=np.float64)))
    x, box = cases[1]
    shift = np.random.default_rng(1).integers(-2, 3, size=x.shape) * box
    cases.append((x + shift, box))
    for pos, box in cases:
        a = M.neighbor_pairs(pos, box)
        b = M.neighbor_pairs_tree(pos, box)
        assert len(a[0]) > 1000
        for u, v in zip(a, b):
            assert np.array_equal(u, v)


def test_force_assembly_over_an_atom_subset(small):
    """Forces assembled from the pairs touching a subset of atoms equal the full assembly on
    those atoms' rows."""
    pos, box, elem = small
    mean, std = _normaliser(pos, box, elem)
    sur = M.Surrogate(_surrogate(32, seed=4), mean, std)
    F, E = sur.forces(pos, box, elem)
    atoms = np.array([0, 5, 17, 100, 191])
    Fs, Es = sur.forces(pos, box, elem, atoms=atoms)
    assert Es == E
    np.testing.assert_allclose(Fs[atoms], F[atoms], rtol=1e-12, atol=1e-12 * np.abs(F).max())
