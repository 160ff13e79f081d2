"""The learned surrogate on the GPU against its float64 oracle (oracle/mlp.py): the MFMA GEMM,
descriptors, forces and energies, force-matching gradients, training, the surrogate inside the
integrator, and the plain-path check with fallback."""
import ctypes
import sys

import numpy as np
import pytest

from oracle import mlp as M

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def bf16_bits(a):
    return (bf16(a).view(np.uint32) >> 16).astype(np.uint16)


def rel_rms(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


@pytest.fixture(scope="module")
def box192():
    from mdpy_amd import systems
    from mdpy_amd.engine import Engine
    from mdpy_amd.mlp import Surrogate
    s = systems.water_box(4, 3.1, seed=3, cutoff=6.0, pme_grid=(12, 12, 12), name="w64")
    e = Engine(s, skin=1.0)
    e.set_positions(s.positions)
    elem = M.elements_of(s.types)
    sur = Surrogate(e, elem, rd=6.0, skin=1.0, seed=1)
    yield s, e, sur, elem
    sur.close()
    e.close()


def _nets(seed, w4_scale=1.0):
    rng = np.random.default_rng(seed)
    nets = []
    for _ in range(2):
        n = M.init_net(rng, 256)
        for k in ("b1", "b2", "b3"):
            n[k] = rng.standard_normal(n[k].shape) * 0.1
        n["w4"] = n["w4"] * w4_scale
        for k in ("W1", "W2", "W3"):
            n[k] = bf16(n[k]).astype(np.float64)  # the device holds these in bf16
        nets.append(n)
    return nets


def test_mfma_gemm_matches_numpy(box192):
    import torch
    from mdpy_amd import _lib as L
    _, _, sur, _ = box192
    say("case: MFMA GEMM 130x70x100")
    rng = np.random.default_rng(0)
    Mm, N, K = 130, 70, 100
    A, B = rng.standard_normal((Mm, K)), rng.standard_normal((K, N))
    dA = torch.tensor(bf16_bits(A).view(np.int16), device="cuda")
    dB = torch.tensor(bf16_bits(B).view(np.int16), device="cuda")
    dC = torch.empty((Mm, N), dtype=torch.float32, device="cuda")
    L.call("mdx_mlp_gemm_check", sur._h, Mm, N, K, ctypes.c_void_p(dA.data_ptr()), ctypes.c_void_p(dB.data_ptr()),
           ctypes.c_void_p(dC.data_ptr()))
    torch.cuda.synchronize()
    ref = bf16(A).astype(np.float64) @ bf16(B).astype(np.float64)
    np.testing.assert_allclose(dC.cpu().numpy(), ref, rtol=0, atol=1e-4 * np.abs(ref).max())


def test_descriptors_match_oracle(box192):
    s, e, sur, elem = box192
    say("case: MLP descriptors, 192 atoms")
    G = sur.descriptors().cpu().numpy()
    Gr, _ = M.descriptors(e.positions(), np.asarray(s.box, dtype=np.float64), elem)
    scale = np.abs(Gr).max(0) + 1e-12
    err = (np.abs(G - Gr) / scale).max()
    say("  max error / feature scale %.2e" % err)
    assert err < 1e-4


def test_forces_and_energy_match_oracle(box192):
    s, e, sur, elem = box192
    say("case: MLP forces vs oracle, bf16 weights")
    x = e.positions()
    box = np.asarray(s.box, dtype=np.float64)
    G, _ = M.descriptors(x, box, elem)
    mean, std = G.mean(0), G.std(0) + 1e-3
    nets = _nets(5)
    sur.set_params(nets)
    sur.set_norm(mean, std)
    F, E = sur.eval()
    F = F.cpu().numpy()
    ref = M.Surrogate(nets, mean.astype(np.float32).astype(np.float64), std.astype(np.float32).astype(np.float64))
    Fr, Er = ref.forces(x, box, elem)
    say("  force rel RMS %.3e, energy %.6e vs %.6e" % (rel_rms(F, Fr), E, Er))
    assert rel_rms(F, Fr) < 3e-2
    assert abs(E - Er) < 1e-2 * abs(Er) + 1e-3 * np.abs(Fr).max()
    assert np.abs(F.sum(0)).max() < 1e-6 * np.abs(F).max() * len(F)


def test_training_gradients_match_oracle(box192):
    import torch
    s, e, sur, elem = box192
    say("case: MLP force-matching gradients vs oracle")
    x = e.positions()
    box = np.asarray(s.box, dtype=np.float64)
    G, _ = M.descriptors(x, box, elem)
    mean, std = G.mean(0), G.std(0) + 1e-3
    nets = _nets(7, w4_scale=0.3)
    sur.set_params(nets)
    sur.set_norm(mean, std)
    rng = np.random.default_rng(3)
    fref = rng.standard_normal(x.shape) * 0.02
    dpos = torch.tensor(x, device="cuda")
    dref = torch.tensor(fref, device="cuda")
    loss = sur.train_step(dpos, dref, 0.0)
    grads = sur.get_grads()
    ref = M.Surrogate(nets, mean.astype(np.float32).astype(np.float64), std.astype(np.float32).astype(np.float64))
    Lr, gr, _ = ref.loss_and_grad(x, box, elem, fref)
    say("  loss %.6e vs %.6e" % (loss, Lr))
    assert loss == pytest.approx(Lr, rel=2e-2)
    for el in range(2):
        for k in ("W1", "b1", "W2", "b2", "W3", "b3", "w4"):
            a, b = grads[el][k].reshape(-1), np.asarray(gr[el][k]).reshape(-1)
            cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-300))
            say("  elem %d %s rel %.3e cos %.5f" % (el, k, rel_rms(a, b), cos))
            assert cos > 0.995 and rel_rms(a, b) < 0.1, (el, k)


def _frames(n_frames=24, every=10, seed=0):
    """Positions and plain LJ + PME forces along a Langevin run of the 648-atom water box."""
    import torch
    from mdpy_amd import _lib as L, systems
    from mdpy_amd.engine import Engine
    from mdpy_amd.replicas import maxwell_velocities
    s = systems.water_box_small()
    e = Engine(s, skin=1.5)
    e.set_positions(s.positions)
    e.set_velocities(maxwell_velocities(s.masses, 300.0, seed))
    e.configure_integrator(L.INT_LANGEVIN, dt=1.0, temperature=300.0, friction=1e-2, seed=seed)
    bonded = L.TERM_BOND | L.TERM_ANGLE
    e.integrate(L.TERM_LJ | L.TERM_PME | bonded, 100, True, 50)  # melt the lattice
    xs, fs = [], []
    for _ in range(n_frames):
        e.integrate(L.TERM_LJ | L.TERM_PME | bonded, every, True, 50)
        f, _ = e.compute(L.TERM_LJ | L.TERM_PME)
        xs.append(e.pos.clone())
        fs.append(f.clone())
    torch.cuda.synchronize()
    return s, e, xs, fs


def test_training_reduces_force_error_and_beats_no_surrogate():
    from mdpy_amd.mlp import Surrogate, force_rmse_kj
    say("case: MLP training on 648-atom water frames")
    s, e, xs, fs = _frames()
    sur = Surrogate(e, M.elements_of(s.types), rd=6.0, skin=1.0, seed=2)
    sur.fit_norm(xs[:8])
    train_x, train_f, test_x, test_f = xs[:20], fs[:20], xs[20:], fs[20:]

    def test_rmse():
        r = []
        for x, f in zip(test_x, test_f):
            e.set_positions(x, wrap=False)
            fp, _ = sur.eval()
            r.append(force_rmse_kj(fp.cpu().numpy(), f.cpu().numpy()))
        return float(np.mean(r))

    base = float(np.mean([force_rmse_kj(np.zeros_like(f.cpu().numpy()), f.cpu().numpy()) for f in test_f]))
    r0 = test_rmse()
    hist = sur.train(train_x, train_f, epochs=8, lr=2e-3, lr_decay=0.8, log_every=5)
    r1 = test_rmse()
    say("  RMSE kJ/mol/nm: no surrogate %.1f, untrained %.1f, trained %.1f; loss by epoch %s"
        % (base, r0, r1, ["%.3g" % h for h in hist]))
    assert hist[-1] < hist[0]
    assert r1 < 0.8 * base and r1 < 0.8 * r0
    sur.close()
    e.close()


def test_surrogate_inside_integrator_and_fallback(box192):
    import torch
    from mdpy_amd import _lib as L
    from mdpy_amd.mlp import SurrogateRunner
    s, e, sur, elem = box192
    say("case: MLP term in the graph step; runner fallback")
    x = e.positions()
    G, _ = M.descriptors(x, np.asarray(s.box, dtype=np.float64), elem)
    sur.set_params(_nets(9, w4_scale=0.05))
    sur.set_norm(G.mean(0), G.std(0) + 1e-3)
    f_eval, _ = sur.eval()
    f_eval = f_eval.cpu().numpy()
    sur.attach(True)
    f_term, en = e.compute(L.TERM_MLP)
    assert np.array_equal(f_term.cpu().numpy(), f_eval)
    assert np.isfinite(en["mlp"])
    e.set_velocities(np.zeros((e.n, 3)))
    e.configure_integrator(L.INT_LANGEVIN, dt=0.5, temperature=300.0, friction=1e-2, seed=1)
    e.integrate(L.TERM_MLP | L.TERM_BOND | L.TERM_ANGLE, 20, True, 10)
    torch.cuda.synchronize()
    assert e.status() == 0 and np.all(np.isfinite(e.positions()))
    sur.attach(False)
    plain = L.TERM_LJ | L.TERM_PME
    strict = SurrogateRunner(e, sur, plain, L.TERM_BOND | L.TERM_ANGLE, check_every=10, rmse_bound=1e-6)
    strict.run(30, dt_fs=0.5, energy_every=10)
    assert not strict.on_surrogate and len(strict.checks) == 1 and not strict.checks[0]["ok"]
    lax = SurrogateRunner(e, sur, plain, L.TERM_BOND | L.TERM_ANGLE, check_every=10, rmse_bound=1e12,
                          drift_bound=1e12)
    lax.run(30, dt_fs=0.5, energy_every=10)
    assert lax.on_surrogate and len(lax.checks) == 3
    sur.attach(False)


def test_descriptors_and_forces_match_oracle_at_23328_atoms():
    """The surrogate at the 23,328-atom water box (its benchmark size): every atom's 96
    descriptors against the oracle, and forces with random bf16-weight networks on 400 sampled
    atoms (the oracle's rows complete for those atoms) plus the total energy."""
    from mdpy_amd import systems
    from mdpy_amd.engine import Engine
    from mdpy_amd.mlp import Surrogate
    say("case: MLP descriptors and forces vs oracle, water23328")
    s = systems.water_box_large(seed=0)
    e = Engine(s, skin=1.0)
    e.set_positions(s.positions)
    elem = M.elements_of(s.types)
    sur = Surrogate(e, elem, rd=6.0, skin=1.0, seed=1)
    try:
        x = e.positions()
        box = np.asarray(s.box, dtype=np.float64)
        G = sur.descriptors().cpu().numpy()
        Gr, _ = M.descriptors(x, box, elem)
        scale = np.abs(Gr).max(0) + 1e-12
        err = (np.abs(G - Gr) / scale).max()
        say("  descriptors: max error / feature scale %.2e" % err)
        assert err < 1e-4
        mean, std = Gr.mean(0), Gr.std(0) + 1e-3
        nets = _nets(5)
        sur.set_params(nets)
        sur.set_norm(mean, std)
        F, E = sur.eval()
        F = F.cpu().numpy()
        ref = M.Surrogate(nets, mean.astype(np.float32).astype(np.float64), std.astype(np.float32).astype(np.float64))
        atoms = np.sort(np.random.default_rng(2).choice(len(x), 400, replace=False))
        Fr, Er = ref.forces(x, box, elem, atoms=atoms)
        say("  forces on 400 atoms: rel RMS %.3e; energy %.6e vs %.6e" % (rel_rms(F[atoms], Fr[atoms]), E, Er))
        assert rel_rms(F[atoms], Fr[atoms]) < 3e-2
        assert abs(E - Er) < 1e-2 * abs(Er) + 1e-3 * np.abs(Fr[atoms]).max()
    finally:
        sur.close()
        e.close()
