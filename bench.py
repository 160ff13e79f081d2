"""Benchmark: ns/day of the plain FP32 force-and-integrate step on the 23,328-atom PME water box.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Workload: 7,776 TIP3P-like waters (23,328 atoms) in a 61.9 A cube, LJ + PME (r_c 12 A,
64^3 grid, order 4, Ewald tolerance 1e-5), harmonic O-H bonds and H-O-H angles, G-JF Langevin
at 300 K with 1 fs steps and 1/ps friction; float32 pair/PME arithmetic, float64 positions and
integrator, int64 fixed-point force accumulation. Synthetic, seeded input (no files).

One "step" is one MD step (force evaluation + integrator update), replayed from a hipGraph.
Prints ONE JSON line (rank 0). The roofline block times the nonbonded kernel with HIP events
on the library stream after the timed region; the CPU baseline times the C restatement of
the reference CPU step (27-cell double-visit LJ, all-pairs Coulomb, three force evaluations
per Langevin step) on one host thread over a bounded sample.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3     # MI355X FP32 vector peak (AMD spec)
PEAK_HBM_GBS = 8000.0
FLOPS_PER_PAIR = 50          # algorithmic flops per unique in-cutoff pair (LJ + erfc Coulomb)


def maxwell_velocities(masses, temperature, seed):
    from mdpy_amd.unit import KB
    rng = np.random.default_rng(seed)
    sd = np.sqrt(KB * temperature / masses)[:, None]
    v = rng.standard_normal((len(masses), 3)) * sd
    v -= (masses[:, None] * v).sum(0) / masses.sum()
    return v


def host_cpu():
    """Model, clock and core count of the host that runs the CPU baseline (from lscpu)."""
    import subprocess
    info = {"nproc": os.cpu_count(), "usable_cores": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "Model name":
                info["model"] = v
            elif k in ("CPU max MHz", "CPU MHz") and "mhz" not in info:
                info["mhz"] = float(v)
            elif k == "Socket(s)":
                info["sockets"] = v
            elif k == "Thread(s) per core":
                info["threads_per_core"] = v
    except (OSError, ValueError, subprocess.SubprocessError) as exc:
        info["lscpu_error"] = str(exc)
    return info


def cpu_baseline(system, steps=2):
    """Time the C restatement of the reference CPU Langevin step (1 thread)."""
    from mdpy_amd.unit import COULOMB_K, KB
    from oracle.refpath import RefPath
    rp = RefPath(system, COULOMB_K, cell_cutoff=12.0)
    v = maxwell_velocities(system.masses, 300.0, 1)
    t0 = time.perf_counter()
    rp.langevin(system.positions, v, steps, 1.0, 1e-3, KB * 300.0, seed=0)
    wall = time.perf_counter() - t0
    evals = 2 + 3 * (steps - 1)            # langevin_integrator.py:51-53,64
    s_per_step = wall / evals * 3          # steady state: three evaluations per step
    progress("CPU baseline: float64 SPME oracle lines")
    return {"value": 0.0864 / s_per_step, "unit": "ns/day", "cores": 1, "kind": "port", "host": host_cpu(),
            "oracle_lines": spme_oracle_lines(system),
            "sample": "%d reference Langevin steps (%d force evaluations) of the same 23,328-atom box, "
                      "%.1f s wall; steady-state step = 3 evaluations" % (steps, evals, wall)}


def spme_oracle_lines(big, c1_steps=100):
    """The float64 numpy oracle (exact pair set, LJ, smooth PME, bonds, angles) timed on the host:
    100 G-JF Langevin steps of the 648-atom water box (r_c 9 A, PME 24^3; one evaluation per
    step), and one force evaluation of the 23,328-atom box. numpy, one process."""
    from mdpy_amd import systems
    from mdpy_amd.unit import KB, ONE_OVER_4PI_EPS0
    from oracle import ewald as EW, mdref as R

    def force_fn(s):
        ep, ei = s.exclusions()
        beta = EW.beta_from_tolerance(s.cutoff, 1e-5)
        ex = np.array(sorted({(a, int(b)) for a in range(s.num_atoms) for b in ei[ep[a]:ep[a + 1]] if b > a}))
        pbc, inv = s.pbc_matrix, np.linalg.inv(s.pbc_matrix)

        def f(x):
            xw = R.wrap_positions(x, pbc, inv)
            pairs = R.exact_pairs(xw, s.box, s.cutoff, ep, ei)
            F = R.lj_pair_terms(xw, s.lj, pairs, s.box)[0]
            F = F + EW.pme_total(xw, s.charges, s.box, pairs, ex, s.pme_grid, 4, beta, ONE_OVER_4PI_EPS0)[0]
            F = F + R.bond_terms(xw, s.bonds, s.bond_params, s.box)[0]
            return F + R.angle_terms(xw, s.angles, s.angle_params, s.box)[0]
        return f

    small = systems.water_box_small()
    f = force_fn(small)
    rng = np.random.default_rng(0)
    x = np.asarray(small.positions, dtype=np.float64)
    v = maxwell_velocities(small.masses, 300.0, 0)
    t0 = time.perf_counter()
    F0 = f(x)
    R.langevin_gjf(x, v, F0, small.masses, f, 1.0, KB * 300.0, 1e-3, rng.standard_normal, c1_steps)
    wall_small = time.perf_counter() - t0
    fb = force_fn(big)
    t0 = time.perf_counter()
    fb(np.asarray(big.positions, dtype=np.float64))
    wall_big = time.perf_counter() - t0
    return [{"workload": "water648 (r_c 9 A, PME 24^3), %d G-JF Langevin steps, float64 SPME oracle" % c1_steps,
             "value": 0.0864 / ((wall_small) / c1_steps), "unit": "ns/day", "wall_s": wall_small, "cores": 1},
            {"workload": "water23328 (r_c 12 A, PME 64^3), one float64 SPME-oracle force evaluation",
             "value": 0.0864 / wall_big, "unit": "ns/day at one evaluation per step", "wall_s": wall_big,
             "cores": 1}]


def drift_fit(etot, dt, every, n_atoms, block=20):
    """Slope of E_total(t) from a least-squares line, per atom, in kJ/mol/ns, with its standard
    error from the residual variance inflated by sqrt(block) for correlated samples."""
    from mdpy_amd.unit import KJ_PER_MOL
    t = np.arange(len(etot)) * every * dt * 1e-6
    A = np.stack([t, np.ones_like(t)], 1)
    coef = np.linalg.lstsq(A, etot, rcond=None)[0]
    r = etot - A @ coef
    se = np.sqrt(r.var() * 12.0 / len(t)) / (t[-1] - t[0]) * np.sqrt(block)
    scale = KJ_PER_MOL * n_atoms
    return float(coef[0] / scale), float(se / scale), float(etot.std() / scale)


def nve_leg(e, terms, masses):
    """Energy drift of velocity Verlet from a thermalised state (G-JF Langevin 3,000 steps at
    gamma 0.05/fs, then 2,000 at 0.01/fs): 1,000 steps at 1 fs with energies every step, then
    20,000 steps at 1 fs and at 0.5 fs with energies every 10 steps, all from the same state."""
    from mdpy_amd import _lib as L
    from mdpy_amd.unit import KB
    for gamma, n in ((0.05, 3000), (0.01, 2000)):
        e.configure_integrator(L.INT_LANGEVIN, dt=1.0, temperature=300.0, friction=gamma, seed=7)
        e.integrate(terms, n, True, 100)
    v = e.velocities()
    t_start = float((masses[:, None] * v * v).sum() / (3 * e.n * KB))
    x0, v0 = e.pos.clone(), e.vel.clone()
    out = {"start_temperature_K": t_start, "runs": []}
    for dt, steps, every in ((1.0, 1000, 1), (1.0, 20000, 10), (0.5, 20000, 10)):
        e.set_positions(x0, wrap=False)
        e.set_velocities(v0)
        e.configure_integrator(L.INT_VELOCITY_VERLET, dt=dt)
        e.integrate(terms, steps, True, every)
        h = e.energy_history()
        etot = h[:, :L.NTERMS].sum(1) + h[:, L.E_KINETIC]
        slope, se, sd = drift_fit(etot, dt, every, e.n)
        out["runs"].append({"dt_fs": dt, "steps": steps, "drift_kj_mol_ns_atom": slope,
                            "stderr_kj_mol_ns_atom": se, "etot_std_kj_mol_atom": sd,
                            "device_flags": e.status()})
    first = out["runs"][0]
    out["drift_kj_mol_ns_atom"] = first["drift_kj_mol_ns_atom"]
    out["note"] = ("a 1,000-step fit resolves the slope only to its standard error (energy "
                   "fluctuations of flexible water); the 20,000-step runs resolve it")
    return out


def chain_leg(steps=1000):
    """ns/day of the solvated chain (1,200-atom chain in 10,000 waters, 68 A, r_c 12 A, PME 72^3).
    Thermalised first (G-JF Langevin, gamma 0.05/fs then 0.01/fs) in 250-step segments that stop
    the leg if the temperature leaves 200-600 K or a device flag is raised; the timed run is
    G-JF at gamma 0.01/fs."""
    import torch
    from mdpy_amd import _lib as L, systems
    from mdpy_amd.engine import Engine
    from mdpy_amd.unit import KB
    s = systems.solvated_chain()
    e = Engine(s, skin=1.5)
    e.set_positions(s.positions)
    e.set_velocities(maxwell_velocities(s.masses, 300.0, 0))
    terms = L.TERM_LJ | L.TERM_PME | L.TERM_BOND | L.TERM_ANGLE | L.TERM_DIHEDRAL | L.TERM_IMPROPER

    def temperature():
        v = e.velocities()
        return float((s.masses[:, None] * v * v).sum() / (3 * s.num_atoms * KB))

    temps = []
    for gamma, n in ((0.05, 2000), (0.01, 1000)):
        e.configure_integrator(L.INT_LANGEVIN, dt=1.0, temperature=300.0, friction=gamma, seed=0)
        for _ in range(n // 250):
            e.integrate(terms, 250, True, 250)
            temps.append(temperature())
            if e.status() or not 200.0 < temps[-1] < 600.0:
                out = {"error": "thermalisation left bounds", "temperatures_K": temps, "device_flags": e.status()}
                e.close()
                return out
    lanes = e.tune_nonbonded(L.TERM_LJ | L.TERM_COUL_DIRECT, 0, reps=10)
    e.integrate(terms, 50, True, 50)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e.integrate(terms, steps, True, 50)
    torch.cuda.synchronize()
    sps = (time.perf_counter() - t0) / steps
    slots, pairs = e.pair_counts()
    nb_ms = e.time_kernel(0, L.TERM_LJ | L.TERM_COUL_DIRECT, reps=20)
    out = {"atoms": s.num_atoms, "ns_per_day": 0.0864 / sps, "ms_per_step": sps * 1e3, "nb_lanes": lanes,
           "nonbonded_ms": nb_ms, "unique_pairs": pairs, "pairs_per_s": pairs / (nb_ms * 1e-3),
           "temperature_K": temperature(), "device_flags": e.status()}
    e.close()
    return out


def replicas_leg(steps=200):
    """64 uneven water boxes (300-3,000 waters at 0.0334 waters/A^3, each its own cube and PME
    grid of ~1 point/A, r_c 9 A, G-JF 300 K): aggregate atom-steps/s of the fused batch (every
    kernel of a step one launch over all replicas, steps replayed from hipGraphs; list skin and
    lanes per atom autotuned on the whole batch) against the sequential per-replica loop (each
    replica's own integrate call, one after another, direct launches). The per-replica graph
    modes are limited to 8 replicas: they are compared with the fused batch on 8 replicas."""
    import torch
    from mdpy_amd import systems
    from mdpy_amd.replicas import ReplicaBatch
    reps = systems.water_replicas(64, seed=0)
    progress("replicas: 64 contexts, %d atoms" % sum(s.num_atoms for s in reps))
    b = ReplicaBatch(reps, energy_every=50)
    best, tune = b.autotune_fused(skins=(1.0, 1.5, 2.0), lanes=(16, 32), rebuild_every=(0, 8, 10, 12), steps=40,
                                  warmup=20)
    progress("replicas: fused batch at skin %.1f A, %d lanes per atom, %d list cells per radius, lists rebuilt "
             "together every %d steps (0: displacement test only)" % best)
    b.integrate(40, "fused")
    fused = b.atom_steps_per_s(steps, "fused")
    rooflines = batch_rooflines(b, reps)
    seq_steps = 50
    for e in b.engines:  # warm: list built at the chosen skin, lanes set
        e.integrate(b.terms, 10, False, 50)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e in b.engines:
        e.integrate(b.terms, seq_steps, False, 50)
    torch.cuda.synchronize()
    sequential = b.n_atoms * seq_steps / (time.perf_counter() - t0)
    flags = max(b.status())
    b.close()
    progress("replicas: 8-replica comparison with the per-replica graph modes")
    b8 = ReplicaBatch(reps[:8], energy_every=50, skins=[best[0]] * 8)
    b8.set_lanes(best[1])
    b8.set_list_cells(best[2])
    b8.set_rebuild_every(best[3])
    r8 = {}
    for mode in ("fused", "streams", "sequential"):
        b8.integrate(20, mode)
        r8[mode] = b8.atom_steps_per_s(steps, mode)
    flags = max(flags, max(b8.status()))
    b8.close()
    return {"metric": "aggregate atom-steps/s, 64 water replicas (fused batch)", "value": fused,
            "atoms": int(sum(s.num_atoms for s in reps)), "steps": steps,
            "sequential_per_replica_loop": sequential, "speedup_vs_sequential": fused / sequential,
            "autotuned": {"skin_A": best[0], "lanes": best[1], "list_cells_per_radius": best[2],
                          "rebuild_every": best[3], "candidates": {"%.1f/%d/%d/%d" % k: v for k, v in tune.items()}},
            "first_8_replicas": r8, "rooflines": rooflines, "device_flags": flags}


def batch_rooflines(b, reps):
    """Rooflines of the fused batch's kernels, each launch covering all 64 replicas (their
    working set, ~3.2e5 atoms and 64 grids, exceeds the single box's cache-resident one), timed
    alone with HIP events on the batch's stream; bytes are the per-replica rules of the single-box
    kernels summed over the replicas, pairs the unique in-cutoff pairs."""
    n = b.n_atoms
    slots = pairs = 0
    for e in b.engines:
        sl, pr = e.pair_counts()
        slots += sl
        pairs += pr
    k3 = fftb = 0
    for s in reps:
        kx, ky, kz = (int(k) for k in s.pme_grid)
        k3 += kx * ky * kz
        fftb += 2 * (4 * kx * ky * kz + 8 * ky * kz * (kx // 2 + 1))
    items = [
        ("k_nonbonded_batch", 0, 20, None),
        ("k_step_pro_batch", 4, 20, 424 * n),
        ("k_build_list_batch", 5, 5, 4 * slots + 24 * n),
        ("k_pme_plane_fwd_batch", 6, 20, 16 * n + 4 * k3),
        ("k_fft_axis_batch", 7, 20, fftb),
        ("k_pme_plane_inv_gather_batch", 8, 20, 40 * n + 4 * k3),
    ]
    out = {"atoms": n, "list_slots": slots, "unique_pairs": pairs, "grid_points": k3}
    for name, which, nrep, nbytes in items:
        ms = b.time_kernel(which, reps=nrep)
        if nbytes is None:
            tf = pairs * FLOPS_PER_PAIR / (ms * 1e-3) / 1e12
            out[name] = {"bound": "valu", "kernel_ms": ms, "achieved": tf, "peak": 157.3, "unit": "TFLOP/s",
                         "frac": tf / 157.3, "flops_per_pair": FLOPS_PER_PAIR}
        else:
            gbs = nbytes / (ms * 1e-3) / 1e9
            out[name] = {"bound": "hbm", "kernel_ms": ms, "bytes": nbytes, "achieved": gbs, "peak": PEAK_HBM_GBPS,
                         "unit": "GB/s", "frac": gbs / PEAK_HBM_GBPS}
    return out


PEAK_HBM_GBPS = 8000.0


def secondary_rooflines(e, s, terms, slots):
    """HBM rooflines of the step's other kernels, each timed alone with HIP events on the library
    stream (mdx_time_kernel): achieved = the kernel's algorithmic bytes / its launch time."""
    n = s.num_atoms
    kx, ky, kz = (int(k) for k in s.pme_grid)
    k3 = kx * ky * kz
    items = [
        ("k_pme_plane_fwd", 6, 50, 16 * n + 4 * k3,
         "spreading: 16 B/atom (packed position + charge) + 4 B per grid point"),
        ("k_fft_axis", 7, 50, 2 * (4 * k3 + 8 * ky * kz * (kx // 2 + 1)),
         "x-axis transform pair + B.C multiply: 2 x (4 K^3 + 8 K^2 (K/2+1)) bytes"),
        ("k_pme_plane_inv_gather", 8, 50, 40 * n + 4 * k3,
         "gather: 40 B/atom (position, charge, force out) + 4 B per grid point"),
        ("k_step_pro", 4, 50, 424 * n,
         "424 B/atom: reads forces 24 + reciprocal planes 96 + x, v, a_prev, xi, build position 5x24 + "
         "masses 16 + ids 8; writes x, v, a_prev, xi, zeroed forces 5x24 + packed copies 40"),
        ("k_build_list_fast", 5, 10, 4 * slots + 24 * n,
         "rebuild step: 4 B per list slot written + 24 B/atom of cell data read"),
    ]
    out = {}
    for name, which, reps, nbytes, rule in items:
        ms = e.time_kernel(which, terms, reps=reps)
        gbs = nbytes / (ms * 1e-3) / 1e9
        out[name] = {"bound": "hbm", "kernel_ms": ms, "bytes": nbytes, "achieved": gbs, "peak": PEAK_HBM_GBPS,
                     "unit": "GB/s", "frac": gbs / PEAK_HBM_GBPS, "bytes_rule": rule}
    return out


def progress(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def mlp_leg(s, e, terms, plain_ns_day, n_frames=1000, every=10, train_budget_s=40.0, max_epochs=8):
    """Train the surrogate on frames of this box (forces of the plain LJ + PME path, 10 steps
    apart) within a wall-clock budget, then measure its force error on held-out frames, its step
    rate alone and inside the plain-path judge (a check every 50 steps, fallback on failure)."""
    import torch
    from mdpy_amd import _lib as L
    from mdpy_amd.mlp import Surrogate, SurrogateRunner, elements_of, force_rmse_kj
    e.configure_integrator(L.INT_LANGEVIN, dt=1.0, temperature=300.0, friction=1e-3, seed=1)
    xs, fs = [], []
    t0 = time.perf_counter()
    for _ in range(n_frames):
        e.integrate(terms, every, True, every)
        f, _ = e.compute(L.TERM_LJ | L.TERM_PME)
        xs.append(e.pos.clone())
        fs.append(f.clone())
    t_frames = time.perf_counter() - t0
    x_keep = e.pos.clone()
    sur = Surrogate(e, elements_of(s.types), rd=6.0, skin=1.0, seed=0)
    sur.fit_norm(xs[:20])
    ntr = int(0.9 * n_frames)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist, lr = [], 1e-3
    while len(hist) < max_epochs and (not hist or time.perf_counter() - t0 < train_budget_s):
        hist += sur.train(xs[:ntr], fs[:ntr], epochs=1, lr=lr, log_every=50, seed=len(hist))
        lr *= 0.7
        torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    rm = []
    for x, f in zip(xs[ntr:], fs[ntr:]):
        e.set_positions(x, wrap=False)
        fp, _ = sur.eval()
        rm.append(force_rmse_kj(fp.cpu().numpy(), f.cpu().numpy()))
    zero = float(np.mean([force_rmse_kj(np.zeros((s.num_atoms, 3)), f.cpu().numpy()) for f in fs[ntr:]]))
    del xs, fs
    e.set_positions(x_keep, wrap=False)
    bonded = L.TERM_BOND | L.TERM_ANGLE
    sur.attach(True)
    mterms = L.TERM_MLP | bonded
    e.integrate(mterms, 100, True, 50)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e.integrate(mterms, 1000, True, 50)
    torch.cuda.synchronize()
    sps = (time.perf_counter() - t0) / 1000
    sur.attach(False)
    # inside the judge: the plain path checks the surrogate every 50 steps and takes over on failure
    e.set_positions(x_keep, wrap=False)
    run = SurrogateRunner(e, sur, L.TERM_LJ | L.TERM_PME, bonded, check_every=50)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    checks = run.run(500, dt_fs=1.0)
    torch.cuda.synchronize()
    sps_judged = (time.perf_counter() - t0) / 500
    sur.close()
    flops = 3 * 311808.0 * s.num_atoms  # dense layers, forward + input gradient, per step
    return {"frames": n_frames, "frame_every_steps": every, "frames_s": t_frames, "train_frames": ntr,
            "epochs": len(hist), "train_s": t_train, "train_ms_per_frame": 1e3 * t_train / (len(hist) * ntr),
            "loss_by_epoch": hist, "force_rmse_kj_mol_nm": float(np.mean(rm)),
            "force_rms_kj_mol_nm_no_surrogate": zero, "rmse_bound_kj_mol_nm": 0.5,
            "ns_per_day": 0.0864 / sps, "ms_per_step": sps * 1e3, "speedup_vs_plain": (0.0864 / sps) / plain_ns_day,
            "judged": {"ns_per_day": 0.0864 / sps_judged, "first_check": checks[0] if checks else None,
                       "fell_back": not run.on_surrogate},
            "mfma_roofline": {"bound": "mfma", "unit": "TFLOP/s", "flops_per_step": flops,
                              "achieved": flops / sps / 1e12, "peak": 2500.0, "frac": flops / sps / 1e12 / 2500.0,
                              "note": "3 x 311,808 dense-layer flops per atom per step over the whole MLP step time "
                                      "(descriptors and force kernels included)"},
            "note": "surrogate of LJ + PME; a 6 A-local model cannot see the force from beyond 6 A (RMS "
                    "~200 kJ/mol/nm on liquid water), so the 0.5 kJ/mol/nm bound is out of reach"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=3000, help="untimed steps before the timed ones")
    # The synthetic box starts from a lattice: over its first few picoseconds it relaxes and its
    # atoms move faster than in the 300 K liquid, so the neighbour list is rebuilt about every 6
    # steps instead of every ~10 (a rebuild step costs about twice a plain one). Preparing the
    # input relaxes it first, whatever --warmup is, so short and long runs time the same liquid.
    ap.add_argument("--equilibrate", type=int, default=3000,
                    help="untimed G-JF steps that relax the synthetic lattice-start box into the 300 K "
                         "liquid before the warm-up (part of preparing the input)")
    ap.add_argument("--skin", type=float, default=0.0,
                    help="neighbour-list skin in A (0 = autotune among 1.25, 1.5 and 1.75)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--energy-every", type=int, default=50,
                    help="potential energies every this many steps (forces every step)")
    ap.add_argument("--lanes", type=int, default=0, help="nonbonded lanes per atom (0 = autotune)")
    ap.add_argument("--cells", type=int, default=0,
                    help="neighbour-list grid cells per list radius, 1-3 (0 = autotune between 2 and 3)")
    ap.add_argument("--persist", type=int, default=-1,
                    help="nonbonded workgroups per compute unit looping over the work (0: one per unit of "
                         "work; -1 = autotune among 2, 3 and 4)")
    ap.add_argument("--phases", default="off",
                    help="'off': nonbonded and PME on two streams joined per step; 'f0,f1': the PME "
                         "launches carry the nonbonded work in shares f0, f1 and the rest, on one stream")
    ap.add_argument("--no-replicas", action="store_true",
                    help="skip the 64 batched water replicas leg (aggregate atom-steps/s of the fused batch "
                         "against the sequential per-replica loop)")
    ap.add_argument("--mlp-frames", type=int, default=1000, help="training + test frames of the MLP leg")
    ap.add_argument("--mlp-train-s", type=float, default=40.0, help="training time budget of the MLP leg (s)")
    ap.add_argument("--no-mlp", action="store_true",
                    help="skip the MLP surrogate leg (trains on 1,000 frames of this box for up to 40 s, then "
                         "measures its force error and step rate)")
    ap.add_argument("--no-nve", action="store_true",
                    help="skip the NVE drift leg (thermalise, then 1,000 + 2 x 20,000 velocity-Verlet steps)")
    ap.add_argument("--chain", action="store_true", help="also measure the solvated 31k-atom chain (72^3 PME)")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="after the run, one device-to-device copy of this many MiB (a kernel of known "
                         "traffic for scaling hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (replicas shard "
              "trivially across GPUs; see the design notes). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    from mdpy_amd import _lib as L
    from mdpy_amd import systems
    from mdpy_amd.engine import Engine

    s = systems.water_box_large(seed=0)
    e = Engine(s, skin=args.skin if args.skin > 0 else 1.5)
    e.set_positions(s.positions)
    e.set_velocities(maxwell_velocities(s.masses, 300.0, 0))
    e.configure_integrator(L.INT_LANGEVIN, dt=1.0, temperature=300.0, friction=1e-3, seed=0)
    terms = L.TERM_LJ | L.TERM_COUL_DIRECT | L.TERM_COUL_RECIP | L.TERM_BOND | L.TERM_ANGLE
    use_graph = not args.no_graph

    e.set_nonbonded_persist(args.persist if args.persist >= 0 else 3)
    if args.equilibrate > 0:
        progress("relaxing the synthetic box: %d steps" % args.equilibrate)
        e.integrate(terms, args.equilibrate, use_graph, args.energy_every)
    phases = None
    if args.phases != "off":
        phases = [float(v) for v in args.phases.split(",")]
        e.set_force_phases(True, phases[0], phases[1])
    e.integrate(terms, args.warmup, use_graph, args.energy_every)
    # autotune the step's settings one after another on the whole step (untimed setup): list skin,
    # list-grid cells per list radius, resident nonbonded workgroups per compute unit. Each
    # candidate integrates 1,000 steps twice, alternating with the others (200 re-warm steps
    # before each); the fastest minimum replaces the default only when it is at least 1 % faster
    # (run-to-run noise of a 1,000-step timing is about 0.5 %). A setting given on the command
    # line is kept.
    def tune(name, setter, candidates, default):
        best_s = {}
        for _ in range(2):
            for k in candidates:
                setter(k)
                e.integrate(terms, 200, use_graph, args.energy_every)
                torch.cuda.synchronize()
                tc = time.perf_counter()
                e.integrate(terms, 1000, use_graph, args.energy_every)
                torch.cuda.synchronize()
                best_s[k] = min(best_s.get(k, 1e9), time.perf_counter() - tc)
        k = min(best_s, key=best_s.get)
        if best_s[k] > 0.99 * best_s[default]:
            k = default
        setter(k)
        tuning[name] = {str(c): round(v * 1e3, 3) for c, v in best_s.items()}
        progress("autotune %s: %s (ms per 1,000 steps %s)" % (name, k, tuning[name]))
        return k

    tuning = {}
    skin = args.skin if args.skin > 0 else tune("list_skin_A", lambda x: e.configure_nonbonded(e.cutoff, skin=x),
                                                (1.25, 1.5, 1.75), 1.5)
    cells = args.cells if args.cells > 0 else tune("list_cells_per_radius", e.set_list_cells, (2, 3), 2)
    persist = args.persist if args.persist >= 0 else tune("nb_resident_per_cu", e.set_nonbonded_persist, (2, 3, 4), 3)
    e.configure_nonbonded(e.cutoff, skin=skin)
    e.set_list_cells(cells)
    e.set_nonbonded_persist(persist)
    e.integrate(terms, 200, use_graph, args.energy_every)
    # autotune the nonbonded kernel's lanes per atom on the current list (each candidate timed
    # over 1,500 launches, ~75 ms: stable choices, and the GPU reaches its steady clock before the
    # timed steps however short the warm-up), then re-warm the graph
    lanes = e.tune_nonbonded(L.TERM_LJ | L.TERM_COUL_DIRECT, 0, reps=1500) if args.lanes == 0 else \
        e.tune_nonbonded(L.TERM_LJ | L.TERM_COUL_DIRECT, args.lanes)
    e.integrate(terms, min(args.warmup, 500), use_graph, args.energy_every)
    torch.cuda.synchronize()
    rebuilds0 = e.stats()["rebuilds"]
    t0 = time.perf_counter()
    e.integrate(terms, args.steps, use_graph, args.energy_every)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    rebuilds_timed = e.stats()["rebuilds"] - rebuilds0
    s_per_step = (t1 - t0) / args.steps
    ns_day = 0.0864 / s_per_step  # 1 fs steps
    status = e.status()
    hist = e.energy_history()
    finite = bool(np.all(np.isfinite(hist)))

    # roofline of the dominant kernel (nonbonded LJ + erfc Coulomb)
    slots, pairs = e.pair_counts()
    nb_ms = e.time_kernel(0, L.TERM_LJ | L.TERM_COUL_DIRECT, reps=50)
    achieved = pairs * FLOPS_PER_PAIR / (nb_ms * 1e-3) / 1e12
    eval_ms = e.time_kernel(1, terms, reps=20)
    pme_ms = e.time_kernel(2, terms, reps=20)
    stats = e.stats()
    # HBM bytes per nonbonded launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on this
    # workload, scaled by the same counters on a 256 MiB copy (profiles/)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "k_nonbonded_water23328_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get("traffic_bytes_per_launch")

    out = {
        "metric": "ns/day @23,328-atom PME water, plain FP32 vs MLP surrogate at force RMSE<=0.5",
        "value": ns_day,
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": s_per_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 pair/PME, fp64 x/v/integrator, int64 fixed-point forces",
        "data": "synthetic seeded water box (no files), relaxed for %d untimed steps before the warm-up"
                % args.equilibrate,
        "config": {"workload": "water23328: 7,776 TIP3P-like waters, 61.9 A cube, r_c 12 A, PME 64^3 "
                               "order 4, G-JF Langevin 300 K, dt 1 fs, plain FP32 path",
                   "atoms": s.num_atoms, "list_skin_A": skin, "list_cells_per_radius": cells, "hipgraph": use_graph, "nb_lanes": lanes,
                   "energy_every": args.energy_every, "force_phases": phases, "nb_resident_per_cu": persist,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "kernel": "k_nonbonded (LJ + erfc Coulomb)",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "kernel_ms": nb_ms, "unique_pairs": pairs, "list_slots": slots,
                     "pairs_per_s": pairs / (nb_ms * 1e-3), "flops_per_pair": FLOPS_PER_PAIR},
        "breakdown_ms": {"force_eval": eval_ms, "pme_reciprocal": pme_ms, "nonbonded": nb_ms},
        "secondary_rooflines": secondary_rooflines(e, s, terms, slots),
        "list_rebuilds": stats["rebuilds"],
        "list_rebuilds_timed": rebuilds_timed,
        "autotune_ms_per_1000_steps": tuning or None,
        "energies_finite": finite,
        "device_flags": status,
    }
    if not args.no_nve:
        progress("NVE drift leg")
        out["nve"] = nve_leg(e, terms, s.masses)
    if args.chain:
        progress("solvated chain leg")
        out["chain"] = chain_leg()
    if not args.no_replicas:
        progress("batched replicas leg")
        out["replicas"] = replicas_leg()
    if not args.no_mlp:
        progress("MLP surrogate leg")
        out["mlp"] = mlp_leg(s, e, terms, ns_day, n_frames=args.mlp_frames, train_budget_s=args.mlp_train_s)
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(s, args.cpu_steps)
        except Exception as ex:  # reported, never silently replaced
            out["cpu_baseline"] = {"value": None, "error": repr(ex)}
    print(json.dumps(out), flush=True)
    if args.calib_copy_mb > 0:
        n = args.calib_copy_mb * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()


if __name__ == "__main__":
    main()
