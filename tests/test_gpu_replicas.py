"""Batched replicas (uneven water boxes, one context each) on the GPU: every batch mode gives
bitwise the same trajectories as advancing each replica on its own."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _systems(k=5):
    from mdpy_amd import systems
    reps = systems.water_replicas(64, seed=0)
    sizes = np.array([s.num_atoms for s in reps])
    pick = list(np.argsort(sizes)[:: max(1, len(reps) // k)][:k])  # spread over the size range
    return [reps[i] for i in pick]


@pytest.mark.skip(reason="its first GPU run ended in a host-side segmentation fault inside mdx_batch_integrate "
                         "(most likely the one-graph mode, which is now disabled); the cause is not identified, so "
                         "this case stays off the GPU")
def test_batch_modes_bitwise_equal_to_single_replica_runs():
    import torch
    from mdpy_amd.replicas import ReplicaBatch
    sys_list = _systems(5)
    say("case: replicas, sizes", [s.num_atoms for s in sys_list], "30 steps per mode")
    results = {}
    for mode in ("one_graph", "streams", "sequential", "single"):
        b = ReplicaBatch(sys_list, energy_every=10)
        if mode == "single":
            for e in b.engines:
                e.integrate(b.terms, 30, True, 10)
        else:
            b.integrate(30, mode)
        torch.cuda.synchronize()
        results[mode] = [e.positions().copy() for e in b.engines]
        assert all(st == 0 for st in b.status()), mode
        hist = [e.energy_history() for e in b.engines]
        assert all(h.shape[0] == 4 and np.all(np.isfinite(h)) for h in hist), mode
        b.close()
    for mode in ("one_graph", "streams", "sequential"):
        for r, (a, c) in enumerate(zip(results[mode], results["single"])):
            assert np.array_equal(a, c), (mode, r, np.abs(a - c).max())
    moved = [np.abs(a - s.positions).max() for a, s in zip(results["single"], sys_list)]
    assert all(m > 1e-3 for m in moved)


def test_stream_and_sequential_batches_bitwise_equal_to_single_runs():
    import torch
    from mdpy_amd.replicas import ReplicaBatch
    sys_list = _systems(4)
    results = {}
    for mode in ("single", "sequential", "streams"):
        say("case: replicas mode %s, sizes %s, 30 steps" % (mode, [s.num_atoms for s in sys_list]))
        b = ReplicaBatch(sys_list, energy_every=10)
        if mode == "single":
            for e in b.engines:
                e.integrate(b.terms, 30, True, 10)
        else:
            b.integrate(30, mode)
        torch.cuda.synchronize()
        results[mode] = [e.positions().copy() for e in b.engines]
        assert all(st == 0 for st in b.status()), mode
        hist = [e.energy_history() for e in b.engines]
        assert all(h.shape[0] == 4 and np.all(np.isfinite(h)) for h in hist), mode
        b.close()
    for mode in ("sequential", "streams"):
        for r, (a, c) in enumerate(zip(results[mode], results["single"])):
            assert np.array_equal(a, c), (mode, r, np.abs(a - c).max())
    with pytest.raises(ValueError):  # the one-graph mode is switched off
        from mdpy_amd import _lib as L
        b = ReplicaBatch(sys_list[:1], energy_every=10)
        try:
            L.call("mdx_batch_integrate", b._batch, b.terms, 10, L.BATCH_ONE_GRAPH)
        finally:
            b.close()


def test_batch_repeated_calls_and_autotune():
    import torch
    from mdpy_amd.replicas import ReplicaBatch
    sys_list = _systems(3)
    say("case: replicas autotune, sizes", [s.num_atoms for s in sys_list])
    b = ReplicaBatch(sys_list, energy_every=10)
    skins = b.autotune_skins(candidates=(1.0, 1.5), steps=20, warmup=10)
    assert len(skins) == 3 and all(s in (1.0, 1.5) for s in skins)
    b._make_batch()
    best, rates = b.autotune_mode(steps=20, warmup=10)
    say("  skins", skins, "rates", {k: "%.3g" % v for k, v in rates.items()}, "best", best)
    assert best in rates and all(v > 0 for v in rates.values())
    for _ in range(3):  # cached per-replica graphs replayed across calls
        b.integrate(20, "streams")
    torch.cuda.synchronize()
    assert all(st == 0 for st in b.status())
    b.close()


def _run_batch(sys_list, calls, mode, every=10, kind=None):
    """positions, velocities and energy histories after each call: the batch in `mode`, or each
    replica alone ('single': its own integrate calls, direct launches)."""
    import torch
    from mdpy_amd import _lib as L
    from mdpy_amd.replicas import ReplicaBatch
    b = ReplicaBatch(sys_list, energy_every=every)
    if kind == "vv":
        for e in b.engines:
            e.configure_integrator(L.INT_VELOCITY_VERLET, dt=1.0)
    out = []
    for n in calls:
        if mode == "single":
            for e in b.engines:
                e.integrate(b.terms, n, False, every)
        else:
            b.integrate(n, mode)
        torch.cuda.synchronize()
        out.append([(e.positions().copy(), e.velocities().copy(), e.energy_history().copy()) for e in b.engines])
    status = b.status()
    b.close()
    return out, status


@pytest.mark.parametrize("kind", ["gjf", "vv"])
def test_fused_batch_bitwise_equal_to_single_runs(kind):
    """MDX_BATCH_FUSED (one launch per kernel per step over all replicas) against each replica
    advanced alone: two calls (30 and 17 steps, energies every 10), positions, velocities and
    energy histories equal bit for bit, for G-JF Langevin and for velocity Verlet."""
    sys_list = _systems(5)
    say("case: fused batch vs single runs (%s), sizes %s, calls 30 + 17" % (kind, [s.num_atoms for s in sys_list]))
    fused, sf = _run_batch(sys_list, (30, 17), "fused", kind=kind)
    single, ss = _run_batch(sys_list, (30, 17), "single", kind=kind)
    assert all(v == 0 for v in sf) and all(v == 0 for v in ss)
    for k in range(2):
        for r in range(len(sys_list)):
            for a, c in zip(fused[k][r], single[k][r]):
                assert np.array_equal(a, c), (k, r)
    assert all(h.shape[0] == 2 and np.all(np.isfinite(h)) for _, _, h in fused[1])
    moved = [np.abs(p - s.positions).max() for (p, _, _), s in zip(fused[1], sys_list)]
    assert all(m > 1e-3 for m in moved)


def test_fused_batch_64_replicas_equal_single_runs():
    """All 64 replicas of the batched-replica workload (300-3,000 waters each, 321,897 atoms) in
    one fused batch: 20 steps equal each replica's own run bit for bit."""
    from mdpy_amd import systems
    reps = systems.water_replicas(64, seed=0)
    say("case: fused batch, 64 replicas, %d atoms, 20 steps" % sum(s.num_atoms for s in reps))
    fused, sf = _run_batch(reps, (20,), "fused")
    single, ss = _run_batch(reps, (20,), "single")
    assert all(v == 0 for v in sf) and all(v == 0 for v in ss)
    for r in range(len(reps)):
        for a, c in zip(fused[0][r], single[0][r]):
            assert np.array_equal(a, c), r


def test_fused_batch_autotune_and_repeated_calls():
    """The fused batch's (skin, lanes) autotuner over a few candidates, then repeated calls that
    replay the captured graphs (and recapture after a setting changes)."""
    import torch
    from mdpy_amd.replicas import ReplicaBatch
    sys_list = _systems(3)
    say("case: fused batch autotune, sizes", [s.num_atoms for s in sys_list])
    b = ReplicaBatch(sys_list, energy_every=10)
    best, rates = b.autotune_fused(skins=(1.0, 1.5), lanes=(16, 32), rebuild_every=(0, 10), steps=20, warmup=10)
    say("  rates", {k: "%.3g" % v for k, v in rates.items()}, "best", best)
    assert best in rates and len(rates) == 6 and all(v > 0 for v in rates.values())
    for n in (20, 33, 1, 0, 12):
        b.integrate(n, "fused")
    torch.cuda.synchronize()
    assert all(st == 0 for st in b.status()), b.status()
    assert all(np.all(np.isfinite(e.positions())) for e in b.engines)
    b.close()


def test_fused_batch_kernel_timing_leaves_the_trajectory_unchanged():
    """Timing the fused batch's nonbonded, prologue (state saved and restored) and PME launches
    between two calls changes nothing: the next call re-evaluates the forces and follows the
    untimed trajectory bit for bit. Timing the list build (a full rebuild on the next call)
    leaves a valid, finite run."""
    import torch
    from mdpy_amd.replicas import ReplicaBatch
    sys_list = _systems(3)
    say("case: fused batch kernel timing between calls, sizes", [s.num_atoms for s in sys_list])
    out = []
    for timed in (False, True):
        b = ReplicaBatch(sys_list, energy_every=10)
        b.integrate(20, "fused")
        if timed:
            for which in (0, 4, 6, 7, 8):
                assert b.time_kernel(which, reps=3) > 0
        b.integrate(15, "fused")
        torch.cuda.synchronize()
        out.append([(e.positions().copy(), e.velocities().copy()) for e in b.engines])
        if timed:
            assert b.time_kernel(5, reps=2) > 0
            b.integrate(10, "fused")
            torch.cuda.synchronize()
            assert all(np.all(np.isfinite(e.positions())) for e in b.engines)
        assert all(st == 0 for st in b.status())
        b.close()
    for r in range(len(sys_list)):
        for a, c in zip(out[0][r], out[1][r]):
            assert np.array_equal(a, c), r


def test_fused_batch_with_scheduled_rebuilds_equal_single_runs():
    """Lists rebuilt together every 7 steps (besides the displacement test): the fused batch and
    each replica alone with the same interval stay bitwise equal, over calls that do not start on
    a multiple of the interval."""
    import torch
    from mdpy_amd.replicas import ReplicaBatch
    sys_list = _systems(4)
    say("case: fused batch with lists rebuilt every 7 steps, sizes", [s.num_atoms for s in sys_list])
    out = []
    for mode in ("fused", "single"):
        b = ReplicaBatch(sys_list, energy_every=10)
        b.set_rebuild_every(7)
        run = []
        for n in (23, 12):
            if mode == "single":
                for e in b.engines:
                    e.integrate(b.terms, n, False, 10)
            else:
                b.integrate(n, "fused")
            torch.cuda.synchronize()
            run.append([(e.positions().copy(), e.velocities().copy(), e.energy_history().copy()) for e in b.engines])
        rebuilds = [e.stats()["rebuilds"] for e in b.engines]
        assert all(st == 0 for st in b.status()), b.status()
        b.close()
        out.append((run, rebuilds))
    assert out[0][1] == out[1][1] and all(r >= 5 for r in out[0][1])  # 35 steps: at least 5 scheduled
    for k in range(2):
        for r in range(len(sys_list)):
            for a, c in zip(out[0][0][k][r], out[1][0][k][r]):
                assert np.array_equal(a, c), (k, r)
