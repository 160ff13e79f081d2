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
