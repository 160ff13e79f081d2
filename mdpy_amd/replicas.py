"""Batched replicas: many independent boxes of uneven size advanced together on one GPU.

Each replica is a full system (its own box, neighbour-list grid, PME grid, integrator and
Philox stream) held by its own libmdx context; a batch (``mdx_batch_*``) advances all of them
in one of these ways:

``fused``       every kernel of a step is one launch over all replicas' workgroups (per-replica
                arguments in device tables), the steps replayed from hipGraphs: the batch mode
                for many replicas (no replica-count limit);
``streams``     every replica replays its own graph on its own stream, launches interleaved;
``sequential``  replica after replica (the per-replica loop the batch is measured against);
``one_graph``   one hipGraph with a branch per replica — disabled in the library for now
                (its first GPU run ended in a host-side crash whose cause is not known).

``streams`` and ``sequential`` capture graphs per replica and are limited to ``MAX_REPLICAS``
(a 64-replica run of them crashed on the host, cause not known). ``autotune_fused`` picks the
list skin (hence the cell grids) and the nonbonded lanes per atom of the fused batch;
``autotune_skins`` picks a skin per replica for the per-replica modes.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import _lib as L
from .engine import Engine
from .unit import KB

TERMS_WATER = L.TERM_LJ | L.TERM_PME | L.TERM_BOND | L.TERM_ANGLE


def maxwell_velocities(masses, temperature, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((len(masses), 3)) * np.sqrt(KB * temperature / np.asarray(masses))[:, None]
    v -= (np.asarray(masses)[:, None] * v).sum(0) / np.sum(masses)
    return v


class ReplicaBatch:
    """Independent replicas on one device, advanced by one batch call."""

    # A 64-replica run of the per-replica modes (bench replicas leg) ended in a host-side
    # segmentation fault whose cause has not been identified; those modes stay at the size the
    # GPU tests cover. The fused mode captures no per-replica graphs and has no such limit.
    MAX_REPLICAS = 8

    def __init__(self, systems, device=0, skins=None, temperature=300.0, friction=1e-3, dt=1.0, seed=0,
                 terms=TERMS_WATER, energy_every=50):
        import torch
        self.torch = torch
        self.systems = list(systems)
        self.terms = int(terms)
        self.energy_every = int(energy_every)
        self.temperature, self.friction, self.dt, self.seed = temperature, friction, dt, seed
        skins = list(skins) if skins is not None else [1.5] * len(self.systems)
        self.engines = []
        for r, s in enumerate(self.systems):
            e = Engine(s, device=device, skin=skins[r])
            e.set_positions(s.positions)
            e.set_velocities(maxwell_velocities(s.masses, temperature, seed * 7919 + r))
            e.configure_integrator(L.INT_LANGEVIN, dt=dt, temperature=temperature, friction=friction,
                                   seed=seed * 7919 + r)
            e.set_energy_interval(self.energy_every)
            self.engines.append(e)
        self.skins = skins
        self.list_cells = 2
        self.n_atoms = int(sum(s.num_atoms for s in self.systems))
        self._batch = None
        self._make_batch()

    def _make_batch(self):
        if self._batch is not None:
            L.call("mdx_batch_destroy", self._batch)
        arr = (ctypes.c_void_p * len(self.engines))(*[e.ctx.value for e in self.engines])
        self._batch = ctypes.c_void_p()
        stream = self.torch.cuda.current_stream().cuda_stream
        L.call("mdx_batch_create", arr, len(self.engines), ctypes.c_void_p(stream), ctypes.byref(self._batch))

    def integrate(self, nsteps, mode="fused"):
        if mode != "fused" and len(self.engines) > self.MAX_REPLICAS:
            raise ValueError("batch mode %r is limited to %d replicas (a 64-replica run crashed on the host; "
                             "cause not known): use mode='fused'" % (mode, self.MAX_REPLICAS))
        L.call("mdx_batch_integrate", self._batch, self.terms, int(nsteps), L.BATCH_MODES[mode])

    def time(self, nsteps, mode="fused"):
        """Wall seconds for nsteps of every replica (synchronised on both sides)."""
        self.torch.cuda.synchronize()
        t0 = time.perf_counter()
        self.integrate(nsteps, mode)
        self.torch.cuda.synchronize()
        return time.perf_counter() - t0

    def atom_steps_per_s(self, nsteps, mode="fused"):
        return self.n_atoms * nsteps / self.time(nsteps, mode)

    def time_kernel(self, which, reps=20):
        """Average ms per launch of one batched kernel of the fused mode over all replicas, alone:
        0 nonbonded (forces only), 4 step prologue (state restored), 5 list build of a rebuild
        step, 6 / 7 / 8 PME plane forward / x-axis pass / inverse + gather. Needs a fused call
        first; the next call re-evaluates the forces."""
        ms = ctypes.c_double()
        L.call("mdx_batch_time_kernel", self._batch, int(which), int(reps), ctypes.byref(ms))
        return ms.value

    def status(self):
        return [e.status() for e in self.engines]

    # ------------------------------------------------------------------ autotuning
    def autotune_skins(self, candidates=(0.5, 1.0, 1.5, 2.0), steps=60, warmup=20):
        """Per replica, the list skin with the shortest step time (each replica timed alone)."""
        best = []
        for r, e in enumerate(self.engines):
            times = []
            for skin in candidates:
                e.configure_nonbonded(e.cutoff, skin=skin)
                e.integrate(self.terms, warmup, True, self.energy_every)
                e.synchronize()
                t0 = time.perf_counter()
                e.integrate(self.terms, steps, True, self.energy_every)
                e.synchronize()
                times.append((time.perf_counter() - t0) / steps)
            k = int(np.argmin(times))
            e.configure_nonbonded(e.cutoff, skin=candidates[k])
            best.append(candidates[k])
        self.skins = best
        return best

    def set_lanes(self, lanes):
        """Nonbonded lanes per atom of every replica (16, 32 or 64; the fused mode needs one value)."""
        for e in self.engines:
            e.tune_nonbonded(L.TERM_LJ | L.TERM_COUL_DIRECT, int(lanes))

    def set_skin(self, skin):
        """The neighbour-list skin of every replica (it sets the cell grids)."""
        for e in self.engines:
            e.configure_nonbonded(e.cutoff, skin=float(skin))
        self.skins = [float(skin)] * len(self.engines)

    def set_rebuild_every(self, every):
        """Rebuild every replica's list together every `every` steps, besides the rebuilds the
        displacement test asks for (0: displacement test only)."""
        for e in self.engines:
            e.set_rebuild_every(every)
        self.rebuild_every = int(every)

    def set_list_cells(self, per_radius):
        """Neighbour-list grid cells per list radius of every replica (1, 2 or 3)."""
        for e in self.engines:
            e.set_list_cells(per_radius)
        self.list_cells = int(per_radius)

    def autotune_fused(self, skins=(1.0, 1.5, 2.0), lanes=(16, 32), cells=(2, 3), rebuild_every=(0, 8, 10, 12),
                       steps=60, warmup=20):
        """The (skin, lanes, list cells per radius, rebuild interval) with the highest aggregate
        atom-steps/s of the fused batch, each candidate timed on the whole batch: (skin, lanes)
        at the first cell and interval candidates, then the cell size, then the interval. The
        batch is left at the best one."""
        rates = {}

        def timed(key):
            self.integrate(warmup, "fused")
            rates[key] = self.atom_steps_per_s(steps, "fused")

        self.set_rebuild_every(rebuild_every[0])
        self.set_list_cells(cells[0])
        for skin in skins:
            self.set_skin(skin)
            for g in lanes:
                self.set_lanes(g)
                timed((float(skin), int(g), int(cells[0]), int(rebuild_every[0])))
        skin, g, _, _ = max(rates, key=rates.get)
        self.set_skin(skin)
        self.set_lanes(g)
        for k in cells[1:]:
            self.set_list_cells(k)
            timed((skin, g, int(k), int(rebuild_every[0])))
        k = max(rates, key=rates.get)[2]
        self.set_list_cells(k)
        for every in rebuild_every[1:]:
            self.set_rebuild_every(every)
            timed((skin, g, k, int(every)))
        best = max(rates, key=rates.get)
        self.set_skin(best[0])
        self.set_lanes(best[1])
        self.set_list_cells(best[2])
        self.set_rebuild_every(best[3])
        return best, rates

    def autotune_mode(self, steps=100, warmup=20):
        """The batch mode with the highest aggregate atom-steps/s."""
        rates = {}
        for mode in ("streams", "sequential"):
            self.integrate(warmup, mode)
            rates[mode] = self.atom_steps_per_s(steps, mode)
        return max(rates, key=rates.get), rates

    def close(self):
        if self._batch is not None:
            L.call("mdx_batch_destroy", self._batch)
            self._batch = None
        for e in self.engines:
            e.close()
