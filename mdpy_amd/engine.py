"""Device-side system: PyTorch-owned HIP buffers plus a libmdx context.

This is the thin layer every mdpy-shaped class in this package drives (Ensemble, the
constraints, the integrators). Positions and velocities are float64 (N,3) tensors on the HIP
device; the library borrows their pointers and works on its own stream joined to PyTorch's.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib as L
from .error import DeviceError
from .unit import KB


def _require_device(device):
    import torch
    if not torch.cuda.is_available():
        raise DeviceError("no HIP device visible: the force-and-integrate path runs on the GPU only")
    return torch.device("cuda", device)


def beta_from_tolerance(cutoff, tol):
    """Ewald splitting parameter beta = sqrt(-ln(2 tol))/rc (the rule OpenMM uses)."""
    return math.sqrt(-math.log(2.0 * tol)) / cutoff


class Engine:
    """A system resident on one GPU.

    Parameters mirror the reference's force-field wiring: cutoff (CharmmNonbondedConstraint
    default 12 A, charmm_nonbonded_constraint.py:22), the cell-list cutoff that defines the
    reference stencil (cell_list.py:20, default 12 A), and the PME grid/tolerance of the new
    electrostatics.
    """

    def __init__(self, system, device=0, cutoff=None, skin=1.5, pair_mode="exact", cell_cutoff=12.0,
                 pme_grid=None, ewald_tol=1e-5, beta=None, dihedral_refcompat=False, velocities=None):
        import torch
        self.torch = torch
        self.device = _require_device(device)
        self.lib = L.load()
        self.system = system
        self.n = system.num_atoms
        self.box = np.asarray(system.box, dtype=np.float64).copy()
        self.cutoff = float(system.cutoff if cutoff is None else cutoff)
        self.pos = torch.tensor(np.ascontiguousarray(system.positions, dtype=np.float64), device=self.device)
        v = np.zeros((self.n, 3)) if velocities is None else np.asarray(velocities, dtype=np.float64)
        self.vel = torch.tensor(np.ascontiguousarray(v), device=self.device)
        self.force_buf = torch.zeros((self.n, 3), dtype=torch.float64, device=self.device)
        self._ctx = ctypes.c_void_p()
        stream = torch.cuda.current_stream(self.device).cuda_stream
        L.call("mdx_ctx_create", self.device.index or 0, ctypes.c_void_p(stream), ctypes.byref(self._ctx))
        self.ctx = self._ctx
        self._bind()
        self.set_masses(system.masses)
        self.set_charges(system.charges)
        self.set_lj(system.lj)
        ep, ei = system.exclusions()
        self.set_exclusions(ep, ei)
        sp, si = system.scaling_pairs()
        self.set_scaling(sp, si)
        self.set_bonded(system, dihedral_refcompat)
        self.pair_mode = L.PAIRS_REFERENCE if pair_mode == "reference" else L.PAIRS_EXACT
        self.skin = 0.0 if self.pair_mode == L.PAIRS_REFERENCE else float(skin)
        self.cell_cutoff = float(cell_cutoff)
        L.call("mdx_nonbonded_config", self.ctx, self.cutoff, self.skin, self.pair_mode, self.cell_cutoff)
        self.pme_grid = tuple(system.pme_grid if pme_grid is None else pme_grid)
        self.beta = float(beta) if beta is not None else beta_from_tolerance(self.cutoff, ewald_tol)
        grid = (ctypes.c_int * 3)(*self.pme_grid)
        L.call("mdx_pme_config", self.ctx, grid, 4, self.beta)
        self._closed = False

    @classmethod
    def bare(cls, n_atoms, box, device=0, positions=None):
        """A context with positions/velocities bound but no force-field parameters yet; the
        mdpy-shaped classes fill it term by term (bind_ensemble of each constraint)."""
        import torch
        self = cls.__new__(cls)
        self.torch = torch
        self.device = _require_device(device)
        self.lib = L.load()
        self.system = None
        self.n = int(n_atoms)
        self.box = np.asarray(box, dtype=np.float64).copy()
        self.cutoff = None
        x = np.zeros((self.n, 3)) if positions is None else np.asarray(positions, dtype=np.float64)
        self.pos = torch.tensor(np.ascontiguousarray(x), device=self.device)
        self.vel = torch.zeros((self.n, 3), dtype=torch.float64, device=self.device)
        self.force_buf = torch.zeros((self.n, 3), dtype=torch.float64, device=self.device)
        self._ctx = ctypes.c_void_p()
        stream = torch.cuda.current_stream(self.device).cuda_stream
        L.call("mdx_ctx_create", self.device.index or 0, ctypes.c_void_p(stream), ctypes.byref(self._ctx))
        self.ctx = self._ctx
        self._bind()
        self.pair_mode, self.skin, self.cell_cutoff = L.PAIRS_EXACT, 1.5, 12.0
        self.beta, self.pme_grid = None, None
        self._closed = False
        return self

    def configure_nonbonded(self, cutoff, skin=1.5, pair_mode="exact", cell_cutoff=12.0):
        self.cutoff = float(cutoff)
        self.pair_mode = L.PAIRS_REFERENCE if pair_mode == "reference" else L.PAIRS_EXACT
        self.skin = 0.0 if self.pair_mode == L.PAIRS_REFERENCE else float(skin)
        self.cell_cutoff = float(cell_cutoff)
        L.call("mdx_nonbonded_config", self.ctx, self.cutoff, self.skin, self.pair_mode, self.cell_cutoff)

    def configure_pme(self, grid, beta, order=4):
        self.pme_grid = tuple(int(k) for k in grid)
        self.beta = float(beta)
        L.call("mdx_pme_config", self.ctx, (ctypes.c_int * 3)(*self.pme_grid), int(order), self.beta)

    # ------------------------------------------------------------------ setup
    def set_pme_fused(self, enable=True):
        """Reciprocal space in one launch or in three (the default); the forces are the same."""
        L.call("mdx_pme_set_fused", self.ctx, 1 if enable else 0)

    def set_force_phases(self, enable=True, frac0=0.26, frac1=0.13):
        """LJ + PME evaluations as three launches on one stream, each PME phase carrying a share
        of the nonbonded work (frac0, frac1, rest), instead of two streams joined per step."""
        L.call("mdx_force_phases_config", self.ctx, 1 if enable else 0, float(frac0), float(frac1))

    def _bind(self):
        box = (ctypes.c_double * 3)(*self.box.tolist())
        L.call("mdx_system_set", self.ctx, self.n, ctypes.c_void_p(self.pos.data_ptr()),
               ctypes.c_void_p(self.vel.data_ptr()), box)

    def set_box(self, box):
        self.box = np.asarray(box, dtype=np.float64).copy()
        L.call("mdx_set_box", self.ctx, (ctypes.c_double * 3)(*self.box.tolist()))

    def set_masses(self, masses):
        m = np.ascontiguousarray(np.asarray(masses, dtype=np.float64).reshape(-1))
        self.masses = m
        L.call("mdx_set_masses", self.ctx, L.dptr(m, ctypes.c_double))

    def set_charges(self, charges):
        q = np.ascontiguousarray(np.asarray(charges, dtype=np.float64).reshape(-1))
        self.charges = q
        L.call("mdx_set_charges", self.ctx, L.dptr(q, ctypes.c_double))

    def set_lj(self, lj):
        a = np.ascontiguousarray(np.asarray(lj, dtype=np.float64).reshape(-1, 4))
        L.call("mdx_set_lj", self.ctx, L.dptr(a, ctypes.c_double))

    def set_exclusions(self, ptr, idx):
        p = np.ascontiguousarray(ptr, dtype=np.int32)
        i = np.ascontiguousarray(idx if len(idx) else [0], dtype=np.int32)
        L.call("mdx_set_exclusions", self.ctx, L.dptr(p, ctypes.c_int), L.dptr(i, ctypes.c_int))

    def set_scaling(self, ptr, idx):
        p = np.ascontiguousarray(ptr, dtype=np.int32)
        i = np.ascontiguousarray(idx if len(idx) else [0], dtype=np.int32)
        L.call("mdx_set_scaling", self.ctx, L.dptr(p, ctypes.c_int), L.dptr(i, ctypes.c_int))

    def _put_terms(self, fn, idx, par, *extra):
        i = np.ascontiguousarray(idx, dtype=np.int32).reshape(-1)
        p = np.ascontiguousarray(par, dtype=np.float64).reshape(-1)
        count = int(len(idx))
        if len(i) == 0:
            i = np.zeros(1, np.int32)
            p = np.zeros(1)
        L.call(fn, self.ctx, count, L.dptr(i, ctypes.c_int), L.dptr(p, ctypes.c_double), *extra)

    def set_bonds(self, idx, par):
        self._put_terms("mdx_set_bonds", idx, par)

    def set_angles(self, idx, par):
        self._put_terms("mdx_set_angles", idx, par)

    def set_dihedrals(self, idx, par, refcompat=False):
        self._put_terms("mdx_set_dihedrals", idx, par, 1 if refcompat else 0)

    def set_impropers(self, idx, par):
        self._put_terms("mdx_set_impropers", idx, par)

    def set_bonded(self, s, dihedral_refcompat=False):
        self.set_bonds(s.bonds, s.bond_params)
        self.set_angles(s.angles, s.angle_params)
        self.set_dihedrals(s.dihedrals, s.dihedral_params, dihedral_refcompat)
        self.set_impropers(s.impropers, s.improper_params)

    # ------------------------------------------------------------------ state
    def set_positions(self, x, wrap=True):
        if isinstance(x, self.torch.Tensor):
            self.pos.copy_(x.to(device=self.device, dtype=self.torch.float64))
        else:
            self.pos.copy_(self.torch.as_tensor(np.asarray(x, dtype=np.float64), device=self.device))
        L.call("mdx_positions_changed", self.ctx)
        if wrap:
            L.call("mdx_wrap_positions", self.ctx)

    def set_velocities(self, v):
        if isinstance(v, self.torch.Tensor):
            self.vel.copy_(v.to(device=self.device, dtype=self.torch.float64))
        else:
            self.vel.copy_(self.torch.as_tensor(np.asarray(v, dtype=np.float64), device=self.device))

    def positions(self):
        return self.pos.cpu().numpy()

    def velocities(self):
        return self.vel.cpu().numpy()

    # ------------------------------------------------------------------ forces
    def compute(self, terms, out=None):
        """Evaluate `terms` (bit mask); returns (forces tensor (N,3) f64, energies dict)."""
        buf = self.force_buf if out is None else out
        L.call("mdx_compute_forces", self.ctx, int(terms), ctypes.c_void_p(buf.data_ptr()))
        return buf, self.energies()

    def compute_forces_only(self, terms, out=None):
        """Forces of `terms` with the kernels of integrator steps that record no energies."""
        buf = self.force_buf if out is None else out
        L.call("mdx_compute_forces_only", self.ctx, int(terms), ctypes.c_void_p(buf.data_ptr()))
        return buf

    def compute_async(self, terms, out=None):
        buf = self.force_buf if out is None else out
        L.call("mdx_compute_forces", self.ctx, int(terms), ctypes.c_void_p(buf.data_ptr() if buf is not None else 0))
        return buf

    def energies(self):
        e = (ctypes.c_double * (L.NTERMS + 1))()
        L.call("mdx_energies_fetch", self.ctx, e)
        return dict(zip(L.TERM_NAMES, list(e)))

    # ------------------------------------------------------------------ cell list / pairs
    def celllist_reference(self, cell_cutoff=None):
        """The reference's particle_cell_index and dense cell table (cell_list.py:83-116)."""
        torch = self.torch
        cc = self.cell_cutoff if cell_cutoff is None else float(cell_cutoff)
        idx = torch.zeros((self.n, 3), dtype=torch.int32, device=self.device)
        n = (ctypes.c_int * 3)()
        maxpc = ctypes.c_int()
        L.call("mdx_celllist_reference", self.ctx, cc, ctypes.c_void_p(idx.data_ptr()), n, ctypes.byref(maxpc))
        dims = tuple(n)
        table = torch.empty(dims + (maxpc.value,), dtype=torch.int32, device=self.device)
        L.call("mdx_celllist_table", self.ctx, ctypes.c_void_p(table.data_ptr()), maxpc.value)
        torch.cuda.synchronize(self.device)
        return idx.cpu().numpy().astype(np.int64), table.cpu().numpy().astype(np.int64), np.array(dims)

    def neighbor_pairs(self):
        """Unique in-cutoff pairs (i<j, original ids) of the configured pair set, sorted."""
        torch = self.torch
        cap = max(1024, int(self.n * 4.0 / 3.0 * math.pi * self.cutoff ** 3 * self.n / np.prod(self.box) * 0.6) + 1024)
        for _ in range(3):
            buf = torch.empty((cap, 2), dtype=torch.int32, device=self.device)
            cnt = ctypes.c_int64()
            L.call("mdx_neighbor_pairs", self.ctx, ctypes.c_void_p(buf.data_ptr()), cap, ctypes.byref(cnt))
            if cnt.value <= cap:
                p = buf[: cnt.value].cpu().numpy().astype(np.int64)
                order = np.lexsort((p[:, 1], p[:, 0]))
                return p[order]
            cap = int(cnt.value * 1.1) + 16
        raise RuntimeError("pair buffer sizing failed")

    def pair_counts(self):
        a, b = ctypes.c_int64(), ctypes.c_int64()
        L.call("mdx_pair_counts", self.ctx, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def tune_nonbonded(self, terms, lanes=0, reps=20):
        """Lanes per atom of the nonbonded kernel (0 = time 16/32/64 on the current list)."""
        out = ctypes.c_int()
        L.call("mdx_nonbonded_tune", self.ctx, int(terms), int(lanes), int(reps), ctypes.byref(out))
        return out.value

    def set_nonbonded_persist(self, per_cu):
        """Nonbonded workgroups per compute unit looping over the work (0: one per unit of work)."""
        L.call("mdx_nonbonded_persist", self.ctx, int(per_cu))

    def set_rebuild_every(self, every):
        """Also rebuild the neighbour list every `every` integrator steps (0: only when the
        displacement test asks for it, which it still does either way)."""
        L.call("mdx_nonbonded_rebuild_every", self.ctx, int(every))

    def set_list_cells(self, per_radius):
        """Neighbour-list grid cells per list radius along each edge (1, 2 = default, 3): the
        same pairs either way, in another row order (forces equal to float32 summation
        rounding); the next evaluation rebuilds."""
        L.call("mdx_nonbonded_cells", self.ctx, int(per_radius))

    def time_kernel(self, which, terms, reps=20):
        """Average ms per launch (HIP events on the library stream): which 0 nonbonded,
        1 force evaluation, 2 PME reciprocal chain, 3 integrator step, 4 step prologue alone
        (state restored), 5 list build of a rebuild step, 6 / 7 / 8 PME plane forward / x-axis
        pass / inverse + gather alone."""
        ms = ctypes.c_double()
        L.call("mdx_time_kernel", self.ctx, int(which), int(terms), int(reps), ctypes.byref(ms))
        return ms.value

    def stats(self):
        a, b = ctypes.c_int64(), ctypes.c_int64()
        L.call("mdx_stats_fetch", self.ctx, ctypes.byref(a), ctypes.byref(b))
        return {"rebuilds": a.value, "evaluations": b.value}

    def status(self):
        f = ctypes.c_uint()
        L.call("mdx_status_fetch", self.ctx, ctypes.byref(f))
        return f.value

    # ------------------------------------------------------------------ integration
    def configure_integrator(self, kind, dt, temperature=0.0, friction=0.0, seed=0):
        kt = KB * float(temperature)
        L.call("mdx_integrator_config", self.ctx, int(kind), float(dt), float(friction), kt, int(seed))

    def reset_integrator(self):
        L.call("mdx_integrator_reset", self.ctx)

    def set_external_noise(self, noise):
        """Standard normals (nsteps, N, 3) used in place of the Philox stream (None = Philox)."""
        if noise is None:
            self._noise = None
            L.call("mdx_set_external_noise", self.ctx, None, 0)
            return
        t = self.torch.as_tensor(np.ascontiguousarray(noise, dtype=np.float64), device=self.device).contiguous()
        self._noise = t
        L.call("mdx_set_external_noise", self.ctx, ctypes.c_void_p(t.data_ptr()), int(t.shape[0]))

    def set_energy_interval(self, every):
        """Potential energies every `every` steps of integrate (a change re-captures the step
        graphs, so an unchanged interval is not sent again)."""
        if getattr(self, "_energy_every", None) != int(every):
            L.call("mdx_set_energy_interval", self.ctx, int(every))
            self._energy_every = int(every)

    def integrate(self, terms, nsteps, use_graph=True, energy_every=1):
        """Advance nsteps; potential energies are evaluated every `energy_every` steps."""
        self.set_energy_interval(energy_every)
        L.call("mdx_integrate", self.ctx, int(terms), int(nsteps), 1 if use_graph else 0)

    def energy_history_async(self, out):
        """Enqueue a copy of the last integrate call's energy-history rows into the pinned host
        tensor `out` ((rows, NTERMS + 1) float64) and return the row count without waiting; the copy
        is ordered on PyTorch's current stream after the steps."""
        cnt = ctypes.c_int()
        L.call("mdx_energy_history_copy_async", self.ctx, ctypes.c_void_p(out.data_ptr()), int(out.shape[0]),
               ctypes.byref(cnt))
        return cnt.value

    def energy_history(self, cap=None):
        cap = cap or 1 << 20
        out = np.zeros((cap, L.NTERMS + 1))
        cnt = ctypes.c_int()
        L.call("mdx_energy_history_fetch", self.ctx, L.dptr(out, ctypes.c_double), cap, ctypes.byref(cnt))
        return out[: cnt.value]

    def synchronize(self):
        self.torch.cuda.synchronize(self.device)

    def close(self):
        if not getattr(self, "_closed", True):
            L.call("mdx_ctx_destroy", self.ctx)
            self._closed = True

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
