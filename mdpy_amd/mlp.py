"""Learned force surrogate on the GPU (no reference counterpart).

A per-atom MLP over radial/angular descriptors stands in for the LJ + PME forces of a water
system; its descriptor, dense-layer (BF16 MFMA) and training kernels are libmdx's
``mdx_mlp_*``. The plain path stays the judge: :class:`SurrogateRunner` integrates with the
surrogate and, every ``check_every`` steps, evaluates the plain forces at the current
positions; when the force RMSE or the energy-error drift exceeds its bound it falls back to the
plain path.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib as L

PER_ELEM = 48
N_FEAT = 96
HIDDEN = 256


def elements_of(types):
    """0 for oxygen-like, 1 for hydrogen-like atom-type names."""
    return np.array([1 if str(t).upper().startswith("H") else 0 for t in types], dtype=np.int32)


def param_layout():
    h, f = HIDDEN, N_FEAT
    sizes = [("W1", (f, h)), ("b1", (h,)), ("W2", (h, h)), ("b2", (h,)), ("W3", (h, h)), ("b3", (h,)),
             ("w4", (h,)), ("b4", ())]
    out, off = {}, 0
    for k, shp in sizes:
        n = int(np.prod(shp)) if shp else 1
        out[k] = (off, shp)
        off += n
    return out, off


def pack(net):
    lay, n = param_layout()
    v = np.zeros(n, dtype=np.float32)
    for k, (off, shp) in lay.items():
        a = np.asarray(net[k], dtype=np.float32).reshape(-1)
        v[off:off + a.size] = a
    return v


def unpack(vec):
    lay, _ = param_layout()
    net = {}
    for k, (off, shp) in lay.items():
        n = int(np.prod(shp)) if shp else 1
        a = np.asarray(vec[off:off + n], dtype=np.float64)
        net[k] = a.reshape(shp) if shp else float(a[0])
    return net


def init_net(rng, hidden=HIDDEN, n_in=N_FEAT):
    sizes = [n_in, hidden, hidden, hidden]
    net = {}
    for k in range(3):
        net["W%d" % (k + 1)] = rng.standard_normal((sizes[k], sizes[k + 1])) * math.sqrt(1.0 / sizes[k])
        net["b%d" % (k + 1)] = np.zeros(sizes[k + 1])
    net["w4"] = rng.standard_normal(hidden) * math.sqrt(1.0 / hidden) * 1e-2
    net["b4"] = 0.0
    return net


class Surrogate:
    """The device surrogate bound to one Engine (its positions, stream and force accumulator)."""

    def __init__(self, engine, elem, rd=6.0, skin=1.0, seed=0):
        import torch
        self.torch = torch
        self.engine = engine
        self.elem = np.ascontiguousarray(np.asarray(elem, dtype=np.int32))
        self.n = engine.n
        self.rd = float(rd)
        self._h = ctypes.c_void_p()
        L.call("mdx_mlp_create", engine.ctx, L.dptr(self.elem, ctypes.c_int), self.rd, float(skin), ctypes.byref(self._h))
        self.nparam = L.load().mdx_mlp_param_count()
        rng = np.random.default_rng(seed)
        self.set_params([init_net(rng), init_net(rng)])
        self.mean = np.zeros(N_FEAT, dtype=np.float32)
        self.std = np.ones(N_FEAT, dtype=np.float32)

    # ------------------------------------------------------------------ parameters
    def set_params(self, nets):
        for e, net in enumerate(nets):
            v = pack(net)
            assert v.size == L.load().mdx_mlp_param_count()
            L.call("mdx_mlp_set_params", self._h, e, v.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))

    def get_params(self):
        out = []
        for e in range(2):
            v = np.zeros(self.nparam, dtype=np.float32)
            L.call("mdx_mlp_get_params", self._h, e, v.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
            out.append(unpack(v))
        return out

    def get_grads(self):
        out = []
        for e in range(2):
            v = np.zeros(self.nparam, dtype=np.float32)
            L.call("mdx_mlp_get_grads", self._h, e, v.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
            out.append(unpack(v))
        return out

    def set_norm(self, mean, std):
        self.mean = np.ascontiguousarray(mean, dtype=np.float32)
        self.std = np.ascontiguousarray(std, dtype=np.float32)
        L.call("mdx_mlp_set_norm", self._h, self.mean.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
               self.std.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))

    # ------------------------------------------------------------------ evaluation
    def descriptors(self):
        out = self.torch.empty((self.n, N_FEAT), dtype=self.torch.float32, device=self.engine.device)
        L.call("mdx_mlp_descriptors", self._h, ctypes.c_void_p(out.data_ptr()))
        return out

    def fit_norm(self, frames):
        """Feature mean/std over the given position frames (N x 3 arrays or tensors)."""
        acc, acc2, cnt = 0.0, 0.0, 0
        for x in frames:
            self.engine.set_positions(x, wrap=False)
            g = self.descriptors().double()
            acc = acc + g.sum(0)
            acc2 = acc2 + (g * g).sum(0)
            cnt += g.shape[0]
        mean = (acc / cnt).cpu().numpy()
        var = (acc2 / cnt).cpu().numpy() - mean ** 2
        self.set_norm(mean, np.sqrt(np.maximum(var, 0.0)) + 1e-3)

    def eval(self):
        f = self.torch.empty((self.n, 3), dtype=self.torch.float64, device=self.engine.device)
        e = ctypes.c_double()
        L.call("mdx_mlp_eval", self._h, ctypes.c_void_p(f.data_ptr()), ctypes.byref(e))
        return f, e.value

    def attach(self, on=True):
        L.call("mdx_ctx_set_mlp", self.engine.ctx, self._h if on else None)

    # ------------------------------------------------------------------ training
    def train_step(self, pos, fref, lr, want_loss=True):
        """One Adam step (lr > 0) or gradients only (lr == 0) on one frame (device tensors)."""
        loss = ctypes.c_double()
        L.call("mdx_mlp_train_step", self._h, ctypes.c_void_p(pos.data_ptr()), ctypes.c_void_p(fref.data_ptr()),
               float(lr), ctypes.byref(loss) if want_loss else None)
        return loss.value if want_loss else None

    def train(self, frames_x, frames_f, epochs=1, lr=1e-3, lr_decay=1.0, log_every=0, seed=0):
        rng = np.random.default_rng(seed)
        hist = []
        for ep in range(epochs):
            order = rng.permutation(len(frames_x))
            losses = []
            for k, idx in enumerate(order):
                want = log_every and (k % log_every == 0)
                l = self.train_step(frames_x[idx], frames_f[idx], lr, want_loss=bool(want))
                if want:
                    losses.append(l)
            hist.append(float(np.mean(losses)) if losses else float("nan"))
            lr *= lr_decay
        return hist

    def close(self):
        if self._h:
            self.attach(False)
            L.call("mdx_mlp_destroy", self._h)
            self._h = ctypes.c_void_p()


def force_rmse_kj(f_pred, f_ref):
    """Root-mean-square force error per component in kJ/mol/nm."""
    from .unit import KJ_PER_MOL_PER_NM
    d = (np.asarray(f_pred) - np.asarray(f_ref)).reshape(-1)
    return float(np.sqrt((d * d).mean()) / KJ_PER_MOL_PER_NM)


class SurrogateRunner:
    """Integrate with the surrogate, checking it against the plain path every `check_every`
    steps; falls back to the plain terms when the force RMSE (kJ/mol/nm) or the drift of the
    surrogate's energy error (kJ/mol/ns/atom) exceeds its bound."""

    def __init__(self, engine, surrogate, plain_terms, extra_terms, check_every=50, rmse_bound=0.5,
                 drift_bound=0.01):
        self.e, self.s = engine, surrogate
        self.plain_terms, self.extra_terms = int(plain_terms), int(extra_terms)
        self.check_every, self.rmse_bound, self.drift_bound = int(check_every), rmse_bound, drift_bound
        self.on_surrogate = True
        self.checks = []
        self._e0 = None
        self.steps_done = 0

    def _check(self):
        from .unit import KJ_PER_MOL
        f_plain, en = self.e.compute(self.plain_terms)
        plain_e = sum(en[k] for k in ("lj", "coulomb_direct", "coulomb_reciprocal"))
        f_sur, e_sur = self.s.eval()
        rmse = force_rmse_kj(f_sur.cpu().numpy(), f_plain.cpu().numpy() if hasattr(f_plain, "cpu") else f_plain)
        err = (e_sur - plain_e) / KJ_PER_MOL / self.e.n
        if self._e0 is None:
            self._e0 = (self.steps_done, err)
        t_ns = (self.steps_done - self._e0[0]) * self.dt_fs * 1e-6
        drift = abs(err - self._e0[1]) / t_ns if t_ns > 0 else 0.0
        ok = rmse <= self.rmse_bound and drift <= self.drift_bound
        self.checks.append({"step": self.steps_done, "rmse_kj": rmse, "energy_error_drift": drift, "ok": ok})
        return ok

    def run(self, nsteps, dt_fs=1.0, energy_every=50):
        self.dt_fs = dt_fs
        done = 0
        while done < nsteps:
            k = min(self.check_every, nsteps - done)
            if self.on_surrogate and not self._check():
                self.on_surrogate = False
                self.s.attach(False)
            if self.on_surrogate:
                self.s.attach(True)
                self.e.integrate(L.TERM_MLP | self.extra_terms, k, True, energy_every)
            else:
                self.e.integrate(self.plain_terms | self.extra_terms, k, True, energy_every)
            done += k
            self.steps_done += k
        return self.checks
