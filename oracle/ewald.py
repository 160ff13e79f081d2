"""Float64 Ewald / smooth-PME oracle (test infrastructure only; see oracle/__init__.py).

mdpy has no PME or Ewald code (mdpy/constraint/__init__.py:22 lists 'PME', 'PPPM' but nothing
uses them), so this module defines the physics of the new ElectrostaticPMEConstraint from the
published methods, in internal units with the Coulomb factor ``ke = 1/(4 pi eps0)``:

    E = E_dir + E_rec + E_self + E_excl + E_bg
    E_dir  = ke sum_{pairs, r <= rc, not excluded} q_i q_j erfc(beta r)/r
    E_rec  = ke/(2 pi V) sum_{m != 0} exp(-pi^2 m^2/beta^2)/m^2 |S(m)|^2      (Ewald 1921)
             approximated by smooth PME, order-p cardinal B-splines (Essmann et al. 1995)
    E_self = -ke beta/sqrt(pi) sum q_i^2
    E_excl = -ke sum_{excluded pairs} q_i q_j erf(beta r)/r   (1-2 and 1-3 partners)
    E_bg   = -ke pi Q^2 / (2 beta^2 V)                         (net charge Q)

beta defaults to the OpenMM rule beta = sqrt(-ln(2 tol))/rc.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erfc, erf

from .mdref import unwrap_diag


def beta_from_tolerance(cutoff, tol):
    return math.sqrt(-math.log(2.0 * tol)) / cutoff


# --------------------------------------------------------------------------- direct space
def direct_terms(positions, charges, pairs, box, beta, ke):
    """erfc-screened Coulomb over a unique pair set: forces (N,3), energy."""
    N = positions.shape[0]
    i, j = pairs[:, 0], pairs[:, 1]
    d = unwrap_diag(positions[j] - positions[i], box)
    r = np.sqrt(np.einsum("ij,ij->i", d, d))
    qq = ke * charges[i] * charges[j]
    e = qq * erfc(beta * r) / r
    # dE/dr = -qq [erfc(br)/r^2 + 2b/sqrt(pi) exp(-b^2 r^2)/r]
    g = qq * (erfc(beta * r) / r + 2 * beta / math.sqrt(math.pi) * np.exp(-(beta * r) ** 2)) / (r * r)
    f = d * g[:, None]  # force on j along +d is +g d ; on i is -g d
    F = np.zeros((N, 3))
    np.add.at(F, i, -f)
    np.add.at(F, j, f)
    return F, float(np.sum(e))


def exclusion_terms(positions, charges, excl_pairs, box, beta, ke):
    """-ke q_i q_j erf(beta r)/r for excluded (1-2, 1-3) pairs; forces and energy."""
    N = positions.shape[0]
    F = np.zeros((N, 3))
    if len(excl_pairs) == 0:
        return F, 0.0
    i, j = excl_pairs[:, 0], excl_pairs[:, 1]
    d = unwrap_diag(positions[j] - positions[i], box)
    r = np.sqrt(np.einsum("ij,ij->i", d, d))
    qq = ke * charges[i] * charges[j]
    e = -qq * erf(beta * r) / r
    # dE/dr = qq [erf(br)/r^2 - 2b/sqrt(pi) exp(-b^2 r^2)/r]
    dEdr = qq * (erf(beta * r) / (r * r) - 2 * beta / math.sqrt(math.pi) * np.exp(-(beta * r) ** 2) / r)
    f = d * (dEdr / r)[:, None]  # -dE/dx_j = -dEdr * d/r
    np.add.at(F, i, f)
    np.add.at(F, j, -f)
    return F, float(np.sum(e))


def self_energy(charges, beta, ke):
    return float(-ke * beta / math.sqrt(math.pi) * np.sum(charges ** 2))


def background_energy(charges, box, beta, ke):
    V = float(np.prod(box))
    Q = float(np.sum(charges))
    return -ke * math.pi * Q * Q / (2 * beta * beta * V)


# --------------------------------------------------------------------------- brute force
def ewald_reciprocal_bruteforce(positions, charges, box, beta, ke, kmax):
    """Direct k-space sum over |m_a| <= kmax (m = n/L): forces and energy."""
    N = positions.shape[0]
    V = float(np.prod(box))
    rng = np.arange(-kmax, kmax + 1)
    mm = np.stack(np.meshgrid(rng, rng, rng, indexing="ij"), -1).reshape(-1, 3)
    mm = mm[np.any(mm != 0, axis=1)]
    m = mm / box[None, :]
    m2 = np.einsum("ij,ij->i", m, m)
    fk = np.exp(-math.pi ** 2 * m2 / beta ** 2) / m2
    F = np.zeros((N, 3))
    E = 0.0
    for a in range(0, len(m), 4096):
        mb, fb = m[a:a + 4096], fk[a:a + 4096]
        ph = 2 * math.pi * positions @ mb.T           # (N, M)
        c, s = np.cos(ph), np.sin(ph)
        Sr, Si = charges @ c, charges @ s
        E += ke / (2 * math.pi * V) * float(np.sum(fb * (Sr * Sr + Si * Si)))
        # Im(exp(i phi_i) S*) = s_i Sr - c_i Si
        im = s * Sr[None, :] - c * Si[None, :]
        F += (2 * ke / V) * charges[:, None] * ((im * fb[None, :]) @ mb)
    return F, E


def ewald_real_bruteforce(positions, charges, box, beta, ke, nimg, excl_keys=None):
    """Real-space Ewald sum over periodic images |n_a| <= nimg (i != j or n != 0), no cutoff.
    Excluded pairs (keys i*N+j) are omitted in the home image only. Energy only."""
    N = positions.shape[0]
    E = 0.0
    rng = np.arange(-nimg, nimg + 1)
    shifts = np.stack(np.meshgrid(rng, rng, rng, indexing="ij"), -1).reshape(-1, 3) * box[None, :]
    for i in range(N):
        d = positions[None, :, :] + shifts[:, None, :] - positions[i][None, None, :]
        r = np.sqrt(np.einsum("snk,snk->sn", d, d))
        home = np.all(shifts == 0, axis=1)
        mask = np.ones_like(r, dtype=bool)
        mask[home, i] = False
        if excl_keys is not None:
            for j in range(N):
                if i * N + j in excl_keys:
                    mask[home, j] = False
        rr = np.where(mask, r, 1.0)
        E += 0.5 * ke * float(np.sum(np.where(mask, charges[i] * charges[None, :] * erfc(beta * rr) / rr, 0.0)))
    return E


# --------------------------------------------------------------------------- smooth PME
def bspline_weights(w, order):
    """Cardinal B-spline values M_p(w + k), k = 0..p-1, and their derivatives, for fractional
    offsets w in [0,1): M_2 = (w, 1-w); M_p(x) = [x M_{p-1}(x) + (p-x) M_{p-1}(x-1)]/(p-1);
    M_p'(x) = M_{p-1}(x) - M_{p-1}(x-1) (Essmann 1995 eqs. 4.1-4.2). Index k is the weight
    of grid point floor(u) - k. Returns (theta (n,p), dtheta (n,p))."""
    w = np.asarray(w, dtype=np.float64)
    vals = [w, 1.0 - w]
    prev = vals
    for p in range(3, order + 1):
        prev = vals
        vals = []
        for k in range(p):
            a = prev[k] if k < p - 1 else 0.0
            b = prev[k - 1] if k >= 1 else 0.0
            vals.append(((w + k) * a + (p - w - k) * b) / (p - 1))
    if order == 2:
        prev = [np.ones_like(w)]
    dvals = []
    for k in range(order):
        a = prev[k] if k < order - 1 else 0.0
        b = prev[k - 1] if k >= 1 else 0.0
        dvals.append(a - b)
    theta = np.stack([np.broadcast_to(v, w.shape) for v in vals], 1)
    dtheta = np.stack([np.broadcast_to(v, w.shape) for v in dvals], 1)
    return theta, dtheta


def bspline_moduli(K, order):
    """|b(m)|^2 of Essmann eq. 4.4 for m = 0..K-1."""
    # M_p at integer points 1..p-1
    w = np.zeros(1)
    th, _ = bspline_weights(w, order)  # M_p(0 + k), k=0..p-1 ; M_p(0)=0
    vals = th[0]  # vals[k] = M_p(k)
    m = np.arange(K)
    den = np.zeros(K, dtype=np.complex128)
    for k in range(order - 1):
        den += vals[k + 1] * np.exp(2j * math.pi * m * k / K)
    mod = np.abs(den) ** 2
    # interpolate zeros (odd orders at m=K/2) as in common practice
    for i in range(K):
        if mod[i] < 1e-10:
            mod[i] = 0.5 * (mod[(i - 1) % K] + mod[(i + 1) % K])
    return 1.0 / mod


def spme_influence(box, K, order, beta):
    """B(m) C(m) on the full K1 x K2 x K3 grid, with C(0) = 0."""
    V = float(np.prod(box))
    Bs = [bspline_moduli(K[a], order) for a in range(3)]
    ms = []
    for a in range(3):
        m = np.arange(K[a])
        m = np.where(m <= K[a] // 2, m, m - K[a]) / box[a]
        ms.append(m)
    mx, my, mz = np.meshgrid(*ms, indexing="ij")
    m2 = mx * mx + my * my + mz * mz
    m2[0, 0, 0] = 1.0
    C = np.exp(-math.pi ** 2 * m2 / beta ** 2) / (math.pi * V * m2)
    C[0, 0, 0] = 0.0
    B = Bs[0][:, None, None] * Bs[1][None, :, None] * Bs[2][None, None, :]
    return B * C


def spme_reciprocal(positions, charges, box, K, order, beta, ke, return_grid=False):
    """Smooth-PME reciprocal energy and forces (Essmann et al. 1995), float64, numpy FFT."""
    N = positions.shape[0]
    K = np.asarray(K, dtype=np.int64)
    s = positions / box[None, :]
    s = s - np.floor(s)
    u = s * K[None, :]
    base = np.floor(u).astype(np.int64)
    w = u - base
    th, dth = [], []
    for a in range(3):
        t, dt = bspline_weights(w[:, a], order)
        th.append(t)
        dth.append(dt)
    Q = np.zeros(tuple(K))
    idx = [(base[:, a][:, None] - np.arange(order)[None, :]) % K[a] for a in range(3)]
    for kx in range(order):
        for ky in range(order):
            wxy = charges * th[0][:, kx] * th[1][:, ky]
            for kz in range(order):
                np.add.at(Q, (idx[0][:, kx], idx[1][:, ky], idx[2][:, kz]), wxy * th[2][:, kz])
    BC = spme_influence(box, K, order, beta)
    FQ = np.fft.fftn(Q)
    E = 0.5 * ke * float(np.sum(BC * (FQ.real ** 2 + FQ.imag ** 2)))
    conv = np.real(np.fft.ifftn(BC * FQ)) * float(np.prod(K))
    F = np.zeros((N, 3))
    for kx in range(order):
        for ky in range(order):
            for kz in range(order):
                c = conv[idx[0][:, kx], idx[1][:, ky], idx[2][:, kz]]
                gx = dth[0][:, kx] * th[1][:, ky] * th[2][:, kz] * K[0] / box[0]
                gy = th[0][:, kx] * dth[1][:, ky] * th[2][:, kz] * K[1] / box[1]
                gz = th[0][:, kx] * th[1][:, ky] * dth[2][:, kz] * K[2] / box[2]
                F[:, 0] -= charges * gx * c
                F[:, 1] -= charges * gy * c
                F[:, 2] -= charges * gz * c
    F *= ke
    if return_grid:
        return F, E, Q, conv
    return F, E


def pme_total(positions, charges, box, pairs, excl_pairs, K, order, beta, ke):
    """All electrostatic terms of the PME constraint: (forces, energy, parts dict)."""
    Fd, Ed = direct_terms(positions, charges, pairs, box, beta, ke)
    Fr, Er = spme_reciprocal(positions, charges, box, K, order, beta, ke)
    Fx, Ex = exclusion_terms(positions, charges, excl_pairs, box, beta, ke)
    Es = self_energy(charges, beta, ke)
    Eb = background_energy(charges, box, beta, ke)
    parts = {"direct": Ed, "reciprocal": Er, "exclusion": Ex, "self": Es, "background": Eb}
    return Fd + Fr + Fx, Ed + Er + Ex + Es + Eb, parts
