"""Float64 numpy oracle of the MLP force surrogate — TEST INFRASTRUCTURE ONLY.

The reference (mdpy) has no learned force field; this module defines the surrogate's
mathematics so the HIP kernels can be checked against it:

descriptors   per atom i and neighbour element e (O = 0, H = 1), neighbours j within r_d:
              radial   G[e, k] = sum_j phi_k(r_ij),  phi_k(r) = exp(-eta (r - mu_k)^2) fc(r), k < 32
              angular  P[e, n, l] = sum_{j,k} g_n(r_ij) g_n(r_ik) (rhat_ij . rhat_ik)^l, n < 4, l < 4
                       = |T_{e,n,l}|^2 with the moment tensors T_l = sum_j g_n(r_ij) rhat_ij^(x l)
                       (O(neighbours) instead of O(neighbours^2); the j = k terms are included)
              fc(r) = (1 + cos(pi r / r_d)) / 2;  96 features = 2 elements x (32 + 16)
network       per central element: x = (G - mean) / std, three SiLU layers of width H, linear output
forces        F = -dE/dx through the network's input gradient and the descriptor Jacobian
training      force matching, L = sum_j |F_hat_j - F_j|^2 / (3 N); dL/dW by a forward-mode pass
              along u = dL/d(input gradient) followed by reverse mode (no double backprop)

Pinning: descriptor gradients and forces against central finite differences of the energy;
dL/dW against finite differences of L; the moment-tensor angular features against the explicit
double sum over neighbour pairs (tests/test_mlp_oracle.py).
"""
from __future__ import annotations

import math

import numpy as np

N_RAD = 32
N_ANG_CH = 4
N_L = 4
N_ELEM = 2
PER_ELEM = N_RAD + N_ANG_CH * N_L   # 48
N_FEAT = N_ELEM * PER_ELEM          # 96
R_D = 6.0
RAD_MU0 = 0.8
RAD_ETA = 1.0 / (2.0 * ((R_D - RAD_MU0) / (N_RAD - 1)) ** 2)
ANG_NU = (1.0, 2.5, 4.0, 5.5)
ANG_ETA = 0.5


def rad_mu():
    return RAD_MU0 + np.arange(N_RAD) * (R_D - RAD_MU0) / (N_RAD - 1)


def cutoff_fn(r, rd=R_D):
    fc = 0.5 * (np.cos(np.pi * r / rd) + 1.0)
    dfc = -0.5 * np.pi / rd * np.sin(np.pi * r / rd)
    return fc, dfc


def gauss(r, mu, eta, rd=R_D):
    """exp(-eta (r - mu)^2) fc(r) and its r-derivative; r (...,1) against mu (K,)."""
    fc, dfc = cutoff_fn(r, rd)
    e = np.exp(-eta * (r - mu) ** 2)
    return e * fc, e * (dfc - 2.0 * eta * (r - mu) * fc)


DENSE_MAX = 4096  # atoms up to which the neighbour search is the dense N x N one


def neighbor_pairs(pos, box, rd=R_D):
    """All ordered pairs (i, j), i != j, with minimum-image |r_ij| < rd; r_ij = x_j - x_i,
    sorted by (i, j). Dense for small systems, a periodic k-d tree beyond DENSE_MAX atoms (the
    same pairs, order and arithmetic: tests/test_mlp_oracle.py)."""
    if len(pos) > DENSE_MAX:
        return neighbor_pairs_tree(pos, box, rd)
    n = len(pos)
    d = pos[None, :, :] - pos[:, None, :]
    d -= box * np.round(d / box)
    r = np.sqrt((d * d).sum(-1))
    i, j = np.nonzero((r < rd) & ~np.eye(n, dtype=bool))
    return i, j, d[i, j], r[i, j]


def neighbor_pairs_tree(pos, box, rd=R_D):
    """neighbor_pairs through scipy's periodic cKDTree: candidates within rd (1 + 1e-9) of the
    wrapped positions, then the minimum-image vector and distance as the dense search forms them
    and its strict r < rd."""
    from scipy.spatial import cKDTree
    box = np.asarray(box, dtype=np.float64)
    w = pos - box * np.floor(pos / box)
    w = np.where(w >= box, w - box, w)
    pr = cKDTree(w, boxsize=box).query_pairs(rd * (1.0 + 1e-9), output_type="ndarray")
    i = np.concatenate([pr[:, 0], pr[:, 1]])
    j = np.concatenate([pr[:, 1], pr[:, 0]])
    order = np.lexsort((j, i))
    i, j = i[order], j[order]
    d = pos[j] - pos[i]
    d -= box * np.round(d / box)
    r = np.sqrt((d * d).sum(-1))
    keep = r < rd
    return i[keep], j[keep], d[keep], r[keep]


# ------------------------------------------------------------------ moment tensors (sym storage)
# rank-2 components xx xy xz yy yz zz, rank-3 xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz
R2 = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
W2 = np.array([1, 2, 2, 1, 2, 1], dtype=np.float64)
R3 = [(0, 0, 0), (0, 0, 1), (0, 0, 2), (0, 1, 1), (0, 1, 2), (0, 2, 2), (1, 1, 1), (1, 1, 2), (1, 2, 2), (2, 2, 2)]
W3 = np.array([1, 3, 3, 3, 6, 3, 1, 3, 3, 1], dtype=np.float64)
N_MOM = 1 + 3 + 6 + 10  # 20 stored components per channel


def monomials(u):
    """(M, 20) moment monomials of unit vectors u (M, 3) in the storage order above."""
    m = [np.ones(len(u))]
    m += [u[:, 0], u[:, 1], u[:, 2]]
    m += [u[:, a] * u[:, b] for a, b in R2]
    m += [u[:, a] * u[:, b] * u[:, c] for a, b, c in R3]
    return np.stack(m, axis=1)


def mom_weights():
    return np.concatenate([[1.0], np.ones(3), W2, W3])


L_SLICES = [slice(0, 1), slice(1, 4), slice(4, 10), slice(10, 20)]


def descriptors(pos, box, elem, rd=R_D):
    """G (N, 96) and moment tensors T (N, 2, 4, 20) of every atom."""
    n = len(pos)
    i, j, d, r = neighbor_pairs(pos, box, rd)
    u = d / r[:, None]
    phi, _ = gauss(r[:, None], rad_mu()[None, :], RAD_ETA, rd)
    g, _ = gauss(r[:, None], np.array(ANG_NU)[None, :], ANG_ETA, rd)
    mono = monomials(u)
    G = np.zeros((n, N_ELEM, PER_ELEM))
    T = np.zeros((n, N_ELEM, N_ANG_CH, N_MOM))
    ej = elem[j]
    for e in range(N_ELEM):
        s = ej == e
        np.add.at(G[:, e, :N_RAD], i[s], phi[s])
        np.add.at(T[:, e], i[s], g[s][:, :, None] * mono[s][:, None, :])
    w = mom_weights()
    for l, sl in enumerate(L_SLICES):
        G[:, :, N_RAD + np.arange(N_ANG_CH) * N_L + l] = (T[..., sl] ** 2 * w[sl]).sum(-1)
    return G.reshape(n, N_FEAT), T


def angular_bruteforce(pos, box, elem, rd=R_D):
    """P[i, e, n, l] by the explicit double sum over neighbour pairs (for pinning)."""
    n = len(pos)
    i, j, d, r = neighbor_pairs(pos, box, rd)
    g, _ = gauss(r[:, None], np.array(ANG_NU)[None, :], ANG_ETA, rd)
    P = np.zeros((n, N_ELEM, N_ANG_CH, N_L))
    for a in range(n):
        for e in range(N_ELEM):
            s = (i == a) & (elem[j] == e)
            if not s.any():
                continue
            uu = d[s] / r[s][:, None]
            c = uu @ uu.T
            gg = g[s]
            for ch in range(N_ANG_CH):
                w = np.outer(gg[:, ch], gg[:, ch])
                for l in range(N_L):
                    P[a, e, ch, l] = (w * c ** l).sum()
    return P


def full2_batch(t6):
    """(..., 6) symmetric storage -> (..., 3, 3)."""
    t = np.zeros(t6.shape[:-1] + (3, 3))
    for c, (a, b) in enumerate(R2):
        t[..., a, b] = t6[..., c]
        t[..., b, a] = t6[..., c]
    return t


def full3_batch(t10):
    t = np.zeros(t10.shape[:-1] + (3, 3, 3))
    for c, (a, b, d) in enumerate(R3):
        for q in {(a, b, d), (a, d, b), (b, a, d), (b, d, a), (d, a, b), (d, b, a)}:
            t[(Ellipsis,) + q] = t10[..., c]
    return t


def pair_grads(T, elem, i, j, d, r):
    """dG_i/dr_ij for every pair: (P, 96, 3), r_ij = x_j - x_i (d = r_ij vectors, r = |d|)."""
    P = len(i)
    u = d / r[:, None]
    out = np.zeros((P, N_ELEM, PER_ELEM, 3))
    _, dphi = gauss(r[:, None], rad_mu()[None, :], RAD_ETA)
    g, dg = gauss(r[:, None], np.array(ANG_NU)[None, :], ANG_ETA)
    ej = elem[j]
    t = T[i, ej]                                   # (P, 4, 20)
    t1 = t[..., 1:4]
    t2 = full2_batch(t[..., 4:10])
    t3 = full3_batch(t[..., 10:20])
    s = np.stack([t[..., 0],
                  np.einsum("pca,pa->pc", t1, u),
                  np.einsum("pcab,pa,pb->pc", t2, u, u),
                  np.einsum("pcabd,pa,pb,pd->pc", t3, u, u, u)], axis=-1)            # (P, 4, 4)
    v = np.stack([np.zeros_like(t1), t1,
                  2.0 * np.einsum("pcab,pb->pca", t2, u),
                  3.0 * np.einsum("pcabd,pb,pd->pca", t3, u, u)], axis=2)            # (P, 4, 4, 3)
    vperp = v - np.einsum("pcla,pa->pcl", v, u)[..., None] * u[:, None, None, :]
    ang = 2.0 * (dg[:, :, None, None] * s[..., None] * u[:, None, None, :]
                 + g[:, :, None, None] * vperp / r[:, None, None, None])              # (P, 4, 4, 3)
    rows = np.arange(P)
    out[rows, ej, :N_RAD] = dphi[:, :, None] * u[:, None, :]
    out[rows, ej, N_RAD:] = ang.reshape(P, N_ANG_CH * N_L, 3)
    return out.reshape(P, N_FEAT, 3)


# ------------------------------------------------------------------ network
def silu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return z * s, s * (1.0 + z * (1.0 - s)), s * (1.0 - s) * (2.0 + z * (1.0 - 2.0 * s))


def init_net(rng, hidden=256, n_in=N_FEAT):
    """Weights of one element's network (He-style scaling for SiLU)."""
    sizes = [n_in, hidden, hidden, hidden]
    net = {}
    for k in range(3):
        net["W%d" % (k + 1)] = rng.standard_normal((sizes[k], sizes[k + 1])) * math.sqrt(1.0 / sizes[k])
        net["b%d" % (k + 1)] = np.zeros(sizes[k + 1])
    net["w4"] = rng.standard_normal(hidden) * math.sqrt(1.0 / hidden)
    net["b4"] = 0.0
    return net


def net_forward(net, x):
    """Per-atom energies and the cache for the backward passes; x (M, n_in) normalised."""
    z1 = x @ net["W1"] + net["b1"]
    h1, d1, dd1 = silu(z1)
    z2 = h1 @ net["W2"] + net["b2"]
    h2, d2, dd2 = silu(z2)
    z3 = h2 @ net["W3"] + net["b3"]
    h3, d3, dd3 = silu(z3)
    e = h3 @ net["w4"] + net["b4"]
    return e, dict(x=x, h1=h1, h2=h2, h3=h3, d1=d1, d2=d2, d3=d3, dd1=dd1, dd2=dd2, dd3=dd3)


def net_input_grad(net, c):
    """dE_i/dx_i (M, n_in) and the intermediate gradients."""
    g3 = net["w4"][None, :] * c["d3"]
    h2b = g3 @ net["W3"].T
    g2 = h2b * c["d2"]
    h1b = g2 @ net["W2"].T
    g1 = h1b * c["d1"]
    gx = g1 @ net["W1"].T
    c.update(g3=g3, g2=g2, g1=g1, h2b=h2b, h1b=h1b)
    return gx


def net_tangent_weight_grad(net, c, u):
    """d/dW of s = sum_i gx_i . u_i (gx = net_input_grad), by forward-mode + reverse mode."""
    x, h1, h2 = c["x"], c["h1"], c["h2"]
    d1, d2, d3, dd1, dd2, dd3 = c["d1"], c["d2"], c["d3"], c["dd1"], c["dd2"], c["dd3"]
    g1, g2, g3, h1b, h2b = c["g1"], c["g2"], c["g3"], c["h1b"], c["h2b"]
    W1, W2, W3, w4 = net["W1"], net["W2"], net["W3"], net["w4"]
    grad = {}
    # gx = g1 W1^T
    dW1 = u.T @ g1                     # through W1 in gx
    a1 = u @ W1                        # ds/dg1
    b1 = a1 * d1                       # ds/dh1b
    c1 = a1 * h1b * dd1                # ds/dz1 via d1
    dW2 = (g2.T @ b1).T                # h1b = g2 W2^T  -> ds/dW2 = b1^T g2 (transposed)
    a2 = b1 @ W2                       # ds/dg2
    b2 = a2 * d2
    c2 = a2 * h2b * dd2
    dW3 = (g3.T @ b2).T
    a3 = b2 @ W3                       # ds/dg3
    dw4 = (a3 * d3).sum(0)
    c3 = a3 * w4[None, :] * dd3
    # reverse through the forward pass z3 = h2 W3 + b3, z2 = h1 W2 + b2, z1 = x W1 + b1
    dz3 = c3
    dW3 = dW3 + h2.T @ dz3
    db3 = dz3.sum(0)
    dz2 = c2 + (dz3 @ W3.T) * d2
    dW2 = dW2 + h1.T @ dz2
    db2 = dz2.sum(0)
    dz1 = c1 + (dz2 @ W2.T) * d1
    dW1 = dW1 + x.T @ dz1
    db1 = dz1.sum(0)
    grad.update(W1=dW1, b1=db1, W2=dW2, b2=db2, W3=dW3, b3=db3, w4=dw4, b4=0.0)
    return grad


# ------------------------------------------------------------------ surrogate
class Surrogate:
    """Two element networks + feature normalisation; energies, forces, loss, dL/dW."""

    def __init__(self, nets, mean, std, rd=R_D):
        self.nets, self.mean, self.std, self.rd = nets, np.asarray(mean), np.asarray(std), rd

    def _per_atom(self, pos, box, elem):
        G, T = descriptors(pos, box, elem, self.rd)
        x = (G - self.mean) / self.std
        E = np.zeros(len(pos))
        gx = np.zeros_like(x)
        caches = {}
        for e in range(N_ELEM):
            m = elem == e
            if not m.any():
                continue
            ee, c = net_forward(self.nets[e], x[m])
            E[m] = ee
            gx[m] = net_input_grad(self.nets[e], c)
            caches[e] = (m, c)
        gG = gx / self.std
        return E, gG, T, caches, x

    def energy(self, pos, box, elem):
        return self._per_atom(pos, box, elem)[0].sum()

    def forces(self, pos, box, elem, atoms=None):
        """Forces (complete rows for `atoms` only, when given) and the total energy."""
        E, gG, T, _, _ = self._per_atom(pos, box, elem)
        return assemble_forces(pos, box, elem, gG, T, self.rd, atoms=atoms), E.sum()

    def loss_and_grad(self, pos, box, elem, f_ref):
        """L = sum |F_hat - F|^2 / (3N) and dL/dW for both element networks."""
        E, gG, T, caches, x = self._per_atom(pos, box, elem)
        F = assemble_forces(pos, box, elem, gG, T, self.rd)
        n = len(pos)
        err = F - f_ref
        L = (err ** 2).sum() / (3 * n)
        # u_i = dL/dgG_i = -(2/3N) sum_j (e_j - e_i) . dG_i/dr_ij ; then dL/dgx = u / std
        uG = descriptor_jvp(pos, box, elem, T, -(2.0 / (3 * n)) * err, self.rd)
        ux = uG / self.std
        grads = {}
        for e, (m, c) in caches.items():
            grads[e] = net_tangent_weight_grad(self.nets[e], c, ux[m])
        return L, grads, F


def assemble_forces(pos, box, elem, gG, T, rd=R_D, atoms=None, chunk=200000):
    """F_j = sum_i dE_j/dr_ji - dE_i/dr_ij over neighbour pairs (gG = dE_i/dG_i). With `atoms`,
    only the pairs that touch those atoms are summed: their rows are complete, the others not.
    Pairs are taken in chunks (the Jacobian of a chunk is chunk x 96 x 3)."""
    i, j, d, r = neighbor_pairs(pos, box, rd)
    if atoms is not None:
        sel = np.zeros(len(pos), dtype=bool)
        sel[np.asarray(atoms)] = True
        m = sel[i] | sel[j]
        i, j, d, r = i[m], j[m], d[m], r[m]
    F = np.zeros_like(pos)
    for a in range(0, len(i), chunk):
        s = slice(a, a + chunk)
        J = pair_grads(T, elem, i[s], j[s], d[s], r[s])        # dG_i/dr_ij
        v = np.einsum("pf,pfa->pa", gG[i[s]], J)                # dE_i/dr_ij
        np.add.at(F, j[s], -v)
        np.add.at(F, i[s], v)
    return F


def descriptor_jvp(pos, box, elem, T, w, rd=R_D):
    """sum_j (w_j - w_i) . dG_i/dr_ij for every atom i: (N, 96)."""
    i, j, d, r = neighbor_pairs(pos, box, rd)
    J = pair_grads(T, elem, i, j, d, r)
    out = np.zeros((len(pos), N_FEAT))
    np.add.at(out, i, np.einsum("pfa,pa->pf", J, w[j] - w[i]))
    return out


def elements_of(types):
    """0 for oxygen-like, 1 for hydrogen-like type names."""
    return np.array([1 if str(t).upper().startswith("H") else 0 for t in types], dtype=np.int32)
