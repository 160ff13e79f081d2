"""Float64 numpy restatement of mdpy v0.2.1's CPU force-and-integrate path (oracle).

Test infrastructure only (see oracle/__init__.py). Every function cites the reference
lines it restates. Arrays: positions (N,3) float64, pbc matrices (3,3) float64, exclusion
tables as CSR (ptr, idx).
"""
from __future__ import annotations

import math

import numpy as np

NUM_NEIGHBOR_CELLS = 27
CELL_LIST_SKIN = 2.0  # mdpy/core/cell_list.py:17


class ParticleLossError(Exception):
    """mdpy/error.py:203"""


class CellListPoorDefinedError(Exception):
    """mdpy/error.py:123"""


# --------------------------------------------------------------------------- PBC helpers
def wrap_positions(positions, pbc_matrix, pbc_inv):
    """mdpy/utils/pbc.py:28-36: x + (-round_half_even(x L^-1)) L; >= 2 images -> error."""
    move_vec = -np.round(np.dot(positions, pbc_inv))
    if np.max(np.abs(move_vec)) >= 2:
        raise ParticleLossError("atom(s) moved beyond 2 PBC images")
    return positions + np.dot(move_vec, pbc_matrix)


def unwrap_vec(vec, pbc_matrix, pbc_inv):
    """mdpy/utils/pbc.py:38-44: s = v L^-1; s -= rint_even(s); return s L (rows of vec)."""
    s = np.dot(vec, pbc_inv)
    s = s - np.round(s)
    return np.dot(s, pbc_matrix)


def unwrap_diag(vec, box):
    """unwrap_vec for a diagonal box, written per component: ((d*inv) - rint(d*inv)) * L.

    With a diagonal L the matrix products of pbc.py:40,44 reduce to these products exactly
    (every other term is an exact zero), so this is bit-identical to unwrap_vec there.
    """
    inv = 1.0 / box
    s = vec * inv
    s = s - np.rint(s)
    return s * box


def get_unit_vec(vec):
    """mdpy/utils/geometry.py:17-20 (rows)."""
    n = np.linalg.norm(vec, axis=-1, keepdims=True)
    return np.where(n != 0, vec / np.where(n != 0, n, 1.0), vec)


def get_pbc_dihedral(p1, p2, p3, p4, box):
    """mdpy/utils/geometry.py:78-90 (rows, diagonal box)."""
    r1 = unwrap_diag(p2 - p1, box)
    r2 = unwrap_diag(p3 - p2, box)
    r3 = unwrap_diag(p4 - p3, box)
    n1 = np.cross(r1, r2)
    n2 = np.cross(r2, r3)
    x = np.einsum("ij,ij->i", np.linalg.norm(r2, axis=1)[:, None] * r1, n2)
    y = np.einsum("ij,ij->i", n1, n2)
    return np.arctan2(x, y)


# --------------------------------------------------------------------------- cell list
def cell_grid(box, cutoff, skin=CELL_LIST_SKIN):
    """mdpy/core/cell_list.py:74-81: n = max(3, floor((L+skin)/rc)); w = (L+skin)/n; inv = 1/w."""
    box = np.asarray(box, dtype=np.float64)
    n = np.floor((box + skin) / (np.ones(3) * cutoff)).astype(np.int64)
    n[n < 3] = 3
    w = (box + skin) / n
    inv = np.diag(np.linalg.inv(np.diag(w)))
    return n, w, inv


def check_cell_cutoff(box, cutoff):
    """mdpy/core/cell_list.py:56-69: zero cutoff or floor(L/rc) < 2 is an error."""
    if cutoff == 0:
        raise CellListPoorDefinedError("cutoff radius is zero")
    if np.any(np.floor(np.asarray(box) / cutoff) < 2):
        raise CellListPoorDefinedError("the cutoff_radius is too large to create cell list")


def cell_list_build(positions, cell_inv, n):
    """mdpy/core/cell_list.py:83-116.

    positions are the wrapped positions; shift by the per-axis minimum, bin with
    floor(x * inv), fill a dense [nx,ny,nz,maxpc] table in ascending atom order, -1 padded.
    Returns (particle_cell_index (N,3) int64, cell_list (nx,ny,nz,maxpc) int64).
    """
    positive = positions - positions.min(0)
    idx = np.floor(np.dot(positive, np.diag(cell_inv))).astype(np.int64)
    counts = np.zeros(tuple(n), dtype=np.int64)
    np.add.at(counts, (idx[:, 0], idx[:, 1], idx[:, 2]), 1)
    maxpc = int(counts.max())
    table = -np.ones(tuple(n) + (maxpc,), dtype=np.int64)
    flag = np.zeros_like(counts)
    for p in range(positions.shape[0]):  # ascending atom order within a cell
        x, y, z = idx[p]
        table[x, y, z, flag[x, y, z]] = p
        flag[x, y, z] += 1
    return idx, table


def cell_list_getitem(cell_index, table, matrix_id):
    """mdpy/core/cell_list.py:45-47: ``x, y, z = particle_cell_index[matrix_id, :]`` then
    ``cell_list[x, y, z, :]``. An int id gives that atom's cell row; a 3-tuple id selects three
    atoms' index rows and indexes the table with them (the form test_cell_list.py:55-58 uses)."""
    x, y, z = cell_index[matrix_id, :]
    return table[x, y, z, :]


def stencil_offsets():
    """mdpy/constraint/__init__.py:12-19: 27 offsets, i-major then j then k over {-1,0,1}."""
    return np.array([[i, j, k] for i in range(-1, 2) for j in range(-1, 2) for k in range(-1, 2)],
                    dtype=np.int64)


# --------------------------------------------------------------------------- exclusions
def csr_rows(ptr, idx, i):
    return idx[ptr[i]:ptr[i + 1]]


def pair_keys(ptr, idx, n):
    """Encode every (i, j) listed in a CSR table as i*n + j."""
    rows = np.repeat(np.arange(len(ptr) - 1), np.diff(ptr))
    return rows.astype(np.int64) * n + idx.astype(np.int64)


# --------------------------------------------------------------------------- neighbour sets
def reference_pairs_literal(positions, box, cutoff, excl_ptr, excl_idx, cell_index, table, n):
    """Literal restatement of the pair enumeration of
    mdpy/constraint/charmm_nonbonded_constraint.py:75-100 (CPU kernel).

    For every i, walk the 27 stencil cells (offset >= n wraps by -n, -1 wraps through
    negative indexing, :78-80), skip -1 padding, i itself and excluded partners (:83), and
    keep j when the minimum-image distance r <= cutoff (:85-90). Returns the set of unique
    pairs (i<j) that the reference visits (each is visited twice there).
    Intended for small systems (pure Python loop over i).
    """
    N = positions.shape[0]
    offs = stencil_offsets()
    out = set()
    for i in range(N):
        ex = set(csr_rows(excl_ptr, excl_idx, i).tolist())
        nc = offs + cell_index[i]
        nc = nc - (nc >= n) * n
        cand = []
        for c in nc:
            cand.extend(int(j) for j in table[c[0], c[1], c[2], :] if j != -1 and j != i and j not in ex)
        if not cand:
            continue
        cand = np.array(cand)
        d = unwrap_diag(positions[cand] - positions[i], box)
        r = np.sqrt(np.einsum("ij,ij->i", d, d))
        for j in cand[r <= cutoff]:
            out.add((min(i, int(j)), max(i, int(j))))
    return out


def exact_pairs(positions, box, cutoff, excl_ptr=None, excl_idx=None):
    """All unique pairs (i<j) with minimum-image distance r <= cutoff, minus exclusions.

    Candidates come from scipy's periodic cKDTree (an algorithm independent of any cell
    grid); the final r <= cutoff decision is re-taken with the reference's float64
    sequence (minimum image as in pbc.py:38-44, then sqrt of the sum of squares).
    Returns an (M,2) int64 array sorted lexicographically.
    """
    from scipy.spatial import cKDTree

    N = positions.shape[0]
    wrapped = positions - box * np.floor(positions / box)
    wrapped = np.where(wrapped >= box, wrapped - box, wrapped)
    tree = cKDTree(wrapped, boxsize=box)
    pairs = tree.query_pairs(cutoff * (1 + 1e-9) + 1e-9, output_type="ndarray").astype(np.int64)
    pairs.sort(axis=1)
    d = unwrap_diag(positions[pairs[:, 1]] - positions[pairs[:, 0]], box)
    r = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
    pairs = pairs[r <= cutoff]
    if excl_ptr is not None and len(excl_idx):
        keys = pairs[:, 0] * N + pairs[:, 1]
        ex = pair_keys(excl_ptr, excl_idx, N)
        pairs = pairs[~np.isin(keys, ex)]
    order = np.lexsort((pairs[:, 1], pairs[:, 0]))
    return pairs[order]


def stencil_adjacent(cell_index, pairs, n):
    """True where the two atoms' reference cells are within one cell on every axis,
    cyclically at n — the pair set the 27-cell stencil can reach (cell_list.py:92-116 +
    charmm_nonbonded_constraint.py:78-80)."""
    d = np.abs(cell_index[pairs[:, 0]] - cell_index[pairs[:, 1]])
    d = np.minimum(d, n[None, :] - d)
    return np.all(d <= 1, axis=1)


def reference_pairs(positions, box, cutoff, excl_ptr, excl_idx, cell_index, n):
    """Vectorised equivalent of reference_pairs_literal: exact pairs restricted to
    stencil-adjacent reference cells (the reference's neighbour set, misses included)."""
    p = exact_pairs(positions, box, cutoff, excl_ptr, excl_idx)
    return p[stencil_adjacent(cell_index, p, np.asarray(n))]


# --------------------------------------------------------------------------- LJ
def lj_pair_terms(positions, lj, pairs, box, scaling_ptr=None, scaling_idx=None):
    """mdpy/constraint/charmm_nonbonded_constraint.py:88-107 for a given unique pair set.

    eps = sqrt(eps_i eps_j), sigma = (sigma_i + sigma_j)/2, with the (eps14, sigma14)
    columns for 1-4 partners (:92-97); s = sigma/r; force_val = -(2 s^12 - s^6)/r * eps * 24;
    F_i += force_val * d/r, F_j -= force_val * d/r with d = mi(x_j - x_i); E = 4 eps (s^12 - s^6).
    (The reference visits each pair twice at half force and half energy; the sum is the same.)
    Returns forces (N,3), energy.
    """
    N = positions.shape[0]
    i, j = pairs[:, 0], pairs[:, 1]
    d = unwrap_diag(positions[j] - positions[i], box)
    r = np.sqrt(np.einsum("ij,ij->i", d, d))
    cols = np.zeros(len(pairs), dtype=bool)
    if scaling_ptr is not None and len(scaling_idx):
        cols = np.isin(i * N + j, pair_keys(scaling_ptr, scaling_idx, N))
    e1 = np.where(cols, lj[i, 2], lj[i, 0])
    s1 = np.where(cols, lj[i, 3], lj[i, 1])
    e2 = np.where(cols, lj[j, 2], lj[j, 0])
    s2 = np.where(cols, lj[j, 3], lj[j, 1])
    eps = np.sqrt(e1 * e2)
    sig = (s1 + s2) / 2
    sr = sig / r
    sr6 = sr ** 6
    sr12 = sr6 * sr6
    fval = -(2 * sr12 - sr6) / r * eps * 24
    f = d / r[:, None] * fval[:, None]
    F = np.zeros((N, 3))
    np.add.at(F, i, f)
    np.add.at(F, j, -f)
    E = float(np.sum(4 * eps * (sr12 - sr6)))
    return F, E


# --------------------------------------------------------------------------- Coulomb
def coulomb_allpairs(positions, charges, excl_ptr, excl_idx, box, coulomb_k, chunk=2048):
    """mdpy/constraint/electrostatic_constraint.py:52-79: all i<j not excluded, minimum image,
    no cutoff. force_val = -q_i q_j / k / r^2 along d/r (d = mi(x_j - x_i)), E = q_i q_j/(k r)."""
    N = positions.shape[0]
    q = np.asarray(charges, dtype=np.float64).reshape(-1)
    F = np.zeros((N, 3))
    E = 0.0
    ex = set(pair_keys(excl_ptr, excl_idx, N).tolist()) if len(excl_idx) else set()
    exk = np.array(sorted(ex), dtype=np.int64)
    for a in range(0, N, chunk):
        ii = np.arange(a, min(N, a + chunk))
        d = unwrap_diag(positions[None, :, :] - positions[ii, None, :], box)  # (c, N, 3): j - i
        r2 = np.einsum("cnk,cnk->cn", d, d)
        mask = np.arange(N)[None, :] > ii[:, None]
        if len(exk):
            keys = ii[:, None] * N + np.arange(N)[None, :]
            mask &= ~np.isin(keys, exk)
        r = np.sqrt(np.where(mask, r2, 1.0))
        qq = q[ii, None] * q[None, :] * mask
        fval = -qq / coulomb_k / r ** 2
        f = d / r[..., None] * fval[..., None]
        F[ii] += f.sum(1)
        F -= f.sum(0)
        E += float(np.sum(qq / coulomb_k / r))
    return F, E


# --------------------------------------------------------------------------- bonded
def bond_terms(positions, bonds, params, box):
    """mdpy/constraint/charmm_bond_constraint.py:52-73: E = k(r-r0)^2; F_1 += 2k(r-r0) d/r,
    F_2 -= 2k(r-r0) d/r with d = mi(x_2 - x_1)."""
    N = positions.shape[0]
    F = np.zeros((N, 3))
    if len(bonds) == 0:
        return F, 0.0
    a, b = bonds[:, 0], bonds[:, 1]
    k, r0 = params[:, 0], params[:, 1]
    d = unwrap_diag(positions[b] - positions[a], box)
    r = np.linalg.norm(d, axis=1)
    f = (2 * k * (r - r0))[:, None] * d / r[:, None]
    np.add.at(F, a, f)
    np.add.at(F, b, -f)
    return F, float(np.sum(k * (r - r0) ** 2))


def angle_terms(positions, angles, params, box):
    """mdpy/constraint/charmm_angle_constraint.py:54-95 (harmonic angle + Urey-Bradley)."""
    N = positions.shape[0]
    F = np.zeros((N, 3))
    if len(angles) == 0:
        return F, 0.0
    i1, i2, i3 = angles[:, 0], angles[:, 1], angles[:, 2]
    k, th0, ku, u0 = params[:, 0], params[:, 1], params[:, 2], params[:, 3]
    r21 = unwrap_diag(positions[i1] - positions[i2], box)
    r23 = unwrap_diag(positions[i3] - positions[i2], box)
    l21 = np.linalg.norm(r21, axis=1)
    l23 = np.linalg.norm(r23, axis=1)
    th = np.arccos(np.einsum("ij,ij->i", r21, r23) / (l21 * l23))
    fv = -2 * k * (th - th0)
    nv = np.cross(r21, r23)
    fv1 = get_unit_vec(np.cross(r21, nv)) / l21[:, None]
    fv3 = get_unit_vec(np.cross(-r23, nv)) / l23[:, None]
    np.add.at(F, i1, fv[:, None] * fv1)
    np.add.at(F, i2, -fv[:, None] * (fv1 + fv3))
    np.add.at(F, i3, fv[:, None] * fv3)
    E = np.sum(k * (th - th0) ** 2)
    r13 = unwrap_diag(positions[i3] - positions[i1], box)
    l13 = np.linalg.norm(r13, axis=1)
    fu = (2 * ku * (l13 - u0))[:, None] * r13 / l13[:, None]
    np.add.at(F, i1, fu)
    np.add.at(F, i3, -fu)
    E += np.sum(ku * (l13 - u0) ** 2)
    return F, float(E)


def _four_body(positions, quads, force_val, box):
    """Force distribution shared by mdpy/constraint/charmm_dihedral_constraint.py:70-92 and
    mdpy/constraint/charmm_improper_constraint.py:78-97."""
    p = [positions[quads[:, c]] for c in range(4)]
    vab = unwrap_diag(p[1] - p[0], box)
    lab = np.linalg.norm(vab, axis=1)
    vbc = unwrap_diag(p[2] - p[1], box)
    lbc = np.linalg.norm(vbc, axis=1)
    voc, loc = vbc / 2, lbc / 2
    vcd = unwrap_diag(p[3] - p[2], box)
    lcd = np.linalg.norm(vcd, axis=1)
    th_abc = np.arccos(np.einsum("ij,ij->i", -vab, vbc) / (lab * lbc))
    th_bcd = np.arccos(np.einsum("ij,ij->i", -vbc, vcd) / (lbc * lcd))
    fa = (force_val / (lab * np.sin(th_abc)))[:, None] * get_unit_vec(np.cross(-vab, vbc))
    fd = (force_val / (lcd * np.sin(th_bcd)))[:, None] * get_unit_vec(np.cross(vcd, -vbc))
    fc = np.cross(-(np.cross(voc, fd) + np.cross(vcd, fd) / 2 + np.cross(-vab, fa) / 2), voc) / (loc ** 2)[:, None]
    fb = -(fa + fc + fd)
    return fa, fb, fc, fd


def dihedral_terms(positions, dihedrals, params, box, refcompat=False):
    """mdpy/constraint/charmm_dihedral_constraint.py:58-95: E = k(1 + cos(n phi - delta)).

    refcompat=True uses the reference's force_val = -k(1 - n sin(n phi - delta)) (:81),
    which is not -dE/dphi; the default uses -dE/dphi = k n sin(n phi - delta).
    """
    N = positions.shape[0]
    F = np.zeros((N, 3))
    if len(dihedrals) == 0:
        return F, 0.0
    p = [positions[dihedrals[:, c]] for c in range(4)]
    phi = get_pbc_dihedral(p[0], p[1], p[2], p[3], box)
    k, n, delta = params[:, 0], params[:, 1], params[:, 2]
    if refcompat:
        fv = -k * (1 - n * np.sin(n * phi - delta))
    else:
        fv = k * n * np.sin(n * phi - delta)
    fa, fb, fc, fd = _four_body(positions, dihedrals, fv, box)
    for c, f in enumerate((fa, fb, fc, fd)):
        np.add.at(F, dihedrals[:, c], f)
    return F, float(np.sum(k * (1 + np.cos(n * phi - delta))))


def improper_terms(positions, impropers, params, box):
    """mdpy/constraint/charmm_improper_constraint.py:56-93: E = k(psi - psi0)^2,
    force_val = -2k(psi - psi0)."""
    N = positions.shape[0]
    F = np.zeros((N, 3))
    if len(impropers) == 0:
        return F, 0.0
    p = [positions[impropers[:, c]] for c in range(4)]
    psi = get_pbc_dihedral(p[0], p[1], p[2], p[3], box)
    k, psi0 = params[:, 0], params[:, 1]
    fa, fb, fc, fd = _four_body(positions, impropers, -2 * k * (psi - psi0), box)
    for c, f in enumerate((fa, fb, fc, fd)):
        np.add.at(F, impropers[:, c], f)
    return F, float(np.sum(k * (psi - psi0) ** 2))


# --------------------------------------------------------------------------- topology tables
class TopologyConflict(Exception):
    """mdpy/error.py ParticleConflictError, as the topology rules below raise it."""


def topology_tables(n, bonds, angles, dihedrals):
    """Per-atom bonded (1-2, 1-3) and scaling (1-4) partner lists in insertion order, as
    Topology.add_bond / add_angle / add_dihedral build them through the particles
    (mdpy/core/topology.py:122-133,155-166,188-199; mdpy/core/particle.py:53-80): a bond adds
    both ends to each other's bonded list, an angle its two ends, a dihedral its two ends to the
    scaling lists; a bonded partner already present or the atom itself raises, a repeated
    scaling partner is skipped."""
    bonded = [[] for _ in range(n)]
    scaling = [[] for _ in range(n)]

    def add_bonded(i, j):
        if j in bonded[i] or j == i:
            raise TopologyConflict("bonded partner %d of %d" % (j, i))
        bonded[i].append(j)

    def add_scaling(i, j):
        if j == i:
            raise TopologyConflict("scaling partner %d of itself" % i)
        if j not in scaling[i]:
            scaling[i].append(j)

    for a, b in bonds:
        add_bonded(int(a), int(b))
        add_bonded(int(b), int(a))
    for a, _, c in angles:
        add_bonded(int(a), int(c))
        add_bonded(int(c), int(a))
    for d in dihedrals:
        add_scaling(int(d[0]), int(d[3]))
        add_scaling(int(d[3]), int(d[0]))
    return bonded, scaling


def kinetic_energy(velocities, masses):
    """mdpy/ensemble.py:63-71: sum(m v^2)/2 (Da A^2/fs^2 is the internal energy unit)."""
    return float(np.sum(np.sum(velocities ** 2, axis=1) * np.asarray(masses).reshape(-1)) / 2)


# --------------------------------------------------------------------------- integrators
class LangevinReference:
    """mdpy/integrator/langevin_integrator.py:17-72, reproduced as written.

    ``force_fn(x)`` returns the ensemble force (N,3) at positions x (the reference evaluates
    it three times per step; the extra evaluations change nothing, so one is made here).
    ``noise_fn(shape)`` supplies standard normals (the reference uses numpy's legacy global
    ``randn``). Positions handed to force_fn are wrapped exactly as State.set_positions does.
    """

    def __init__(self, dt, kbt, gamma):
        self.dt, self.gamma, self.kbt = dt, gamma, kbt
        self.sigma = math.sqrt(2 * kbt * gamma)
        self.a = (1 - gamma * dt / 2) / (1 + gamma * dt / 2)
        self.b = 1 / (1 + gamma * dt / 2)
        self.cached = False

    def integrate(self, x_state, v_state, forces_at_entry, masses, force_fn, noise_fn, pbc, pbc_inv, num_steps):
        m = np.asarray(masses).reshape(-1, 1)
        dt = self.dt
        cur_acc = forces_at_entry / m
        if not self.cached:
            self.cur_v = v_state.copy(); self.pre_v = v_state.copy()
            self.cur_x = x_state.copy(); self.pre_x = x_state.copy()
            self.cached = True
        state_x = x_state
        for _ in range(num_steps):
            xi = noise_fn(self.cur_x.shape) / np.sqrt(m)
            pre_acc = cur_acc
            new_x = (self.pre_x + self.b * dt * self.pre_v - self.b * dt ** 2 / 2 * pre_acc
                     + self.b * self.sigma * dt ** 1.5 / 2 * xi)
            self.cur_x, self.pre_x = new_x, self.cur_x
            state_x = wrap_positions(self.cur_x, pbc, pbc_inv)
            cur_acc = force_fn(state_x) / m
            new_v = (self.a * self.pre_v - dt / 2 * (self.a * pre_acc + cur_acc)
                     + self.b * self.sigma * math.sqrt(dt) * xi)
            self.cur_v, self.pre_v = new_v, self.cur_v
        return state_x, self.cur_v.copy(), cur_acc * m


class VerletReference:
    """mdpy/integrator/verlet_integrator.py:15-50, reproduced as written (start-up uses
    +a dt^2, the output velocity is mi(x_n - x_{n-1})/(2 dt))."""

    def __init__(self, dt):
        self.dt = dt
        self.cached = False

    def integrate(self, x_state, v_state, masses, force_fn, pbc, pbc_inv, num_steps):
        m = np.asarray(masses).reshape(-1, 1)
        dt = self.dt
        acc = force_fn(x_state) / m
        if not self.cached:
            self.cur_x = x_state.copy()
            self.pre_x = self.cur_x - v_state * dt + acc * dt ** 2
            self.cached = True
        state_x = x_state
        for step in range(num_steps):
            if step != 0:
                acc = force_fn(state_x) / m
            self.cur_x, self.pre_x = 2 * self.cur_x - self.pre_x + acc * dt ** 2, self.cur_x
            state_x = wrap_positions(self.cur_x, pbc, pbc_inv)
        v = unwrap_vec(self.cur_x - self.pre_x, pbc, pbc_inv) / 2 / dt
        return state_x, v


def velocity_verlet(x, v, f, masses, force_fn, dt, num_steps):
    """Textbook velocity Verlet (exact mode; not the reference's position Verlet)."""
    m = np.asarray(masses).reshape(-1, 1)
    for _ in range(num_steps):
        v = v + 0.5 * dt * f / m
        x = x + dt * v
        f = force_fn(x)
        v = v + 0.5 * dt * f / m
    return x, v, f


def langevin_gjf(x, v, f, masses, force_fn, dt, kbt, gamma, noise_fn, num_steps):
    """Gronbech-Jensen & Farago (Mol. Phys. 111, 983 (2013)) Langevin update with the
    reference's a, b and sigma = sqrt(2 kT gamma) (langevin_integrator.py:22-28), forces with
    the physical sign and starting from (x_n, v_n)."""
    m = np.asarray(masses).reshape(-1, 1)
    a = (1 - gamma * dt / 2) / (1 + gamma * dt / 2)
    b = 1 / (1 + gamma * dt / 2)
    sigma = math.sqrt(2 * kbt * gamma)
    for _ in range(num_steps):
        xi = noise_fn(x.shape) / np.sqrt(m)
        acc = f / m
        x = x + b * dt * v + b * dt ** 2 / 2 * acc + b * sigma * dt ** 1.5 / 2 * xi
        f_new = force_fn(x)
        v = a * v + dt / 2 * (a * acc + f_new / m) + b * sigma * math.sqrt(dt) * xi
        f = f_new
    return x, v, f
