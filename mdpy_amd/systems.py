"""Seeded synthetic systems for the force-and-integrate path.

Four workloads are generated here, all in internal units (A, Da, fs, e):

* ``water_box_small``   216 TIP3P-like waters (648 atoms), 6^3 lattice at 3.1 A, 18.6 A cube.
* ``water_box_large``   7,776 waters (23,328 atoms), 7,776 sites drawn from a 20^3 lattice at
                        3.095 A by a seeded permutation, 61.9 A cube.
* ``solvated_chain``    1,200-atom self-avoiding random-walk chain with CHARMM-form bonds,
                        angles (+Urey-Bradley), periodic dihedrals and impropers, embedded in
                        10,000 waters (30,000 atoms) in a 68 A cube.
* ``water_replicas``    independent water boxes of 300-3,000 waters at 0.0334 waters/A^3.

Water model: r_OH 0.9572 A, theta_HOH 104.52 deg, q_O -0.834 e, q_H +0.417 e, harmonic
O-H bonds (k 450 kcal/mol/A^2) and H-O-H angles (k 55 kcal/mol/rad^2), LJ on O only
(eps 0.1521 kcal/mol, Rmin/2 1.7682 A), masses O 15.9994 / H 1.008 Da
(mdpy/test/data/toppar_water_ions_namd.str:38,40). Parameter conventions follow the
reference parser: eps is stored positive, sigma = 2 * Rmin/2 * 2^(-1/6), angles in radians
(mdpy/io/charmm_toppar_parser.py:16,160-222).

Molecule orientations are uniform random rotations from Shoemake's quaternion method
(K. Shoemake, "Uniform random rotations", Graphics Gems III, 1992); jitter is U(-0.1, 0.1) A
per axis; all randomness comes from ``numpy.random.default_rng(seed)``.
"""
from __future__ import annotations

import hashlib
import math
from dataclasses import dataclass, field

import numpy as np

from .unit import KCAL_PER_MOL, RMIN_TO_SIGMA_FACTOR

R_OH = 0.9572
THETA_HOH = math.radians(104.52)
Q_O, Q_H = -0.834, 0.417
K_BOND_OH = 450.0 * KCAL_PER_MOL
K_ANGLE_HOH = 55.0 * KCAL_PER_MOL
EPS_O = 0.1521 * KCAL_PER_MOL
SIGMA_O = 2.0 * 1.7682 * RMIN_TO_SIGMA_FACTOR
SIGMA_H = 2.0 * 0.2245 * RMIN_TO_SIGMA_FACTOR   # irrelevant: eps_H = 0
M_O, M_H = 15.9994, 1.008


@dataclass
class System:
    """A system as flat arrays (the layout every layer of this package consumes)."""

    name: str
    positions: np.ndarray          # (N,3) float64
    box: np.ndarray                # (3,) float64, orthorhombic edge lengths
    masses: np.ndarray             # (N,) float64
    charges: np.ndarray            # (N,) float64
    lj: np.ndarray                 # (N,4) float64: eps, sigma, eps14, sigma14
    types: list                    # N atom-type names (keys of the parameter tables)
    bonds: np.ndarray              # (nb,2) int64
    bond_params: np.ndarray        # (nb,2) k, r0
    angles: np.ndarray             # (na,3) int64
    angle_params: np.ndarray       # (na,4) k, theta0, k_ub, u0
    dihedrals: np.ndarray          # (nd,4) int64, one row per cosine term
    dihedral_params: np.ndarray    # (nd,3) k, n, delta
    impropers: np.ndarray          # (ni,4) int64
    improper_params: np.ndarray    # (ni,2) k, psi0
    cutoff: float = 12.0
    pme_grid: tuple = (64, 64, 64)
    meta: dict = field(default_factory=dict)

    @property
    def num_atoms(self) -> int:
        return int(self.positions.shape[0])

    @property
    def pbc_matrix(self) -> np.ndarray:
        return np.diag(self.box.astype(np.float64))

    def sha256(self) -> str:
        """Digest of the float64 arrays that define the system."""
        h = hashlib.sha256()
        for a in (self.positions, self.box, self.masses, self.charges, self.lj,
                  self.bond_params, self.angle_params, self.dihedral_params, self.improper_params):
            h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        for a in (self.bonds, self.angles, self.dihedrals, self.impropers):
            h.update(np.ascontiguousarray(a, dtype=np.int64).tobytes())
        return h.hexdigest()

    def exclusions(self):
        """Per-atom excluded partners: 1-2 from bonds plus (end, end) of every angle.

        Same rule as Topology.add_bond/add_angle (mdpy/core/topology.py:122-167).
        Returns CSR (ptr, idx) with each row sorted ascending.
        """
        n = self.num_atoms
        rows = [[] for _ in range(n)]
        pairs = [self.bonds[:, [0, 1]]] if len(self.bonds) else []
        if len(self.angles):
            pairs.append(self.angles[:, [0, 2]])
        for p in pairs:
            for a, b in p:
                rows[a].append(int(b))
                rows[b].append(int(a))
        return _to_csr(rows)

    def scaling_pairs(self):
        """Per-atom 1-4 partners from the ends of every dihedral (mdpy/core/topology.py:188-200)."""
        n = self.num_atoms
        rows = [set() for _ in range(n)]
        for d in self.dihedrals:
            a, b = int(d[0]), int(d[3])
            rows[a].add(b)
            rows[b].add(a)
        return _to_csr([sorted(r) for r in rows])


def _to_csr(rows):
    ptr = np.zeros(len(rows) + 1, dtype=np.int64)
    for i, r in enumerate(rows):
        ptr[i + 1] = ptr[i] + len(r)
    idx = np.zeros(int(ptr[-1]), dtype=np.int64)
    for i, r in enumerate(rows):
        idx[ptr[i]:ptr[i + 1]] = np.sort(np.asarray(r, dtype=np.int64))
    return ptr, idx


def random_rotations(rng, n):
    """n uniform random rotation matrices (Shoemake quaternions)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    x, y = a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2)
    z, w = b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def _water_arrays(rng, sites):
    """Place one water per site (O at the jittered site, random orientation)."""
    nw = sites.shape[0]
    s, c = math.sin(THETA_HOH / 2), math.cos(THETA_HOH / 2)
    frame = np.array([[0.0, 0.0, 0.0], [R_OH * s, 0.0, R_OH * c], [-R_OH * s, 0.0, R_OH * c]])
    jitter = rng.uniform(-0.1, 0.1, size=(nw, 3))
    R = random_rotations(rng, nw)
    o = sites + jitter
    pos = o[:, None, :] + np.einsum("nij,aj->nai", R, frame)
    return pos.reshape(-1, 3)


def _water_topology(nw, offset=0):
    base = offset + 3 * np.arange(nw)
    bonds = np.stack([np.stack([base, base + 1], 1), np.stack([base, base + 2], 1)], 1).reshape(-1, 2)
    angles = np.stack([base + 1, base, base + 2], 1)
    bond_params = np.tile([K_BOND_OH, R_OH], (2 * nw, 1))
    angle_params = np.tile([K_ANGLE_HOH, THETA_HOH, 0.0, 0.0], (nw, 1))
    return bonds.astype(np.int64), bond_params, angles.astype(np.int64), angle_params


def _water_atom_arrays(nw):
    masses = np.tile([M_O, M_H, M_H], nw)
    charges = np.tile([Q_O, Q_H, Q_H], nw)
    lj_o = [EPS_O, SIGMA_O, EPS_O, SIGMA_O]
    lj_h = [0.0, SIGMA_H, 0.0, SIGMA_H]
    lj = np.tile(np.array([lj_o, lj_h, lj_h]), (nw, 1))
    types = ["OT", "HT", "HT"] * nw
    return masses, charges, lj, types


def _empty(cols, dtype=np.float64):
    return np.zeros((0, cols), dtype=dtype)


def water_box(lattice_n, spacing, n_waters=None, seed=0, cutoff=12.0, pme_grid=(64, 64, 64), name="water"):
    """Waters on a lattice_n^3 cubic lattice (edge lattice_n*spacing), optionally a seeded subset."""
    rng = np.random.default_rng(seed)
    g = np.arange(lattice_n) * spacing
    sites = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    if n_waters is not None and n_waters < sites.shape[0]:
        sel = np.sort(rng.permutation(sites.shape[0])[:n_waters])
        sites = sites[sel]
    nw = sites.shape[0]
    pos = _water_arrays(rng, sites)
    masses, charges, lj, types = _water_atom_arrays(nw)
    bonds, bp, angles, ap = _water_topology(nw)
    L = lattice_n * spacing
    return System(name=name, positions=pos, box=np.array([L, L, L]), masses=masses, charges=charges,
                  lj=lj, types=types, bonds=bonds, bond_params=bp, angles=angles, angle_params=ap,
                  dihedrals=_empty(4, np.int64), dihedral_params=_empty(3),
                  impropers=_empty(4, np.int64), improper_params=_empty(2),
                  cutoff=cutoff, pme_grid=tuple(pme_grid), meta={"n_waters": nw, "seed": seed})


def water_box_small(seed=0):
    """216 waters, 18.6 A cube, r_c 9 A, PME grid 24^3."""
    return water_box(6, 3.1, seed=seed, cutoff=9.0, pme_grid=(24, 24, 24), name="water648")


def water_box_large(seed=0):
    """7,776 waters (23,328 atoms), 61.9 A cube, r_c 12 A, PME grid 64^3."""
    return water_box(20, 3.095, n_waters=7776, seed=seed, cutoff=12.0, pme_grid=(64, 64, 64),
                     name="water23328")


def fft_friendly(n):
    """Smallest integer >= n whose prime factors are all in {2, 3, 5}."""
    m = max(int(n), 1)
    while True:
        k = m
        for p in (2, 3, 5):
            while k % p == 0:
                k //= p
        if k == 1:
            return m
        m += 1


def water_replicas(n_replicas=64, seed=0, min_waters=300, max_waters=3000, density=0.0334, cutoff=9.0):
    """Independent water boxes of seeded uneven size, each with its own edge and PME grid."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(min_waters, max_waters + 1, size=n_replicas)
    out = []
    for r, nw in enumerate(sizes):
        nw = int(nw)
        L = (nw / density) ** (1.0 / 3.0)
        m = int(math.ceil(nw ** (1.0 / 3.0)))
        k = fft_friendly(int(math.ceil(L)))
        s = water_box(m, L / m, n_waters=nw, seed=seed * 1000 + r + 1, cutoff=cutoff, pme_grid=(k, k, k),
                      name="replica%d" % r)
        out.append(s)
    return out


def _min_image(d, L):
    return d - L * np.round(d / L)


def _chain_molecule(rng, n_chain, L):
    """Positions and CHARMM-form parameters of a self-avoiding random-walk chain in a cube of
    edge L (periodic).

    Geometry: bond lengths U(1.0, 1.5) A; successive bond directions 50-75 deg apart (bond
    angles 105-130 deg, so no angle approaches 180 deg, where the dihedral normals vanish);
    every non-bonded pair (|i-j| > 2) at least 2.5 A apart under the minimum image.
    Parameters: bonds k U(200, 400) kcal/mol/A^2 at their generated length; angles k U(40, 80)
    kcal/mol/rad^2 at their generated angle, 30% with a Urey-Bradley term k U(5, 20); 1-3
    periodic dihedral terms per quad, k U(0, 2) kcal/mol, n in {1, 2, 3}, delta in {0, 180};
    every tenth quad (moved forward to the next quad whose angle is within 90 deg of 0) carries a
    harmonic improper k U(5, 20) near its generated angle. The improper energy k (psi - psi0)^2 is
    not periodic (mdpy/constraint/charmm_improper_constraint.py:70,90), so psi0 is kept away from
    +-180 deg, where a crossing would jump the energy. Charges U(-0.6, 0.6) e shifted to a
    neutral chain; LJ eps U(0.05, 0.2) kcal/mol, Rmin/2 U(1.7, 2.1) A, 1-4 eps halved.
    """
    pos = np.zeros((n_chain, 3))
    pos[0] = rng.uniform(0, L, 3)
    dirs = np.zeros((n_chain, 3))
    c_lo, c_hi = math.cos(math.radians(75)), math.cos(math.radians(50))
    i, fails, tries = 1, 0, 0
    while i < n_chain:
        tries += 1
        if tries > 400000:
            raise RuntimeError("chain generation failed")
        if i == 1:
            v = rng.normal(size=3)
            v /= np.linalg.norm(v)
        else:  # uniform on the cone band around the previous direction
            p = dirs[i - 1]
            u = np.cross(p, [1.0, 0.0, 0.0] if abs(p[0]) < 0.9 else [0.0, 1.0, 0.0])
            u /= np.linalg.norm(u)
            w = np.cross(p, u)
            c = rng.uniform(c_lo, c_hi)
            az = rng.uniform(0, 2 * math.pi)
            v = c * p + math.sqrt(1 - c * c) * (math.cos(az) * u + math.sin(az) * w)
        b = rng.uniform(1.0, 1.5)
        cand = pos[i - 1] + b * v
        if i > 2:
            d = _min_image(pos[: i - 2] - cand, L)
            if np.min(np.einsum("ij,ij->i", d, d)) < 2.5 ** 2:
                fails += 1
                if fails > 200:  # trapped: step back a few monomers
                    i = max(1, i - 5)
                    fails = 0
                continue
        pos[i] = cand
        dirs[i] = v
        fails = 0
        i += 1
    idx = np.arange(n_chain)
    bonds = np.stack([idx[:-1], idx[1:]], 1)
    blen = np.linalg.norm(pos[1:] - pos[:-1], axis=1)
    bond_params = np.stack([rng.uniform(200, 400, n_chain - 1) * KCAL_PER_MOL, blen], 1)
    angles = np.stack([idx[:-2], idx[1:-1], idx[2:]], 1)
    v1 = pos[:-2] - pos[1:-1]
    v2 = pos[2:] - pos[1:-1]
    th = np.arccos(np.einsum("ij,ij->i", v1, v2) / (np.linalg.norm(v1, axis=1) * np.linalg.norm(v2, axis=1)))
    ku = np.where(rng.random(n_chain - 2) < 0.3, rng.uniform(5, 20, n_chain - 2), 0.0) * KCAL_PER_MOL
    u0 = np.linalg.norm(pos[2:] - pos[:-2], axis=1)
    angle_params = np.stack([rng.uniform(40, 80, n_chain - 2) * KCAL_PER_MOL, th, ku,
                             np.where(ku > 0, u0 + rng.uniform(-0.05, 0.05, n_chain - 2), 0.0)], 1)
    dih_rows, dih_params = [], []
    for a in range(n_chain - 3):
        nterms = int(rng.integers(1, 4))
        for n in rng.choice([1, 2, 3], size=nterms, replace=False):
            dih_rows.append([a, a + 1, a + 2, a + 3])
            dih_params.append([rng.uniform(0, 2) * KCAL_PER_MOL, float(n), math.radians(rng.choice([0.0, 180.0]))])
    dihedrals = np.array(dih_rows, dtype=np.int64).reshape(-1, 4)
    dihedral_params = np.array(dih_params).reshape(-1, 3)
    quads = np.array([[a, a + 1, a + 2, a + 3] for a in range(n_chain - 3)], dtype=np.int64).reshape(-1, 4)
    psi_all = _dihedral_angles(pos, quads, L) if len(quads) else np.zeros(0)
    imp_rows = []
    for a0 in range(0, n_chain - 3, 10):
        for a in range(a0, min(a0 + 10, n_chain - 3)):
            if abs(psi_all[a]) < 0.5 * math.pi:
                imp_rows.append(a)
                break
    impropers = quads[imp_rows].reshape(-1, 4)
    improper_params = np.stack([rng.uniform(5, 20, len(imp_rows)) * KCAL_PER_MOL,
                                psi_all[imp_rows] + rng.uniform(-0.1, 0.1, len(imp_rows))], 1).reshape(-1, 2)
    q = rng.uniform(-0.6, 0.6, n_chain)
    q -= q.mean()
    eps = rng.uniform(0.05, 0.2, n_chain) * KCAL_PER_MOL
    sig = 2.0 * rng.uniform(1.7, 2.1, n_chain) * RMIN_TO_SIGMA_FACTOR
    lj = np.stack([eps, sig, 0.5 * eps, sig], 1)
    return dict(pos=pos, bonds=bonds, bond_params=bond_params, angles=angles, angle_params=angle_params,
                dihedrals=dihedrals, dihedral_params=dihedral_params, impropers=impropers,
                improper_params=improper_params, charges=q, lj=lj)


def chain_in_vacuum(seed=0, n_chain=120, L=200.0, cutoff=12.0):
    """The chain molecule alone in a large periodic cube (no solvent)."""
    rng = np.random.default_rng(seed)
    c = _chain_molecule(rng, n_chain, L)
    positions = c["pos"] - L * np.floor(c["pos"] / L)
    return System(
        name="chain%d_vacuum" % n_chain, positions=positions, box=np.array([L, L, L]),
        masses=np.full(n_chain, 12.011), charges=c["charges"], lj=c["lj"],
        types=["C%d" % a for a in range(n_chain)], bonds=c["bonds"], bond_params=c["bond_params"],
        angles=c["angles"], angle_params=c["angle_params"], dihedrals=c["dihedrals"],
        dihedral_params=c["dihedral_params"], impropers=c["impropers"], improper_params=c["improper_params"],
        cutoff=cutoff, pme_grid=(64, 64, 64), meta={"n_chain": n_chain, "seed": seed})


def solvated_chain(seed=0, n_chain=1200, n_waters=10000, L=68.0, cutoff=12.0, pme_grid=(72, 72, 72)):
    """Self-avoiding random-walk chain with CHARMM-form bonded terms, embedded in water."""
    rng = np.random.default_rng(seed)
    c = _chain_molecule(rng, n_chain, L)
    pos = c["pos"]
    # --- waters on a 23^3 lattice, minus sites that overlap the chain, then a seeded subset
    m = 23
    g = (np.arange(m) + 0.5) * (L / m)
    sites = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    keep = np.ones(len(sites), dtype=bool)
    for c0 in range(0, n_chain, 200):
        d = _min_image(sites[:, None, :] - pos[None, c0:c0 + 200, :], L)
        keep &= np.min(np.einsum("ijk,ijk->ij", d, d), axis=1) >= 3.0 ** 2
    free = np.nonzero(keep)[0]
    if len(free) < n_waters:
        raise RuntimeError("not enough free water sites: %d" % len(free))
    sel = np.sort(rng.permutation(free)[:n_waters])
    wpos = _water_arrays(rng, sites[sel])
    wm, wq, wlj, wtypes = _water_atom_arrays(n_waters)
    wb, wbp, wa, wap = _water_topology(n_waters, offset=n_chain)
    positions = np.concatenate([pos, wpos])
    positions -= L * np.floor(positions / L)  # the walk can leave the box: store periodic images in [0, L)
    return System(
        name="chain31200", positions=positions, box=np.array([L, L, L]),
        masses=np.concatenate([np.full(n_chain, 12.011), wm]),
        charges=np.concatenate([c["charges"], wq]), lj=np.concatenate([c["lj"], wlj]),
        types=["C%d" % a for a in range(n_chain)] + wtypes,
        bonds=np.concatenate([c["bonds"], wb]).astype(np.int64),
        bond_params=np.concatenate([c["bond_params"], wbp]),
        angles=np.concatenate([c["angles"], wa]).astype(np.int64),
        angle_params=np.concatenate([c["angle_params"], wap]),
        dihedrals=c["dihedrals"], dihedral_params=c["dihedral_params"],
        impropers=c["impropers"], improper_params=c["improper_params"],
        cutoff=cutoff, pme_grid=tuple(pme_grid),
        meta={"n_chain": n_chain, "n_waters": n_waters, "water_sites_deleted": int((~keep).sum()), "seed": seed})


def _dihedral_angles(pos, quads, L):
    """atan2 dihedral of each quad, sign convention of mdpy/utils/geometry.py:78-90."""
    p = pos[quads]
    r1 = _min_image(p[:, 1] - p[:, 0], L)
    r2 = _min_image(p[:, 2] - p[:, 1], L)
    r3 = _min_image(p[:, 3] - p[:, 2], L)
    n1 = np.cross(r1, r2)
    n2 = np.cross(r2, r3)
    x = np.einsum("ij,ij->i", np.linalg.norm(r2, axis=1)[:, None] * r1, n2)
    y = np.einsum("ij,ij->i", n1, n2)
    return np.arctan2(x, y)


def make(name, seed=0):
    """Look a workload up by name: water648, water23328, chain31200."""
    table = {"water648": water_box_small, "water23328": water_box_large, "chain31200": solvated_chain}
    return table[name](seed=seed)
