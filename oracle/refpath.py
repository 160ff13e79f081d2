"""Build recipe and ctypes wrapper for refpath.c (the C restatement of the reference CPU step).

Test infrastructure / CPU baseline only (see oracle/__init__.py).

    python -m oracle.refpath      # builds oracle/librefpath.so with gcc
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "refpath.c")
LIB = os.path.join(HERE, "librefpath.so")
# x86-64-v3 (AVX2/FMA) runs on every EPYC host; no FMA contraction so the arithmetic follows
# the reference's operation order
CFLAGS = ["-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fPIC", "-shared", "-std=gnu11"]


def build():
    tmp = LIB + ".tmp%d" % os.getpid()
    subprocess.run(["gcc"] + CFLAGS + [SRC, "-o", tmp, "-lm"], check=True)
    os.replace(tmp, LIB)
    return LIB


class RefSys(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int),
        ("m", ctypes.c_void_p), ("q", ctypes.c_void_p), ("lj", ctypes.c_void_p),
        ("ex_ptr", ctypes.c_void_p), ("ex_idx", ctypes.c_void_p),
        ("s14_ptr", ctypes.c_void_p), ("s14_idx", ctypes.c_void_p),
        ("nb", ctypes.c_int), ("bonds", ctypes.c_void_p), ("bpar", ctypes.c_void_p),
        ("na", ctypes.c_int), ("angles", ctypes.c_void_p), ("apar", ctypes.c_void_p),
        ("nd", ctypes.c_int), ("dih", ctypes.c_void_p), ("dpar", ctypes.c_void_p),
        ("ni", ctypes.c_int), ("imp", ctypes.c_void_p), ("ipar", ctypes.c_void_p),
        ("L", ctypes.c_double * 3),
        ("rc", ctypes.c_double), ("cell_rc", ctypes.c_double), ("ke_inv", ctypes.c_double),
    ]


class RefPath:
    """A System bound to the C reference path."""

    def __init__(self, system, coulomb_k, cell_cutoff=12.0):
        if not os.path.exists(LIB):
            raise ImportError("oracle/librefpath.so not built (python -m oracle.refpath)")
        self.lib = ctypes.CDLL(LIB)
        self.lib.refpath_forces.restype = ctypes.c_int
        self.lib.refpath_langevin.restype = ctypes.c_int
        s = system
        self.n = s.num_atoms
        keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(np.asarray(a, dtype=dt).reshape(-1))
            if a.size == 0:
                a = np.zeros(1, dtype=dt)
            keep.append(a)
            return a.ctypes.data

        ep, ei = s.exclusions()
        sp, si = s.scaling_pairs()
        r = RefSys()
        r.n = self.n
        r.m, r.q, r.lj = arr(s.masses, np.float64), arr(s.charges, np.float64), arr(s.lj, np.float64)
        r.ex_ptr, r.ex_idx = arr(ep, np.int32), arr(ei, np.int32)
        r.s14_ptr, r.s14_idx = arr(sp, np.int32), arr(si, np.int32)
        r.nb, r.bonds, r.bpar = len(s.bonds), arr(s.bonds, np.int32), arr(s.bond_params, np.float64)
        r.na, r.angles, r.apar = len(s.angles), arr(s.angles, np.int32), arr(s.angle_params, np.float64)
        r.nd, r.dih, r.dpar = len(s.dihedrals), arr(s.dihedrals, np.int32), arr(s.dihedral_params, np.float64)
        r.ni, r.imp, r.ipar = len(s.impropers), arr(s.impropers, np.int32), arr(s.improper_params, np.float64)
        r.L = (ctypes.c_double * 3)(*[float(v) for v in s.box])
        r.rc = float(s.cutoff)
        r.cell_rc = float(cell_cutoff)
        r.ke_inv = float(coulomb_k)
        self._keep = keep
        self.sys = r

    def forces(self, positions, ref_dihedral=False):
        """(wrapped x, per-term forces (5,N,3): LJ, Coulomb, bond, angle, dihedral+improper, energies (5,))"""
        x = np.ascontiguousarray(positions, dtype=np.float64).copy()
        F = np.zeros((5, self.n, 3))
        E = np.zeros(5)
        rc = self.lib.refpath_forces(ctypes.byref(self.sys), x.ctypes.data_as(ctypes.c_void_p),
                                     F.ctypes.data_as(ctypes.c_void_p), E.ctypes.data_as(ctypes.c_void_p),
                                     ctypes.c_int(1 if ref_dihedral else 0))
        if rc:
            raise RuntimeError("refpath_forces failed: %d" % rc)
        return x, F, E

    def langevin(self, positions, velocities, nsteps, dt, gamma, kbt, seed=0):
        x = np.ascontiguousarray(positions, dtype=np.float64).copy()
        v = np.ascontiguousarray(velocities, dtype=np.float64).copy()
        e = np.zeros(1)
        rc = self.lib.refpath_langevin(ctypes.byref(self.sys), x.ctypes.data_as(ctypes.c_void_p),
                                       v.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(nsteps),
                                       ctypes.c_double(dt), ctypes.c_double(gamma), ctypes.c_double(kbt),
                                       ctypes.c_uint64(seed), e.ctypes.data_as(ctypes.c_void_p))
        if rc:
            raise RuntimeError("refpath_langevin failed: %d" % rc)
        return x, v, float(e[0])


if __name__ == "__main__":
    print(build())
