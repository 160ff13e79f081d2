// CHARMM bonded terms in float64 (device functions shared by the bonded launch and the
// nonbonded launch that carries them), one lane per term, forces added to the int64 fixed-point
// accumulator (deterministic). Restates:
//   bond      mdpy/constraint/charmm_bond_constraint.py:52-73
//   angle+UB  mdpy/constraint/charmm_angle_constraint.py:54-95
//   dihedral  mdpy/constraint/charmm_dihedral_constraint.py:58-95 (dihedral angle from
//             mdpy/utils/geometry.py:78-90)
//   improper  mdpy/constraint/charmm_improper_constraint.py:56-93
// Minimum image per vector as mdpy/utils/pbc.py:38-44 (orthorhombic box).
#pragma once
#include "mdx_internal.h"

#define BD_BLOCK 256

struct Box3 {
  double L[3], inv[3];
};

__device__ __forceinline__ void mi(const Box3& b, const double* v, double* out) {
  for (int q = 0; q < 3; ++q) {
    double s = v[q] * b.inv[q];
    s -= rint(s);
    out[q] = s * b.L[q];
  }
}
__device__ __forceinline__ void sub_mi(const Box3& b, const double* __restrict__ pos, int j, int i, double* out) {
  double v[3] = {pos[3 * j] - pos[3 * i], pos[3 * j + 1] - pos[3 * i + 1], pos[3 * j + 2] - pos[3 * i + 2]};
  mi(b, v, out);
}
__device__ __forceinline__ double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__device__ __forceinline__ double norm3(const double* a) { return sqrt(dot3(a, a)); }
__device__ __forceinline__ void cross3(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ void unit3(double* v) {  // geometry.py:17-20
  double n = norm3(v);
  if (n != 0) {
    v[0] /= n;
    v[1] /= n;
    v[2] /= n;
  }
}
__device__ __forceinline__ void add_force(long long* __restrict__ ffix, int n, int i, double s, const double* f) {
  atomicAdd((unsigned long long*)&ffix[i], (unsigned long long)to_fix(s * f[0]));
  atomicAdd((unsigned long long*)&ffix[n + i], (unsigned long long)to_fix(s * f[1]));
  atomicAdd((unsigned long long*)&ffix[2 * n + i], (unsigned long long)to_fix(s * f[2]));
}

static __device__ void block_energy(double e, long long* __restrict__ slot) {
  __shared__ double sm[16];  // up to 1024 lanes
  for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off, 64);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sm[w];
    atomicAdd((unsigned long long*)slot, (unsigned long long)to_fix(s));
  }
}

// ---------------------------------------------------------------- per-term device functions
static __device__ double bond_term(int t, int n, const int* __restrict__ idx, const double* __restrict__ par, const Box3& b,
                            const double* __restrict__ pos, long long* __restrict__ ffix) {
  int i = idx[2 * t], j = idx[2 * t + 1];
  double k = par[2 * t], r0 = par[2 * t + 1];
  double d[3];
  sub_mi(b, pos, j, i, d);
  double r = norm3(d);
  double fv = 2 * k * (r - r0);
  double f[3] = {fv * d[0] / r, fv * d[1] / r, fv * d[2] / r};
  add_force(ffix, n, i, 1.0, f);
  add_force(ffix, n, j, -1.0, f);
  return k * (r - r0) * (r - r0);
}

static __device__ double angle_term(int t, int n, const int* __restrict__ idx, const double* __restrict__ par, const Box3& b,
                             const double* __restrict__ pos, long long* __restrict__ ffix) {
  int i1 = idx[3 * t], i2 = idx[3 * t + 1], i3 = idx[3 * t + 2];
  double k = par[4 * t], th0 = par[4 * t + 1], ku = par[4 * t + 2], u0 = par[4 * t + 3];
  double r21[3], r23[3];
  sub_mi(b, pos, i1, i2, r21);
  sub_mi(b, pos, i3, i2, r23);
  double l21 = norm3(r21), l23 = norm3(r23);
  double th = acos(dot3(r21, r23) / (l21 * l23));
  double fv = -2 * k * (th - th0);
  double nv[3], v1[3], v3[3], mr23[3] = {-r23[0], -r23[1], -r23[2]};
  cross3(r21, r23, nv);
  cross3(r21, nv, v1);
  unit3(v1);
  cross3(mr23, nv, v3);
  unit3(v3);
  double f1[3], f3[3], f2[3];
  for (int q = 0; q < 3; ++q) {
    f1[q] = fv * v1[q] / l21;
    f3[q] = fv * v3[q] / l23;
    f2[q] = -(f1[q] + f3[q]);
  }
  double e = k * (th - th0) * (th - th0);
  // Urey-Bradley
  double r13[3];
  sub_mi(b, pos, i3, i1, r13);
  double l13 = norm3(r13);
  double fu = 2 * ku * (l13 - u0);
  for (int q = 0; q < 3; ++q) {
    f1[q] += fu * r13[q] / l13;
    f3[q] -= fu * r13[q] / l13;
  }
  e += ku * (l13 - u0) * (l13 - u0);
  add_force(ffix, n, i1, 1.0, f1);
  add_force(ffix, n, i2, 1.0, f2);
  add_force(ffix, n, i3, 1.0, f3);
  return e;
}

// dihedral angle (geometry.py:78-90) and the four-body force distribution shared by
// charmm_dihedral_constraint.py:70-92 and charmm_improper_constraint.py:78-97
static __device__ double four_body_angle(const Box3& b, const double* __restrict__ pos, int a, int bb, int c, int d) {
  double r1[3], r2[3], r3[3], n1[3], n2[3];
  sub_mi(b, pos, bb, a, r1);
  sub_mi(b, pos, c, bb, r2);
  sub_mi(b, pos, d, c, r3);
  cross3(r1, r2, n1);
  cross3(r2, r3, n2);
  double l2 = norm3(r2);
  double s1[3] = {l2 * r1[0], l2 * r1[1], l2 * r1[2]};
  return atan2(dot3(s1, n2), dot3(n1, n2));
}

static __device__ void four_body_forces(const Box3& b, const double* __restrict__ pos, int a, int bb, int c, int d,
                                 double fv, long long* __restrict__ ffix, int n) {
  double vab[3], vbc[3], vcd[3];
  sub_mi(b, pos, bb, a, vab);
  sub_mi(b, pos, c, bb, vbc);
  sub_mi(b, pos, d, c, vcd);
  double lab = norm3(vab), lbc = norm3(vbc), lcd = norm3(vcd);
  double voc[3] = {vbc[0] / 2, vbc[1] / 2, vbc[2] / 2}, loc = lbc / 2;
  double mab[3] = {-vab[0], -vab[1], -vab[2]}, mbc[3] = {-vbc[0], -vbc[1], -vbc[2]};
  double th_abc = acos(dot3(mab, vbc) / (lab * lbc));
  double th_bcd = acos(dot3(mbc, vcd) / (lbc * lcd));
  double ua[3], ud[3];
  cross3(mab, vbc, ua);
  unit3(ua);
  cross3(vcd, mbc, ud);
  unit3(ud);
  double fa[3], fd[3], fc[3], fb[3];
  double sa = fv / (lab * sin(th_abc)), sd = fv / (lcd * sin(th_bcd));
  for (int q = 0; q < 3; ++q) {
    fa[q] = sa * ua[q];
    fd[q] = sd * ud[q];
  }
  double t1[3], t2[3], t3[3], acc[3];
  cross3(voc, fd, t1);
  cross3(vcd, fd, t2);
  cross3(mab, fa, t3);
  for (int q = 0; q < 3; ++q) acc[q] = -(t1[q] + t2[q] / 2 + t3[q] / 2);
  cross3(acc, voc, fc);
  for (int q = 0; q < 3; ++q) {
    fc[q] /= loc * loc;
    fb[q] = -(fa[q] + fc[q] + fd[q]);
  }
  add_force(ffix, n, a, 1.0, fa);
  add_force(ffix, n, bb, 1.0, fb);
  add_force(ffix, n, c, 1.0, fc);
  add_force(ffix, n, d, 1.0, fd);
}

static __device__ double dihedral_term(int t, int n, const int* __restrict__ idx, const double* __restrict__ par,
                                const Box3& b, const double* __restrict__ pos, long long* __restrict__ ffix,
                                int refcompat) {
  int a = idx[4 * t], bb = idx[4 * t + 1], c = idx[4 * t + 2], d = idx[4 * t + 3];
  double k = par[3 * t], nn = par[3 * t + 1], delta = par[3 * t + 2];
  double phi = four_body_angle(b, pos, a, bb, c, d);
  double fv = refcompat ? -k * (1 - nn * sin(nn * phi - delta))  // charmm_dihedral_constraint.py:81
                        : k * nn * sin(nn * phi - delta);         // -dE/dphi
  four_body_forces(b, pos, a, bb, c, d, fv, ffix, n);
  return k * (1 + cos(nn * phi - delta));
}

static __device__ double improper_term(int t, int n, const int* __restrict__ idx, const double* __restrict__ par,
                                const Box3& b, const double* __restrict__ pos, long long* __restrict__ ffix) {
  int a = idx[4 * t], bb = idx[4 * t + 1], c = idx[4 * t + 2], d = idx[4 * t + 3];
  double k = par[2 * t], psi0 = par[2 * t + 1];
  double psi = four_body_angle(b, pos, a, bb, c, d);
  four_body_forces(b, pos, a, bb, c, d, -2 * k * (psi - psi0), ffix, n);
  return k * (psi - psi0) * (psi - psi0);
}

// PME correction of one excluded pair: -ke q_i q_j erf(beta r)/r (float64)
static __device__ double pme_excl_term(int t, int n, const int* __restrict__ pairs, const double* __restrict__ pos,
                                const double* __restrict__ q, double ke, double beta, const Box3& b,
                                long long* __restrict__ ffix) {
  int i = pairs[2 * t], j = pairs[2 * t + 1];
  double d[3];
  sub_mi(b, pos, j, i, d);
  double r = norm3(d);
  double qq = ke * q[i] * q[j];
  double er = erf(beta * r);
  double dEdr = qq * (er / (r * r) - 2.0 * beta / sqrt(M_PI) * exp(-beta * beta * r * r) / r);
  double g = dEdr / r;  // force on i = +g d, on j = -g d
  double f[3] = {g * d[0], g * d[1], g * d[2]};
  add_force(ffix, n, i, 1.0, f);
  add_force(ffix, n, j, -1.0, f);
  return -qq * er / r;
}

// ---------------------------------------------------------------- one launch for all terms
// Consecutive block ranges handle bonds, angles, dihedrals, impropers and the PME exclusion
// correction (whose first lane also adds the self and background energies); every block
// belongs to one term, so its energy goes to that term's slot.
struct BondedArgs {
  int n;
  Box3 b;
  const double* pos;
  long long* ffix;
  long long* efix;
  int cnt[5], blocks[5];
  const int* idx[5];
  const double* par[5];
  int refcompat;
  const double* q;
  double ke, beta, e_const;
};

__device__ __forceinline__ void bonded_block(const BondedArgs& a, int blk) {
  int seg = 0;
  while (seg < 4 && blk >= a.blocks[seg]) blk -= a.blocks[seg++];
  const int t = blk * (int)blockDim.x + threadIdx.x;  // workgroups of `block` lanes (mdx_bonded_args)
  double e = 0.0;
  if (t < a.cnt[seg]) {
    switch (seg) {
      case 0: e = bond_term(t, a.n, a.idx[0], a.par[0], a.b, a.pos, a.ffix); break;
      case 1: e = angle_term(t, a.n, a.idx[1], a.par[1], a.b, a.pos, a.ffix); break;
      case 2: e = dihedral_term(t, a.n, a.idx[2], a.par[2], a.b, a.pos, a.ffix, a.refcompat); break;
      case 3: e = improper_term(t, a.n, a.idx[3], a.par[3], a.b, a.pos, a.ffix); break;
      default: e = pme_excl_term(t, a.n, a.idx[4], a.pos, a.q, a.ke, a.beta, a.b, a.ffix); break;
    }
  }
  if (seg == 4 && t == 0) e += a.e_const;
  block_energy(e, a.efix + (seg < 4 ? 3 + seg : 2));
}

// arguments of the bonded terms of `terms` (and the PME exclusion part with MDX_TERM_COUL_RECIP);
// returns the number of workgroups of `block` lanes that bonded_block expects
int mdx_bonded_args(const mdx_ctx* c, unsigned terms, BondedArgs* out, int block = BD_BLOCK);
