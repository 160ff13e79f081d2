/*
 * refpath.c — C restatement of mdpy v0.2.1's CPU force-and-integrate step, used ONLY as the
 * timed CPU baseline and as a second oracle (test infrastructure; see oracle/__init__.py).
 *
 * Same loops and data structures as the reference, single threaded (numba njit without
 * parallel runs on one core):
 *   State.set_positions: wrap (mdpy/utils/pbc.py:28-36) + CellList rebuild (cell_list.py:83-116)
 *   CharmmNonbondedConstraint.cpu_kernel  (charmm_nonbonded_constraint.py:64-108): 27-cell
 *       stencil, excluded/self/-1 skipped, every pair visited twice at half force and energy
 *   ElectrostaticConstraint.cpu_kernel    (electrostatic_constraint.py:52-79): all i<j, no cutoff
 *   bond / angle+UB / dihedral / improper kernels (charmm_*_constraint.py)
 *   LangevinIntegrator.integrate          (langevin_integrator.py:37-72): three ensemble
 *       updates per step (two on the first), restart from x_{n-1}, v_{n-1} as written
 * Normal deviates come from a local xorshift + Box-Muller generator (the reference uses numpy's
 * legacy randn; the stream only affects the trajectory, not the cost).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int n;
  const double *m, *q, *lj; /* lj: n x 4 (eps, sigma, eps14, sigma14) */
  const int *ex_ptr, *ex_idx, *s14_ptr, *s14_idx;
  int nb;
  const int* bonds;
  const double* bpar;
  int na;
  const int* angles;
  const double* apar;
  int nd;
  const int* dih;
  const double* dpar;
  int ni;
  const int* imp;
  const double* ipar;
  double L[3];
  double rc;      /* nonbonded cutoff */
  double cell_rc; /* cell-list cutoff */
  double ke_inv;  /* 1 / (4 pi eps0) divisor k of electrostatic_constraint.py:60 */
} refsys;

/* ---------------------------------------------------------------- PBC (pbc.py:28-44) */
static void unwrap_vec(const double* L, double* v) {
  for (int d = 0; d < 3; ++d) {
    double s = v[d] * (1.0 / L[d]);
    s -= nearbyint(s);
    v[d] = s * L[d];
  }
}
static int wrap_positions(const refsys* S, double* x) {
  for (int i = 0; i < S->n; ++i)
    for (int d = 0; d < 3; ++d) {
      double mv = -nearbyint(x[3 * i + d] * (1.0 / S->L[d]));
      if (fabs(mv) >= 2) return -4;
    }
  for (int i = 0; i < S->n; ++i)
    for (int d = 0; d < 3; ++d) {
      double mv = -nearbyint(x[3 * i + d] * (1.0 / S->L[d]));
      x[3 * i + d] += mv * S->L[d];
    }
  return 0;
}
static double norm3(const double* v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

/* ---------------------------------------------------------------- cell list (cell_list.py:74-116) */
typedef struct {
  int nc[3];
  int maxpc;
  int* idx;   /* n x 3 */
  int* table; /* nc0*nc1*nc2*maxpc, -1 padded */
} cell_list;

static void cell_list_update(const refsys* S, const double* x, cell_list* C) {
  double inv[3], mn[3] = {INFINITY, INFINITY, INFINITY};
  for (int d = 0; d < 3; ++d) {
    int k = (int)floor((S->L[d] + 2.0) / S->cell_rc);
    if (k < 3) k = 3;
    C->nc[d] = k;
    inv[d] = 1.0 / ((S->L[d] + 2.0) / k);
  }
  for (int i = 0; i < S->n; ++i)
    for (int d = 0; d < 3; ++d)
      if (x[3 * i + d] < mn[d]) mn[d] = x[3 * i + d];
  int ncell = C->nc[0] * C->nc[1] * C->nc[2];
  int* cnt = calloc(ncell, sizeof(int));
  free(C->idx);
  C->idx = malloc(sizeof(int) * 3 * S->n);
  for (int i = 0; i < S->n; ++i) {
    for (int d = 0; d < 3; ++d) C->idx[3 * i + d] = (int)floor((x[3 * i + d] - mn[d]) * inv[d]);
    cnt[(C->idx[3 * i] * C->nc[1] + C->idx[3 * i + 1]) * C->nc[2] + C->idx[3 * i + 2]]++;
  }
  int maxpc = 0;
  for (int c = 0; c < ncell; ++c) maxpc = cnt[c] > maxpc ? cnt[c] : maxpc;
  C->maxpc = maxpc;
  free(C->table);
  C->table = malloc(sizeof(int) * (size_t)ncell * maxpc);
  for (size_t k = 0; k < (size_t)ncell * maxpc; ++k) C->table[k] = -1;
  memset(cnt, 0, sizeof(int) * ncell);
  for (int i = 0; i < S->n; ++i) {
    int c = (C->idx[3 * i] * C->nc[1] + C->idx[3 * i + 1]) * C->nc[2] + C->idx[3 * i + 2];
    C->table[(size_t)c * maxpc + cnt[c]++] = i;
  }
  free(cnt);
}

static int is_in(const int* a, int b, int e, int v) {
  for (int k = b; k < e; ++k)
    if (a[k] == v) return 1;
  return 0;
}

/* ---------------------------------------------------------------- LJ (charmm_nonbonded_constraint.py:64-108) */
static double lj_kernel(const refsys* S, const double* x, const cell_list* C, double* F) {
  double E = 0.0;
  for (int i = 0; i < S->n; ++i) {
    int e0 = S->ex_ptr[i], e1 = S->ex_ptr[i + 1];
    int t0 = S->s14_ptr[i], t1 = S->s14_ptr[i + 1];
    for (int a = -1; a <= 1; ++a)
      for (int b = -1; b <= 1; ++b)
        for (int cc = -1; cc <= 1; ++cc) {
          int c3[3] = {C->idx[3 * i] + a, C->idx[3 * i + 1] + b, C->idx[3 * i + 2] + cc};
          for (int d = 0; d < 3; ++d) {
            if (c3[d] >= C->nc[d]) c3[d] -= C->nc[d];
            if (c3[d] < 0) c3[d] += C->nc[d]; /* Python negative indexing */
          }
          const int* row = C->table + ((size_t)(c3[0] * C->nc[1] + c3[1]) * C->nc[2] + c3[2]) * C->maxpc;
          for (int k = 0; k < C->maxpc; ++k) {
            int j = row[k];
            if (j == -1 || j == i || is_in(S->ex_idx, e0, e1, j)) continue;
            double v[3] = {x[3 * j] - x[3 * i], x[3 * j + 1] - x[3 * i + 1], x[3 * j + 2] - x[3 * i + 2]};
            unwrap_vec(S->L, v);
            double r = norm3(v);
            if (r <= S->rc) {
              int sc = is_in(S->s14_idx, t0, t1, j);
              const double* p1 = S->lj + 4 * i + (sc ? 2 : 0);
              const double* p2 = S->lj + 4 * j + (sc ? 2 : 0);
              double eps = sqrt(p1[0] * p2[0]), sig = (p1[1] + p2[1]) / 2;
              double sr = sig / r;
              double sr6 = pow(sr, 6), sr12 = pow(sr, 12);
              double fval = -(2 * sr12 - sr6) / r * eps * 24;
              for (int d = 0; d < 3; ++d) {
                double f = v[d] / r * fval / 2;
                F[3 * i + d] += f;
                F[3 * j + d] -= f;
              }
              E += 4 * eps * (sr12 - sr6) / 2;
            }
          }
        }
  }
  return E;
}

/* ---------------------------------------------------------------- Coulomb (electrostatic_constraint.py:52-79) */
static double coulomb_kernel(const refsys* S, const double* x, double* F) {
  double E = 0.0;
  const double k = S->ke_inv;
  for (int i = 0; i < S->n; ++i) {
    int e0 = S->ex_ptr[i], e1 = S->ex_ptr[i + 1];
    for (int j = i + 1; j < S->n; ++j) {
      if (is_in(S->ex_idx, e0, e1, j)) continue;
      double v[3] = {x[3 * j] - x[3 * i], x[3 * j + 1] - x[3 * i + 1], x[3 * j + 2] - x[3 * i + 2]};
      unwrap_vec(S->L, v);
      double r = norm3(v);
      double fval = -S->q[i] * S->q[j] / k / (r * r);
      for (int d = 0; d < 3; ++d) {
        double f = v[d] / r * fval;
        F[3 * i + d] += f;
        F[3 * j + d] -= f;
      }
      E += S->q[i] * S->q[j] / k / r;
    }
  }
  return E;
}

/* ---------------------------------------------------------------- bonded */
static void sub_mi(const refsys* S, const double* x, int j, int i, double* v) {
  for (int d = 0; d < 3; ++d) v[d] = x[3 * j + d] - x[3 * i + d];
  unwrap_vec(S->L, v);
}
static void cross(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
static void unit(double* v) {
  double nn = norm3(v);
  if (nn != 0)
    for (int d = 0; d < 3; ++d) v[d] /= nn;
}
static double dot(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static double bond_kernel(const refsys* S, const double* x, double* F) {
  double E = 0;
  for (int t = 0; t < S->nb; ++t) {
    int i = S->bonds[2 * t], j = S->bonds[2 * t + 1];
    double k = S->bpar[2 * t], r0 = S->bpar[2 * t + 1], v[3];
    sub_mi(S, x, j, i, v);
    double r = norm3(v), fv = 2 * k * (r - r0);
    for (int d = 0; d < 3; ++d) {
      F[3 * i + d] += fv * v[d] / r;
      F[3 * j + d] -= fv * v[d] / r;
    }
    E += k * (r - r0) * (r - r0);
  }
  return E;
}

static double angle_kernel(const refsys* S, const double* x, double* F) {
  double E = 0;
  for (int t = 0; t < S->na; ++t) {
    int i1 = S->angles[3 * t], i2 = S->angles[3 * t + 1], i3 = S->angles[3 * t + 2];
    const double* p = S->apar + 4 * t;
    double r21[3], r23[3], nv[3], f1[3], f3[3], m23[3], r13[3];
    sub_mi(S, x, i1, i2, r21);
    sub_mi(S, x, i3, i2, r23);
    double l21 = norm3(r21), l23 = norm3(r23);
    double th = acos(dot(r21, r23) / (l21 * l23));
    double fv = -2 * p[0] * (th - p[1]);
    cross(r21, r23, nv);
    cross(r21, nv, f1);
    unit(f1);
    for (int d = 0; d < 3; ++d) m23[d] = -r23[d];
    cross(m23, nv, f3);
    unit(f3);
    for (int d = 0; d < 3; ++d) {
      F[3 * i1 + d] += fv * f1[d] / l21;
      F[3 * i2 + d] -= fv * (f1[d] / l21 + f3[d] / l23);
      F[3 * i3 + d] += fv * f3[d] / l23;
    }
    E += p[0] * (th - p[1]) * (th - p[1]);
    sub_mi(S, x, i3, i1, r13);
    double l13 = norm3(r13), fu = 2 * p[2] * (l13 - p[3]);
    for (int d = 0; d < 3; ++d) {
      F[3 * i1 + d] += fu * r13[d] / l13;
      F[3 * i3 + d] -= fu * r13[d] / l13;
    }
    E += p[2] * (l13 - p[3]) * (l13 - p[3]);
  }
  return E;
}

static double four_body(const refsys* S, const double* x, const int* q, int ref_dihedral, const double* par,
                        int improper, double* F) {
  int a = q[0], b = q[1], c = q[2], dd = q[3];
  double r1[3], r2[3], r3[3], n1[3], n2[3], s1[3];
  sub_mi(S, x, b, a, r1);
  sub_mi(S, x, c, b, r2);
  sub_mi(S, x, dd, c, r3);
  cross(r1, r2, n1);
  cross(r2, r3, n2);
  double l2 = norm3(r2);
  for (int d = 0; d < 3; ++d) s1[d] = l2 * r1[d];
  double phi = atan2(dot(s1, n2), dot(n1, n2));
  double fv, E;
  if (improper) {
    fv = -2 * par[0] * (phi - par[1]);
    E = par[0] * (phi - par[1]) * (phi - par[1]);
  } else {
    fv = ref_dihedral ? -par[0] * (1 - par[1] * sin(par[1] * phi - par[2])) : par[0] * par[1] * sin(par[1] * phi - par[2]);
    E = par[0] * (1 + cos(par[1] * phi - par[2]));
  }
  double vab[3], vbc[3], vcd[3], voc[3], mab[3], mbc[3], ua[3], ud[3], fa[3], fd[3], fc[3], t1[3], t2[3], t3[3], acc[3];
  sub_mi(S, x, b, a, vab);
  sub_mi(S, x, c, b, vbc);
  sub_mi(S, x, dd, c, vcd);
  double lab = norm3(vab), lbc = norm3(vbc), lcd = norm3(vcd), loc = lbc / 2;
  for (int d = 0; d < 3; ++d) {
    voc[d] = vbc[d] / 2;
    mab[d] = -vab[d];
    mbc[d] = -vbc[d];
  }
  double th_abc = acos(dot(mab, vbc) / (lab * lbc)), th_bcd = acos(dot(mbc, vcd) / (lbc * lcd));
  cross(mab, vbc, ua);
  unit(ua);
  cross(vcd, mbc, ud);
  unit(ud);
  for (int d = 0; d < 3; ++d) {
    fa[d] = fv / (lab * sin(th_abc)) * ua[d];
    fd[d] = fv / (lcd * sin(th_bcd)) * ud[d];
  }
  cross(voc, fd, t1);
  cross(vcd, fd, t2);
  cross(mab, fa, t3);
  for (int d = 0; d < 3; ++d) acc[d] = -(t1[d] + t2[d] / 2 + t3[d] / 2);
  cross(acc, voc, fc);
  for (int d = 0; d < 3; ++d) {
    fc[d] /= loc * loc;
    F[3 * a + d] += fa[d];
    F[3 * b + d] += -(fa[d] + fc[d] + fd[d]);
    F[3 * c + d] += fc[d];
    F[3 * dd + d] += fd[d];
  }
  return E;
}

/* ---------------------------------------------------------------- Ensemble.update (ensemble.py:53-61) */
/* F_terms: 5 x n x 3 (LJ, Coulomb, bond, angle, dihedral+improper) or NULL; F: total */
static void ensemble_update(const refsys* S, const double* x, const cell_list* C, double* F, double* E5,
                            double* F_terms, int ref_dihedral) {
  size_t n3 = 3 * (size_t)S->n;
  double* tmp = calloc(n3, sizeof(double));
  memset(F, 0, n3 * sizeof(double));
  for (int term = 0; term < 5; ++term) {
    memset(tmp, 0, n3 * sizeof(double));
    double e = 0;
    if (term == 0) e = lj_kernel(S, x, C, tmp);
    else if (term == 1) e = coulomb_kernel(S, x, tmp);
    else if (term == 2) e = bond_kernel(S, x, tmp);
    else if (term == 3) e = angle_kernel(S, x, tmp);
    else {
      for (int t = 0; t < S->nd; ++t) e += four_body(S, x, S->dih + 4 * t, ref_dihedral, S->dpar + 3 * t, 0, tmp);
      for (int t = 0; t < S->ni; ++t) e += four_body(S, x, S->imp + 4 * t, 0, S->ipar + 2 * t, 1, tmp);
    }
    if (E5) E5[term] = e;
    for (size_t k = 0; k < n3; ++k) F[k] += tmp[k];
    if (F_terms) memcpy(F_terms + term * n3, tmp, n3 * sizeof(double));
  }
  free(tmp);
}

/* ---------------------------------------------------------------- public entry points */
/* One force evaluation at positions x (wrapped in place, cell list rebuilt) */
int refpath_forces(const refsys* S, double* x, double* F_terms, double* E5, int ref_dihedral) {
  int rc = wrap_positions(S, x);
  if (rc) return rc;
  cell_list C = {{0, 0, 0}, 0, NULL, NULL};
  cell_list_update(S, x, &C);
  double* F = malloc(sizeof(double) * 3 * S->n);
  ensemble_update(S, x, &C, F, E5, F_terms, ref_dihedral);
  free(F);
  free(C.idx);
  free(C.table);
  return 0;
}

static uint64_t xs_state;
static double urand(void) {
  xs_state ^= xs_state << 13;
  xs_state ^= xs_state >> 7;
  xs_state ^= xs_state << 17;
  return ((xs_state >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
static double nrand(void) {
  double u1 = urand(), u2 = urand();
  return sqrt(-2 * log(u1)) * cos(2 * M_PI * u2);
}

/* LangevinIntegrator.integrate as written (langevin_integrator.py:37-72), nsteps steps.
 * x, v: state positions/velocities (n x 3), updated. e_out[0] = last potential energy. */
int refpath_langevin(const refsys* S, double* x, double* v, int nsteps, double dt, double gamma, double kbt,
                     uint64_t seed, double* e_out) {
  size_t n3 = 3 * (size_t)S->n;
  double sigma = sqrt(2 * kbt * gamma);
  double a = (1 - gamma * dt / 2) / (1 + gamma * dt / 2), b = 1 / (1 + gamma * dt / 2);
  double *cx = malloc(n3 * 8), *px = malloc(n3 * 8), *cv = malloc(n3 * 8), *pv = malloc(n3 * 8);
  double *ca = calloc(n3, 8), *pa = malloc(n3 * 8), *F = malloc(n3 * 8), *xi = malloc(n3 * 8), *xs = malloc(n3 * 8);
  double E5[5] = {0};
  xs_state = seed * 2685821657736338717ull + 1;
  memcpy(cx, x, n3 * 8);
  memcpy(px, x, n3 * 8);
  memcpy(cv, v, n3 * 8);
  memcpy(pv, v, n3 * 8);
  memcpy(xs, x, n3 * 8);
  int rc = wrap_positions(S, xs);
  cell_list C = {{0, 0, 0}, 0, NULL, NULL};
  cell_list_update(S, xs, &C);
  for (int step = 0; step < nsteps && !rc; ++step) {
    if (step != 0) ensemble_update(S, xs, &C, F, E5, NULL, 0);
    ensemble_update(S, xs, &C, F, E5, NULL, 0);
    for (int i = 0; i < S->n; ++i)
      for (int d = 0; d < 3; ++d) xi[3 * i + d] = nrand() / sqrt(S->m[i]);
    memcpy(pa, ca, n3 * 8);
    for (size_t k = 0; k < n3; ++k) {
      double nx = px[k] + b * dt * pv[k] - b * dt * dt / 2 * pa[k] + b * sigma * pow(dt, 1.5) / 2 * xi[k];
      px[k] = cx[k];
      cx[k] = nx;
    }
    memcpy(xs, cx, n3 * 8);
    rc = wrap_positions(S, xs);
    if (rc) break;
    cell_list_update(S, xs, &C);
    ensemble_update(S, xs, &C, F, E5, NULL, 0);
    for (int i = 0; i < S->n; ++i)
      for (int d = 0; d < 3; ++d) ca[3 * i + d] = F[3 * i + d] / S->m[i];
    for (size_t k = 0; k < n3; ++k) {
      double nv = a * pv[k] - dt / 2 * (a * pa[k] + ca[k]) + b * sigma * sqrt(dt) * xi[k];
      pv[k] = cv[k];
      cv[k] = nv;
    }
  }
  memcpy(x, xs, n3 * 8);
  memcpy(v, cv, n3 * 8);
  if (e_out) e_out[0] = E5[0] + E5[1] + E5[2] + E5[3] + E5[4];
  free(cx); free(px); free(cv); free(pv); free(ca); free(pa); free(F); free(xi); free(xs);
  free(C.idx);
  free(C.table);
  return rc;
}
