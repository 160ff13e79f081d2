// Nonbonded launch (the pair evaluation of mdx_nonbonded_body.h, with the bonded terms as
// trailing workgroups), the Ewald-correction table, and the plain all-pairs Coulomb kernel
// (mdpy/constraint/electrostatic_constraint.py:52-79), gfx950.
#include "mdx_bonded_terms.h"
#include "mdx_nonbonded_body.h"

#define NB_BLOCK 256

// G(z) of the forces-only Coulomb path and its derivative, float64, for the table
static void coulomb_correction(double beta, double z, double* G, double* dG) {
  const double k = 2.0 * beta * beta * beta / sqrt(M_PI);
  const double x2 = beta * beta * z;
  if (x2 < 2.0) {
    // G = k sum_{n>=1} (-1)^(n+1) 2n / (n! (2n+1)) x^(2n-2): the series of erf(x)/x^3 minus
    // that of exp(-x^2)/x^2, without their cancellation at small z
    double s = 0, sp = 0, fact = 1, xp = 1, xpm = 0;  // xp = x2^(n-1), xpm = x2^(n-2)
    for (int n = 1; n <= 40; ++n) {
      fact *= n;
      const double cn = ((n & 1) ? 2.0 : -2.0) * n / (fact * (2 * n + 1));
      s += cn * xp;
      if (n >= 2) sp += cn * (n - 1) * beta * beta * xpm;
      xpm = xp;
      xp *= x2;
    }
    *G = k * s;
    *dG = k * sp;
    return;
  }
  const double r = sqrt(z), e = exp(-x2), er = erf(beta * r), c = 2.0 * beta / sqrt(M_PI);
  *G = er / (z * r) - c * e / z;
  *dG = beta * e / (sqrt(M_PI) * z * z) - 1.5 * er / (z * z * r) + c * (beta * beta * e / z + e / (z * z));
}

int mdx_coulomb_table_update(mdx_ctx* c) {
  if (!(c->beta > 0 && c->rc > 0) || (c->ctab.p && c->ctab_beta == c->beta && c->ctab_rc == c->rc)) return 0;
  const int N = MDX_CTAB_N;
  const double zmax = c->rc * c->rc * (1.0 + 1e-3);  // the float64 cutoff re-decision may accept r2 a hair above rc^2
  const double h = zmax / N;
  std::vector<float4> tab(N);
  double g0, d0;
  coulomb_correction(c->beta, 0.0, &g0, &d0);
  for (int k = 0; k < N; ++k) {
    double g1, d1;
    coulomb_correction(c->beta, (k + 1) * h, &g1, &d1);
    const double a0 = g0, a1 = d0 * h, a2 = 3.0 * (g1 - g0) - (2.0 * d0 + d1) * h, a3 = 2.0 * (g0 - g1) + (d0 + d1) * h;
    tab[k] = make_float4((float)a0, (float)a1, (float)a2, (float)a3);
    g0 = g1;
    d0 = d1;
  }
  int rc = c->ctab.ensure(N);
  if (rc) return rc;
  MDX_HIP(hipMemcpy(c->ctab.p, tab.data(), N * sizeof(float4), hipMemcpyHostToDevice));
  c->ctab_beta = c->beta;
  c->ctab_rc = c->rc;
  c->ctab_scale = (float)(N / zmax);
  mdx_invalidate_graph(c);
  return 0;
}

// Workgroups [0, nres) loop over the nb_blocks nonbonded workgroup indices (nres == nb_blocks:
// one each); the trailing ones evaluate the bonded terms. Blocks are numbered in cell order, so
// a contiguous range of them is a compact region of the box.
//
// At most 128 VGPRs (4 waves per SIMD) instead of the 132 the compiler picks by itself: the cap
// costs two 8-byte spills outside the pair loop (reloaded once per atom), leaves the kernel's own
// time unchanged and measured the whole PME-concurrent step ~1.3 % faster on water23328.
template <bool LJ, bool COUL, bool ENERGY, int G, bool S14>
__global__ __launch_bounds__(NB_BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_nonbonded(NbArgs a, NbPtrs P, BondedArgs ba, int nb_blocks, int nres) {
  if ((int)blockIdx.x >= nres) {  // trailing workgroups: bonded terms
    bonded_block(ba, blockIdx.x - nres);
    return;
  }
  if (nres % 8 == 0 && nres < nb_blocks) {
    // workgroups are dealt to the 8 XCDs round-robin: give each XCD a contiguous (spatially
    // compact) range of the blocks, its nres/8 workgroups striding through it
    const int x = blockIdx.x & 7, per = nres >> 3;
    const int b1 = (int)((long long)nb_blocks * (x + 1) / 8);
    for (int vb = (int)((long long)nb_blocks * x / 8) + (blockIdx.x >> 3); vb < b1; vb += per)
      nb_block<LJ, COUL, ENERGY, G, NB_BLOCK, S14>(a, P, vb);
    return;
  }
  for (int vb = blockIdx.x; vb < nb_blocks; vb += nres) nb_block<LJ, COUL, ENERGY, G, NB_BLOCK, S14>(a, P, vb);
}

NbArgs nb_args(const mdx_ctx* c) {
  NbArgs a;
  a.n = c->n;
  a.cap = c->nbr_cap;
  a.Lx = (float)c->box[0];
  a.Ly = (float)c->box[1];
  a.Lz = (float)c->box[2];
  a.ix = 1.0f / a.Lx;
  a.iy = 1.0f / a.Ly;
  a.iz = 1.0f / a.Lz;
  a.rc2 = (float)(c->rc * c->rc);
  // float32 r^2 carries ~2 * r * 4e-6 A^2 of error from the float positions: re-decide within
  // ~20x that band in float64
  a.guard2 = (float)(c->rc * 2e-4);
  a.beta = (float)c->beta;
  a.twob_sqrtpi = (float)(2.0 * c->beta / sqrt(M_PI));
  a.ct_scale = c->ctab_scale;
  a.ctab = c->ctab.p;
  a.Lxd = c->box[0];
  a.Lyd = c->box[1];
  a.Lzd = c->box[2];
  a.rcd = c->rc;
  return a;
}

NbPtrs nb_ptrs(const mdx_ctx* c) {
  NbPtrs P;
  P.xq = c->xq.p;
  P.ljs = c->ljs.p;
  P.ljs14 = c->ljs14.p;
  P.nbr_count = c->nbr_count.p;
  P.nbr_list = c->nbr_list.p;
  P.sorted_orig = c->sorted_orig.p;
  P.pos = c->pos;
  P.ffix = c->ffix.p;
  P.efix = c->efix.p;
  P.st = c->dstate.p;
  return P;
}

int mdx_enqueue_nonbonded(mdx_ctx* c, bool lj, bool coul, bool energy, unsigned bonded_terms) {
  const NbArgs a = nb_args(c);
  const NbPtrs P = nb_ptrs(c);
  MDX_CHECK(!coul || (a.ctab && c->ctab_beta == c->beta && c->ctab_rc == c->rc), MDX_ERR_INVALID,
            "Ewald-correction table not built for the current beta and cutoff");
  const int lanes = c->nb_lanes;
  static_assert(NB_BLOCK == BD_BLOCK, "bonded workgroups share the nonbonded launch");
  BondedArgs ba{};
  const int nbb = (int)div_up((long long)c->n * lanes, NB_BLOCK);
  const int nbo = bonded_terms ? mdx_bonded_args(c, bonded_terms, &ba) : 0;
  // resident workgroups: all of them, or nb_persist per compute unit looping over the rest (so
  // that the concurrent PME kernels find room on every compute unit)
  const int nres = c->nb_persist > 0 ? std::min(nbb, c->nb_persist * c->ncu) : nbb;
  dim3 grid(nres + nbo);
#define NB_LAUNCH4(L_, C_, E_, S_)                                                                     \
  do {                                                                                                 \
    if (lanes == 16)                                                                                   \
      k_nonbonded<L_, C_, E_, 16, S_><<<grid, NB_BLOCK, 0, c->stream>>>(NB_ARGS);                      \
    else if (lanes == 64)                                                                              \
      k_nonbonded<L_, C_, E_, 64, S_><<<grid, NB_BLOCK, 0, c->stream>>>(NB_ARGS);                      \
    else                                                                                               \
      k_nonbonded<L_, C_, E_, 32, S_><<<grid, NB_BLOCK, 0, c->stream>>>(NB_ARGS);                      \
  } while (0)
  // the list holds 1-4 pairs only if the system has scaled pairs
#define NB_LAUNCH(L_, C_, E_)                                                                          \
  do {                                                                                                 \
    if (c->has_s14) NB_LAUNCH4(L_, C_, E_, true); else NB_LAUNCH4(L_, C_, E_, false);                  \
  } while (0)
#define NB_ARGS a, P, ba, nbb, nres
  if (lj && coul) {
    if (energy) NB_LAUNCH(true, true, true); else NB_LAUNCH(true, true, false);
  } else if (lj) {
    if (energy) NB_LAUNCH(true, false, true); else NB_LAUNCH(true, false, false);
  } else {
    if (energy) NB_LAUNCH(false, true, true); else NB_LAUNCH(false, true, false);
  }
#undef NB_LAUNCH
#undef NB_LAUNCH4
#undef NB_ARGS
  MDX_LAUNCH_CHECK();
  return 0;
}

// ------------------------------------------------------------------------------------------
// Plain Coulomb (mdpy ElectrostaticConstraint): every i < j not excluded, minimum image, no
// cutoff, E = q_i q_j / (4 pi eps0 r). float64, one lane per atom over all j (full double loop,
// each pair from both ends), j tiles staged in LDS.
#define PC_BLOCK 256
__global__ __launch_bounds__(PC_BLOCK) void k_coulomb_plain(int n, const double* __restrict__ pos,
                                                          const double* __restrict__ q, double ke, double Lx,
                                                          double Ly, double Lz, const int* __restrict__ excl_ptr,
                                                          const int* __restrict__ excl_idx, long long* __restrict__ ffix,
                                                          long long* __restrict__ efix) {
  __shared__ double4 tile[PC_BLOCK];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double xi = 0, yi = 0, zi = 0, qi = 0;
  int e0 = 0, e1 = 0;
  if (i < n) {
    xi = pos[3 * i];
    yi = pos[3 * i + 1];
    zi = pos[3 * i + 2];
    qi = q[i];
    if (excl_ptr) {
      e0 = excl_ptr[i];
      e1 = excl_ptr[i + 1];
    }
  }
  double fx = 0, fy = 0, fz = 0, en = 0;
  const double ix = 1.0 / Lx, iy = 1.0 / Ly, iz = 1.0 / Lz;
  for (int base = 0; base < n; base += PC_BLOCK) {
    int j = base + threadIdx.x;
    tile[threadIdx.x] = j < n ? make_double4(pos[3 * j], pos[3 * j + 1], pos[3 * j + 2], q[j]) : make_double4(0, 0, 0, 0);
    __syncthreads();
    int lim = min(PC_BLOCK, n - base);
    if (i < n) {
      for (int t = 0; t < lim; ++t) {
        int jj = base + t;
        double4 pj = tile[t];
        double dx = pj.x - xi, dy = pj.y - yi, dz = pj.z - zi;
        double sx = dx * ix, sy = dy * iy, sz = dz * iz;
        dx = (sx - rint(sx)) * Lx;
        dy = (sy - rint(sy)) * Ly;
        dz = (sz - rint(sz)) * Lz;
        double r2 = dx * dx + dy * dy + dz * dz;
        bool skip = (jj == i) || pj.w == 0.0 || qi == 0.0;
        for (int q2 = e0; q2 < e1; ++q2) skip |= (excl_idx[q2] == jj);
        if (skip) continue;
        double rinv = 1.0 / sqrt(r2);
        double qq = qi * pj.w * ke;
        double fr = -qq * rinv * rinv * rinv;  // force on i = -qq/r^2 * d/r
        fx += fr * dx;
        fy += fr * dy;
        fz += fr * dz;
        en += qq * rinv;
      }
    }
    __syncthreads();
  }
  if (i < n) {
    atomicAdd((unsigned long long*)&ffix[i], (unsigned long long)to_fix(fx));
    atomicAdd((unsigned long long*)&ffix[n + i], (unsigned long long)to_fix(fy));
    atomicAdd((unsigned long long*)&ffix[2 * n + i], (unsigned long long)to_fix(fz));
  }
  __shared__ double sm[PC_BLOCK / 64];
  en = wave_sum(en);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = en;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
    for (int w = 0; w < PC_BLOCK / 64; ++w) s += sm[w];
    atomicAdd((unsigned long long*)&efix[7], (unsigned long long)to_fix(0.5 * s));
  }
}

int mdx_enqueue_coulomb_plain(mdx_ctx* c) {
  const double ke = 1.0 / (4.0 * M_PI * 0.5727652348434076);  // 1/(4 pi eps0), internal units
  k_coulomb_plain<<<div_up(c->n, PC_BLOCK), PC_BLOCK, 0, c->stream>>>(
      c->n, c->pos, c->charge64.p, ke, c->box[0], c->box[1], c->box[2], c->has_excl ? c->excl_ptr.p : nullptr,
      c->excl_idx.p, c->ffix.p, c->efix.p);
  MDX_LAUNCH_CHECK();
  return 0;
}

// ------------------------------------------------------------------------------------------
// Batched replicas: every replica's nonbonded workgroups (one per unit of work, no persistent
// loop) followed by its bonded ones, in one launch.
struct NbCall {
  NbArgs a;
  NbPtrs P;
  BondedArgs ba;
  int nb_blocks;  // the replica's nonbonded workgroups; the rest of its workgroups are bonded
};

template <bool LJ, bool COUL, bool ENERGY, int G, bool S14>
__global__ __launch_bounds__(NB_BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_nonbonded_batch(
    BatchMap map, const NbCall* __restrict__ calls) {
  int b;
  const NbCall& k = calls[batch_replica(map, &b)];
  if (b >= k.nb_blocks) {
    bonded_block(k.ba, b - k.nb_blocks);
    return;
  }
  nb_block<LJ, COUL, ENERGY, G, NB_BLOCK, S14>(k.a, k.P, b);
}

int mdx_batch_prepare_nb(mdx_ctx* const* cs, int n, unsigned bonded_terms, BatchKind* out) {
  std::vector<NbCall> calls;
  std::vector<int> blocks;
  for (int r = 0; r < n; ++r) {
    mdx_ctx* c = cs[r];
    NbCall k;
    k.a = nb_args(c);
    k.P = nb_ptrs(c);
    MDX_CHECK(k.a.ctab && c->ctab_beta == c->beta && c->ctab_rc == c->rc, MDX_ERR_INVALID,
              "replica %d: Ewald-correction table not built for the current beta and cutoff", r);
    MDX_CHECK(c->nb_lanes == cs[0]->nb_lanes && c->has_s14 == cs[0]->has_s14, MDX_ERR_INVALID,
              "replica %d: batched replicas need the same lanes per atom and 1-4 pair mode", r);
    k.ba = BondedArgs{};
    k.nb_blocks = (int)div_up((long long)c->n * c->nb_lanes, NB_BLOCK);
    const int nbo = bonded_terms ? mdx_bonded_args(c, bonded_terms, &k.ba) : 0;
    calls.push_back(k);
    blocks.push_back(k.nb_blocks + nbo);
  }
  return batch_kind_upload(out, calls, blocks, 0);
}

// LJ + direct-space Coulomb, the batched step's pair terms
int mdx_batch_launch_nb(const BatchKind& k, int lanes, bool s14, bool energy, hipStream_t s) {
  const NbCall* calls = (const NbCall*)k.calls;
#define NBB(E_, G_, S_) k_nonbonded_batch<true, true, E_, G_, S_><<<k.blocks, NB_BLOCK, 0, s>>>(k.map, calls)
#define NBB_G(E_, S_)                 \
  do {                                \
    if (lanes == 16) NBB(E_, 16, S_); \
    else if (lanes == 64) NBB(E_, 64, S_); \
    else NBB(E_, 32, S_);             \
  } while (0)
  if (energy) {
    if (s14) NBB_G(true, true); else NBB_G(true, false);
  } else {
    if (s14) NBB_G(false, true); else NBB_G(false, false);
  }
#undef NBB_G
#undef NBB
  MDX_LAUNCH_CHECK();
  return 0;
}
