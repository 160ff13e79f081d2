// libmdx C-ABI: context, system description, force evaluation, integrators, hipGraph replay.
//
// Boundary replaced (mdpy v0.2.1):
//   Constraint.update / Ensemble.update      mdpy/constraint/constraint.py:45-46, mdpy/ensemble.py:53-71
//   Integrator.integrate                     mdpy/integrator/integrator.py:36-39,
//                                            verlet_integrator.py:20-50, langevin_integrator.py:37-72
//   State.set_positions / CellList.update    mdpy/core/state.py:56-61, mdpy/core/cell_list.py:83-116
// The context runs on its own non-blocking stream (default priority) and joins the caller's
// stream with events at every entry and exit, so callers may pass PyTorch's current stream and
// the step can still be captured into a hipGraph.
#include "mdx_internal.h"
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <set>

static thread_local std::string g_err;

void mdx_set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

extern "C" const char* mdx_last_error(void) { return g_err.c_str(); }
extern "C" int mdx_version(void) { return 100; }
// 1/(4 pi eps0) in internal units (eps0 = 8.85418e-12 C^2 s^2/(kg m^3), mdpy/unit/__init__.py:56-58)
extern "C" double mdx_coulomb_constant(void) { return 1.0 / (4.0 * M_PI * 0.5727652348434076); }

// ------------------------------------------------------------------ stream plumbing
struct ExtSync {
  hipStream_t ext = nullptr;
  hipEvent_t in = nullptr, out = nullptr;
};
static std::vector<std::pair<mdx_ctx*, ExtSync>> g_sync;  // per context (small)

static ExtSync* sync_of(mdx_ctx* c) {
  for (auto& p : g_sync)
    if (p.first == c) return &p.second;
  return nullptr;
}

int mdx_enter(mdx_ctx* c) {
  ExtSync* s = sync_of(c);
  if (!s) return 0;
  MDX_HIP(hipSetDevice(c->device));
  MDX_HIP(hipEventRecord(s->in, s->ext));
  MDX_HIP(hipStreamWaitEvent(c->stream, s->in, 0));
  return 0;
}
int mdx_leave(mdx_ctx* c) {
  ExtSync* s = sync_of(c);
  if (!s) return 0;
  MDX_HIP(hipEventRecord(s->out, c->stream));
  MDX_HIP(hipStreamWaitEvent(s->ext, s->out, 0));
  return 0;
}
#define ENTER()                   \
  do {                            \
    int _r = mdx_enter(c);            \
    if (_r) return _r;            \
  } while (0)
#define LEAVE_RET(v)              \
  do {                            \
    int _v = (v);                 \
    int _r = mdx_leave(c);            \
    return _v ? _v : _r;          \
  } while (0)

// ------------------------------------------------------------------ small kernels
// Zero the force/energy accumulators. With record != 0 this also closes the previous step:
// when its index (relative to the call's first step) is a multiple of `every`, its energies go
// to history row index/every; then the step counter advances.
__global__ void k_zero_acc(long long* __restrict__ ffix, long long n3, long long* __restrict__ efix, int ne,
                           double* __restrict__ ehist, int ehist_cap, DevState* st, int record, int every) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n3) ffix[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (record) {
      long long k = st->step - st->hist_base;
      if (k >= 0 && k % every == 0 && k / every < ehist_cap)
        for (int t = 0; t < ne; ++t) ehist[(k / every) * ne + t] = (double)efix[t] * kFixInv;
      st->step += 1;
    }
    for (int t = 0; t < ne; ++t) efix[t] = 0;
  }
}

// start of an integrate call: its energy-history rows count from the current step, and the
// kinetic-energy slot is recomputed from the current velocities
__global__ void k_record(const long long* __restrict__ efix, int ne, double* __restrict__ ehist, int ehist_cap,
                         DevState* st, int every) {
  long long k = st->step - st->hist_base;
  if (k >= 0 && k % every == 0 && k / every < ehist_cap)
    for (int t = 0; t < ne; ++t) ehist[(k / every) * ne + t] = (double)efix[t] * kFixInv;
}

// total force on atom i along d: the fixed-point accumulator plus, when the PME plane kernels
// ran, the four per-plane reciprocal contributions (each written once, summed in fixed order)
__device__ __forceinline__ double total_force(const long long* __restrict__ ffix, const double* __restrict__ fpart,
                                             int n, int i, int d) {
  double f = (double)ffix[(long long)d * n + i] * kFixInv;
  if (fpart) {
    const long long n3 = 3LL * n;
    const double* p = fpart + 3LL * i + d;
    f += (p[0] + p[n3]) + (p[2 * n3] + p[3 * n3]);
  }
  return f;
}

// total_force for the three components, every load issued before the sums (a branch per
// component would wait for memory three times); the same sums in the same order
__device__ __forceinline__ void total_forces_loaded(const long long* __restrict__ ffix,
                                                    const double* __restrict__ fpart, int n, int i, double* f) {
  long long q[3];
  double p[3][4];
  for (int d = 0; d < 3; ++d) q[d] = ffix[(long long)d * n + i];
  if (fpart) {
    const long long n3 = 3LL * n;
    for (int d = 0; d < 3; ++d)
      for (int k = 0; k < 4; ++k) p[d][k] = fpart[k * n3 + 3LL * i + d];
  }
  for (int d = 0; d < 3; ++d) {
    f[d] = (double)q[d] * kFixInv;
    if (fpart) f[d] += (p[d][0] + p[d][1]) + (p[d][2] + p[d][3]);
  }
}

__global__ void k_forces_out(const long long* __restrict__ ffix, const double* __restrict__ fpart, int n,
                             double* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int d = 0; d < 3; ++d) out[3 * i + d] = total_force(ffix, fpart, n, i, d);
}

__device__ void block_add_energy(double e, long long* slot) {
  __shared__ double sm[16];
  for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off, 64);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sm[w];
    atomicAdd((unsigned long long*)slot, (unsigned long long)to_fix(s));
  }
}

__global__ void k_kinetic(int n, const double* __restrict__ v, const double* __restrict__ m, long long* __restrict__ slot) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0;
  if (i < n) e = 0.5 * m[i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
  block_add_energy(e, slot);
}

// ------------------------------------------------------------------ Philox4x32-10
// Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11.
__device__ __forceinline__ uint4 philox4x32_10(uint4 ctr, uint2 key) {
  const unsigned M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r) {
    unsigned hi0 = __umulhi(M0, ctr.x), lo0 = M0 * ctr.x;
    unsigned hi1 = __umulhi(M1, ctr.z), lo1 = M1 * ctr.z;
    ctr = make_uint4(hi1 ^ ctr.y ^ key.x, lo1, hi0 ^ ctr.w ^ key.y, lo0);
    key.x += W0;
    key.y += W1;
  }
  return ctr;
}
// three standard normals for (atom, step) by Box-Muller
__device__ void normal3(unsigned long long seed, long long step, int atom, double* g) {
  uint2 key = make_uint2((unsigned)seed, (unsigned)(seed >> 32));
  uint4 a = philox4x32_10(make_uint4((unsigned)atom, (unsigned)step, (unsigned)((unsigned long long)step >> 32), 0u), key);
  uint4 b = philox4x32_10(make_uint4((unsigned)atom, (unsigned)step, (unsigned)((unsigned long long)step >> 32), 1u), key);
  const double s = 2.3283064365386963e-10;  // 2^-32
  double u1 = (a.x + 0.5) * s, u2 = (a.y + 0.5) * s, u3 = (a.z + 0.5) * s, u4 = (a.w + 0.5) * s;
  double r1 = sqrt(-2.0 * log(u1)), r2 = sqrt(-2.0 * log(u3));
  double s1, c1, s2, c2;
  sincospi(2.0 * u2, &s1, &c1);
  sincospi(2.0 * u4, &s2, &c2);
  g[0] = r1 * c1;
  g[1] = r1 * s1;
  g[2] = r2 * c2;
  (void)b;
}

struct NoiseSrc {
  unsigned long long seed;
  const double* ext;
  long long ext_base;
  int ext_steps;
};
__device__ __forceinline__ void noise3(const NoiseSrc& ns, long long step, int atom, int n, double* g) {
  long long k = step - ns.ext_base;
  if (ns.ext && k >= 0 && k < ns.ext_steps) {
    const double* p = ns.ext + (k * n + atom) * 3;
    g[0] = p[0];
    g[1] = p[1];
    g[2] = p[2];
  } else {
    normal3(ns.seed, step, atom, g);
  }
}

// ------------------------------------------------------------------ integrator kernels
struct IntArgs {
  int n;
  double dt, a, b, sigma;
  double L[3], inv[3];
  NoiseSrc ns;
  const double* fpart;  // per-plane reciprocal forces of the last evaluation, or null
  double* gsave;        // G-JF: the position half's xi_{n+1} (3 per atom), reused by the velocity half
};

__device__ __forceinline__ double fix_force(const long long* ffix, const IntArgs& A, int i, int d) {
  return total_force(ffix, A.fpart, A.n, i, d);
}

// velocity Verlet, first half: v += dt/2 a; x += dt v
__global__ void k_vv_a(IntArgs A, double* __restrict__ x, double* __restrict__ v, const double* __restrict__ im,
                       const long long* __restrict__ ffix) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  for (int d = 0; d < 3; ++d) {
    double vv = v[3 * i + d] + 0.5 * A.dt * fix_force(ffix, A, i, d) * im[i];
    v[3 * i + d] = vv;
    x[3 * i + d] += A.dt * vv;
  }
}
// velocity Verlet, second half: v += dt/2 a; kinetic energy
__device__ __forceinline__ double vv_half(const IntArgs& A, double* __restrict__ v, const double* __restrict__ im,
                                          const double* __restrict__ m, const long long* __restrict__ ffix, int i) {
  double e = 0;
  for (int d = 0; d < 3; ++d) {
    double vv = v[3 * i + d] + 0.5 * A.dt * fix_force(ffix, A, i, d) * im[i];
    v[3 * i + d] = vv;
    e += vv * vv;
  }
  return e * (0.5 * m[i]);
}
__global__ void k_vv_b(IntArgs A, double* __restrict__ v, const double* __restrict__ im, const double* __restrict__ m,
                       const long long* __restrict__ ffix, long long* __restrict__ ke) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0;
  if (i < A.n) e = vv_half(A, v, im, m, ffix, i);
  block_add_energy(e, ke);
}

// G-JF Langevin (Gronbech-Jensen & Farago, Mol. Phys. 111, 983 (2013)) with mdpy's a, b, sigma
// (langevin_integrator.py:22-28). Position half: uses a_n and xi_{n+1}; stores a_n.
__global__ void k_gjf_a(IntArgs A, double* __restrict__ x, const double* __restrict__ v, const double* __restrict__ im,
                        const long long* __restrict__ ffix, double* __restrict__ acc_prev, const DevState* st) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  double g[3];
  noise3(A.ns, st->step + 1, i, A.n, g);
  for (int d = 0; d < 3; ++d) A.gsave[3 * i + d] = g[d];
  double sm = sqrt(im[i]);
  for (int d = 0; d < 3; ++d) {
    double acc = fix_force(ffix, A, i, d) * im[i];
    acc_prev[3 * i + d] = acc;
    x[3 * i + d] += A.b * A.dt * v[3 * i + d] + A.b * A.dt * A.dt * 0.5 * acc +
                    A.b * A.sigma * A.dt * sqrt(A.dt) * 0.5 * g[d] * sm;
  }
}
// Fused step prologue (G-JF position half or velocity-Verlet first half) for contexts with a
// neighbour-list grid: besides the integrator update each lane zeroes its atom's force
// accumulator, measures the displacement since the list build and writes the atom's packed
// sorted-order copy (valid unless the list is rebuilt, which repacks). The last workgroup to
// finish records the previous step's energies, advances the step counter, clears the energy
// accumulators and, when some atom moved more than skin/2, zeroes the cell counts for the
// rebuild kernels that follow.
struct StepPro {
  long long* ffix;
  const double* pos_build;
  const int* orig_sorted;
  const float* qs;
  float4* xq;
  double* pos_sorted;
  long long* efix;
  int ne;
  double* ehist;
  int ehist_cap, every, record;
  int* cell_count;
  int ncell;
  double thr2;
  int rebuild_every;  // > 0: also rebuild on steps whose counter is a multiple of it
  // binning by the last workgroup (bin != 0): every lane stores its atom's list-grid cell; on a
  // rebuild step the last workgroup counts, scans and scatters the atoms into cell order
  int bin;
  int* atom_cell;
  int* cell_start;
  int* sorted_orig;
  int gn[3];
  double gw[3];
  int n;
};

#define PRO_THREADS 512
#define PRO_SORT_MAX 1024         // largest cell population the per-cell sort accepts (as k_cell_scan)
#define PRO_BIN_LDS_MAX (60 * 1024)  // dynamic LDS for last-workgroup binning (2 ints per cell)

// An atom's loads for the list bookkeeping, issued with the integrator's loads before any store
struct ProAtomIn {
  double pbuild[3];
  int s;
  float q;
};
__device__ __forceinline__ ProAtomIn prologue_load(const StepPro& P, int i) {
  ProAtomIn a;
  for (int d = 0; d < 3; ++d) a.pbuild[d] = P.pos_build[3 * i + d];
  a.s = P.orig_sorted[i];
  a.q = P.qs ? P.qs[i] : 0.f;
  return a;
}

__device__ void prologue_atom(const IntArgs& A, const StepPro& P, int i, const ProAtomIn& in, const double* xn,
                              double* d2, bool* bad) {
  for (int d = 0; d < 3; ++d) P.ffix[(long long)d * A.n + i] = 0;
  for (int d = 0; d < 3; ++d) {
    if (!isfinite(xn[d])) *bad = true;
    double dd = xn[d] - in.pbuild[d];
    dd -= A.L[d] * rint(dd * A.inv[d]);
    *d2 += dd * dd;
  }
  pack_atom(xn[0], xn[1], xn[2], in.q, A.L[0], A.L[1], A.L[2], &P.xq[in.s], &P.pos_sorted[3 * in.s]);
  if (P.bin) {  // the list-grid cell, as k_grid_bin computes it
    int c[3];
    for (int d = 0; d < 3; ++d) {
      double x = xn[d];
      x -= A.L[d] * floor(x / A.L[d]);
      c[d] = clamped_floor_index(x / P.gw[d], P.gn[d]);
    }
    P.atom_cell[i] = (c[0] * P.gn[1] + c[1]) * P.gn[2] + c[2];
  }
}

// rebuild step, last workgroup: count the atoms per cell in LDS, exclusive scan into cell_start,
// scatter the atom ids into their cells (the per-cell sort that follows makes the order
// deterministic). LDS: 2 ncell ints.
__device__ void last_block_bin(const StepPro& P, DevState* st) {
  extern __shared__ int pro_lds[];
  int* cnt = pro_lds;
  int* start = pro_lds + P.ncell;
  __shared__ int wsum[PRO_THREADS / 64];
  constexpr int U = 8;  // independent loads in flight per lane
  const int tid = threadIdx.x, bs = blockDim.x;
  for (int k = tid; k < P.ncell; k += bs) cnt[k] = 0;
  __syncthreads();
  for (int j0 = tid; j0 < P.n; j0 += U * bs) {
    int c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) c[u] = j0 + u * bs < P.n ? P.atom_cell[j0 + u * bs] : -1;
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (c[u] >= 0) atomicAdd(&cnt[c[u]], 1);
  }
  __syncthreads();
  // exclusive scan: per-lane chunk sums, wave-level inclusive scan, then the wave totals
  const int per = (P.ncell + bs - 1) / bs;
  const int b = tid * per, e = min(P.ncell, b + per);
  int sum = 0;
  for (int k = b; k < e; ++k) sum += cnt[k];
  int incl = sum;
  const int lane = tid & 63;
  for (int off = 1; off < 64; off <<= 1) {
    const int v = __shfl_up(incl, off, 64);
    if (lane >= off) incl += v;
  }
  if (lane == 63) wsum[tid >> 6] = incl;
  __syncthreads();
  int wbase = 0;
  for (int w = 0; w < (tid >> 6); ++w) wbase += wsum[w];
  int run = wbase + incl - sum;
  bool over = false;
  for (int k = b; k < e; ++k) {
    const int v = cnt[k];
    over |= v > PRO_SORT_MAX;
    start[k] = run;
    P.cell_start[k] = run;
    P.cell_count[k] = v;
    run += v;
  }
  if (over) atomicOr(&st->flags, MDX_FLAG_OVERFLOW);
  __syncthreads();
  for (int k = tid; k < P.ncell; k += bs) cnt[k] = 0;  // scatter cursors
  __syncthreads();
  for (int j0 = tid; j0 < P.n; j0 += U * bs) {
    int c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) c[u] = j0 + u * bs < P.n ? P.atom_cell[j0 + u * bs] : -1;
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (c[u] >= 0) P.sorted_orig[start[c[u]] + atomicAdd(&cnt[c[u]], 1)] = j0 + u * bs;
  }
}

__device__ void prologue_block_end(const StepPro& P, DevState* st, double d2, bool bad, int nblocks) {
  __shared__ double sm[16];
  __shared__ int last;
  if (bad) atomicOr(&st->flags, MDX_FLAG_NONFINITE);
  for (int off = 32; off > 0; off >>= 1) d2 = fmax(d2, __shfl_xor(d2, off, 64));
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = d2;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) v = fmax(v, sm[w]);
    atomicMax(&st->max_disp2_bits, enc_double_order(v));
    __threadfence();
    last = atomicAdd(&st->step_ticket, 1u) == (unsigned)nblocks - 1;
  }
  __syncthreads();
  if (!last) return;
  if (threadIdx.x == 0) __threadfence();  // acquire for the block (device-scope fences act on whole caches)
  __syncthreads();
  const bool rebuild = atomicAdd(&st->rebuild, 0) != 0 ||
                       !(dec_double_order(atomicAdd(&st->max_disp2_bits, 0ull)) <= P.thr2) ||
                       (P.rebuild_every > 0 && st->step % P.rebuild_every == 0);
  if (rebuild) {
    if (P.bin) last_block_bin(P, st);
    else
      for (int k = threadIdx.x; k < P.ncell; k += blockDim.x) P.cell_count[k] = 0;
    // the sort/fill and list-build kernels that follow test the flag and the displacement
    // themselves; the flag carries this decision to them whatever made it (a scheduled rebuild
    // included). The nonbonded kernel clears it.
    if (threadIdx.x == 0) atomicExch(&st->rebuild, 1);
  }
  if (threadIdx.x == 0) {
    if (P.record) {
      const long long k = st->step - st->hist_base;
      if (k >= 0 && k % P.every == 0 && k / P.every < P.ehist_cap)
        for (int t = 0; t < P.ne; ++t) P.ehist[(k / P.every) * P.ne + t] = (double)P.efix[t] * kFixInv;
      atomicAdd((unsigned long long*)&st->step, 1ull);
    }
    for (int t = 0; t < P.ne; ++t) atomicExch((unsigned long long*)&P.efix[t], 0ull);
    if (rebuild) atomicAdd((unsigned long long*)&st->nrebuild, 1ull);
    atomicExch(&st->step_ticket, 0u);
  }
}

// The arguments of one context's step prologue launch (passed by value to k_step_pro, or one
// replica's record of a batched launch).
struct ProCall {
  IntArgs A;
  StepPro P;
  double* x;
  double* v;
  const double* im;
  const double* m;
  double* acc_prev;
  long long* ke_slot;
  DevState* st;
  int blocks;  // workgroups of the context's launch: the last-workgroup ticket counts to it
};

// The step prologue: with FUSE_B it first completes the previous step (the velocity half of
// G-JF or velocity Verlet with the forces of the last evaluation, plus the kinetic energy; the
// same arithmetic as k_gjf_b / k_vv_b), then runs this step's position half (as k_gjf_a /
// k_vv_a) and the list bookkeeping of prologue_atom / prologue_block_end. blk: the context's
// workgroup.
template <bool GJF, bool FUSE_B>
__device__ __forceinline__ void step_pro_body(const ProCall& c, int blk) {
  const IntArgs& A = c.A;
  const StepPro& P = c.P;
  double* __restrict__ x = c.x;
  double* __restrict__ v = c.v;
  const double* __restrict__ im = c.im;
  const double* __restrict__ m = c.m;
  double* __restrict__ acc_prev = c.acc_prev;
  long long* __restrict__ ke_slot = c.ke_slot;
  DevState* st = c.st;
  const int i = blk * blockDim.x + threadIdx.x;
  double d2 = 0.0, ke = 0.0;
  bool bad = false;
  if (i < A.n) {
    // every load of the atom first: the stores below cannot then hold them back (the arrays
    // are distinct, but the compiler cannot know it), so the workgroup waits for memory once
    double vv[3], xn[3], f[3], xo[3], ap[3], gs[3];
    for (int d = 0; d < 3; ++d) {
      vv[d] = v[3 * i + d];
      xo[d] = x[3 * i + d];
      ap[d] = GJF ? acc_prev[3 * i + d] : 0.0;
      gs[d] = GJF && FUSE_B ? A.gsave[3 * i + d] : 0.0;
    }
    const double imi = im[i];
    const double mi = FUSE_B ? m[i] : 0.0;
    const ProAtomIn pin = prologue_load(P, i);
    total_forces_loaded(P.ffix, A.fpart, A.n, i, f);  // last: its sums wait for all of these
    if (FUSE_B) {
      if (GJF) {
        const double sm = sqrt(imi);  // gs: xi_{n+1}, drawn by the previous step's position half
        for (int d = 0; d < 3; ++d) {
          const double acc = f[d] * imi;
          vv[d] = A.a * vv[d] + 0.5 * A.dt * (A.a * ap[d] + acc) + A.b * A.sigma * sqrt(A.dt) * gs[d] * sm;
          ke += vv[d] * vv[d];
        }
      } else {
        for (int d = 0; d < 3; ++d) {
          vv[d] = vv[d] + 0.5 * A.dt * f[d] * imi;
          ke += vv[d] * vv[d];
        }
      }
      ke *= 0.5 * mi;
    }
    if (GJF) {
      double g[3];
      noise3(A.ns, st->step + 1, i, A.n, g);
      for (int d = 0; d < 3; ++d) A.gsave[3 * i + d] = g[d];
      const double sm = sqrt(imi);
      for (int d = 0; d < 3; ++d) {
        const double acc = f[d] * imi;
        acc_prev[3 * i + d] = acc;
        xn[d] = xo[d] + A.b * A.dt * vv[d] + A.b * A.dt * A.dt * 0.5 * acc +
                A.b * A.sigma * A.dt * sqrt(A.dt) * 0.5 * g[d] * sm;
        x[3 * i + d] = xn[d];
        if (FUSE_B) v[3 * i + d] = vv[d];
      }
    } else {
      for (int d = 0; d < 3; ++d) {
        const double vh = vv[d] + 0.5 * A.dt * f[d] * imi;
        v[3 * i + d] = vh;
        xn[d] = xo[d] + A.dt * vh;
        x[3 * i + d] = xn[d];
      }
    }
    prologue_atom(A, P, i, pin, xn, &d2, &bad);
  }
  if (FUSE_B) block_add_energy(ke, ke_slot);  // before this block's ticket: the last block records it
  prologue_block_end(P, st, d2, bad, c.blocks);
}

template <bool GJF, bool FUSE_B>
__global__ __launch_bounds__(PRO_THREADS) void k_step_pro(ProCall c) {
  step_pro_body<GJF, FUSE_B>(c, blockIdx.x);
}
// every replica's prologue in one launch
template <bool GJF, bool FUSE_B>
__global__ __launch_bounds__(PRO_THREADS) void k_step_pro_batch(BatchMap map, const ProCall* __restrict__ calls) {
  int blk;
  const int r = batch_replica(map, &blk);
  step_pro_body<GJF, FUSE_B>(calls[r], blk);
}

// velocity half: v_{n+1} = a v_n + dt/2 (a a_n + a_{n+1}) + b sigma sqrt(dt) xi_{n+1}/sqrt(m)
__device__ __forceinline__ double gjf_half(const IntArgs& A, double* __restrict__ v, const double* __restrict__ im,
                                           const double* __restrict__ m, const long long* __restrict__ ffix,
                                           const double* __restrict__ acc_prev, int i) {
  double e = 0;
  double g[3];  // xi_{n+1}: drawn by the position half
  for (int d = 0; d < 3; ++d) g[d] = A.gsave[3 * i + d];
  double sm = sqrt(im[i]);
  for (int d = 0; d < 3; ++d) {
    double acc = fix_force(ffix, A, i, d) * im[i];
    double vv = A.a * v[3 * i + d] + 0.5 * A.dt * (A.a * acc_prev[3 * i + d] + acc) + A.b * A.sigma * sqrt(A.dt) * g[d] * sm;
    v[3 * i + d] = vv;
    e += vv * vv;
  }
  return e * (0.5 * m[i]);
}
__global__ void k_gjf_b(IntArgs A, double* __restrict__ v, const double* __restrict__ im, const double* __restrict__ m,
                        const long long* __restrict__ ffix, const double* __restrict__ acc_prev, const DevState* st,
                        long long* __restrict__ ke) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0;
  if (i < A.n) e = gjf_half(A, v, im, m, ffix, acc_prev, i);
  block_add_energy(e, ke);
}

// reference wrap (pbc.py:28-36) of the integrator's unwrapped positions into the state array
__device__ __forceinline__ void wrap_into(const IntArgs& A, const double* xin, double* xout, DevState* st) {
  bool lost = false;
  double o[3];
  for (int d = 0; d < 3; ++d) {
    double m = -rint(xin[d] * A.inv[d]);
    if (fabs(m) >= 2.0) lost = true;
    o[d] = xin[d] + m * A.L[d];
  }
  if (lost) atomicOr(&st->flags, MDX_FLAG_LOSS);
  for (int d = 0; d < 3; ++d) xout[d] = o[d];
}

// reference position Verlet step (verlet_integrator.py:37-43): new = 2 cur - pre + a dt^2
__global__ void k_vref(IntArgs A, double* __restrict__ cur, double* __restrict__ pre, double* __restrict__ pos,
                       const double* __restrict__ im, const long long* __restrict__ ffix, DevState* st) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  double nx[3];
  for (int d = 0; d < 3; ++d) {
    double acc = fix_force(ffix, A, i, d) * im[i];
    nx[d] = 2 * cur[3 * i + d] - pre[3 * i + d] + acc * A.dt * A.dt;
    pre[3 * i + d] = cur[3 * i + d];
    cur[3 * i + d] = nx[d];
  }
  wrap_into(A, nx, pos + 3 * i, st);
}

// reference Langevin, position half (langevin_integrator.py:54-61)
__global__ void k_lref_a(IntArgs A, double* __restrict__ cx, double* __restrict__ px, const double* __restrict__ pv,
                         double* __restrict__ pos, const double* __restrict__ im, const double* __restrict__ acc_cur,
                         double* __restrict__ acc_pre, DevState* st) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  double g[3];
  noise3(A.ns, st->step + 1, i, A.n, g);
  double sm = sqrt(im[i]);
  double nx[3];
  for (int d = 0; d < 3; ++d) {
    double pa = acc_cur[3 * i + d];
    acc_pre[3 * i + d] = pa;
    double xi = g[d] * sm;
    nx[d] = px[3 * i + d] + A.b * A.dt * pv[3 * i + d] - A.b * A.dt * A.dt / 2 * pa +
            A.b * A.sigma * pow(A.dt, 1.5) / 2 * xi;
    px[3 * i + d] = cx[3 * i + d];
    cx[3 * i + d] = nx[d];
  }
  wrap_into(A, nx, pos + 3 * i, st);
}

// reference Langevin, velocity half (langevin_integrator.py:64-70)
__global__ void k_lref_b(IntArgs A, double* __restrict__ cv, double* __restrict__ pv, const double* __restrict__ im,
                         const double* __restrict__ m, const long long* __restrict__ ffix, double* __restrict__ acc_cur,
                         const double* __restrict__ acc_pre, const DevState* st, long long* __restrict__ ke) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0;
  if (i < A.n) {
    double g[3];
    noise3(A.ns, st->step, i, A.n, g);
    double sm = sqrt(im[i]);
    for (int d = 0; d < 3; ++d) {
      double ca = fix_force(ffix, A, i, d) * im[i];
      acc_cur[3 * i + d] = ca;
      double nv = A.a * pv[3 * i + d] - A.dt / 2 * (A.a * acc_pre[3 * i + d] + ca) + A.b * A.sigma * sqrt(A.dt) * g[d] * sm;
      pv[3 * i + d] = cv[3 * i + d];
      cv[3 * i + d] = nv;
      e += nv * nv;
    }
    e *= 0.5 * m[i];
  }
  block_add_energy(e, ke);
}

// prime kernels
__global__ void k_vref_prime(IntArgs A, const double* __restrict__ x, const double* __restrict__ v,
                             const double* __restrict__ im, const long long* __restrict__ ffix, double* __restrict__ cur,
                             double* __restrict__ pre, int cached) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n || cached) return;
  for (int d = 0; d < 3; ++d) {
    double acc = fix_force(ffix, A, i, d) * im[i];
    cur[3 * i + d] = x[3 * i + d];
    pre[3 * i + d] = x[3 * i + d] - v[3 * i + d] * A.dt + acc * A.dt * A.dt;  // verlet_integrator.py:31-34
  }
}
__global__ void k_lref_prime(IntArgs A, const double* __restrict__ x, const double* __restrict__ v,
                             const double* __restrict__ im, const long long* __restrict__ ffix, double* __restrict__ cx,
                             double* __restrict__ px, double* __restrict__ cv, double* __restrict__ pv,
                             double* __restrict__ acc_cur, int cached) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  for (int d = 0; d < 3; ++d) {
    acc_cur[3 * i + d] = fix_force(ffix, A, i, d) * im[i];  // langevin_integrator.py:42
    if (!cached) {
      cx[3 * i + d] = px[3 * i + d] = x[3 * i + d];
      cv[3 * i + d] = pv[3 * i + d] = v[3 * i + d];
    }
  }
}
// end of a reference Verlet call: v = mi(cur - pre)/(2 dt) (verlet_integrator.py:47-50)
__global__ void k_vref_finish(IntArgs A, const double* __restrict__ cur, const double* __restrict__ pre,
                              double* __restrict__ v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  for (int d = 0; d < 3; ++d) {
    double s = (cur[3 * i + d] - pre[3 * i + d]) * A.inv[d];
    s -= rint(s);
    v[3 * i + d] = s * A.L[d] / 2 / A.dt;
  }
}
__global__ void k_copy3(const double* __restrict__ src, double* __restrict__ dst, long long n3) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n3) dst[i] = src[i];
}
// wrap the state positions at the end of a velocity-Verlet / G-JF call (State.set_positions)
__global__ void k_wrap_state(IntArgs A, double* __restrict__ pos, DevState* st) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  double x[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
  wrap_into(A, x, pos + 3 * i, st);
}

// true in one lane (thread 0) of the grid's last workgroup to finish; every workgroup's earlier
// global writes are visible to it. The ticket is left at 0 for the next launch.
__device__ bool last_workgroup(unsigned* ticket) {
  __shared__ int last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    last = atomicAdd(ticket, 1u) == gridDim.x - 1;
    if (last) {
      __threadfence();
      atomicExch(ticket, 0u);
    }
  }
  __syncthreads();
  return last && threadIdx.x == 0;
}

// An integrate call's opening kernel: the kinetic energy of the incoming velocities (k_kinetic's
// sum), written (not added) into its slot by the last workgroup, which also sets the call's
// energy-history base (rows are device step counter - base: no host round trip).
__global__ void k_call_start(int n, const double* __restrict__ v, const double* __restrict__ m,
                             long long* __restrict__ ke_slot, DevState* st) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0;
  if (i < n) e = 0.5 * m[i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
  block_add_energy(e, (long long*)&st->ke_acc);
  if (last_workgroup(&st->call_ticket)) {
    *ke_slot = (long long)atomicExch(&st->ke_acc, 0ull);
    st->hist_base = st->step;
  }
}

// An integrate call's closing kernel for G-JF / velocity Verlet (the closing velocity half as
// k_gjf_b / k_vv_b, k_wrap_state and k_record in one launch): the last velocity half and its kinetic energy, the
// positions wrapped as the reference leaves them, then the last workgroup records the energies.
template <bool GJF>
__global__ void k_close(IntArgs A, double* __restrict__ v, double* __restrict__ pos, const double* __restrict__ im,
                        const double* __restrict__ m, const long long* __restrict__ ffix,
                        const double* __restrict__ acc_prev, DevState* st, long long* __restrict__ efix, int ne,
                        double* __restrict__ ehist, int ehist_cap, int every) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0;
  if (i < A.n) {
    e = GJF ? gjf_half(A, v, im, m, ffix, acc_prev, i) : vv_half(A, v, im, m, ffix, i);
    double x[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    wrap_into(A, x, pos + 3 * i, st);
  }
  block_add_energy(e, efix + MDX_E_KINETIC);
  if (last_workgroup(&st->call_ticket)) {
    const long long k = st->step - st->hist_base;
    if (k >= 0 && k % every == 0 && k / every < ehist_cap)
      for (int t = 0; t < ne; ++t) {
        // read through the atomic path (the other workgroups' sums are L2 atomics); the slots
        // hold signed fixed-point values
        const long long v = (long long)atomicAdd((unsigned long long*)&efix[t], 0ull);
        ehist[(k / every) * ne + t] = (double)v * kFixInv;
      }
  }
}

// ------------------------------------------------------------------ context API
extern "C" int mdx_ctx_create(int device, void* hip_stream, mdx_ctx** out) {
  MDX_CHECK(out, MDX_ERR_INVALID, "mdx_ctx_create: out is NULL");
  int ndev = 0;
  MDX_HIP(hipGetDeviceCount(&ndev));
  MDX_CHECK(device >= 0 && device < ndev, MDX_ERR_INVALID, "device %d not present (%d devices)", device, ndev);
  MDX_HIP(hipSetDevice(device));
  mdx_ctx* c = new mdx_ctx();
  c->device = device;
  if (hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) c->ncu = 0;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->side_stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess) {
    delete c;
    mdx_set_error("stream/event creation failed");
    return MDX_ERR_HIP;
  }
  ExtSync s;
  s.ext = (hipStream_t)hip_stream;
  if (hipEventCreateWithFlags(&s.in, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&s.out, hipEventDisableTiming) != hipSuccess) {
    if (s.in) hipEventDestroy(s.in);
    mdx_ctx_destroy(c);
    mdx_set_error("event creation failed");
    return MDX_ERR_HIP;
  }
  g_sync.push_back({c, s});
  int rc;
  if ((rc = c->dstate.ensure(1)) || (rc = c->efix.ensure(MDX_NTERMS + 1))) {
    mdx_ctx_destroy(c);
    return rc;
  }
  MDX_HIP(hipMemset(c->dstate.p, 0, sizeof(DevState)));
  MDX_HIP(hipMemset(c->efix.p, 0, sizeof(long long) * (MDX_NTERMS + 1)));
  int one = 1;
  MDX_HIP(hipMemcpy(&c->dstate.p->rebuild, &one, sizeof(int), hipMemcpyHostToDevice));
  *out = c;
  return 0;
}

extern "C" int mdx_ctx_set_stream(mdx_ctx* c, void* hip_stream) {
  MDX_CHECK(c, MDX_ERR_INVALID, "null context");
  ExtSync* s = sync_of(c);
  if (s) s->ext = (hipStream_t)hip_stream;
  return 0;
}

extern "C" int mdx_ctx_destroy(mdx_ctx* c) {
  if (!c) return 0;
  hipSetDevice(c->device);
  mdx_invalidate_graph(c);  // synchronises the context's streams first
  for (size_t k = 0; k < g_sync.size(); ++k)
    if (g_sync[k].first == c) {
      hipEventDestroy(g_sync[k].second.in);
      hipEventDestroy(g_sync[k].second.out);
      g_sync.erase(g_sync.begin() + k);
      break;
    }
  c->inv_mass.release(); c->mass.release(); c->charge64.release(); c->charge_s.release(); c->lj64.release();
  c->excl_ptr.release(); c->excl_idx.release(); c->s14_ptr.release(); c->s14_idx.release();
  c->bond_idx.release(); c->angle_idx.release(); c->dih_idx.release(); c->imp_idx.release();
  c->bond_par.release(); c->angle_par.release(); c->dih_par.release(); c->imp_par.release();
  c->ffix.release(); c->efix.release(); c->dstate.release();
  c->ref_cell_lin.release(); c->ref_counts.release(); c->ref_cursor.release();
  c->cell_count.release(); c->cell_start.release(); c->cell_cursor.release(); c->atom_cell.release();
  c->sorted_orig.release(); c->orig_sorted.release(); c->xq.release(); c->ljs.release(); c->ljs14.release();
  c->ref_cell_sorted.release(); c->pos_at_build.release(); c->nbr_count.release(); c->nbr_list.release();
  c->d_stencil.release(); c->qfix.release(); c->cgrid.release(); c->conv.release(); c->bc.release();
  c->twiddle.release(); c->excl_pairs.release(); c->xa.release(); c->xb.release(); c->va.release();
  c->vb.release(); c->acc_prev.release(); c->acc_cur.release(); c->ehist.release(); c->tmp64.release();
  c->pos_sorted.release(); c->fpart.release(); c->pme_grec.release(); c->pme_grec_n.release();
  c->ex_win.release(); c->s4_win.release(); c->win_bad.release(); c->ctab.release(); c->gsave.release();
  if (c->owns_stream) hipStreamDestroy(c->stream);
  if (c->side_stream) hipStreamDestroy(c->side_stream);
  if (c->ev_fork) hipEventDestroy(c->ev_fork);
  if (c->ev_join) hipEventDestroy(c->ev_join);
  delete c;
  return 0;
}

void mdx_invalidate_graph(mdx_ctx* c) {
  c->graph_gen++;
  c->forces_current = false;
  bool any = c->graph_e != nullptr || c->graph_head != nullptr || c->graph_tail != nullptr;
  for (hipGraphExec_t g : c->graph_f) any |= g != nullptr;
  if (any) {
    // a replay of these graphs may still be in flight on the context's streams (launches return
    // before the work ends): let it finish before the executable graphs are destroyed
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->side_stream) (void)hipStreamSynchronize(c->side_stream);
    for (hipGraphExec_t& g : c->graph_f) {
      if (g) (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
    for (hipGraphExec_t* g : {&c->graph_e, &c->graph_head, &c->graph_tail}) {
      if (*g) (void)hipGraphExecDestroy(*g);
      *g = nullptr;
    }
  }
  c->graph_kind = -1;
}

extern "C" int mdx_positions_changed(mdx_ctx* c) {
  MDX_CHECK(c, MDX_ERR_INVALID, "null context");
  c->forces_current = false;
  return 0;
}

extern "C" int mdx_system_set(mdx_ctx* c, int n, double* d_pos, double* d_vel, const double* h_box) {
  MDX_CHECK(c && n > 0 && d_pos && h_box, MDX_ERR_INVALID, "mdx_system_set: bad arguments");
  MDX_CHECK(n < (1 << 28) - 1, MDX_ERR_INVALID, "too many atoms (at most 2^28 - 2)");
  MDX_HIP(hipSetDevice(c->device));
  mdx_invalidate_graph(c);
  bool resized = c->n != n;
  c->n = n;
  c->pos = d_pos;
  c->vel = d_vel;
  for (int d = 0; d < 3; ++d) {
    MDX_CHECK(h_box[d] > 0, MDX_ERR_INVALID, "box edges must be positive");
    c->box[d] = h_box[d];
  }
  int rc;
  if ((rc = c->ffix.ensure(3 * (size_t)n))) return rc;
  if (resized) {
    c->nb_configured = false;
    c->pme_configured = false;
    c->int_primed = false;
  }
  MDX_HIP(hipMemset(c->ffix.p, 0, 3 * sizeof(long long) * n));
  int one = 1;
  MDX_HIP(hipMemcpy(&c->dstate.p->rebuild, &one, sizeof(int), hipMemcpyHostToDevice));
  return 0;
}

extern "C" int mdx_set_box(mdx_ctx* c, const double* h_box) {
  MDX_CHECK(c && h_box, MDX_ERR_INVALID, "mdx_set_box: bad arguments");
  for (int d = 0; d < 3; ++d) {
    MDX_CHECK(h_box[d] > 0, MDX_ERR_INVALID, "box edges must be positive");
    c->box[d] = h_box[d];
  }
  mdx_invalidate_graph(c);
  int rc;
  if (c->nb_configured && (rc = mdx_nlist_prepare(c))) return rc;
  if (c->pme_configured) {
    int K[3] = {c->K[0], c->K[1], c->K[2]};
    if ((rc = mdx_pme_config(c, K, c->order, c->beta))) return rc;
  }
  return 0;
}

template <typename T>
static int upload(DevBuf<T>& b, const T* h, size_t count) {
  int rc = b.ensure(count);
  if (rc) return rc;
  if (count) MDX_HIP(hipMemcpy(b.p, h, count * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

extern "C" int mdx_set_masses(mdx_ctx* c, const double* h_mass) {
  MDX_CHECK(c && c->n > 0 && h_mass, MDX_ERR_INVALID, "mdx_set_masses: bad arguments");
  mdx_invalidate_graph(c);  // parameter buffers may move
  std::vector<double> im(c->n);
  for (int i = 0; i < c->n; ++i) {
    MDX_CHECK(h_mass[i] > 0, MDX_ERR_INVALID, "mass of atom %d is not positive", i);
    im[i] = 1.0 / h_mass[i];
  }
  int rc;
  if ((rc = upload(c->mass, h_mass, c->n)) || (rc = upload(c->inv_mass, im.data(), c->n))) return rc;
  c->has_mass = true;
  return 0;
}

extern "C" int mdx_set_charges(mdx_ctx* c, const double* h_q) {
  MDX_CHECK(c && c->n > 0 && h_q, MDX_ERR_INVALID, "mdx_set_charges: bad arguments");
  mdx_invalidate_graph(c);  // parameter buffers may move
  const double sk = sqrt(mdx_coulomb_constant());
  std::vector<float> qs(c->n);
  c->qsum = 0;
  c->q2sum = 0;
  for (int i = 0; i < c->n; ++i) {
    qs[i] = (float)(h_q[i] * sk);
    c->qsum += h_q[i];
    c->q2sum += h_q[i] * h_q[i];
  }
  int rc;
  if ((rc = upload(c->charge64, h_q, c->n)) || (rc = upload(c->charge_s, qs.data(), c->n))) return rc;
  c->has_charges = true;
  return 0;
}

extern "C" int mdx_set_lj(mdx_ctx* c, const double* h_lj) {
  MDX_CHECK(c && c->n > 0 && h_lj, MDX_ERR_INVALID, "mdx_set_lj: bad arguments");
  mdx_invalidate_graph(c);  // parameter buffers may move
  for (int i = 0; i < 4 * c->n; ++i)
    MDX_CHECK(!((i % 2 == 0) && h_lj[i] < 0), MDX_ERR_INVALID, "LJ epsilon must be >= 0 (atom %d)", i / 4);
  int rc = upload(c->lj64, h_lj, 4 * (size_t)c->n);
  if (rc) return rc;
  c->has_lj = true;
  int one = 1;
  MDX_HIP(hipMemcpy(&c->dstate.p->rebuild, &one, sizeof(int), hipMemcpyHostToDevice));
  return 0;
}

static int set_csr(mdx_ctx* c, const int* ptr, const int* idx, DevBuf<int>& dp, DevBuf<int>& di, int* maxrow) {
  MDX_CHECK(ptr && ptr[0] == 0, MDX_ERR_INVALID, "CSR ptr must start at 0");
  int nnz = ptr[c->n];
  *maxrow = 0;
  for (int i = 0; i < c->n; ++i) {
    MDX_CHECK(ptr[i + 1] >= ptr[i], MDX_ERR_INVALID, "CSR ptr not monotone at %d", i);
    *maxrow = std::max(*maxrow, ptr[i + 1] - ptr[i]);
  }
  for (int k = 0; k < nnz; ++k) MDX_CHECK(idx[k] >= 0 && idx[k] < c->n, MDX_ERR_INVALID, "CSR index out of range");
  int rc;
  if ((rc = upload(dp, ptr, c->n + 1)) || (rc = upload(di, idx, std::max(nnz, 1)))) return rc;
  return 0;
}

// index-window masks of one partner relation (bit = 0 exclusions, 1 scaled 1-4 pairs)
static int set_window_masks(mdx_ctx* c, const int* h_ptr, const int* h_idx, int bit) {
  const int n = c->n;
  std::vector<unsigned long long>& m = bit == 0 ? c->h_ex_win : c->h_s4_win;
  m.assign(n, 0ull);
  if (c->h_win_bad.size() != (size_t)n) c->h_win_bad.assign(n, 0);
  if (bit == 0 && c->h_s4_win.size() != (size_t)n) c->h_s4_win.assign(n, 0ull);
  if (bit == 1 && c->h_ex_win.size() != (size_t)n) c->h_ex_win.assign(n, 0ull);
  // window: partners i-32 .. i+30 (bits 0..62); bit 63 stays clear, so a clamped out-of-window
  // offset reads a zero bit. Bit 32 of the exclusion mask excludes the atom itself.
  for (int i = 0; i < n; ++i) {
    c->h_win_bad[i] &= (unsigned char)~(1u << bit);
    for (int k = h_ptr[i]; k < h_ptr[i + 1]; ++k) {
      const int d = h_idx[k] - i + 32;
      if (d >= 0 && d < 63) m[i] |= 1ull << d;
      else c->h_win_bad[i] |= (unsigned char)(1u << bit);
    }
  }
  for (int i = 0; i < n; ++i) c->h_ex_win[i] |= 1ull << 32;
  int rc;
  if ((rc = upload(c->ex_win, c->h_ex_win.data(), (size_t)n)) || (rc = upload(c->s4_win, c->h_s4_win.data(), (size_t)n)) ||
      (rc = upload(c->win_bad, c->h_win_bad.data(), (size_t)n)))
    return rc;
  c->win_ready = true;
  return 0;
}

extern "C" int mdx_set_exclusions(mdx_ctx* c, const int* h_ptr, const int* h_idx) {
  MDX_CHECK(c && c->n > 0, MDX_ERR_INVALID, "mdx_set_exclusions: no system");
  mdx_invalidate_graph(c);  // parameter buffers may move
  int rc = set_csr(c, h_ptr, h_idx, c->excl_ptr, c->excl_idx, &c->max_excl);
  if (rc) return rc;
  c->has_excl = h_ptr[c->n] > 0;
  if ((rc = set_window_masks(c, h_ptr, h_idx, 0))) return rc;
  // unique excluded pairs (i < j) for the PME erf correction
  std::vector<int> pairs;
  for (int i = 0; i < c->n; ++i)
    for (int k = h_ptr[i]; k < h_ptr[i + 1]; ++k)
      if (h_idx[k] > i) {
        pairs.push_back(i);
        pairs.push_back(h_idx[k]);
      }
  c->n_excl_pairs = (int)pairs.size() / 2;
  if (pairs.empty()) pairs.assign(2, 0);
  if ((rc = upload(c->excl_pairs, pairs.data(), pairs.size()))) return rc;
  int one = 1;
  MDX_HIP(hipMemcpy(&c->dstate.p->rebuild, &one, sizeof(int), hipMemcpyHostToDevice));
  return 0;
}

extern "C" int mdx_set_scaling(mdx_ctx* c, const int* h_ptr, const int* h_idx) {
  MDX_CHECK(c && c->n > 0, MDX_ERR_INVALID, "mdx_set_scaling: no system");
  mdx_invalidate_graph(c);  // parameter buffers may move
  int rc = set_csr(c, h_ptr, h_idx, c->s14_ptr, c->s14_idx, &c->max_s14);
  if (rc) return rc;
  c->has_s14 = h_ptr[c->n] > 0;
  if ((rc = set_window_masks(c, h_ptr, h_idx, 1))) return rc;
  int one = 1;
  MDX_HIP(hipMemcpy(&c->dstate.p->rebuild, &one, sizeof(int), hipMemcpyHostToDevice));
  return 0;
}

static int set_terms(mdx_ctx* c, int count, int width_i, int width_p, const int* idx, const double* par,
                     DevBuf<int>& di, DevBuf<double>& dp, int* out_count) {
  MDX_CHECK(c && c->n > 0 && count >= 0, MDX_ERR_INVALID, "bad bonded-term arguments");
  mdx_invalidate_graph(c);  // parameter buffers may move
  for (long long k = 0; k < (long long)count * width_i; ++k)
    MDX_CHECK(idx[k] >= 0 && idx[k] < c->n, MDX_ERR_INVALID, "bonded-term atom index out of range");
  int rc;
  if (count > 0) {
    if ((rc = upload(di, idx, (size_t)count * width_i)) || (rc = upload(dp, par, (size_t)count * width_p))) return rc;
  }
  *out_count = count;
  return 0;
}
extern "C" int mdx_set_bonds(mdx_ctx* c, int nb, const int* i, const double* p) { return set_terms(c, nb, 2, 2, i, p, c->bond_idx, c->bond_par, &c->nb); }
extern "C" int mdx_set_angles(mdx_ctx* c, int na, const int* i, const double* p) { return set_terms(c, na, 3, 4, i, p, c->angle_idx, c->angle_par, &c->na); }
extern "C" int mdx_set_dihedrals(mdx_ctx* c, int nd, const int* i, const double* p, int refcompat) {
  int rc = set_terms(c, nd, 4, 3, i, p, c->dih_idx, c->dih_par, &c->nd);
  if (!rc) c->dih_refcompat = refcompat;
  return rc;
}
extern "C" int mdx_set_impropers(mdx_ctx* c, int ni, const int* i, const double* p) { return set_terms(c, ni, 4, 2, i, p, c->imp_idx, c->imp_par, &c->ni); }

extern "C" int mdx_nonbonded_config(mdx_ctx* c, double cutoff, double skin, int pair_mode, double cell_cutoff) {
  MDX_CHECK(c && c->n > 0, MDX_ERR_INVALID, "mdx_nonbonded_config: no system");
  MDX_CHECK(cutoff > 0 && skin >= 0, MDX_ERR_INVALID, "cutoff must be > 0 and skin >= 0");
  MDX_CHECK(pair_mode == MDX_PAIRS_EXACT || pair_mode == MDX_PAIRS_REFERENCE, MDX_ERR_INVALID, "bad pair mode");
  for (int d = 0; d < 3; ++d)
    MDX_CHECK(cutoff <= 0.5 * c->box[d], MDX_ERR_CELL_CUTOFF,
              "cutoff %.3f exceeds half the box edge %.3f (minimum image)", cutoff, c->box[d]);
  c->rc = cutoff;
  // the reference pair set depends on the minimum position, which moves every step
  c->skin = pair_mode == MDX_PAIRS_REFERENCE ? 0.0 : skin;
  c->pair_mode = pair_mode;
  c->cell_cutoff = cell_cutoff > 0 ? cell_cutoff : 12.0;
  mdx_invalidate_graph(c);
  int rc = mdx_nlist_prepare(c);
  if (rc || (rc = mdx_coulomb_table_update(c))) return rc;
  c->nb_configured = true;
  return 0;
}

// ------------------------------------------------------------------ force evaluation
// reciprocal forces held outside the fixed-point accumulator by the last evaluation
static const double* live_fpart(const mdx_ctx* c) { return c->fpart_live ? c->fpart.p : nullptr; }
// whether an evaluation of `terms` leaves per-plane reciprocal forces in fpart (the PME plane
// kernels, with or without the phase launches; mdx_pme_enqueue's condition)
static bool evaluation_leaves_fpart(const mdx_ctx* c, unsigned terms) {
  return (terms & MDX_TERM_COUL_RECIP) && c->pme_configured && c->pme_planes && c->nb_configured;
}

int mdx_enqueue_zero_accumulators(mdx_ctx* c) {
  long long n3 = 3LL * c->n;
  k_zero_acc<<<div_up(n3, 256), 256, 0, c->stream>>>(c->ffix.p, n3, c->efix.p, MDX_NTERMS + 1, nullptr, 0,
                                                     c->dstate.p, 0, 1);
  MDX_LAUNCH_CHECK();
  return 0;
}

int mdx_enqueue_forces(mdx_ctx* c, unsigned terms, bool energy) {
  int rc;
  bool lj = (terms & MDX_TERM_LJ) != 0, cd = (terms & MDX_TERM_COUL_DIRECT) != 0;
  if (lj || cd) {
    MDX_CHECK(c->nb_configured, MDX_ERR_INVALID, "nonbonded terms requested but mdx_nonbonded_config not called");
    MDX_CHECK(!lj || c->has_lj, MDX_ERR_INVALID, "LJ requested but no LJ parameters set");
    MDX_CHECK(!cd || (c->has_charges && c->pme_configured), MDX_ERR_INVALID,
              "PME direct space requested but charges/PME not configured");
    if ((rc = mdx_nlist_update_order(c, false))) return rc;
  }
  const unsigned bonded = terms & (MDX_TERM_BOND | MDX_TERM_ANGLE | MDX_TERM_DIHEDRAL | MDX_TERM_IMPROPER |
                                   MDX_TERM_COUL_RECIP);
  if (lj && cd && (terms & MDX_TERM_COUL_RECIP) && mdx_force_phases_ok(c)) {
    // one stream: list build, then the PME phases each carrying a share of the nonbonded work
    if ((rc = mdx_nlist_update_list(c)) || (rc = mdx_force_phases_enqueue(c, energy, bonded))) return rc;
    if (terms & MDX_TERM_COUL_PLAIN) {
      MDX_CHECK(c->has_charges, MDX_ERR_INVALID, "Coulomb requested but no charges set");
      if ((rc = mdx_enqueue_coulomb_plain(c))) return rc;
    }
    if (terms & MDX_TERM_MLP) {
      if ((rc = mdx_mlp_enqueue(c, energy))) return rc;
    }
    return 0;
  }
  // the PME plane kernels read only the cell-sorted order and packed positions, final after the
  // first part of the list update: run them on the side stream, concurrently with the list build
  // (all of a rebuild step's ~160 us), the nonbonded and the bonded kernels; join before returning
  const bool fork_pme = (terms & MDX_TERM_COUL_RECIP) && (lj || cd) && c->pme_planes && c->side_stream &&
                        c->has_charges && c->pme_configured;
  if (fork_pme) {
    MDX_HIP(hipEventRecord(c->ev_fork, c->stream));
    MDX_HIP(hipStreamWaitEvent(c->side_stream, c->ev_fork, 0));
    hipStream_t main_stream = c->stream;
    c->stream = c->side_stream;
    rc = mdx_pme_enqueue(c, energy);
    c->stream = main_stream;
    if (rc) return rc;
    MDX_HIP(hipEventRecord(c->ev_join, c->side_stream));
  }
  // the bonded terms ride in the nonbonded launch as extra workgroups (off the critical path:
  // they need only positions and add into the fixed-point accumulator)
  if (lj || cd) {
    if ((rc = mdx_nlist_update_list(c)) || (rc = mdx_enqueue_nonbonded(c, lj, cd, energy, bonded))) return rc;
  }
  if (terms & MDX_TERM_COUL_RECIP) {
    if (!fork_pme) {
      MDX_CHECK(c->has_charges && c->pme_configured, MDX_ERR_INVALID, "PME requested but not configured");
      // the PME plane kernels walk the cell-sorted atom order of the list grid
      const bool grid_here = c->nb_configured && !(lj || cd);
      if (grid_here && (rc = mdx_nlist_update(c, false))) return rc;
      if ((rc = mdx_pme_enqueue(c, energy))) return rc;
      if (grid_here && (rc = mdx_nlist_enqueue_clear(c))) return rc;
    }
  } else {
    c->fpart_live = false;
  }
  if (terms & MDX_TERM_COUL_PLAIN) {
    MDX_CHECK(c->has_charges, MDX_ERR_INVALID, "Coulomb requested but no charges set");
    if ((rc = mdx_enqueue_coulomb_plain(c))) return rc;
  }
  if (terms & MDX_TERM_MLP) {
    if ((rc = mdx_mlp_enqueue(c, energy))) return rc;
  }
  if (!(lj || cd) && (rc = mdx_bonded_enqueue(c, terms))) return rc;
  if (fork_pme) MDX_HIP(hipStreamWaitEvent(c->stream, c->ev_join, 0));
  return 0;
}

static int compute_forces(mdx_ctx* c, unsigned terms, double* d_force_out, bool energy) {
  MDX_CHECK(c && c->n > 0, MDX_ERR_INVALID, "mdx_compute_forces: no system");
  ENTER();
  c->forces_current = false;
  int rc = mdx_enqueue_zero_accumulators(c);
  if (!rc) rc = mdx_enqueue_forces(c, terms, energy);
  if (!rc && energy && c->vel && c->has_mass) {
    k_kinetic<<<div_up(c->n, 256), 256, 0, c->stream>>>(c->n, c->vel, c->mass.p, c->efix.p + MDX_E_KINETIC);
    if (hipGetLastError() != hipSuccess) rc = MDX_ERR_HIP;
  }
  if (!rc && d_force_out) {
    k_forces_out<<<div_up(c->n, 256), 256, 0, c->stream>>>(c->ffix.p, live_fpart(c), c->n, d_force_out);
    if (hipGetLastError() != hipSuccess) rc = MDX_ERR_HIP;
  }
  c->evals++;
  LEAVE_RET(rc);
}

extern "C" int mdx_compute_forces(mdx_ctx* c, unsigned terms, double* d_force_out) {
  return compute_forces(c, terms, d_force_out, true);
}

extern "C" int mdx_compute_forces_only(mdx_ctx* c, unsigned terms, double* d_force_out) {
  return compute_forces(c, terms, d_force_out, false);
}

extern "C" int mdx_energies_fetch(mdx_ctx* c, double* h_out) {
  MDX_CHECK(c && h_out, MDX_ERR_INVALID, "mdx_energies_fetch: bad arguments");
  long long e[MDX_NTERMS + 1];
  MDX_HIP(hipMemcpyAsync(e, c->efix.p, sizeof(e), hipMemcpyDeviceToHost, c->stream));
  MDX_HIP(hipStreamSynchronize(c->stream));
  for (int t = 0; t <= MDX_NTERMS; ++t) h_out[t] = (double)e[t] * kFixInv;
  return 0;
}

extern "C" int mdx_status_fetch(mdx_ctx* c, unsigned* h_flags) {
  MDX_CHECK(c && h_flags, MDX_ERR_INVALID, "mdx_status_fetch: bad arguments");
  MDX_HIP(hipMemcpyAsync(h_flags, &c->dstate.p->flags, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
  MDX_HIP(hipMemsetAsync(&c->dstate.p->flags, 0, sizeof(unsigned), c->stream));
  MDX_HIP(hipStreamSynchronize(c->stream));
  if (*h_flags & MDX_FLAG_SYNC) {  // a barrier gave up: re-arm its arrival count
    MDX_HIP(hipMemset(&c->dstate.p->bar_count, 0, sizeof(unsigned)));
  }
  return 0;
}

extern "C" int mdx_stats_fetch(mdx_ctx* c, int64_t* h_rebuilds, int64_t* h_evals) {
  MDX_CHECK(c, MDX_ERR_INVALID, "null context");
  long long nr = 0;
  MDX_HIP(hipMemcpyAsync(&nr, &c->dstate.p->nrebuild, sizeof(long long), hipMemcpyDeviceToHost, c->stream));
  MDX_HIP(hipStreamSynchronize(c->stream));
  if (h_rebuilds) *h_rebuilds = nr;
  if (h_evals) *h_evals = c->evals;
  return 0;
}

// ------------------------------------------------------------------ integration
extern "C" int mdx_integrator_config(mdx_ctx* c, int kind, double dt, double gamma, double kT, uint64_t seed) {
  MDX_CHECK(c && c->n > 0, MDX_ERR_INVALID, "mdx_integrator_config: no system");
  MDX_CHECK(kind >= 0 && kind <= 3, MDX_ERR_INVALID, "unknown integrator kind %d", kind);
  MDX_CHECK(dt > 0 && gamma >= 0 && kT >= 0, MDX_ERR_INVALID, "dt > 0, gamma >= 0, kT >= 0 required");
  if (kind != c->int_kind) c->int_primed = false;
  c->int_kind = kind;
  c->dt = dt;
  c->gamma = gamma;
  c->kT = kT;
  c->seed = seed;
  const size_t n3 = 3 * (size_t)c->n;  // step kernels write these from the first step on
  int rc;
  if ((rc = c->acc_prev.ensure(n3)) || (rc = c->acc_cur.ensure(n3)) || (rc = c->gsave.ensure(n3))) return rc;
  c->int_configured = true;
  mdx_invalidate_graph(c);
  return 0;
}

extern "C" int mdx_set_energy_interval(mdx_ctx* c, int every) {
  MDX_CHECK(c && every >= 1, MDX_ERR_INVALID, "energy interval must be >= 1");
  if (every != c->energy_every) {
    // the captured step graphs hold the interval (energy-history rows): capture again; the
    // forces of the last step stay current
    const bool cur = c->forces_current;
    mdx_invalidate_graph(c);
    c->forces_current = cur;
  }
  c->energy_every = every;
  return 0;
}

extern "C" int mdx_integrator_reset(mdx_ctx* c) {
  MDX_CHECK(c, MDX_ERR_INVALID, "null context");
  c->int_primed = false;
  return 0;
}

extern "C" int mdx_set_external_noise(mdx_ctx* c, const double* d_noise, int nsteps) {
  MDX_CHECK(c, MDX_ERR_INVALID, "null context");
  long long step = 0;
  MDX_HIP(hipMemcpy(&step, &c->dstate.p->step, sizeof(long long), hipMemcpyDeviceToHost));
  c->ext_noise = d_noise;
  c->ext_noise_steps = d_noise ? nsteps : 0;
  c->ext_noise_base = step + 1;  // first noise row is used by the next step
  mdx_invalidate_graph(c);
  return 0;
}

static IntArgs int_args(mdx_ctx* c) {
  IntArgs A;
  A.n = c->n;
  A.dt = c->dt;
  A.a = (1 - c->gamma * c->dt / 2) / (1 + c->gamma * c->dt / 2);
  A.b = 1 / (1 + c->gamma * c->dt / 2);
  A.sigma = sqrt(2 * c->kT * c->gamma);
  for (int d = 0; d < 3; ++d) {
    A.L[d] = c->box[d];
    A.inv[d] = 1.0 / c->box[d];
  }
  A.ns.seed = c->seed;
  A.ns.ext = c->ext_noise;
  A.ns.ext_base = c->ext_noise_base;
  A.ns.ext_steps = c->ext_noise_steps;
  A.fpart = live_fpart(c);
  A.gsave = c->gsave.p;
  return A;
}

static int zero_and_record(mdx_ctx* c) {
  long long n3 = 3LL * c->n;
  k_zero_acc<<<div_up(n3, 256), 256, 0, c->stream>>>(c->ffix.p, n3, c->efix.p, MDX_NTERMS + 1, c->ehist.p,
                                                     c->ehist_cap, c->dstate.p, 1, c->energy_every);
  MDX_LAUNCH_CHECK();
  return 0;
}

// One MD step; the launch sequence depends only on (kind, terms), so it can be captured.
static StepPro step_pro(mdx_ctx* c) {
  StepPro P;
  P.ffix = c->ffix.p;
  P.pos_build = c->pos_at_build.p;
  P.orig_sorted = c->orig_sorted.p;
  P.qs = c->has_charges ? c->charge_s.p : nullptr;
  P.xq = c->xq.p;
  P.pos_sorted = c->pos_sorted.p;
  P.efix = c->efix.p;
  P.ne = MDX_NTERMS + 1;
  P.ehist = c->ehist.p;
  P.ehist_cap = c->ehist_cap;
  P.every = c->energy_every;
  P.record = 1;
  P.cell_count = c->cell_count.p;
  P.ncell = c->ncell;
  const double half = 0.5 * c->skin;
  P.thr2 = c->skin > 0 ? half * half : -1.0;
  P.rebuild_every = c->rebuild_every;
  // the last workgroup bins the atoms itself when the grid's counts fit its LDS
  P.bin = 2 * (size_t)c->ncell * sizeof(int) <= PRO_BIN_LDS_MAX ? 1 : 0;
  P.atom_cell = c->atom_cell.p;
  P.cell_start = c->cell_start.p;
  P.sorted_orig = c->sorted_orig.p;
  for (int d = 0; d < 3; ++d) {
    P.gn[d] = c->gn[d];
    P.gw[d] = c->box[d] / c->gn[d];
  }
  P.n = c->n;
  return P;
}

// The step prologue applies (G-JF or velocity Verlet) when the step's forces walk the list grid
// (pair terms or the PME plane kernels) of an exact-pair-set context.
static bool fused_steps(const mdx_ctx* c, unsigned terms) {
  const bool grid_used = c->nb_configured && c->pair_mode == MDX_PAIRS_EXACT &&
                         (terms & (MDX_TERM_LJ | MDX_TERM_COUL_DIRECT | MDX_TERM_COUL_RECIP)) != 0 &&
                         ((terms & (MDX_TERM_LJ | MDX_TERM_COUL_DIRECT)) || c->pme_planes);
  return grid_used && (c->int_kind == MDX_INT_LANGEVIN || c->int_kind == MDX_INT_VELOCITY_VERLET);
}

// the step prologue launch of a context (A.fpart: the reciprocal forces of the last evaluation)
static ProCall pro_call(mdx_ctx* c) {
  ProCall k;
  k.A = int_args(c);
  k.A.fpart = live_fpart(c);
  k.P = step_pro(c);
  k.x = c->pos;
  k.v = c->vel;
  k.im = c->inv_mass.p;
  k.m = c->mass.p;
  k.acc_prev = c->acc_prev.p;
  k.ke_slot = c->efix.p + MDX_E_KINETIC;
  k.st = c->dstate.p;
  k.blocks = div_up(c->n, PRO_THREADS);
  return k;
}
// its dynamic LDS: the last workgroup's binning counts and starts
static size_t pro_lds_bytes(const mdx_ctx* c, const ProCall& k) { return k.P.bin ? 2 * (size_t)c->ncell * sizeof(int) : 0; }

// One integrator step. fuse_prev_b: this step's prologue also completes the previous step's
// velocity half (the previous call to this function had emit_b = false); emit_b: finish this
// step with the standalone velocity half.
int mdx_integrator_enqueue_step(mdx_ctx* c, unsigned terms, bool energy, bool fuse_prev_b, bool emit_b) {
  MDX_CHECK(c->int_configured && c->acc_prev.p && c->gsave.p, MDX_ERR_INVALID, "integrator not configured");
  IntArgs A = int_args(c);
  dim3 g(div_up(c->n, 256)), b(256);
  int rc;
  long long* ke = c->efix.p + MDX_E_KINETIC;
  if (fused_steps(c, terms)) {
    const ProCall pc = pro_call(c);  // the forces the prologue reads are the last evaluation's
    const size_t lds = pro_lds_bytes(c, pc);
    dim3 gp(pc.blocks), bp(PRO_THREADS);
    const bool gjf = c->int_kind == MDX_INT_LANGEVIN;
#define PRO_L(G, F) k_step_pro<G, F><<<gp, bp, lds, c->stream>>>(pc)
    if (gjf && fuse_prev_b) PRO_L(true, true);
    else if (gjf) PRO_L(true, false);
    else if (fuse_prev_b) PRO_L(false, true);
    else PRO_L(false, false);
#undef PRO_L
    MDX_LAUNCH_CHECK();
    c->prologue_done = true;
    c->prologue_binned = pc.P.bin != 0;
    rc = mdx_enqueue_forces(c, terms, energy);
    c->prologue_done = false;
    c->prologue_binned = false;
    if (rc) return rc;
    if (emit_b) {
      A.fpart = live_fpart(c);
      if (gjf)
        k_gjf_b<<<g, b, 0, c->stream>>>(A, c->vel, c->inv_mass.p, c->mass.p, c->ffix.p, c->acc_prev.p, c->dstate.p, ke);
      else
        k_vv_b<<<g, b, 0, c->stream>>>(A, c->vel, c->inv_mass.p, c->mass.p, c->ffix.p, ke);
      MDX_LAUNCH_CHECK();
    }
    return 0;
  }
  switch (c->int_kind) {
    case MDX_INT_VELOCITY_VERLET:
      k_vv_a<<<g, b, 0, c->stream>>>(A, c->pos, c->vel, c->inv_mass.p, c->ffix.p);
      MDX_LAUNCH_CHECK();
      if ((rc = zero_and_record(c)) || (rc = mdx_enqueue_forces(c, terms, energy))) return rc;
      A.fpart = live_fpart(c);
      k_vv_b<<<g, b, 0, c->stream>>>(A, c->vel, c->inv_mass.p, c->mass.p, c->ffix.p, ke);
      MDX_LAUNCH_CHECK();
      break;
    case MDX_INT_LANGEVIN:
      k_gjf_a<<<g, b, 0, c->stream>>>(A, c->pos, c->vel, c->inv_mass.p, c->ffix.p, c->acc_prev.p, c->dstate.p);
      MDX_LAUNCH_CHECK();
      if ((rc = zero_and_record(c)) || (rc = mdx_enqueue_forces(c, terms, energy))) return rc;
      A.fpart = live_fpart(c);
      k_gjf_b<<<g, b, 0, c->stream>>>(A, c->vel, c->inv_mass.p, c->mass.p, c->ffix.p, c->acc_prev.p, c->dstate.p, ke);
      MDX_LAUNCH_CHECK();
      break;
    case MDX_INT_VERLET_REF:
      k_vref<<<g, b, 0, c->stream>>>(A, c->xa.p, c->xb.p, c->pos, c->inv_mass.p, c->ffix.p, c->dstate.p);
      MDX_LAUNCH_CHECK();
      if ((rc = zero_and_record(c)) || (rc = mdx_enqueue_forces(c, terms, energy))) return rc;
      A.fpart = live_fpart(c);
      k_kinetic<<<g, b, 0, c->stream>>>(c->n, c->vel, c->mass.p, ke);
      MDX_LAUNCH_CHECK();
      break;
    case MDX_INT_LANGEVIN_REF:
      k_lref_a<<<g, b, 0, c->stream>>>(A, c->xa.p, c->xb.p, c->vb.p, c->pos, c->inv_mass.p, c->acc_cur.p,
                                       c->acc_prev.p, c->dstate.p);
      MDX_LAUNCH_CHECK();
      if ((rc = zero_and_record(c)) || (rc = mdx_enqueue_forces(c, terms, energy))) return rc;
      A.fpart = live_fpart(c);
      k_lref_b<<<g, b, 0, c->stream>>>(A, c->va.p, c->vb.p, c->inv_mass.p, c->mass.p, c->ffix.p, c->acc_cur.p,
                                       c->acc_prev.p, c->dstate.p, ke);
      MDX_LAUNCH_CHECK();
      break;
  }
  return 0;
}

// ---- integrate = prologue (prime forces, energy-history base) + steps + epilogue
// buffers of an integrate call (before any graph is captured: growing the energy history
// invalidates captured steps, which hold its pointer)
static int integ_alloc(mdx_ctx* c, int nsteps) {
  MDX_CHECK(c && c->n > 0 && c->vel, MDX_ERR_INVALID, "mdx_integrate: system with velocities required");
  MDX_CHECK(c->int_configured && c->has_mass, MDX_ERR_INVALID, "integrator or masses not configured");
  MDX_CHECK(nsteps >= 0, MDX_ERR_INVALID, "nsteps must be >= 0");
  int rc = 0;
  size_t n3 = 3 * (size_t)c->n;
  if ((rc = c->acc_prev.ensure(n3)) || (rc = c->acc_cur.ensure(n3)) || (rc = c->gsave.ensure(n3))) return rc;
  if (c->int_kind >= MDX_INT_VERLET_REF &&
      ((rc = c->xa.ensure(n3)) || (rc = c->xb.ensure(n3)) || (rc = c->va.ensure(n3)) || (rc = c->vb.ensure(n3))))
    return rc;
  const int ne = MDX_NTERMS + 1;
  const int nrows = nsteps / c->energy_every + 1;
  if (nrows > c->ehist_cap) {
    const int cap = std::max(nrows, 4096);
    if ((rc = c->ehist.ensure((size_t)cap * ne))) return rc;
    c->ehist_cap = cap;
    mdx_invalidate_graph(c);  // captured steps hold the old history pointer
  }
  return 0;
}

static bool prime_skipped(const mdx_ctx* c, unsigned terms) {
  return c->forces_current && c->forces_terms == terms && fused_steps(c, terms);
}

static int integ_begin(mdx_ctx* c, unsigned terms, int nsteps) {
  int rc = integ_alloc(c, nsteps);
  if (rc) return rc;
  dim3 g(div_up(c->n, 256)), b(256);
  c->ehist_count = 0;
  // prime: forces at the current positions (every mdpy integrate call starts with ensemble.update,
  // integrator.py:26-28). The step kernels of G-JF and velocity Verlet keep the last step's forces
  // and potential energies (the last step of a call always evaluates energies); when nothing has
  // changed since, they are the forces at the current positions and the prime is skipped.
  const bool reuse = prime_skipped(c, terms);
  if (!reuse && ((rc = mdx_enqueue_zero_accumulators(c)) || (rc = mdx_enqueue_forces(c, terms, true)))) return rc;
  IntArgs A = int_args(c);
  k_call_start<<<g, b, 0, c->stream>>>(c->n, c->vel, c->mass.p, c->efix.p + MDX_E_KINETIC, c->dstate.p);
  if (c->int_kind == MDX_INT_VERLET_REF) {
    k_vref_prime<<<g, b, 0, c->stream>>>(A, c->pos, c->vel, c->inv_mass.p, c->ffix.p, c->xa.p, c->xb.p,
                                         c->int_primed ? 1 : 0);
  } else if (c->int_kind == MDX_INT_LANGEVIN_REF) {
    k_lref_prime<<<g, b, 0, c->stream>>>(A, c->pos, c->vel, c->inv_mass.p, c->ffix.p, c->xa.p, c->xb.p, c->va.p,
                                         c->vb.p, c->acc_cur.p, c->int_primed ? 1 : 0);
  }
  MDX_LAUNCH_CHECK();
  c->int_primed = true;
  return 0;
}

// the closing kernels of a call: the last energy record, then the state as the reference leaves it
static int integ_end_kernels(mdx_ctx* c) {
  const int ne = MDX_NTERMS + 1;
  dim3 g(div_up(c->n, 256)), b(256);
  IntArgs A = int_args(c);
  k_record<<<1, 1, 0, c->stream>>>(c->efix.p, ne, c->ehist.p, c->ehist_cap, c->dstate.p, c->energy_every);
  MDX_LAUNCH_CHECK();
  // finish: state velocities / wrapped positions as the reference leaves them
  if (c->int_kind == MDX_INT_VERLET_REF) {
    k_vref_finish<<<g, b, 0, c->stream>>>(A, c->xa.p, c->xb.p, c->vel);
  } else if (c->int_kind == MDX_INT_LANGEVIN_REF) {
    k_copy3<<<div_up(3LL * c->n, 256), 256, 0, c->stream>>>(c->va.p, c->vel, 3LL * c->n);
  } else {
    k_wrap_state<<<g, b, 0, c->stream>>>(A, c->pos, c->dstate.p);
  }
  MDX_LAUNCH_CHECK();
  return 0;
}

static int integ_end(mdx_ctx* c, int nsteps, bool kernels) {
  if (kernels) {
    int rc = integ_end_kernels(c);
    if (rc) return rc;
  }
  c->ehist_count = nsteps / c->energy_every + 1;
  c->evals += nsteps + 1;
  // The wrap moves atoms by whole box edges: the list's displacement test is a minimum-image
  // distance and every kernel re-wraps positions, so the list stays valid. The reference-
  // compatible integrators wrap every step and keep their own history; they always prime.
  if (c->int_kind >= MDX_INT_VERLET_REF) return mdx_enqueue_request_rebuild(c);
  return 0;
}

// Potential energies are evaluated on the call's steps whose index + 1 is a multiple of the
// energy interval (the prime counts as step 0 of the history) and always on the call's last step,
// so the next call can start from its forces and energies.
static bool energy_step(const mdx_ctx* c, int local, int nsteps) {
  return (local + 1) % c->energy_every == 0 || local == nsteps - 1;
}

// steps local0 .. local0+count-1 of the current call. With the step prologue every step but
// the call's first folds the previous step's velocity half into its prologue (fuse_first for the
// first of these), and the call ends with one standalone velocity half (integ_close), so graph
// replays and direct launches run the same kernels in the same order.
static int integ_steps(mdx_ctx* c, unsigned terms, int local0, int count, bool fuse_first, int nsteps) {
  int rc = 0;
  for (int k = 0; k < count && !rc; ++k)
    rc = mdx_integrator_enqueue_step(c, terms, energy_step(c, local0 + k, nsteps), k > 0 || fuse_first, false);
  return rc;
}

// the closing kernels of a call: for G-JF / velocity Verlet with the step prologue, one launch
// (the closing velocity half, the wrap and the last energy record); otherwise the energy record
// and the reference integrators' finishing kernels
static int integ_close(mdx_ctx* c, unsigned terms, int nsteps) {
  if (nsteps <= 0 || !fused_steps(c, terms)) return integ_end_kernels(c);
  IntArgs A = int_args(c);
  dim3 g(div_up(c->n, 256)), b(256);
  const int ne = MDX_NTERMS + 1;
  if (c->int_kind == MDX_INT_LANGEVIN)
    k_close<true><<<g, b, 0, c->stream>>>(A, c->vel, c->pos, c->inv_mass.p, c->mass.p, c->ffix.p, c->acc_prev.p,
                                          c->dstate.p, c->efix.p, ne, c->ehist.p, c->ehist_cap, c->energy_every);
  else
    k_close<false><<<g, b, 0, c->stream>>>(A, c->vel, c->pos, c->inv_mass.p, c->mass.p, c->ffix.p, c->acc_prev.p,
                                           c->dstate.p, c->efix.p, ne, c->ehist.p, c->ehist_cap, c->energy_every);
  MDX_LAUNCH_CHECK();
  return 0;
}

// Graphs of kGraphSizes[k] forces-only steps and one of a single energy step, each following a
// step of the same call (its prologue folds the previous velocity half). A call's steps after
// its first are a sequence of forces-only runs separated by energy steps; each run is covered
// by the largest graphs that fit, so the same executable graphs serve every call length and are
// captured once per (terms, integrator kind).
static const int kGraphSizes[MDX_NGRAPH_SIZES] = {32, 16, 8, 4, 2, 1};
static const int kGraphEnergy = MDX_NGRAPH_SIZES, kGraphTail = MDX_NGRAPH_SIZES + 1;  // graph_of indices

template <class Enqueue>
static int capture(mdx_ctx* c, Enqueue enqueue, hipGraphExec_t* out) {
  hipGraph_t graph;
  MDX_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = enqueue();
  hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
  if (rc) return rc;
  if (e2 != hipSuccess) { mdx_set_error("graph capture failed: %s", hipGetErrorString(e2)); return MDX_ERR_HIP; }
  hipError_t e3 = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (e3 != hipSuccess) { *out = nullptr; mdx_set_error("graph instantiate failed: %s", hipGetErrorString(e3)); return MDX_ERR_HIP; }
  return 0;
}

static int capture_steps(mdx_ctx* c, unsigned terms, int count, bool energy, hipGraphExec_t* out) {
  return capture(c, [&]() {
    int rc = 0;
    for (int k = 0; k < count && !rc; ++k) rc = mdx_integrator_enqueue_step(c, terms, energy, true, false);
    return rc;
  }, out);
}

// a call's set-up and first step when the prime is skipped (integ_begin + integ_steps(0, 1))
static int capture_head(mdx_ctx* c, unsigned terms, hipGraphExec_t* out) {
  return capture(c, [&]() {
    dim3 g(div_up(c->n, 256)), b(256);
    k_call_start<<<g, b, 0, c->stream>>>(c->n, c->vel, c->mass.p, c->efix.p + MDX_E_KINETIC, c->dstate.p);
    MDX_LAUNCH_CHECK();
    return mdx_integrator_enqueue_step(c, terms, false, false, false);
  }, out);
}

// a call's last step (an energy step), its closing velocity half and the closing kernels
static int capture_tail(mdx_ctx* c, unsigned terms, hipGraphExec_t* out) {
  return capture(c, [&]() {
    int rc = mdx_integrator_enqueue_step(c, terms, true, true, false);
    if (!rc) rc = integ_close(c, terms, 1);
    return rc;
  }, out);
}

static int ensure_graphs(mdx_ctx* c, unsigned terms) {
  if (c->graph_e && c->graph_terms == terms && c->graph_kind == c->int_kind) return 0;
  const bool cur = c->forces_current;  // capturing enqueues nothing that runs
  const unsigned cur_terms = c->forces_terms;
  const bool cur_fpart = c->fpart_live;
  mdx_invalidate_graph(c);
  // every captured step follows an evaluation of `terms` when it replays: its prologue must read
  // the reciprocal forces that evaluation leaves, whatever the context evaluated last (a fresh
  // context has evaluated nothing yet)
  c->fpart_live = evaluation_leaves_fpart(c, terms);
  int rc = 0;
  for (int k = 0; k < MDX_NGRAPH_SIZES && !rc; ++k) rc = capture_steps(c, terms, kGraphSizes[k], false, &c->graph_f[k]);
  if (!rc) rc = capture_steps(c, terms, 1, true, &c->graph_e);
  if (!rc && fused_steps(c, terms)) rc = capture_head(c, terms, &c->graph_head);
  if (!rc && fused_steps(c, terms)) rc = capture_tail(c, terms, &c->graph_tail);
  c->fpart_live = cur_fpart;
  if (rc) {
    mdx_invalidate_graph(c);
    return rc;
  }
  c->graph_terms = terms;
  c->graph_kind = c->int_kind;
  c->graph_lanes = c->nb_lanes;
  c->forces_current = cur;
  c->forces_terms = cur_terms;
  return 0;
}

// the launches covering steps 1 .. nsteps-1 of a call (graph_of indices: a forces-only size, or
// kGraphEnergy for an energy step)
static void graph_schedule(const mdx_ctx* c, int nsteps, std::vector<int>& seq) {
  seq.clear();
  int s = 1;
  while (s < nsteps) {
    if (energy_step(c, s, nsteps)) {
      seq.push_back(kGraphEnergy);
      ++s;
      continue;
    }
    int run = 0;
    while (s + run < nsteps && !energy_step(c, s + run, nsteps)) ++run;
    s += run;
    for (int k = 0; k < MDX_NGRAPH_SIZES; ++k)
      for (; run >= kGraphSizes[k]; run -= kGraphSizes[k]) seq.push_back(k);
  }
}

static hipGraphExec_t graph_of(const mdx_ctx* c, int k) {
  return k == kGraphTail ? c->graph_tail : (k == kGraphEnergy ? c->graph_e : c->graph_f[k]);
}

extern "C" int mdx_integrate(mdx_ctx* c, unsigned terms, int nsteps, int use_graph) {
  MDX_CHECK(c, MDX_ERR_INVALID, "null context");
  ENTER();
  int rc = integ_alloc(c, nsteps);
  if (!rc && use_graph && nsteps > 1) rc = ensure_graphs(c, terms);
  if (rc) LEAVE_RET(rc);
  // a call of n >= 2 steps whose first step needs no prime and no energies: the head graph
  // (set-up + step 0), the step graphs, then the tail graph (step n-1 + the closing kernels)
  if (use_graph && nsteps > 1 && c->graph_head && c->graph_tail && prime_skipped(c, terms) &&
      !energy_step(c, 0, nsteps)) {
    c->ehist_count = 0;
    c->int_primed = true;
    std::vector<int> seq;
    graph_schedule(c, nsteps, seq);  // ends with step n-1, an energy step: the tail graph's
    seq.back() = kGraphTail;
    hipError_t e = hipGraphLaunch(c->graph_head, c->stream);
    for (size_t k = 0; k < seq.size() && e == hipSuccess; ++k) e = hipGraphLaunch(graph_of(c, seq[k]), c->stream);
    if (e != hipSuccess) { mdx_set_error("hipGraphLaunch: %s", hipGetErrorString(e)); LEAVE_RET(MDX_ERR_HIP); }
    rc = integ_end(c, nsteps, false);
    if (!rc) {
      c->forces_current = true;
      c->forces_terms = terms;
    }
    LEAVE_RET(rc);
  }
  if (!rc) rc = integ_begin(c, terms, nsteps);
  if (rc) LEAVE_RET(rc);
  // step 0 of the call is launched directly (its prologue has no previous velocity half), the
  // following steps are graph replays, then the final velocity half
  if (use_graph && nsteps > 1) {
    if ((rc = integ_steps(c, terms, 0, 1, false, nsteps))) LEAVE_RET(rc);
    std::vector<int> seq;
    graph_schedule(c, nsteps, seq);
    for (int k : seq) {
      hipError_t e = hipGraphLaunch(graph_of(c, k), c->stream);
      if (e != hipSuccess) { mdx_set_error("hipGraphLaunch: %s", hipGetErrorString(e)); LEAVE_RET(MDX_ERR_HIP); }
    }
  } else {
    rc = integ_steps(c, terms, 0, nsteps, false, nsteps);
  }
  if (!rc) rc = integ_close(c, terms, nsteps);
  if (rc) LEAVE_RET(rc);
  rc = integ_end(c, nsteps, false);
  if (!rc && fused_steps(c, terms)) {
    c->forces_current = true;
    c->forces_terms = terms;
  }
  LEAVE_RET(rc);
}

// ------------------------------------------------------------------ batched replicas
// Independent systems (one context each, same device) advanced together. The contexts keep
// their own streams, lists, grids and graphs; the batch joins the caller's stream to all of
// them and back.
// MDX_BATCH_FUSED: each kernel of a step is one launch over every replica's workgroups, the
// replicas' arguments in device tables (BatchKind); steps 1 .. n-1 of a call replay graphs of
// such launches captured on the lead stream, the PME planes forked onto the batch's side stream.
struct BatchFused {
  BatchKind pro, fill, lfast, lcell, pfwd, pfft, pgat, nb;
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipGraphExec_t gf[MDX_NGRAPH_SIZES] = {};
  hipGraphExec_t ge = nullptr;
  std::vector<unsigned long long> gens;  // the contexts' graph generations the tables were built at
  unsigned terms = 0;
  bool ready = false;
  bool gjf = false, s14 = false;
  int lanes = 0;
};

struct mdx_batch {
  std::vector<mdx_ctx*> ctx;
  hipStream_t ext = nullptr;     // caller's stream
  hipStream_t lead = nullptr;    // capture / launch stream of the one-graph and fused modes
  hipEvent_t ev_in = nullptr, ev_fork = nullptr;
  std::vector<hipEvent_t> ev_out;
  BatchFused fz;
};

// drop the fused mode's graphs (after the batch's streams have finished with them)
static void fused_drop_graphs(mdx_batch* b) {
  BatchFused& f = b->fz;
  bool any = f.ge != nullptr;
  for (hipGraphExec_t g : f.gf) any |= g != nullptr;
  if (any) {
    if (b->lead) (void)hipStreamSynchronize(b->lead);
    if (f.side) (void)hipStreamSynchronize(f.side);
    for (hipGraphExec_t& g : f.gf) {
      if (g) (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
    if (f.ge) (void)hipGraphExecDestroy(f.ge);
    f.ge = nullptr;
  }
  f.ready = false;
}

extern "C" int mdx_batch_create(mdx_ctx** ctxs, int n, void* hip_stream, mdx_batch** out) {
  MDX_CHECK(ctxs && n > 0 && out, MDX_ERR_INVALID, "mdx_batch_create: bad arguments");
  for (int i = 0; i < n; ++i) {
    MDX_CHECK(ctxs[i], MDX_ERR_INVALID, "mdx_batch_create: context %d is NULL", i);
    MDX_CHECK(ctxs[i]->device == ctxs[0]->device, MDX_ERR_INVALID, "all contexts of a batch share one device");
    for (int j = 0; j < i; ++j) MDX_CHECK(ctxs[j] != ctxs[i], MDX_ERR_INVALID, "context %d repeated", i);
  }
  MDX_HIP(hipSetDevice(ctxs[0]->device));
  mdx_batch* b = new mdx_batch();
  b->ctx.assign(ctxs, ctxs + n);
  // the one-launch PME kernel needs all its blocks resident at once; many contexts advancing
  // concurrently could together exceed the device, so batched contexts use the three-kernel path
  for (int i = 0; i < n; ++i) {
    ctxs[i]->in_batch = true;
    mdx_invalidate_graph(ctxs[i]);
  }
  b->ext = (hipStream_t)hip_stream;
  b->ev_out.resize(n, nullptr);
  bool ok = hipStreamCreateWithFlags(&b->lead, hipStreamNonBlocking) == hipSuccess &&
            hipStreamCreateWithFlags(&b->fz.side, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&b->fz.ev_fork, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&b->fz.ev_join, hipEventDisableTiming) == hipSuccess;
  for (int i = 0; i < n && ok; ++i) ok = hipEventCreateWithFlags(&b->ev_out[i], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    mdx_batch_destroy(b);
    mdx_set_error("mdx_batch_create: stream/event creation failed");
    return MDX_ERR_HIP;
  }
  *out = b;
  return 0;
}

extern "C" int mdx_batch_destroy(mdx_batch* b) {
  if (!b) return 0;
  fused_drop_graphs(b);
  for (mdx_ctx* c : b->ctx) {  // standalone again
    c->in_batch = false;
    mdx_invalidate_graph(c);
  }
  if (b->fz.ev_fork) hipEventDestroy(b->fz.ev_fork);
  if (b->fz.ev_join) hipEventDestroy(b->fz.ev_join);
  if (b->fz.side) hipStreamDestroy(b->fz.side);
  for (BatchKind* k : {&b->fz.pro, &b->fz.fill, &b->fz.lfast, &b->fz.lcell, &b->fz.pfwd, &b->fz.pfft, &b->fz.pgat,
                       &b->fz.nb})
    k->buf.release();
  for (hipEvent_t e : b->ev_out)
    if (e) hipEventDestroy(e);
  if (b->ev_in) hipEventDestroy(b->ev_in);
  if (b->ev_fork) hipEventDestroy(b->ev_fork);
  if (b->lead) hipStreamDestroy(b->lead);
  delete b;
  return 0;
}

// stream `s` waits for all work enqueued so far on `from`
static int stream_after(hipStream_t s, hipStream_t from, hipEvent_t ev) {
  MDX_HIP(hipEventRecord(ev, from));
  MDX_HIP(hipStreamWaitEvent(s, ev, 0));
  return 0;
}

// ---- MDX_BATCH_FUSED
// The terms a fused batch step evaluates: LJ and PME (direct and reciprocal), plus bonded terms.
static const unsigned kFusedBonded = MDX_TERM_BOND | MDX_TERM_ANGLE | MDX_TERM_DIHEDRAL | MDX_TERM_IMPROPER;

// Build (or keep) the per-replica argument tables of every batched kernel. They hold the
// contexts' buffers and settings, so they are rebuilt whenever a context's captured graphs went
// stale (every setting that changes a launch invalidates them) or the terms change.
static int fused_prepare(mdx_batch* b, unsigned terms) {
  BatchFused& f = b->fz;
  const int n = (int)b->ctx.size();
  bool fresh = f.ready && f.terms == terms && f.gens.size() == (size_t)n && f.lanes == b->ctx[0]->nb_lanes &&
               f.gjf == (b->ctx[0]->int_kind == MDX_INT_LANGEVIN);
  for (int r = 0; r < n && fresh; ++r) fresh = f.gens[r] == b->ctx[r]->graph_gen;
  if (fresh) return 0;
  fused_drop_graphs(b);
  // the tables are rewritten in place: no earlier batched launch may still be reading them
  MDX_HIP(hipStreamSynchronize(b->lead));
  MDX_HIP(hipStreamSynchronize(f.side));
  const unsigned need = MDX_TERM_LJ | MDX_TERM_COUL_DIRECT | MDX_TERM_COUL_RECIP;
  MDX_CHECK((terms & need) == need && !(terms & ~(need | kFusedBonded)), MDX_ERR_INVALID,
            "MDX_BATCH_FUSED evaluates LJ + PME (+ bond, angle, dihedral, improper terms) only");
  mdx_ctx* c0 = b->ctx[0];
  std::vector<ProCall> pro;
  std::vector<int> pblocks;
  size_t plds = 0;
  for (int r = 0; r < n; ++r) {
    mdx_ctx* c = b->ctx[r];
    MDX_CHECK(fused_steps(c, terms) && c->int_kind == c0->int_kind, MDX_ERR_INVALID,
              "replica %d: MDX_BATCH_FUSED needs G-JF or velocity Verlet (the same in every replica) on a "
              "neighbour-list context", r);
    MDX_CHECK(c->has_lj && c->has_charges && c->pme_configured && c->skin > 0, MDX_ERR_INVALID,
              "replica %d: MDX_BATCH_FUSED needs LJ parameters, charges, PME and a list skin > 0", r);
    // the steps' prologues read the reciprocal forces of the batched evaluation before them
    const bool keep = c->fpart_live;
    c->fpart_live = true;
    const ProCall pc = pro_call(c);
    c->fpart_live = keep;
    MDX_CHECK(pc.P.bin, MDX_ERR_INVALID, "replica %d: list grid too large for the prologue's binning", r);
    pro.push_back(pc);
    pblocks.push_back(pc.blocks);
    plds = std::max(plds, pro_lds_bytes(c, pc));
  }
  int rc;
  if ((rc = batch_kind_upload(&f.pro, pro, pblocks, plds)) ||
      (rc = mdx_batch_prepare_list(b->ctx.data(), n, &f.fill, &f.lfast, &f.lcell)) ||
      (rc = mdx_batch_prepare_pme(b->ctx.data(), n, &f.pfwd, &f.pfft, &f.pgat)) ||
      (rc = mdx_batch_prepare_nb(b->ctx.data(), n, terms & (kFusedBonded | MDX_TERM_COUL_RECIP), &f.nb)))
    return rc;
  f.gjf = c0->int_kind == MDX_INT_LANGEVIN;
  f.lanes = c0->nb_lanes;
  f.s14 = c0->has_s14;
  f.terms = terms;
  f.gens.resize(n);
  for (int r = 0; r < n; ++r) f.gens[r] = b->ctx[r]->graph_gen;
  f.ready = true;
  return 0;
}

// one step of every replica on the lead stream (fuse_prev: the prologue completes the previous
// step's velocity half)
static int fused_step(mdx_batch* b, bool energy, bool fuse_prev) {
  BatchFused& f = b->fz;
  hipStream_t s = b->lead;
  const ProCall* pc = (const ProCall*)f.pro.calls;
#define PRO_B(G, F) k_step_pro_batch<G, F><<<f.pro.blocks, PRO_THREADS, f.pro.lds, s>>>(f.pro.map, pc)
  if (f.gjf && fuse_prev) PRO_B(true, true);
  else if (f.gjf) PRO_B(true, false);
  else if (fuse_prev) PRO_B(false, true);
  else PRO_B(false, false);
#undef PRO_B
  MDX_LAUNCH_CHECK();
  int rc;
  if ((rc = mdx_batch_launch_fill(f.fill, s))) return rc;
  // PME planes beside the list build and the pair kernel, as in a single context's step
  MDX_HIP(hipEventRecord(f.ev_fork, s));
  MDX_HIP(hipStreamWaitEvent(f.side, f.ev_fork, 0));
  if ((rc = mdx_batch_launch_pme(f.pfwd, f.pfft, f.pgat, f.side))) return rc;
  MDX_HIP(hipEventRecord(f.ev_join, f.side));
  if ((rc = mdx_batch_launch_list(f.lfast, f.lcell, s)) || (rc = mdx_batch_launch_nb(f.nb, f.lanes, f.s14, energy, s)))
    return rc;
  MDX_HIP(hipStreamWaitEvent(s, f.ev_join, 0));
  return 0;
}

static int fused_capture(mdx_batch* b, int count, bool energy, hipGraphExec_t* out) {
  hipGraph_t graph;
  MDX_HIP(hipStreamBeginCapture(b->lead, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  for (int k = 0; k < count && !rc; ++k) rc = fused_step(b, energy, true);
  hipError_t e2 = hipStreamEndCapture(b->lead, &graph);
  if (rc) return rc;
  if (e2 != hipSuccess) { mdx_set_error("graph capture failed: %s", hipGetErrorString(e2)); return MDX_ERR_HIP; }
  hipError_t e3 = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (e3 != hipSuccess) { *out = nullptr; mdx_set_error("graph instantiate failed: %s", hipGetErrorString(e3)); return MDX_ERR_HIP; }
  return 0;
}

static int fused_graphs(mdx_batch* b) {
  BatchFused& f = b->fz;
  if (f.ge) return 0;
  int rc = 0;
  for (int k = 0; k < MDX_NGRAPH_SIZES && !rc; ++k) rc = fused_capture(b, kGraphSizes[k], false, &f.gf[k]);
  if (!rc) rc = fused_capture(b, 1, true, &f.ge);
  if (rc) fused_drop_graphs(b);
  return rc;
}

// a replica's own call set-up / closing kernels, enqueued on the batch's lead stream
template <class Fn>
static int on_lead(mdx_batch* b, mdx_ctx* c, Fn fn) {
  hipStream_t own = c->stream;
  c->stream = b->lead;
  const int rc = fn();
  c->stream = own;
  return rc;
}

static int fused_integrate(mdx_batch* b, unsigned terms, int nsteps) {
  int rc = 0;
  for (mdx_ctx* c : b->ctx)
    if ((rc = integ_alloc(c, nsteps))) return rc;
  if ((rc = fused_prepare(b, terms))) return rc;
  if (nsteps > 1 && (rc = fused_graphs(b))) return rc;
  MDX_HIP(hipEventRecord(b->ev_in, b->ext));
  MDX_HIP(hipStreamWaitEvent(b->lead, b->ev_in, 0));
  auto steps = [&]() -> int {
    // per replica: the prime (forces at the current positions, unless the last call left them)
    // and the call's opening kernel
    for (mdx_ctx* c : b->ctx)
      if ((rc = on_lead(b, c, [&]() { return integ_begin(c, terms, nsteps); }))) return rc;
    mdx_ctx* c0 = b->ctx[0];
    if (nsteps > 0 && (rc = fused_step(b, energy_step(c0, 0, nsteps), false))) return rc;
    std::vector<int> seq;
    graph_schedule(c0, nsteps, seq);
    for (int k : seq) {
      hipError_t e = hipGraphLaunch(k == kGraphEnergy ? b->fz.ge : b->fz.gf[k], b->lead);
      if (e != hipSuccess) { mdx_set_error("hipGraphLaunch: %s", hipGetErrorString(e)); return MDX_ERR_HIP; }
    }
    for (mdx_ctx* c : b->ctx) {
      if ((rc = on_lead(b, c, [&]() { return integ_close(c, terms, nsteps); })) || (rc = integ_end(c, nsteps, false)))
        return rc;
      if (nsteps > 0) {
        c->fpart_live = true;
        c->forces_current = true;
        c->forces_terms = terms;
      }
    }
    return 0;
  };
  rc = steps();
  // the caller's stream waits for whatever was enqueued, also when a step above failed
  if (rc)
    for (mdx_ctx* c : b->ctx) c->forces_current = false;
  MDX_HIP(hipEventRecord(b->ev_out[0], b->lead));
  MDX_HIP(hipStreamWaitEvent(b->ext, b->ev_out[0], 0));
  return rc;
}

extern "C" int mdx_batch_integrate(mdx_batch* b, unsigned terms, int nsteps, int mode) {
  MDX_CHECK(b && !b->ctx.empty(), MDX_ERR_INVALID, "mdx_batch_integrate: empty batch");
  MDX_CHECK(mode >= MDX_BATCH_ONE_GRAPH && mode <= MDX_BATCH_FUSED, MDX_ERR_INVALID, "unknown batch mode %d", mode);
  // The one-graph mode (every replica's steps captured into one graph from a lead stream, each
  // replica's step forking its PME onto a side stream inside that capture) ended its first GPU
  // run in a host-side segmentation fault whose cause has not been identified; it is switched
  // off until it is.
  MDX_CHECK(mode != MDX_BATCH_ONE_GRAPH, MDX_ERR_INVALID,
            "MDX_BATCH_ONE_GRAPH is disabled (host crash under investigation); use MDX_BATCH_STREAMS");
  MDX_CHECK(nsteps >= 0, MDX_ERR_INVALID, "nsteps must be >= 0");
  const int every = b->ctx[0]->energy_every;
  for (mdx_ctx* c : b->ctx)
    MDX_CHECK(c->energy_every == every, MDX_ERR_INVALID, "all replicas of a batch need the same energy interval");
  MDX_HIP(hipSetDevice(b->ctx[0]->device));
  if (mode == MDX_BATCH_FUSED) return fused_integrate(b, terms, nsteps);
  // join: every replica stream (and the lead stream) waits for the caller's stream
  MDX_HIP(hipEventRecord(b->ev_in, b->ext));
  MDX_HIP(hipStreamWaitEvent(b->lead, b->ev_in, 0));
  for (mdx_ctx* c : b->ctx) MDX_HIP(hipStreamWaitEvent(c->stream, b->ev_in, 0));
  int rc = 0;
  // same schedule as mdx_integrate: step 0 directly, fixed-size graphs for the following steps,
  // final velocity half
  for (mdx_ctx* c : b->ctx) {
    if ((rc = integ_alloc(c, nsteps))) return rc;
    if (nsteps > 1 && (rc = ensure_graphs(c, terms))) return rc;
  }
  for (mdx_ctx* c : b->ctx)
    if ((rc = integ_begin(c, terms, nsteps))) return rc;
  if (nsteps > 0)
    for (mdx_ctx* c : b->ctx)
      if ((rc = integ_steps(c, terms, 0, 1, false, nsteps))) return rc;
  std::vector<int> seq;
  graph_schedule(b->ctx[0], nsteps, seq);
  auto launch = [](mdx_ctx* c, int k) -> int {
    hipError_t e = hipGraphLaunch(graph_of(c, k), c->stream);
    if (e != hipSuccess) { mdx_set_error("hipGraphLaunch: %s", hipGetErrorString(e)); return MDX_ERR_HIP; }
    return 0;
  };
  if (mode == MDX_BATCH_STREAMS) {  // interleaved launches: the replicas' graphs run side by side
    for (int k : seq)
      for (mdx_ctx* c : b->ctx)
        if ((rc = launch(c, k))) return rc;
  } else {  // sequential: one replica after another, on one stream order
    for (size_t i = 0; i < b->ctx.size(); ++i) {
      mdx_ctx* c = b->ctx[i];
      if (i > 0 && (rc = stream_after(c->stream, b->ctx[i - 1]->stream, b->ev_out[i - 1]))) return rc;
      for (int k : seq)
        if ((rc = launch(c, k))) return rc;
    }
  }
  for (mdx_ctx* c : b->ctx) {
    if ((rc = integ_close(c, terms, nsteps))) return rc;
    if ((rc = integ_end(c, nsteps, false))) return rc;
    if (fused_steps(c, terms)) {
      c->forces_current = true;
      c->forces_terms = terms;
    }
  }
  // leave: the caller's stream waits for every replica
  for (size_t i = 0; i < b->ctx.size(); ++i) {
    MDX_HIP(hipEventRecord(b->ev_out[i], b->ctx[i]->stream));
    MDX_HIP(hipStreamWaitEvent(b->ext, b->ev_out[i], 0));
  }
  return 0;
}

// ------------------------------------------------------------------ energy minimisation
// SteepestDescentMinimizer and ConjugateGradientMinimizer (mdpy/minimizer/*.py) on the device
// force path: the position updates and the reductions run on the GPU, every force evaluation
// is the fused one of mdx_compute_forces, and per iteration only a few scalars (the potential
// energy, the conjugate-gradient sums) come back to decide when to stop.
#define MIN_BLOCK 256

// x_i <- wrap(x_i + alpha F_i / |F_i|) (steepest_descent_minimizer.py:37-40). An atom with no
// force at all stays where it is (the reference divides by its zero norm).
__global__ void k_sd_move(IntArgs A, double* __restrict__ pos, const long long* __restrict__ ffix, double alpha,
                          DevState* st) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  double f[3], x[3];
  for (int d = 0; d < 3; ++d) f[d] = fix_force(ffix, A, i, d);
  const double nrm = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  const double s = nrm > 0.0 ? alpha / nrm : 0.0;
  for (int d = 0; d < 3; ++d) x[d] = pos[3 * i + d] + s * f[d];
  wrap_into(A, x, pos + 3 * i, st);
}

// conjugate-gradient probe (conjugate_gradient_minimizer.py:46-53): t = cur_f = F(x), kept in
// cf; x <- wrap(x + theta t); x0 (optional) keeps the starting positions of the outer iteration
__global__ void k_cg_probe(IntArgs A, double* __restrict__ pos, const long long* __restrict__ ffix, double theta,
                           double* __restrict__ cf, double* __restrict__ x0, DevState* st) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  double x[3];
  for (int d = 0; d < 3; ++d) {
    const double f = fix_force(ffix, A, i, d);
    cf[3 * i + d] = f;
    if (x0) x0[3 * i + d] = pos[3 * i + d];
    x[d] = pos[3 * i + d] + theta * f;
  }
  wrap_into(A, x, pos + 3 * i, st);
}

// per-block partial sums (fixed order inside the block) of sum cf^2, sum omega.cf, sum omega^2
// with omega = -(F(x + theta t) - cf)/theta (:54); with cf == NULL, sum F^2 only
__global__ void k_cg_partials(IntArgs A, const long long* __restrict__ ffix, const double* __restrict__ cf, double theta,
                              double* __restrict__ part) {
  __shared__ double sm[3][MIN_BLOCK];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double s1 = 0, s2 = 0, s3 = 0;
  if (i < A.n)
    for (int d = 0; d < 3; ++d) {
      const double f = fix_force(ffix, A, i, d);
      if (cf) {
        const double c = cf[3 * i + d], w = -(f - c) / theta;
        s1 += c * c;
        s2 += w * c;
        s3 += w * w;
      } else {
        s1 += f * f;
      }
    }
  sm[0][threadIdx.x] = s1;
  sm[1][threadIdx.x] = s2;
  sm[2][threadIdx.x] = s3;
  __syncthreads();
  for (int h = MIN_BLOCK / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
      for (int k = 0; k < 3; ++k) sm[k][threadIdx.x] += sm[k][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x < 3) part[3 * blockIdx.x + threadIdx.x] = sm[threadIdx.x][0];
}

// sum of the block partials in block order (deterministic) into out[0..2]
__global__ void k_cg_total(const double* __restrict__ part, int nblocks, double* __restrict__ out) {
  if (threadIdx.x >= 3) return;
  double s = 0;
  for (int b = 0; b < nblocks; ++b) s += part[3 * b + threadIdx.x];
  out[threadIdx.x] = s;
}

// x <- wrap(x0 + alpha cf) (the last sub-iteration's d = alpha t, :66-69)
__global__ void k_cg_place(IntArgs A, double* __restrict__ pos, const double* __restrict__ x0,
                           const double* __restrict__ cf, double alpha, DevState* st) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  double x[3];
  for (int d = 0; d < 3; ++d) x[d] = x0[3 * i + d] + alpha * cf[3 * i + d];
  wrap_into(A, x, pos + 3 * i, st);
}

// one fused evaluation with energies; returns the potential energy (sum of the term slots) and
// fails on a particle-loss or non-finite flag
static int min_evaluate(mdx_ctx* c, unsigned terms, double* h_epot) {
  int rc = mdx_enqueue_zero_accumulators(c);
  if (!rc) rc = mdx_enqueue_forces(c, terms, true);
  if (rc) return rc;
  c->evals++;
  long long e[MDX_NTERMS + 1];
  unsigned flags = 0;
  MDX_HIP(hipMemcpyAsync(e, c->efix.p, sizeof(e), hipMemcpyDeviceToHost, c->stream));
  MDX_HIP(hipMemcpyAsync(&flags, &c->dstate.p->flags, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
  MDX_HIP(hipStreamSynchronize(c->stream));
  MDX_CHECK(!(flags & MDX_FLAG_LOSS), MDX_ERR_PARTICLE_LOSS, "minimisation moved atom(s) beyond 2 PBC images");
  double s = 0;
  for (int t = 0; t < MDX_NTERMS; ++t) s += (double)e[t] * kFixInv;
  MDX_CHECK(std::isfinite(s), MDX_ERR_NONFINITE, "minimisation reached a non-finite potential energy");
  *h_epot = s;
  return 0;
}

static int min_prepare(mdx_ctx* c) {
  MDX_CHECK(c && c->n > 0 && c->pos, MDX_ERR_INVALID, "minimisation needs a bound system");
  c->forces_current = false;
  int rc;
  if ((rc = c->xa.ensure(3 * (size_t)c->n)) || (rc = c->xb.ensure(3 * (size_t)c->n))) return rc;
  if ((rc = c->tmp64.ensure(3 * (size_t)div_up(c->n, MIN_BLOCK) + 4))) return rc;
  c->int_primed = false;  // the reference-compatible integrators keep history in xa / xb
  return 0;
}

extern "C" int mdx_minimize_sd(mdx_ctx* c, unsigned terms, double alpha, double energy_tol, int max_iter,
                               int* h_iters, double* h_energy) {
  MDX_CHECK(c && alpha > 0 && max_iter >= 0, MDX_ERR_INVALID, "mdx_minimize_sd: bad arguments");
  ENTER();
  int rc = min_prepare(c);
  double cur = 0, pre = 0;
  int it = 0;
  if (!rc) rc = min_evaluate(c, terms, &cur);
  pre = cur;
  const dim3 g(div_up(c->n, MIN_BLOCK)), b(MIN_BLOCK);
  while (!rc && it < max_iter) {
    IntArgs A = int_args(c);
    k_sd_move<<<g, b, 0, c->stream>>>(A, c->pos, c->ffix.p, alpha, c->dstate.p);
    if (hipGetLastError() != hipSuccess) { rc = MDX_ERR_HIP; break; }
    if ((rc = min_evaluate(c, terms, &cur))) break;
    ++it;
    const double err = std::fabs((cur - pre) / pre);
    if (err < energy_tol) break;
    pre = cur;
  }
  if (h_iters) *h_iters = it;
  if (h_energy) *h_energy = cur;
  LEAVE_RET(rc);
}

extern "C" int mdx_minimize_cg(mdx_ctx* c, unsigned terms, double theta, double force_tol, int max_sub_iter,
                               double energy_tol, int max_iter, int* h_iters, double* h_energy) {
  MDX_CHECK(c && theta > 0 && max_sub_iter >= 1 && max_iter >= 0, MDX_ERR_INVALID, "mdx_minimize_cg: bad arguments");
  ENTER();
  int rc = min_prepare(c);
  double cur = 0;
  int it = 0;
  if (!rc) rc = min_evaluate(c, terms, &cur);
  const int nbk = div_up(c->n, MIN_BLOCK);
  const dim3 g(nbk), b(MIN_BLOCK);
  double* part = (double*)c->tmp64.p;
  double* tot = part + 3 * (size_t)nbk;
  double* x0 = c->xa.p;
  double* cf = c->xb.p;
  auto sums = [&](const double* cfp, double* h) -> int {
    IntArgs A = int_args(c);
    k_cg_partials<<<g, b, 0, c->stream>>>(A, c->ffix.p, cfp, theta, part);
    k_cg_total<<<1, 64, 0, c->stream>>>(part, nbk, tot);
    MDX_LAUNCH_CHECK();
    MDX_HIP(hipMemcpyAsync(h, tot, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    MDX_HIP(hipStreamSynchronize(c->stream));
    return 0;
  };
  while (!rc && it < max_iter) {
    const double pre = cur;
    double s0[3];
    if ((rc = sums(nullptr, s0))) break;  // sum f0^2
    double alpha = 0;
    for (int sub = 0; sub < max_sub_iter && !rc; ++sub) {
      IntArgs A = int_args(c);
      k_cg_probe<<<g, b, 0, c->stream>>>(A, c->pos, c->ffix.p, theta, cf, sub == 0 ? x0 : nullptr, c->dstate.p);
      if (hipGetLastError() != hipSuccess) { rc = MDX_ERR_HIP; break; }
      double e_probe = 0;
      if ((rc = min_evaluate(c, terms, &e_probe))) break;
      double s[3];
      if ((rc = sums(cf, s))) break;
      alpha = s[0] / s[1];  // (cur_f . cur_f) / (omega . t) with t = cur_f
      // |cur_f + alpha omega|^2 / |f0|^2 < force_tol ends the sub-iterations (:57-60)
      const double fn2 = s[0] + 2 * alpha * s[1] + alpha * alpha * s[2];
      if (fn2 / s0[0] < force_tol) break;
    }
    if (rc) break;
    IntArgs A = int_args(c);
    k_cg_place<<<g, b, 0, c->stream>>>(A, c->pos, x0, cf, alpha, c->dstate.p);
    if (hipGetLastError() != hipSuccess) { rc = MDX_ERR_HIP; break; }
    if ((rc = min_evaluate(c, terms, &cur))) break;
    ++it;
    if (std::fabs((cur - pre) / pre) < energy_tol) break;
  }
  if (h_iters) *h_iters = it;
  if (h_energy) *h_energy = cur;
  LEAVE_RET(rc);
}

extern "C" int mdx_energy_history_copy_async(mdx_ctx* c, double* h_out, int h_cap, int* h_count) {
  MDX_CHECK(c && h_out && h_count, MDX_ERR_INVALID, "mdx_energy_history_copy_async: bad arguments");
  ENTER();
  const int cnt = std::min(h_cap, c->ehist_count);
  const int ne = MDX_NTERMS + 1;
  if (cnt > 0)
    MDX_HIP(hipMemcpyAsync(h_out, c->ehist.p, (size_t)cnt * ne * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  *h_count = cnt;
  LEAVE_RET(0);
}

extern "C" int mdx_energy_history_fetch(mdx_ctx* c, double* h_out, int h_cap, int* h_count) {
  MDX_CHECK(c && h_out && h_count, MDX_ERR_INVALID, "mdx_energy_history_fetch: bad arguments");
  int cnt = std::min(h_cap, c->ehist_count);
  const int ne = MDX_NTERMS + 1;
  if (cnt > 0)
    MDX_HIP(hipMemcpyAsync(h_out, c->ehist.p, (size_t)cnt * ne * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  MDX_HIP(hipStreamSynchronize(c->stream));
  *h_count = cnt;
  return 0;
}

__global__ void k_count_pairs(int n, int cap, const int* __restrict__ nbr_count, const unsigned* __restrict__ nbr_list,
                              const float4* __restrict__ xq, float Lx, float Ly, float Lz, float rc2,
                              unsigned long long* __restrict__ out) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long slots = 0, inside = 0;
  if (s < n) {
    int cnt = nbr_count[s];
    slots = cnt;
    float4 xi = xq[s];
    for (int k = 0; k < cnt; ++k) {
      int t = (int)(nbr_list[(long long)s * cap + k] & 0x7fffffffu);
      float dx = xq[t].x - xi.x, dy = xq[t].y - xi.y, dz = xq[t].z - xi.z;
      dx -= Lx * rintf(dx / Lx);
      dy -= Ly * rintf(dy / Ly);
      dz -= Lz * rintf(dz / Lz);
      if (dx * dx + dy * dy + dz * dz <= rc2) inside++;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    slots += __shfl_xor(slots, off, 64);
    inside += __shfl_xor(inside, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&out[0], slots);
    atomicAdd(&out[1], inside);
  }
}

extern "C" int mdx_pair_counts(mdx_ctx* c, int64_t* h_list_slots, int64_t* h_unique_in_cutoff) {
  MDX_CHECK(c && c->nb_configured, MDX_ERR_INVALID, "mdx_pair_counts: nonbonded not configured");
  int rc;
  if ((rc = c->tmp64.ensure(2))) return rc;
  MDX_HIP(hipMemsetAsync(c->tmp64.p, 0, 2 * sizeof(long long), c->stream));
  k_count_pairs<<<div_up(c->n, 256), 256, 0, c->stream>>>(c->n, c->nbr_cap, c->nbr_count.p, c->nbr_list.p, c->xq.p,
                                                         (float)c->box[0], (float)c->box[1], (float)c->box[2],
                                                         (float)(c->rc * c->rc), (unsigned long long*)c->tmp64.p);
  MDX_LAUNCH_CHECK();
  long long h[2];
  MDX_HIP(hipMemcpyAsync(h, c->tmp64.p, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  MDX_HIP(hipStreamSynchronize(c->stream));
  *h_list_slots = h[0];
  *h_unique_in_cutoff = h[1] / 2;
  return 0;
}

// ------------------------------------------------------------------ kernel timing
// Average duration (ms) of one launch of a hot kernel, timed with HIP events on the context's
// own stream (the stream the kernel runs on). which: 0 nonbonded (LJ + erfc Coulomb, forces
// only, on the current list), 1 full force evaluation of `terms`, 2 PME reciprocal chain,
// 3 one integrator step of `terms`.
extern "C" int mdx_nonbonded_tune(mdx_ctx* c, unsigned terms, int lanes, int reps, int* h_lanes) {
  MDX_CHECK(c && c->nb_configured, MDX_ERR_INVALID, "mdx_nonbonded_tune: nonbonded not configured");
  MDX_CHECK(lanes == 0 || lanes == 16 || lanes == 32 || lanes == 64, MDX_ERR_INVALID, "lanes must be 0, 16, 32 or 64");
  int best = lanes;
  if (lanes == 0) {
    double best_ms = 1e300;
    const int cand[3] = {16, 32, 64};
    for (int k = 0; k < 3; ++k) {
      c->nb_lanes = cand[k];
      double ms = 0;
      int rc = mdx_time_kernel(c, 0, terms, reps, &ms);
      if (rc) return rc;
      if (ms < best_ms) {
        best_ms = ms;
        best = cand[k];
      }
    }
  }
  // the timing runs ended with the last candidate: captured steps hold the launch shape
  // they were captured with, so they are dropped only when that differs from the choice
  if (best != c->graph_lanes) mdx_invalidate_graph(c);
  c->nb_lanes = best;
  if (h_lanes) *h_lanes = best;
  return 0;
}

extern "C" int mdx_nonbonded_rebuild_every(mdx_ctx* c, int every) {
  MDX_CHECK(c && every >= 0, MDX_ERR_INVALID, "mdx_nonbonded_rebuild_every: bad arguments");
  if (every != c->rebuild_every) {
    c->rebuild_every = every;
    mdx_invalidate_graph(c);  // captured prologues hold the interval
  }
  return 0;
}

extern "C" int mdx_nonbonded_cells(mdx_ctx* c, int per_radius) {
  MDX_CHECK(c && per_radius >= 1 && per_radius <= 3, MDX_ERR_INVALID,
            "mdx_nonbonded_cells: cells per list radius must be 1, 2 or 3");
  if (per_radius == c->cells_per_rl) return 0;
  c->cells_per_rl = per_radius;
  mdx_invalidate_graph(c);  // captured launches hold the grid's size
  // a new grid: atoms are binned again and the list rebuilt by the next evaluation
  return c->nb_configured ? mdx_nlist_prepare(c) : 0;
}

extern "C" int mdx_nonbonded_persist(mdx_ctx* c, int per_cu) {
  MDX_CHECK(c && per_cu >= 0, MDX_ERR_INVALID, "mdx_nonbonded_persist: bad arguments");
  c->nb_persist = c->ncu > 0 ? per_cu : 0;
  mdx_invalidate_graph(c);
  return 0;
}

// The step prologue (G-JF or velocity Verlet, folding the previous velocity half) timed alone:
// the state it writes is saved first and put back after, no rebuild is decided and nothing is
// recorded, so the trajectory continues as if it had not run.
// The device state a timed prologue launch overwrites (saved before the timing, restored after)
struct StatePart {
  void* p;
  size_t bytes;
};
static const int kProStateParts = 9;
static int prologue_state_parts(mdx_ctx* c, StatePart* parts) {
  MDX_CHECK(c->int_configured && c->acc_prev.p && c->gsave.p && c->xq.p && c->pos_sorted.p &&
                (c->int_kind == MDX_INT_LANGEVIN || c->int_kind == MDX_INT_VELOCITY_VERLET),
            MDX_ERR_INVALID, "prologue timing needs a G-JF or velocity-Verlet context with a list grid");
  const size_t n3 = 3 * (size_t)c->n;
  const StatePart q[kProStateParts] = {
      {c->pos, n3 * 8},        {c->vel, n3 * 8},  {c->acc_prev.p, n3 * 8},
      {c->gsave.p, n3 * 8},    {c->ffix.p, n3 * 8}, {c->efix.p, (MDX_NTERMS + 1) * 8},
      {c->dstate.p, sizeof(DevState)}, {c->xq.p, (size_t)c->n * sizeof(float4)}, {c->pos_sorted.p, n3 * 8}};
  for (int k = 0; k < kProStateParts; ++k) parts[k] = q[k];
  return 0;
}
static size_t state_bytes(const StatePart* parts) {
  size_t total = 0;
  for (int k = 0; k < kProStateParts; ++k) total += (parts[k].bytes + 255) & ~(size_t)255;
  return total;
}
// copy the parts into `save` (or back from it) on stream s
static int state_copy(const StatePart* parts, char* save, bool restore, hipStream_t s) {
  size_t off = 0;
  for (int k = 0; k < kProStateParts; ++k) {
    if (restore) MDX_HIP(hipMemcpyAsync(parts[k].p, save + off, parts[k].bytes, hipMemcpyDeviceToDevice, s));
    else MDX_HIP(hipMemcpyAsync(save + off, parts[k].p, parts[k].bytes, hipMemcpyDeviceToDevice, s));
    off += (parts[k].bytes + 255) & ~(size_t)255;
  }
  return 0;
}

static int time_prologue(mdx_ctx* c, int reps, double* h_avg_ms) {
  StatePart parts[kProStateParts];
  int rc0 = prologue_state_parts(c, parts);
  if (rc0) return rc0;
  const size_t total = state_bytes(parts);
  struct Save {  // freed on every return (hipFree waits for the device)
    DevBuf<char> b;
    ~Save() { b.release(); }
  } keep;
  DevBuf<char>& save = keep.b;
  int rc = save.ensure(total);
  if (rc || (rc = state_copy(parts, save.p, false, c->stream))) return rc;
  ProCall pc = pro_call(c);
  pc.P.bin = 0;
  pc.P.record = 0;
  pc.P.thr2 = HUGE_VAL;
  pc.P.rebuild_every = 0;
  const bool gjf = c->int_kind == MDX_INT_LANGEVIN;
  dim3 gp(pc.blocks), bp(PRO_THREADS);
  hipEvent_t e0, e1;
  MDX_HIP(hipEventCreate(&e0));
  MDX_HIP(hipEventCreate(&e1));
  MDX_HIP(hipEventRecord(e0, c->stream));
  for (int k = 0; k < reps; ++k) {
    if (gjf)
      k_step_pro<true, true><<<gp, bp, 0, c->stream>>>(pc);
    else
      k_step_pro<false, true><<<gp, bp, 0, c->stream>>>(pc);
  }
  MDX_LAUNCH_CHECK();
  MDX_HIP(hipEventRecord(e1, c->stream));
  if ((rc = state_copy(parts, save.p, true, c->stream))) return rc;
  MDX_HIP(hipStreamSynchronize(c->stream));
  float ms = 0;
  MDX_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *h_avg_ms = ms / reps;
  return 0;
}

// The neighbour-list build of a rebuild step timed alone: each repetition requests a rebuild and
// builds from the current cell order; the next evaluation rebuilds in full (binning included).
static int time_list_build(mdx_ctx* c, int reps, double* h_avg_ms) {
  MDX_CHECK(c->nb_configured && c->list_valid, MDX_ERR_INVALID, "list-build timing needs a built neighbour list");
  hipEvent_t e0, e1;
  MDX_HIP(hipEventCreate(&e0));
  MDX_HIP(hipEventCreate(&e1));
  double total = 0;
  int rc = 0;
  for (int k = 0; k < reps && !rc; ++k) {
    if ((rc = mdx_enqueue_request_rebuild(c))) break;
    MDX_HIP(hipEventRecord(e0, c->stream));
    if ((rc = mdx_nlist_update_list(c))) break;
    MDX_HIP(hipEventRecord(e1, c->stream));
    MDX_HIP(hipEventSynchronize(e1));
    float ms = 0;
    MDX_HIP(hipEventElapsedTime(&ms, e0, e1));
    total += ms;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (!rc) rc = mdx_enqueue_request_rebuild(c);
  *h_avg_ms = total / reps;
  return rc;
}

extern "C" int mdx_time_kernel(mdx_ctx* c, int which, unsigned terms, int reps, double* h_avg_ms) {
  MDX_CHECK(c && h_avg_ms && reps > 0, MDX_ERR_INVALID, "mdx_time_kernel: bad arguments");
  MDX_ENTER();
  c->forces_current = false;  // the timed launches accumulate into the force buffers
  int rc = 0;
  if (which == 4) MDX_LEAVE_RET(time_prologue(c, reps, h_avg_ms));
  if (which == 5) MDX_LEAVE_RET(time_list_build(c, reps, h_avg_ms));
  bool lj = (terms & MDX_TERM_LJ) != 0, cd = (terms & MDX_TERM_COUL_DIRECT) != 0;
  if (which == 0) {  // make sure the list is current
    MDX_CHECK(c->nb_configured && (lj || cd), MDX_ERR_INVALID, "nonbonded timing needs LJ or direct Coulomb");
    if ((rc = mdx_nlist_update(c, false)) || (rc = mdx_enqueue_nonbonded(c, lj, cd, true))) MDX_LEAVE_RET(rc);
  }
  hipEvent_t e0, e1;
  MDX_HIP(hipEventCreate(&e0));
  MDX_HIP(hipEventCreate(&e1));
  MDX_HIP(hipEventRecord(e0, c->stream));
  for (int k = 0; k < reps && !rc; ++k) {
    if (which == 0) rc = mdx_enqueue_nonbonded(c, lj, cd, false);
    else if (which == 1) {
      rc = mdx_enqueue_zero_accumulators(c);
      if (!rc) rc = mdx_enqueue_forces(c, terms, true);
    } else if (which == 2) rc = mdx_pme_enqueue(c, true);
    else if (which == 3) rc = mdx_integrator_enqueue_step(c, terms, false, false, true);
    else if (which >= 6 && which <= 8) rc = mdx_pme_enqueue_part(c, which - 6);
    else rc = MDX_ERR_INVALID;
  }
  MDX_HIP(hipEventRecord(e1, c->stream));
  MDX_HIP(hipEventSynchronize(e1));
  float ms = 0;
  MDX_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *h_avg_ms = ms / reps;
  MDX_LEAVE_RET(rc);
}

// One batched kernel kind of MDX_BATCH_FUSED timed alone over every replica (HIP events on the
// lead stream): which 0 nonbonded (forces only), 4 step prologue (state saved and restored),
// 5 list build of a rebuild step, 6 / 7 / 8 PME plane forward / x-axis pass / inverse + gather.
// Uses the tables of the last fused call; the replicas' forces are re-evaluated by the next call.
extern "C" int mdx_batch_time_kernel(mdx_batch* b, int which, int reps, double* h_avg_ms) {
  MDX_CHECK(b && h_avg_ms && reps > 0, MDX_ERR_INVALID, "mdx_batch_time_kernel: bad arguments");
  MDX_CHECK(b->fz.ready, MDX_ERR_INVALID, "mdx_batch_time_kernel: run a MDX_BATCH_FUSED integrate call first");
  MDX_CHECK(which == 0 || (which >= 4 && which <= 8), MDX_ERR_INVALID, "mdx_batch_time_kernel: unknown kernel %d",
            which);
  MDX_HIP(hipSetDevice(b->ctx[0]->device));
  BatchFused& f = b->fz;
  hipStream_t s = b->lead;
  MDX_HIP(hipEventRecord(b->ev_in, b->ext));
  MDX_HIP(hipStreamWaitEvent(s, b->ev_in, 0));
  for (mdx_ctx* c : b->ctx) c->forces_current = false;  // the timed launches accumulate into the forces
  struct Keep {  // freed on every return (hipFree waits for the device)
    DevBuf<char> save;
    BatchKind pro;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Keep() {
      save.release();
      pro.buf.release();
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
    }
  } k;
  MDX_HIP(hipEventCreate(&k.e0));
  MDX_HIP(hipEventCreate(&k.e1));
  int rc = 0;
  double total_ms = 0;
  std::vector<StatePart> parts;
  if (which == 4) {  // the prologue alone: no binning, no energy record, no rebuild decision
    const int n = (int)b->ctx.size();
    parts.resize((size_t)n * kProStateParts);
    size_t bytes = 0;
    std::vector<ProCall> calls;
    std::vector<int> blocks;
    for (int r = 0; r < n && !rc; ++r) {
      mdx_ctx* c = b->ctx[r];
      if ((rc = prologue_state_parts(c, &parts[(size_t)r * kProStateParts]))) break;
      bytes += state_bytes(&parts[(size_t)r * kProStateParts]);
      ProCall pc = pro_call(c);
      pc.P.bin = 0;
      pc.P.record = 0;
      pc.P.thr2 = HUGE_VAL;
      pc.P.rebuild_every = 0;
      calls.push_back(pc);
      blocks.push_back(pc.blocks);
    }
    if (rc || (rc = k.save.ensure(bytes)) || (rc = batch_kind_upload(&k.pro, calls, blocks, 0))) return rc;
    size_t off = 0;
    for (int r = 0; r < n && !rc; ++r) {
      rc = state_copy(&parts[(size_t)r * kProStateParts], k.save.p + off, false, s);
      off += state_bytes(&parts[(size_t)r * kProStateParts]);
    }
    if (rc) return rc;
  }
  if (which != 5) MDX_HIP(hipEventRecord(k.e0, s));
  for (int rep = 0; rep < reps && !rc; ++rep) {
    if (which == 0) {
      rc = mdx_batch_launch_nb(f.nb, f.lanes, f.s14, false, s);
    } else if (which == 4) {
      const ProCall* pc = (const ProCall*)k.pro.calls;
      if (f.gjf) k_step_pro_batch<true, true><<<k.pro.blocks, PRO_THREADS, 0, s>>>(k.pro.map, pc);
      else k_step_pro_batch<false, true><<<k.pro.blocks, PRO_THREADS, 0, s>>>(k.pro.map, pc);
      MDX_LAUNCH_CHECK();
    } else if (which == 5) {  // each repetition rebuilds every replica's list from its cell order
      for (mdx_ctx* c : b->ctx)
        if ((rc = on_lead(b, c, [&]() { return mdx_enqueue_request_rebuild(c); }))) break;
      if (rc) break;
      MDX_HIP(hipEventRecord(k.e0, s));
      if ((rc = mdx_batch_launch_list(f.lfast, f.lcell, s))) break;
      MDX_HIP(hipEventRecord(k.e1, s));
      MDX_HIP(hipEventSynchronize(k.e1));
      float ms = 0;
      MDX_HIP(hipEventElapsedTime(&ms, k.e0, k.e1));
      total_ms += ms;
    } else {
      rc = mdx_batch_launch_pme(f.pfwd, f.pfft, f.pgat, s, 1u << (which - 6));
    }
  }
  if (rc) return rc;
  if (which != 5) {
    MDX_HIP(hipEventRecord(k.e1, s));
    MDX_HIP(hipEventSynchronize(k.e1));
    float ms = 0;
    MDX_HIP(hipEventElapsedTime(&ms, k.e0, k.e1));
    total_ms = ms;
  }
  if (which == 4) {
    size_t off = 0;
    for (size_t r = 0; r < b->ctx.size() && !rc; ++r) {
      rc = state_copy(&parts[r * kProStateParts], k.save.p + off, true, s);
      off += state_bytes(&parts[r * kProStateParts]);
    }
    if (rc) return rc;
  }
  if (which == 5)  // the next evaluation rebuilds in full (binning included)
    for (mdx_ctx* c : b->ctx)
      if ((rc = on_lead(b, c, [&]() { return mdx_enqueue_request_rebuild(c); }))) return rc;
  MDX_HIP(hipStreamSynchronize(s));
  MDX_HIP(hipEventRecord(b->ev_out[0], s));
  MDX_HIP(hipStreamWaitEvent(b->ext, b->ev_out[0], 0));
  *h_avg_ms = total_ms / reps;
  return 0;
}
