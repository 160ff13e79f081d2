// Internal declarations shared by the libmdx translation units (gfx950 / CDNA4 only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include <cstring>
#include "../../include/mdx.h"

// ------------------------------------------------------------------ fixed point
// Forces and energies are accumulated as int64 with 40 fractional bits: integer addition is
// associative, so every sum is bitwise reproducible whatever the order of the atomics.
// Range +-8.4e6 internal units (forces ~1e-2, energies ~1e0 in practice), resolution 9.1e-13.
#define MDX_FIX_SHIFT 40
static constexpr double kFixScale = 1099511627776.0;          // 2^40
static constexpr double kFixInv = 1.0 / 1099511627776.0;

__device__ __forceinline__ long long to_fix(double v) { return __double2ll_rn(v * kFixScale); }
__device__ __forceinline__ long long to_fix_f(float v) { return __double2ll_rn((double)v * kFixScale); }

// Orders one wave's LDS writes before its later LDS reads (and the reads before later writes):
// a wave's LDS operations complete in program order, so only compiler motion must be stopped.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// status flag bits (device word)
#define MDX_FLAG_OVERFLOW 1u
#define MDX_FLAG_LOSS 2u
#define MDX_FLAG_NONFINITE 4u
#define MDX_FLAG_PME_COVERAGE 8u   // reciprocal-space gather missed an atom (internal invariant)
#define MDX_FLAG_SYNC 16u          // a grid barrier gave up waiting (blocks not co-resident)

// ------------------------------------------------------------------ error handling
void mdx_set_error(const char* fmt, ...);
#define MDX_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t _e = (call);                                                            \
    if (_e != hipSuccess) {                                                            \
      mdx_set_error("%s failed at %s:%d: %s", #call, __FILE__, __LINE__,               \
                    hipGetErrorString(_e));                                            \
      return MDX_ERR_HIP;                                                              \
    }                                                                                  \
  } while (0)
#define MDX_LAUNCH_CHECK()                                                             \
  do {                                                                                 \
    hipError_t _e = hipGetLastError();                                                 \
    if (_e != hipSuccess) {                                                            \
      mdx_set_error("kernel launch failed at %s:%d: %s", __FILE__, __LINE__,           \
                    hipGetErrorString(_e));                                            \
      return MDX_ERR_HIP;                                                              \
    }                                                                                  \
  } while (0)
#define MDX_CHECK(cond, code, ...)                                                     \
  do {                                                                                 \
    if (!(cond)) {                                                                     \
      mdx_set_error(__VA_ARGS__);                                                      \
      return code;                                                                     \
    }                                                                                  \
  } while (0)

// ------------------------------------------------------------------ device buffers
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  int ensure(size_t count) {  // grow-only
    if (count <= n && p) return 0;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (count == 0) return 0;
    if (hipMalloc((void**)&p, count * sizeof(T)) != hipSuccess) {
      mdx_set_error("hipMalloc of %zu bytes failed", count * sizeof(T));
      return MDX_ERR_HIP;
    }
    n = count;
    return 0;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

// Per-context counters and flags kept in one small device block (zeroed at creation).
#define MDX_NGRAPH_SIZES 6  // forces-only step graphs of 32, 16, 8, 4, 2 and 1 steps

struct DevState {
  unsigned flags;           // MDX_FLAG_*
  int rebuild;              // 1 when the neighbour list must be rebuilt before use
  unsigned long long max_disp2_bits;  // max squared displacement since build (double bits)
  long long step;           // integrator step counter (drives the Philox stream)
  long long nrebuild;       // number of list rebuilds
  long long maxcount;       // largest neighbour count of the last build
  unsigned long long minpos[3];  // order-preserving encoded float64 minima (reference binning)
  int maxpc;                // largest cell population (reference binning)
  int pad;
  long long hist_base;      // step counter value at the start of the current integrate call
  int pme_slots;            // reciprocal force slots written in the current evaluation
  unsigned pme_ticket;      // finished blocks of the PME gather kernel
  unsigned step_ticket;     // finished blocks of the fused step prologue
  unsigned bar_count;       // grid barrier of the fused PME kernel: blocks arrived
  unsigned bar_gen;         // and its generation (advanced by the last arrival)
  unsigned call_ticket;     // finished blocks of an integrate call's opening / closing kernel
  unsigned pad2;
  unsigned long long ke_acc;  // kinetic-energy partial sums of the opening kernel (0 between calls)
};

// neighbour-list entries allocated past the last row (at least the longest read-ahead: one
// chunk of 64 lanes x 4 entries)
#define MDX_LIST_SLACK 512

// ------------------------------------------------------------------ batched launches
// One launch of a kernel covers the workgroups of many contexts (replicas): workgroup b belongs
// to replica rep[b] and is that replica's workgroup b - base[rep[b]]; the replica's arguments
// are a record in device memory (the same record a single-context launch passes by value).
struct BatchMap {
  const int* rep;   // replica of each workgroup of the launch
  const int* base;  // first workgroup of each replica
};
__device__ __forceinline__ int batch_replica(const BatchMap& m, int* local) {
  const int r = m.rep[blockIdx.x];
  *local = (int)blockIdx.x - m.base[r];
  return r;
}

// Host side of one batched kernel: the per-replica argument records, the workgroup map and the
// dynamic LDS of the launch (the largest replica's), in one device allocation.
struct BatchKind {
  DevBuf<char> buf;
  BatchMap map{};
  const void* calls = nullptr;
  int blocks = 0;  // workgroups of the launch, all replicas
  size_t lds = 0;
  int threads = 0;  // lanes per workgroup, where the kind chooses it
};
template <class Call>
int batch_kind_upload(BatchKind* k, const std::vector<Call>& calls, const std::vector<int>& blocks, size_t lds) {
  const size_t n = calls.size();
  std::vector<int> base(n);
  int total = 0;
  for (size_t r = 0; r < n; ++r) {
    base[r] = total;
    total += blocks[r];
  }
  std::vector<int> rep((size_t)total);
  for (size_t r = 0; r < n; ++r)
    for (int b = 0; b < blocks[r]; ++b) rep[(size_t)base[r] + b] = (int)r;
  const size_t cbytes = (n * sizeof(Call) + 255) & ~(size_t)255;
  const size_t bytes = cbytes + (n + (size_t)total) * sizeof(int);
  int rc = k->buf.ensure(bytes);
  if (rc) return rc;
  std::vector<char> h(bytes);
  memcpy(h.data(), calls.data(), n * sizeof(Call));
  memcpy(h.data() + cbytes, base.data(), n * sizeof(int));
  if (total) memcpy(h.data() + cbytes + n * sizeof(int), rep.data(), (size_t)total * sizeof(int));
  MDX_HIP(hipMemcpy(k->buf.p, h.data(), bytes, hipMemcpyHostToDevice));
  k->calls = k->buf.p;
  k->map.base = (const int*)(k->buf.p + cbytes);
  k->map.rep = k->map.base + n;
  k->blocks = total;
  k->lds = lds;
  return 0;
}

// floor(u) clamped to [0, n) for a cell or grid coordinate u >= 0. The range is decided in
// floating point before the conversion: a float-to-int conversion of NaN, or of a value outside
// int's range, has no defined result (LLVM treats it as poison, which a later integer clamp does
// not repair), so a NaN or runaway coordinate must land in a valid cell, never at an arbitrary
// index. NaN and negative u give 0, u >= n gives n - 1.
__device__ __forceinline__ int clamped_floor_index(double u, int n) {
  return u >= 1.0 ? (u < (double)n ? (int)u : n - 1) : 0;
}

// float32 sorted-order copy of one atom: wrapped, centred position and q*sqrt(ke) (the layout
// the pair kernels read), plus its float64 position
__device__ __forceinline__ void pack_atom(double x, double y, double z, float q, double Lx, double Ly, double Lz,
                                          float4* __restrict__ xq_s, double* __restrict__ ps_s) {
  ps_s[0] = x;
  ps_s[1] = y;
  ps_s[2] = z;
  x -= Lx * floor(x / Lx);
  y -= Ly * floor(y / Ly);
  z -= Lz * floor(z / Lz);
  *xq_s = make_float4((float)(x - 0.5 * Lx), (float)(y - 0.5 * Ly), (float)(z - 0.5 * Lz), q);
}

struct mdx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool owns_stream = true;       // false for internal contexts that run on another context's stream
  hipStream_t side_stream = nullptr;  // concurrent branch (PME) of a force evaluation
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // ---- system (borrowed)
  int n = 0;
  double* pos = nullptr;
  double* vel = nullptr;
  double box[3] = {0, 0, 0};
  // ---- per-atom data (owned)
  DevBuf<double> inv_mass, mass, charge64;
  DevBuf<float> charge_s;         // q * sqrt(ke), float
  DevBuf<double> lj64;            // N x 4 (eps, sigma, eps14, sigma14)
  DevBuf<int> excl_ptr, excl_idx, s14_ptr, s14_idx;
  int max_excl = 0, max_s14 = 0;
  bool has_excl = false, has_s14 = false;
  std::vector<int> h_excl_ptr, h_excl_idx;
  double qsum = 0, q2sum = 0;
  bool has_charges = false, has_lj = false, has_mass = false;
  // ---- bonded (owned)
  int nb = 0, na = 0, nd = 0, ni = 0;
  DevBuf<int> bond_idx, angle_idx, dih_idx, imp_idx;
  DevBuf<double> bond_par, angle_par, dih_par, imp_par;
  int dih_refcompat = 0;
  // ---- fixed-point accumulators
  DevBuf<long long> ffix;        // 3 x N (SoA)
  DevBuf<long long> efix;        // MDX_NTERMS + 1
  DevBuf<DevState> dstate;
  // ---- reference cell list
  DevBuf<int> ref_cell_lin;      // N (linear cell id of the last reference binning)
  DevBuf<int> ref_counts, ref_cursor;
  int ref_n[3] = {0, 0, 0};
  double ref_w[3] = {0, 0, 0};
  double ref_inv[3] = {0, 0, 0};
  // ---- nonbonded configuration
  bool nb_configured = false;
  double rc = 12.0, skin = 0.0, cell_cutoff = 12.0;
  int pair_mode = MDX_PAIRS_EXACT;
  // list grid
  int gn[3] = {0, 0, 0};
  int reach[3] = {0, 0, 0};
  int ncell = 0;
  DevBuf<int> cell_count, cell_start, cell_cursor, atom_cell;
  DevBuf<int> sorted_orig;       // sorted slot -> original atom
  DevBuf<int> orig_sorted;       // original atom -> sorted slot
  DevBuf<float4> xq;             // sorted: x, y, z (centred, float), q*sqrt(ke)
  DevBuf<float2> ljs, ljs14;     // sorted: sigma/2, sqrt(eps)
  DevBuf<float4> ctab;           // Ewald correction G(r^2) as cubic pieces (MDX_CTAB_N of them)
  double ctab_beta = 0, ctab_rc = 0;
  float ctab_scale = 0.f;        // pieces per A^2
  DevBuf<int> ref_cell_sorted;   // sorted: packed reference cell coords (REFERENCE mode)
  DevBuf<double> pos_at_build;   // original order, 3N
  DevBuf<double> pos_sorted;     // sorted order, 3N, refreshed with xq every evaluation
  DevBuf<int> nbr_count;         // sorted
  DevBuf<unsigned> nbr_list;     // column major: [k * stride + s]
  int nbr_cap = 0;               // capacity per atom
  int nb_lanes = 32;             // lanes per atom in the nonbonded kernel (16, 32 or 64)
  // nonbonded workgroups per compute unit looping over the work (0: one per unit of work); 3 (12
  // waves of 123 VGPRs) leaves a quarter of every SIMD's registers to the PME kernels beside it
  int nb_persist = 3;
  int rebuild_every = 0;  // scheduled list rebuilds (step counter multiple), 0 = displacement test only
  int cells_per_rl = 2;   // list-grid cells per list radius r_c + skin along each axis (1..3)
  int ncu = 0;                   // compute units of the device
  int list_stride = 0;
  bool list_valid = false;
  bool prologue_done = false;    // the integrator's fused kernel already checked/zeroed/packed this step
  bool prologue_binned = false;  // ... and (on rebuild steps) counted, scanned and scattered the atoms
  // exclusion / 1-4 partners as 64-bit masks over the index window i-32 .. i+31 (list build);
  // win_bad[i] bit 0 / 1: atom i has an exclusion / 1-4 partner outside the window
  DevBuf<unsigned long long> ex_win, s4_win;
  DevBuf<unsigned char> win_bad;
  std::vector<unsigned long long> h_ex_win, h_s4_win;
  std::vector<unsigned char> h_win_bad;
  bool win_ready = false;
  std::vector<int> stencil;      // packed stencil offsets (dx,dy,dz) of the list grid
  DevBuf<int> d_stencil;
  // ---- PME
  bool pme_configured = false;
  int K[3] = {0, 0, 0};
  int order = 4;
  double beta = 0.0;
  DevBuf<long long> qfix;        // spread grid (fixed point)
  DevBuf<float2> cgrid;          // complex work grid
  DevBuf<float> conv;            // real convolved grid
  DevBuf<float> bc;              // influence function B*C (float)
  DevBuf<float2> twiddle;        // per axis, K entries each (3 * maxK)
  DevBuf<int> excl_pairs;        // unique excluded pairs (i<j), 2 x n_excl_pairs
  int n_excl_pairs = 0;
  DevBuf<double> fpart;          // plane path: reciprocal force per (kx, atom), 4 x 3N
  bool pme_planes = false;       // plane kernels usable (list grid present, plane fits LDS)
  int pme_wf = 0;                // candidate window of the forward plane kernel
  int pme_rec_cap = 0;           // gather records per plane
  DevBuf<int> pme_grec;          // gather records (6 words each), pme_rec_cap per plane
  DevBuf<int> pme_grec_n;        // records per plane
  int pme_fused_B = 0;           // x-axis lines per chunk of the one-launch PME kernel (0: not usable)
  size_t pme_fused_lds = 0;      // its dynamic LDS
  bool pme_fused_on = false;     // one-launch PME kernel selected (never inside a batch)
  // phase launches: the three PME reciprocal kernels each carry a share of the nonbonded
  // workgroups (and the last one the bonded terms), all on one stream (no fork and join)
  bool phases_on = false;
  double phase_frac[2] = {0.26, 0.13};  // nonbonded shares of the first two launches
  int pme_wf_phase = 0;                 // forward-plane candidate window under the phase kernels' LDS
  bool in_batch = false;         // advanced by an mdx_batch
  bool fpart_live = false;       // fpart belongs to the current evaluation
  // ---- integrator
  bool int_configured = false;
  int int_kind = MDX_INT_VELOCITY_VERLET;
  double dt = 1.0, gamma = 0.0, kT = 0.0;
  uint64_t seed = 0;
  bool int_primed = false;       // history arrays valid
  DevBuf<double> xa, xb, va, vb, acc_prev, acc_cur;   // integrator history (3N each)
  DevBuf<double> gsave;          // G-JF: xi_{n+1} of the last position half (3N), read by the velocity half
  const double* ext_noise = nullptr;
  int ext_noise_steps = 0;
  long long ext_noise_base = 0;
  DevBuf<double> ehist;          // per-step energies
  int ehist_cap = 0;
  int ehist_count = 0;
  long long evals = 0;
  // step graphs: a few fixed sizes of forces-only steps plus one energy step, captured once per
  // (terms, integrator kind) and replayed whatever the call's length (mdx_api.hip kGraphSizes)
  hipGraphExec_t graph_f[MDX_NGRAPH_SIZES] = {};
  hipGraphExec_t graph_e = nullptr;
  // a call's first step with its set-up (energy-history base, kinetic energy) when the last
  // call's forces are current, and its last step with the closing velocity half, energy record
  // and wrap: a call of n >= 2 steps is then only graph launches
  hipGraphExec_t graph_head = nullptr;
  hipGraphExec_t graph_tail = nullptr;
  unsigned graph_terms = 0;
  int graph_kind = -1;
  int graph_lanes = 0;           // nonbonded lanes per atom the graphs were captured with
  int energy_every = 1;          // potential energies every this many steps during integrate
  unsigned long long graph_gen = 0;  // bumped whenever captured graphs of this context go stale
  // the forces in the accumulators (and the potential energies in efix) belong to the current
  // positions: the last step of the last mdx_integrate call, with nothing changed since
  bool forces_current = false;
  unsigned forces_terms = 0;
  // scratch
  DevBuf<long long> tmp64;
  // learned surrogate (optional)
  mdx_mlp* mlp = nullptr;
};

// ------------------------------------------------------------------ shared helpers
__device__ __forceinline__ unsigned long long enc_double_order(double v) {
  unsigned long long b = (unsigned long long)__double_as_longlong(v);
  return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double dec_double_order(unsigned long long e) {
  unsigned long long b = (e & 0x8000000000000000ull) ? (e & 0x7fffffffffffffffull) : ~e;
  return __longlong_as_double((long long)b);
}
static inline unsigned long long enc_double_order_host(double v) {
  unsigned long long b;
  std::memcpy(&b, &v, 8);
  return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
static inline double dec_double_order_host(unsigned long long e) {
  unsigned long long b = (e & 0x8000000000000000ull) ? (e & 0x7fffffffffffffffull) : ~e;
  double v;
  std::memcpy(&v, &b, 8);
  return v;
}

// ---- list rebuild, per-cell sort and fill (shared by the fill kernel and the fill + bonded launch)
// decision shared by every rebuild kernel: rebuild requested, or an atom moved > skin/2
// (a NaN displacement, from non-finite positions, also rebuilds: written as !(d <= thr2))
__device__ __forceinline__ bool need_rebuild(const DevState* st, double thr2) {
  return st->rebuild != 0 || !(dec_double_order(st->max_disp2_bits) <= thr2);
}

struct FillArgs {
  const int* cell_start;
  const int* cell_count;
  int* sorted_orig;
  int* orig_sorted;
  const double* pos;
  double* pos_build;
  const double* lj64;
  float2* ljs;
  float2* ljs14;
  const int* ref_lin;
  int* ref_cell_sorted;
  int rny, rnz;
  const DevState* st;
  double thr2;
  const float* qs;
  float4* xq;
  double* pos_sorted;
  double L[3];
  int ncell;
};

#ifndef SORT_MAX
#define SORT_MAX 1024
#endif
// slot s of the sorted order holds atom o: its position at the build, packed copies and LJ
__device__ inline void fill_slot(const FillArgs& f, int s, int o) {
  f.orig_sorted[o] = s;
  for (int d = 0; d < 3; ++d) f.pos_build[3 * o + d] = f.pos[3 * o + d];
  pack_atom(f.pos[3 * o], f.pos[3 * o + 1], f.pos[3 * o + 2], f.qs ? f.qs[o] : 0.f, f.L[0], f.L[1], f.L[2], &f.xq[s],
            &f.pos_sorted[3 * s]);
  if (f.lj64) {
    double e = f.lj64[4 * o + 0], sg = f.lj64[4 * o + 1], e14 = f.lj64[4 * o + 2], s14 = f.lj64[4 * o + 3];
    // (sigma/2, sqrt(24 eps)): a pair's 24 eps_ij is one product
    f.ljs[s] = make_float2((float)(0.5 * sg), (float)sqrt(24.0 * e));
    f.ljs14[s] = make_float2((float)(0.5 * s14), (float)sqrt(24.0 * e14));
  }
  if (f.ref_lin) {
    int lin = f.ref_lin[o];
    int cz = lin % f.rnz, cy = (lin / f.rnz) % f.rny, cx = lin / (f.rnz * f.rny);
    f.ref_cell_sorted[s] = (cx << 20) | (cy << 10) | cz;
  }
}

// per cell: sort members ascending (deterministic order), then fill the per-slot arrays
__device__ inline void grid_sort_fill_block(const FillArgs& f, int cell) {
  __shared__ int buf[SORT_MAX];
  int* sorted_orig = f.sorted_orig;
  int start = f.cell_start[cell], cnt = f.cell_count[cell];
  // A cell beyond the sort's capacity (flagged as an overflow where the cells were counted) keeps
  // the scatter's order: every one of its slots is still filled, so the sorted order stays a
  // permutation and no slot is left stale for the list build or the PME kernels.
  const bool sortable = cnt <= SORT_MAX;
  int P = 1;
  if (!sortable) {
    for (int t = threadIdx.x; t < cnt; t += blockDim.x) fill_slot(f, start + t, sorted_orig[start + t]);
    return;
  }
  while (P < cnt) P <<= 1;
  for (int t = threadIdx.x; t < P; t += blockDim.x) buf[t] = t < cnt ? sorted_orig[start + t] : 0x7fffffff;
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = threadIdx.x; t < P; t += blockDim.x) {
        int u = t ^ j;
        if (u > t) {
          bool up = (t & k) == 0;
          int a = buf[t], b = buf[u];
          if ((a > b) == up) {
            buf[t] = b;
            buf[u] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
    int o = buf[t];
    int s = start + t;
    sorted_orig[s] = o;
    fill_slot(f, s, o);
  }
}

// lanes cooperating on one atom's neighbour list (nonbonded kernel and list build)
#define NB_G 32

static inline int div_up(long long a, long long b) { return (int)((a + b - 1) / b); }

// ------------------------------------------------------------------ module entry points
// join the caller's stream on entry / make it wait for the context stream on exit
int mdx_enter(mdx_ctx* c);
int mdx_leave(mdx_ctx* c);
#define MDX_ENTER()                   \
  do {                                \
    int _r = mdx_enter(c);            \
    if (_r) return _r;                \
  } while (0)
#define MDX_LEAVE_RET(v)              \
  do {                                \
    int _v = (v);                     \
    int _r = mdx_leave(c);            \
    return _v ? _v : _r;              \
  } while (0)
// drop the context's captured step graph (and mark graphs of batches holding it stale)
void mdx_invalidate_graph(mdx_ctx* c);
// celllist / neighbour list
// reference grid of a cell-list cutoff: n = max(3, floor((L + 2)/rc)), w = (L + 2)/n, 1/w
void mdx_ref_grid(const double* box, double rc, int* n, double* w, double* inv);
int mdx_nlist_prepare(mdx_ctx* c);                    // allocate grid/list buffers for current config
int mdx_nlist_update(mdx_ctx* c, bool force_rebuild);  // enqueue displacement check + conditional rebuild + pack
int mdx_nlist_update_order(mdx_ctx* c, bool force_rebuild);  // its first part: up to the sorted arrays and pack
int mdx_nlist_update_list(mdx_ctx* c);                       // its second part: the list build
// the PME reciprocal chain as three launches that also run the nonbonded kernel (LJ + direct
// Coulomb, in shares) and the bonded terms of `bonded_terms`; true if the context can
bool mdx_force_phases_ok(const mdx_ctx* c);
int mdx_force_phases_enqueue(mdx_ctx* c, bool energy, unsigned bonded_terms);
int mdx_nlist_enqueue_pack(mdx_ctx* c);
int mdx_nlist_enqueue_clear(mdx_ctx* c);
int mdx_enqueue_request_rebuild(mdx_ctx* c);
// forces
// nonbonded launch; the bonded terms in `bonded_terms` (bond/angle/dihedral/improper bits, and the
// PME exclusion correction with MDX_TERM_COUL_RECIP) ride along as extra workgroups
int mdx_enqueue_nonbonded(mdx_ctx* c, bool lj, bool coul, bool energy, unsigned bonded_terms = 0u);
// batched replicas (MDX_BATCH_FUSED): each kind's per-replica records, and its launches
int mdx_batch_prepare_list(mdx_ctx* const* cs, int n, BatchKind* fill, BatchKind* fast, BatchKind* cell);
int mdx_batch_launch_fill(const BatchKind& k, hipStream_t s);
int mdx_batch_launch_list(const BatchKind& fast, const BatchKind& cell, hipStream_t s);
int mdx_batch_prepare_pme(mdx_ctx* const* cs, int n, BatchKind* fwd, BatchKind* fft, BatchKind* gat);
int mdx_batch_launch_pme(const BatchKind& fwd, const BatchKind& fft, const BatchKind& gat, hipStream_t s,
                         unsigned parts = 7u);  // bits: forward, x-axis pass, inverse + gather
int mdx_batch_prepare_nb(mdx_ctx* const* cs, int n, unsigned bonded_terms, BatchKind* out);
int mdx_batch_launch_nb(const BatchKind& k, int lanes, bool s14, bool energy, hipStream_t s);
// Rebuild the nonbonded kernel's Ewald-correction table when beta or the cutoff changed.
#define MDX_CTAB_N 256
int mdx_coulomb_table_update(mdx_ctx* c);
int mdx_enqueue_coulomb_plain(mdx_ctx* c);
int mdx_pme_enqueue(mdx_ctx* c, bool energy);
int mdx_pme_enqueue_part(mdx_ctx* c, int part);  // one plane-path launch alone (timing)
int mdx_bonded_enqueue(mdx_ctx* c, unsigned terms);
int mdx_enqueue_zero_accumulators(mdx_ctx* c);
int mdx_enqueue_forces_out(mdx_ctx* c, double* d_out);
int mdx_enqueue_forces(mdx_ctx* c, unsigned terms, bool energy);
#define MDX_E_MLP 8  // energy slot of MDX_TERM_MLP (its bit index)
int mdx_mlp_enqueue(mdx_ctx* c, bool energy);
// integrator
int mdx_integrator_enqueue_step(mdx_ctx* c, unsigned terms, bool energy, bool fuse_prev_b, bool emit_b);
FillArgs fill_args(mdx_ctx* c, double thr2);
int mdx_integrator_prime(mdx_ctx* c, unsigned terms);
int mdx_integrator_finish(mdx_ctx* c);
