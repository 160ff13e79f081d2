// Wrapping, mdpy-compatible cell binning, and the per-atom neighbour list (gfx950).
//
// Reference behaviour restated here:
//   wrap_positions            mdpy/utils/pbc.py:28-36
//   CellList grid + binning   mdpy/core/cell_list.py:56-116
//   27-cell stencil           mdpy/constraint/__init__.py:12-19 and
//                             mdpy/constraint/charmm_nonbonded_constraint.py:75-100
// The force kernels do not walk the reference's dense cell table: atoms are binned into a
// periodic grid of cells at least (rc + skin)/2 wide, sorted by cell (ascending original id
// inside a cell, so every order is deterministic), and each atom gets a full neighbour list
// (column-major, one lane per atom in the force kernel). In REFERENCE mode the list also drops
// pairs whose mdpy cells are not stencil neighbours, reproducing the reference's pair set.
#include "mdx_internal.h"
#include <algorithm>
#include <cmath>

#define NL_BLOCK 256

// ============================================================== wrap / reference binning
__global__ void k_wrap(double* __restrict__ pos, int n, double Lx, double Ly, double Lz, double ix,
                       double iy, double iz, DevState* st) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double L[3] = {Lx, Ly, Lz}, inv[3] = {ix, iy, iz};
  bool lost = false;
  double out[3];
  for (int d = 0; d < 3; ++d) {
    double x = pos[3 * i + d];
    double m = -rint(x * inv[d]);  // np.round: half to even
    if (fabs(m) >= 2.0) lost = true;
    out[d] = x + m * L[d];
  }
  if (lost) {
    atomicOr(&st->flags, MDX_FLAG_LOSS);
    return;
  }
  for (int d = 0; d < 3; ++d) pos[3 * i + d] = out[d];
}

__global__ void k_minpos(const double* __restrict__ pos, int n, DevState* st) {
  __shared__ unsigned long long sm[3][NL_BLOCK / 64];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long e[3] = {~0ull, ~0ull, ~0ull};
  if (i < n)
    for (int d = 0; d < 3; ++d) e[d] = enc_double_order(pos[3 * i + d]);
  for (int d = 0; d < 3; ++d) {
    unsigned long long v = e[d];
    for (int off = 32; off > 0; off >>= 1) {
      unsigned long long o = __shfl_xor(v, off, 64);
      v = o < v ? o : v;
    }
    if ((threadIdx.x & 63) == 0) sm[d][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    unsigned long long v = ~0ull;
    for (int w = 0; w < NL_BLOCK / 64; ++w) v = sm[threadIdx.x][w] < v ? sm[threadIdx.x][w] : v;
    atomicMin(&st->minpos[threadIdx.x], v);
  }
}

__global__ void k_ref_bin(const double* __restrict__ pos, int n, int nx, int ny, int nz, double ix,
                          double iy, double iz, DevState* st, int* __restrict__ cell_index,
                          int* __restrict__ cell_lin, int* __restrict__ counts) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double mn[3] = {dec_double_order(st->minpos[0]), dec_double_order(st->minpos[1]),
                  dec_double_order(st->minpos[2])};
  double inv[3] = {ix, iy, iz};
  int nn[3] = {nx, ny, nz};
  int c[3];
  for (int d = 0; d < 3; ++d) {
    double s = (pos[3 * i + d] - mn[d]) * inv[d];  // (x - min x) . cell_inv, diagonal
    const double f = floor(s);
    // range decided before the conversion (NaN or runaway positions flag and clamp)
    if (!(f >= 0.0 && f < (double)nn[d])) atomicOr(&st->flags, MDX_FLAG_OVERFLOW);
    const int k = clamped_floor_index(f, nn[d]);
    c[d] = k;
    if (cell_index) cell_index[3 * i + d] = k;
  }
  int lin = (c[0] * ny + c[1]) * nz + c[2];
  if (cell_lin) cell_lin[i] = lin;
  int cnt = atomicAdd(&counts[lin], 1) + 1;
  atomicMax(&st->maxpc, cnt);
}

__global__ void k_ref_scatter(const int* __restrict__ cell_lin, int n, int* __restrict__ cursor,
                              int* __restrict__ table, int maxpc) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c = cell_lin[i];
  int slot = atomicAdd(&cursor[c], 1);
  if (slot < maxpc) table[(long long)c * maxpc + slot] = i;
}

// Sort each row ascending (bitonic sort in LDS); entries >= valid are padding -1.
__global__ void k_sort_rows(int* __restrict__ table, const int* __restrict__ counts, int width) {
  __shared__ int buf[SORT_MAX];
  int row = blockIdx.x;
  int cnt = counts[row];
  int P = 1;
  while (P < cnt) P <<= 1;
  for (int t = threadIdx.x; t < P; t += blockDim.x)
    buf[t] = t < cnt ? table[(long long)row * width + t] : 0x7fffffff;
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
  for (int t = threadIdx.x; t < width; t += blockDim.x)
    table[(long long)row * width + t] = t < cnt ? buf[t] : -1;
}

__global__ void k_fill_i32(int* __restrict__ p, int v, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// reset the reference-binning reductions (minimum position per axis, largest population,
// flags) and the per-cell counts
__global__ void k_ref_reset(DevState* st, int* __restrict__ counts, int nc) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nc) counts[i] = 0;
  if (i == 0) {
    st->minpos[0] = st->minpos[1] = st->minpos[2] = ~0ull;
    st->maxpc = 0;
    st->flags = 0;
  }
}

extern "C" int mdx_wrap_positions(mdx_ctx* c) {
  MDX_CHECK(c && c->pos, MDX_ERR_INVALID, "mdx_wrap_positions: no system bound");
  MDX_ENTER();
  c->forces_current = false;
  double inv[3];
  for (int d = 0; d < 3; ++d) inv[d] = 1.0 / c->box[d];
  MDX_HIP(hipMemsetAsync(&c->dstate.p->flags, 0, sizeof(unsigned), c->stream));
  k_wrap<<<div_up(c->n, NL_BLOCK), NL_BLOCK, 0, c->stream>>>(c->pos, c->n, c->box[0], c->box[1], c->box[2],
                                                            inv[0], inv[1], inv[2], c->dstate.p);
  MDX_LAUNCH_CHECK();
  unsigned flags = 0;
  MDX_HIP(hipMemcpyAsync(&flags, &c->dstate.p->flags, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
  MDX_HIP(hipStreamSynchronize(c->stream));
  int h1 = 1;
  MDX_HIP(hipMemcpy(&c->dstate.p->rebuild, &h1, sizeof(int), hipMemcpyHostToDevice));
  MDX_CHECK(!(flags & MDX_FLAG_LOSS), MDX_ERR_PARTICLE_LOSS, "atom(s) moved beyond 2 PBC images");
  MDX_LEAVE_RET(0);
}

// Reference grid for a cell-list cutoff (cell_list.py:74-81).
void mdx_ref_grid(const double* box, double rc, int* n, double* w, double* inv) {
  for (int d = 0; d < 3; ++d) {
    int k = (int)std::floor((box[d] + 2.0) / rc);
    if (k < 3) k = 3;
    n[d] = k;
    w[d] = (box[d] + 2.0) / (double)k;
    inv[d] = 1.0 / w[d];
  }
}

static int ref_bin_enqueue(mdx_ctx* c, double cell_cutoff, int* d_cell_index) {
  mdx_ref_grid(c->box, cell_cutoff, c->ref_n, c->ref_w, c->ref_inv);
  int nc = c->ref_n[0] * c->ref_n[1] * c->ref_n[2];
  int rc;
  if ((rc = c->ref_cell_lin.ensure(c->n))) return rc;
  if ((rc = c->ref_counts.ensure(nc))) return rc;
  if ((rc = c->ref_cursor.ensure(nc))) return rc;
  k_ref_reset<<<div_up(nc, NL_BLOCK), NL_BLOCK, 0, c->stream>>>(c->dstate.p, c->ref_counts.p, nc);
  MDX_LAUNCH_CHECK();
  k_minpos<<<div_up(c->n, NL_BLOCK), NL_BLOCK, 0, c->stream>>>(c->pos, c->n, c->dstate.p);
  MDX_LAUNCH_CHECK();
  k_ref_bin<<<div_up(c->n, NL_BLOCK), NL_BLOCK, 0, c->stream>>>(
      c->pos, c->n, c->ref_n[0], c->ref_n[1], c->ref_n[2], c->ref_inv[0], c->ref_inv[1], c->ref_inv[2],
      c->dstate.p, d_cell_index, c->ref_cell_lin.p, c->ref_counts.p);
  MDX_LAUNCH_CHECK();
  return 0;
}

extern "C" int mdx_celllist_reference(mdx_ctx* c, double cell_cutoff, int* d_cell_index, int* h_n, int* h_maxpc) {
  MDX_CHECK(c && c->pos, MDX_ERR_INVALID, "mdx_celllist_reference: no system bound");
  MDX_CHECK(cell_cutoff > 0, MDX_ERR_CELL_CUTOFF, "cutoff radius is poor defined");
  for (int d = 0; d < 3; ++d)
    MDX_CHECK(std::floor(c->box[d] / cell_cutoff) >= 2, MDX_ERR_CELL_CUTOFF,
              "the cutoff_radius is too large to create cell list");
  MDX_ENTER();
  MDX_HIP(hipMemsetAsync(&c->dstate.p->flags, 0, sizeof(unsigned), c->stream));
  int rc = ref_bin_enqueue(c, cell_cutoff, d_cell_index);
  if (rc) return rc;
  int maxpc = 0;
  unsigned flags = 0;
  MDX_HIP(hipMemcpyAsync(&maxpc, &c->dstate.p->maxpc, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  MDX_HIP(hipMemcpyAsync(&flags, &c->dstate.p->flags, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
  MDX_HIP(hipStreamSynchronize(c->stream));
  MDX_CHECK(!(flags & MDX_FLAG_OVERFLOW), MDX_ERR_OVERFLOW, "reference binning produced an out-of-range cell");
  for (int d = 0; d < 3; ++d) h_n[d] = c->ref_n[d];
  *h_maxpc = maxpc;
  MDX_LEAVE_RET(0);
}

extern "C" int mdx_celllist_table(mdx_ctx* c, int* d_table, int maxpc) {
  MDX_CHECK(c && d_table && maxpc > 0, MDX_ERR_INVALID, "mdx_celllist_table: bad arguments");
  MDX_CHECK(maxpc <= SORT_MAX * 8, MDX_ERR_OVERFLOW, "cell population %d too large", maxpc);
  int nc = c->ref_n[0] * c->ref_n[1] * c->ref_n[2];
  MDX_CHECK(nc > 0, MDX_ERR_INVALID, "mdx_celllist_table before mdx_celllist_reference");
  MDX_CHECK(maxpc <= SORT_MAX, MDX_ERR_OVERFLOW, "cell population %d exceeds %d", maxpc, SORT_MAX);
  MDX_ENTER();
  MDX_HIP(hipMemsetAsync(c->ref_cursor.p, 0, nc * sizeof(int), c->stream));
  MDX_HIP(hipMemsetAsync(d_table, 0xff, (size_t)nc * maxpc * sizeof(int), c->stream));
  k_ref_scatter<<<div_up(c->n, NL_BLOCK), NL_BLOCK, 0, c->stream>>>(c->ref_cell_lin.p, c->n, c->ref_cursor.p,
                                                                   d_table, maxpc);
  MDX_LAUNCH_CHECK();
  k_sort_rows<<<nc, 256, 0, c->stream>>>(d_table, c->ref_counts.p, maxpc);
  MDX_LAUNCH_CHECK();
  MDX_LEAVE_RET(0);
}

// ============================================================== neighbour-list grid

__global__ void k_check_disp(const double* __restrict__ pos, const double* __restrict__ pos_build, int n,
                             double Lx, double Ly, double Lz, DevState* st) {
  __shared__ double sm[NL_BLOCK / 64];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double d2 = 0.0;
  bool bad = false;
  if (i < n) {
    double L[3] = {Lx, Ly, Lz};
    for (int d = 0; d < 3; ++d) {
      double x = pos[3 * i + d];
      if (!isfinite(x)) bad = true;
      double dd = x - pos_build[3 * i + d];
      dd -= L[d] * rint(dd / L[d]);
      d2 += dd * dd;
    }
  }
  if (bad) atomicOr(&st->flags, MDX_FLAG_NONFINITE);
  for (int off = 32; off > 0; off >>= 1) d2 = fmax(d2, __shfl_xor(d2, off, 64));
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = d2;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = 0;
    for (int w = 0; w < NL_BLOCK / 64; ++w) v = fmax(v, sm[w]);
    atomicMax(&st->max_disp2_bits, enc_double_order(v));
  }
}

struct GridArgs {
  int nx, ny, nz;
  double Lx, Ly, Lz;
  double wx, wy, wz;  // cell widths
};

__global__ void k_grid_bin(const double* __restrict__ pos, int n, GridArgs g, const DevState* st, double thr2,
                           int* __restrict__ atom_cell, int* __restrict__ cell_count) {
  if (!need_rebuild(st, thr2)) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double L[3] = {g.Lx, g.Ly, g.Lz}, w[3] = {g.wx, g.wy, g.wz};
  int nn[3] = {g.nx, g.ny, g.nz};
  int c[3];
  for (int d = 0; d < 3; ++d) {
    double x = pos[3 * i + d];
    x -= L[d] * floor(x / L[d]);
    c[d] = clamped_floor_index(x / w[d], nn[d]);
  }
  int lin = (c[0] * g.ny + c[1]) * g.nz + c[2];
  atom_cell[i] = lin;
  atomicAdd(&cell_count[lin], 1);
}

// single-block exclusive scan over cells; also resets cursors and checks the sort limit
__global__ void k_cell_scan(const int* __restrict__ cell_count, int* __restrict__ cell_start,
                            int* __restrict__ cursor, int ncell, const DevState* st_c, DevState* st, double thr2) {
  if (!need_rebuild(st_c, thr2)) return;
  __shared__ int part[1024];
  int per = (ncell + blockDim.x - 1) / blockDim.x;
  int b = threadIdx.x * per, e = min(ncell, b + per);
  int s = 0;
  for (int k = b; k < e; ++k) {
    int v = cell_count[k];
    if (v > SORT_MAX) atomicOr(&st->flags, MDX_FLAG_OVERFLOW);
    s += v;
  }
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int t = 0; t < (int)blockDim.x; ++t) {
      int v = part[t];
      part[t] = acc;
      acc += v;
    }
  }
  __syncthreads();
  int acc = part[threadIdx.x];
  for (int k = b; k < e; ++k) {
    cell_start[k] = acc;
    cursor[k] = 0;
    acc += cell_count[k];
  }
}

__global__ void k_grid_scatter(const int* __restrict__ atom_cell, int n, const int* __restrict__ cell_start,
                               int* __restrict__ cursor, int* __restrict__ sorted_orig, const DevState* st,
                               double thr2) {
  if (!need_rebuild(st, thr2)) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c = atom_cell[i];
  int slot = atomicAdd(&cursor[c], 1);
  sorted_orig[cell_start[c] + slot] = i;
}

__global__ void k_grid_sort_fill(FillArgs f) {
  if (!need_rebuild(f.st, f.thr2)) return;
  grid_sort_fill_block(f, blockIdx.x);
}

FillArgs fill_args(mdx_ctx* c, double thr2) {
  FillArgs f;
  f.cell_start = c->cell_start.p;
  f.cell_count = c->cell_count.p;
  f.sorted_orig = c->sorted_orig.p;
  f.orig_sorted = c->orig_sorted.p;
  f.pos = c->pos;
  f.pos_build = c->pos_at_build.p;
  f.lj64 = c->has_lj ? c->lj64.p : nullptr;
  f.ljs = c->ljs.p;
  f.ljs14 = c->ljs14.p;
  const bool ref = c->pair_mode == MDX_PAIRS_REFERENCE;
  f.ref_lin = ref ? c->ref_cell_lin.p : nullptr;
  f.ref_cell_sorted = c->ref_cell_sorted.p;
  f.rny = c->ref_n[1];
  f.rnz = c->ref_n[2];
  f.st = c->dstate.p;
  f.thr2 = thr2;
  f.qs = c->has_charges ? c->charge_s.p : nullptr;
  f.xq = c->xq.p;
  f.pos_sorted = c->pos_sorted.p;
  f.L[0] = c->box[0];
  f.L[1] = c->box[1];
  f.L[2] = c->box[2];
  f.ncell = c->ncell;
  return f;
}

// every evaluation: float32 positions (wrapped, centred) and scaled charges in sorted order,
// plus the float64 positions as they are (sorted order, read by the PME plane kernels)
__global__ void k_pack(const double* __restrict__ pos, const int* __restrict__ sorted_orig,
                       const float* __restrict__ qs, int n, double Lx, double Ly, double Lz,
                       float4* __restrict__ xq, double* __restrict__ pos_sorted) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int o = sorted_orig[s];
  pack_atom(pos[3 * o + 0], pos[3 * o + 1], pos[3 * o + 2], qs ? qs[o] : 0.f, Lx, Ly, Lz, &xq[s], &pos_sorted[3 * s]);
}

struct ListArgs {
  GridArgs g;
  int n;
  int cap;
  float rl2;          // list radius squared (float pre-filter)
  int exact64;        // decide r <= rc in float64 (skin == 0)
  double rc;
  int ref_mode;       // REFERENCE stencil filter
  int rnx, rny, rnz;  // reference grid
  int nstencil;
  int shifted;        // staged candidates are shifted into the home frame (no minimum image)
};

__device__ __forceinline__ int cyc_dist(int a, int b, int n) {
  int d = a > b ? a - b : b - a;
  return min(d, n - d);
}

// One workgroup per home cell: the candidates of every stencil cell (identical for all atoms of
// the home cell) are staged in LDS in chunks, then each NB_G-lane group scans them for one home
// atom at a time and compacts the accepted ones with a wave ballot. Same candidate order, hence
// the same lists, as a per-atom walk of the stencil; global memory is read once per cell. The
// next chunk's candidates are fetched into registers while the current one is scanned.
// Staged candidates are shifted into the home cell's periodic frame (when the stencil holds each
// periodic cell once, on no more than one side), so the scan needs no minimum-image arithmetic.
// Exclusions and 1-4 partners within 32 atom indices of the home atom are bits of two 64-bit
// masks; an atom with a partner outside that window scans its CSR lists (staged in LDS) instead.
// Each lane tests BL_U candidates per pass, all LDS reads issued first. FAST instantiations
// (shifted candidates, exact pair mode, masks present) compile out the reference-mode filter and
// the minimum image; the float64 decision of skin-0 lists is the EXACT64 instantiation.
#ifndef BL_CH
#define BL_CH 1024  // candidates staged per chunk (a multiple of the fast kernel's pass)
#endif
#define BL_MAXST 512
#define BL_MAXEX 32  // exclusions (and 1-4 partners) per atom
#define BL_HMAX 192  // home atoms staged in LDS (more: their data is read from global memory)
#define BL_POOL 1024 // exclusion / 1-4 entries of a cell's home atoms staged in LDS
#define BL_THREADS 512
#define BL_SLOTS ((BL_CH + BL_THREADS - 1) / BL_THREADS)  // staged candidates per lane per chunk
#define BL_U 4       // candidates per lane per pass
#define BL_WIN 32    // index window of the exclusion masks: partners i-32 .. i+31

// The arguments of one context's list-build launch (passed by value to the single-context
// kernels, or one replica's record of a batched launch).
struct ListCall {
  ListArgs a;
  const float4* xq;
  const int* sorted_orig;
  const int* cell_start;
  const int* cell_count;
  const int* stencil;
  const int* excl_ptr;
  const int* excl_idx;
  const int* s14_ptr;
  const int* s14_idx;
  const unsigned long long* ex_win;
  const unsigned long long* s4_win;
  const unsigned char* win_bad;
  const int* ref_cell_sorted;
  const double* pos;
  int* nbr_count;
  unsigned* nbr_list;
  DevState* st;
  double thr2;
};

// the build of home cell `cell`
template <bool EXACT64, bool FAST>
__device__ __forceinline__ void build_list_cell_body(const ListCall& L, const int cell) {
  const ListArgs& a = L.a;
  const float4* __restrict__ xq = L.xq;
  const int* __restrict__ sorted_orig = L.sorted_orig;
  const int* __restrict__ cell_start = L.cell_start;
  const int* __restrict__ cell_count = L.cell_count;
  const int* __restrict__ stencil = L.stencil;
  const int* __restrict__ excl_ptr = L.excl_ptr;
  const int* __restrict__ excl_idx = L.excl_idx;
  const int* __restrict__ s14_ptr = L.s14_ptr;
  const int* __restrict__ s14_idx = L.s14_idx;
  const unsigned long long* __restrict__ ex_win = L.ex_win;
  const unsigned long long* __restrict__ s4_win = L.s4_win;
  const unsigned char* __restrict__ win_bad = L.win_bad;
  const int* __restrict__ ref_cell_sorted = L.ref_cell_sorted;
  const double* __restrict__ pos = L.pos;
  int* __restrict__ nbr_count = L.nbr_count;
  unsigned* __restrict__ nbr_list = L.nbr_list;
  DevState* st = L.st;
  const double thr2 = L.thr2;
  if (!need_rebuild(st, thr2)) return;
  __shared__ float4 cxq[BL_CH];
  __shared__ int2 cto[BL_CH];      // (sorted slot, original atom id)
  __shared__ int cr[FAST ? 1 : BL_CH];
  __shared__ int sb[BL_MAXST], sp[BL_MAXST + 1];
  __shared__ unsigned char shc[BL_MAXST];  // periodic shift of each stencil cell: 2 bits per axis
  __shared__ int acnt[SORT_MAX];
  __shared__ int exl[BL_THREADS / NB_G][BL_MAXEX], s4l[BL_THREADS / NB_G][BL_MAXEX];
  // home atoms staged once per block: position, atom id, window masks, exclusion / 1-4 lists
  __shared__ float4 hxi[BL_HMAX];
  __shared__ int ho[BL_HMAX], hex0[BL_HMAX + 1], hs40[BL_HMAX + 1];
  __shared__ unsigned long long hexm[BL_HMAX], hs4m[BL_HMAX];
  __shared__ unsigned char hbad[BL_HMAX];
  __shared__ int pool_ex[BL_POOL], pool_s4[BL_POOL];
  __shared__ int fast;
  const GridArgs& gr = a.g;
  const int hb = cell_start[cell], hn = cell_count[cell];
  if (hn > SORT_MAX) {  // oversize cells are flagged where the cells are counted: empty lists
    for (int t = threadIdx.x; t < hn; t += blockDim.x) nbr_count[hb + t] = 0;
    return;
  }
  if (hn == 0) return;
  const int tid = threadIdx.x;
  const int hz = cell % gr.nz, hy = (cell / gr.nz) % gr.ny, hx = cell / (gr.nz * gr.ny);
  const bool shifted = FAST || a.shifted;
  for (int k = tid; k < a.nstencil; k += blockDim.x) {
    int p = stencil[k];
    int nx = hx + ((p >> 16) & 0xff) - 128, ny = hy + ((p >> 8) & 0xff) - 128, nz = hz + (p & 0xff) - 128;
    // a neighbour reached across the box edge is the periodic image one box length away
    // (code per axis: 0 = -L, 1 = 0, 2 = +L)
    const int cx = nx < 0 ? 0 : (nx >= gr.nx ? 2 : 1);
    const int cy = ny < 0 ? 0 : (ny >= gr.ny ? 2 : 1);
    const int cz = nz < 0 ? 0 : (nz >= gr.nz ? 2 : 1);
    shc[k] = (unsigned char)(cx | (cy << 2) | (cz << 4));
    nx += nx < 0 ? gr.nx : (nx >= gr.nx ? -gr.nx : 0);
    ny += ny < 0 ? gr.ny : (ny >= gr.ny ? -gr.ny : 0);
    nz += nz < 0 ? gr.nz : (nz >= gr.nz ? -gr.nz : 0);
    const int nc = (nx * gr.ny + ny) * gr.nz + nz;
    sb[k] = cell_start[nc];
    sp[k + 1] = cell_count[nc];
  }
  for (int j = tid; j < hn; j += blockDim.x) acnt[j] = 0;
  const bool fit = hn <= BL_HMAX;
  if (fit) {
    for (int h = tid; h < hn; h += blockDim.x) {
      const int o = sorted_orig[hb + h];
      ho[h] = o;
      hxi[h] = xq[hb + h];
      hex0[h + 1] = excl_ptr ? excl_ptr[o + 1] - excl_ptr[o] : 0;
      hs40[h + 1] = s14_ptr ? s14_ptr[o + 1] - s14_ptr[o] : 0;
      hexm[h] = ex_win ? ex_win[o] : 0ull;
      hs4m[h] = s4_win ? s4_win[o] : 0ull;
      hbad[h] = win_bad ? win_bad[o] : (unsigned char)1;
    }
  }
  __syncthreads();
  if (tid == 0) {
    sp[0] = 0;
    for (int k = 0; k < a.nstencil; ++k) sp[k + 1] += sp[k];
    if (fit) {
      hex0[0] = 0;
      hs40[0] = 0;
      for (int h = 0; h < hn; ++h) {
        hex0[h + 1] += hex0[h];
        hs40[h + 1] += hs40[h];
      }
    }
    fast = fit && hex0[fit ? hn : 0] <= BL_POOL && hs40[fit ? hn : 0] <= BL_POOL;
  }
  __syncthreads();
  if (fast) {
    for (int h = tid; h < hn; h += blockDim.x) {
      const int o = ho[h];
      if (excl_ptr)
        for (int r = 0; r < hex0[h + 1] - hex0[h]; ++r) pool_ex[hex0[h] + r] = excl_idx[excl_ptr[o] + r];
      if (s14_ptr)
        for (int r = 0; r < hs40[h + 1] - hs40[h]; ++r) pool_s4[hs40[h] + r] = s14_idx[s14_ptr[o] + r];
    }
  }
  const int total = sp[a.nstencil];
  const float Lx = (float)gr.Lx, Ly = (float)gr.Ly, Lz = (float)gr.Lz;
  // a lane's share of one chunk's candidates, fetched from global memory into registers
  float4 pv[BL_SLOTS];
  int2 pt[BL_SLOTS];
  int pr[BL_SLOTS];
  auto fetch = [&](int base) {
    const int nch = min(BL_CH, total - base);
#pragma unroll
    for (int r = 0; r < BL_SLOTS; ++r) {
      const int c = tid + r * BL_THREADS;
      if (c < nch) {
        const int idx = base + c;
        int lo = 0, hi = a.nstencil - 1;  // last stencil cell whose prefix <= idx
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (sp[mid] <= idx) lo = mid;
          else hi = mid - 1;
        }
        const int t = sb[lo] + (idx - sp[lo]);
        float4 v = xq[t];
        if (shifted) {
          const int code = shc[lo];
          v.x += (float)((code & 3) - 1) * Lx;
          v.y += (float)(((code >> 2) & 3) - 1) * Ly;
          v.z += (float)(((code >> 4) & 3) - 1) * Lz;
        }
        pv[r] = v;
        pt[r] = make_int2(t, sorted_orig[t]);
        if (!FAST && a.ref_mode) pr[r] = ref_cell_sorted[t];
      }
    }
  };
  auto store = [&](int base) {
    const int nch = min(BL_CH, total - base);
#pragma unroll
    for (int r = 0; r < BL_SLOTS; ++r) {
      const int c = tid + r * BL_THREADS;
      if (c < nch) {
        cxq[c] = pv[r];
        cto[c] = pt[r];
        if (!FAST && a.ref_mode) cr[c] = pr[r];
      }
    }
  };
  __syncthreads();  // sp, pools
  if (total > 0) fetch(0);
  store(0);
  __syncthreads();
  const int lane = tid & 63;
  const int q = tid / NB_G, g = tid & (NB_G - 1);
  const int ngroups = blockDim.x / NB_G;
  const unsigned long long gmask = (NB_G == 64 ? ~0ull : ((1ull << NB_G) - 1ull)) << (lane & ~(NB_G - 1));
  const unsigned long long below = (1ull << lane) - 1ull;
  const float ix = 1.0f / Lx, iy = 1.0f / Ly, iz = 1.0f / Lz;
  int* my_ex = exl[q];
  int* my_s4 = s4l[q];
  for (int base = 0; base < total; base += BL_CH) {
    const int nch = min(BL_CH, total - base);
    const bool more = base + BL_CH < total;
    if (more) fetch(base + BL_CH);  // in flight during this chunk's scan
    for (int h = q; h < hn; h += ngroups) {  // uniform per group
      const int s = hb + h;
      int o, nex_all, ns4_all;
      const int* src_ex;
      const int* src_s4;
      float4 xi;
      bool masked;
      unsigned long long exm, s4m;
      if (fast) {  // everything from LDS
        o = ho[h];
        xi = hxi[h];
        nex_all = hex0[h + 1] - hex0[h];
        ns4_all = hs40[h + 1] - hs40[h];
        src_ex = pool_ex + hex0[h];
        src_s4 = pool_s4 + hs40[h];
        masked = hbad[h] == 0;
        exm = hexm[h];
        s4m = hs4m[h];
      } else {     // crowded cell: read the home atom's data from global memory
        o = sorted_orig[s];
        xi = xq[s];
        const int e0 = excl_ptr ? excl_ptr[o] : 0, t0s = s14_ptr ? s14_ptr[o] : 0;
        nex_all = excl_ptr ? excl_ptr[o + 1] - e0 : 0;
        ns4_all = s14_ptr ? s14_ptr[o + 1] - t0s : 0;
        src_ex = excl_idx + e0;
        src_s4 = s14_idx + t0s;
        masked = win_bad && win_bad[o] == 0;
        exm = ex_win ? ex_win[o] : 0ull;
        s4m = s4_win ? s4_win[o] : 0ull;
      }
      // index-window masks, or the CSR lists for an atom with a partner outside the window
      const int nex = masked ? 0 : min(nex_all, BL_MAXEX), ns4 = masked ? 0 : min(ns4_all, BL_MAXEX);
      if (!masked) {
        if (g < nex) my_ex[g] = src_ex[g];
        if (g < ns4) my_s4[g] = src_s4[g];
        if (g == 0 && (nex_all > BL_MAXEX || ns4_all > BL_MAXEX)) atomicOr(&st->flags, MDX_FLAG_OVERFLOW);
      }
      const int rci = (!FAST && a.ref_mode) ? ref_cell_sorted[s] : 0;
      double pi[3] = {0.0, 0.0, 0.0};
      if (EXACT64) {
        pi[0] = pos[3 * o];
        pi[1] = pos[3 * o + 1];
        pi[2] = pos[3 * o + 2];
      }
      wave_lds_sync();  // the group's lanes are in one wave
      unsigned* row = nbr_list + (long long)s * a.cap;
      int cnt = acnt[h];
      for (int cb = 0; cb < nch; cb += BL_U * NB_G) {
        float4 xj[BL_U];
        int2 tj[BL_U];
        int rcj[BL_U];
#pragma unroll
        for (int u = 0; u < BL_U; ++u) {
          const int ci = min(cb + u * NB_G + g, nch - 1);
          xj[u] = cxq[ci];
          tj[u] = cto[ci];
          rcj[u] = (!FAST && a.ref_mode) ? cr[ci] : 0;
        }
#pragma unroll
        for (int u = 0; u < BL_U; ++u) {
          const int jo = tj[u].y;
          float dx = xj[u].x - xi.x, dy = xj[u].y - xi.y, dz = xj[u].z - xi.z;
          if (!shifted) {
            dx -= Lx * rintf(dx * ix);
            dy -= Ly * rintf(dy * iy);
            dz -= Lz * rintf(dz * iz);
          }
          bool acc = (cb + u * NB_G + g < nch) & (tj[u].x != s) & (dx * dx + dy * dy + dz * dz < a.rl2);
          bool s14;
          if (masked) {
            const int dw = jo - o + BL_WIN;
            const bool win = (unsigned)dw < 2u * BL_WIN;
            const int sh = dw & (2 * BL_WIN - 1);
            acc &= !(win && ((exm >> sh) & 1ull));
            s14 = win && ((s4m >> sh) & 1ull);
          } else {
            s14 = false;
            for (int r = 0; r < nex; ++r) acc &= (my_ex[r] != jo);
            for (int r = 0; r < ns4; ++r) s14 |= (my_s4[r] == jo);
          }
          if (!FAST && a.ref_mode) {
            const int ax = cyc_dist(rci >> 20, rcj[u] >> 20, a.rnx);
            const int ay = cyc_dist((rci >> 10) & 1023, (rcj[u] >> 10) & 1023, a.rny);
            const int az = cyc_dist(rci & 1023, rcj[u] & 1023, a.rnz);
            acc &= (ax <= 1) & (ay <= 1) & (az <= 1);
          }
          if (EXACT64 && acc) {  // float64 decision with the reference's sequence (pbc.py:40-44)
            double d[3];
            const double L[3] = {gr.Lx, gr.Ly, gr.Lz};
            for (int r = 0; r < 3; ++r) {
              double v = pos[3 * jo + r] - pi[r];
              double sv = v * (1.0 / L[r]);
              sv -= rint(sv);
              d[r] = sv * L[r];
            }
            acc = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) <= a.rc;
          }
          const unsigned long long m = __ballot(acc) & gmask;
          const int slot = cnt + __popcll(m & below);
          if (acc && slot < a.cap) row[slot] = (unsigned)tj[u].x | (s14 ? 0x80000000u : 0u);
          cnt += __popcll(m);
        }
      }
      if (g == 0) acnt[h] = cnt;
      wave_lds_sync();  // the next home atom of this group overwrites my_ex / my_s4
    }
    __syncthreads();
    if (more) {
      store(base + BL_CH);
      __syncthreads();
    }
  }
  for (int h = tid; h < hn; h += blockDim.x) {
    const int cnt = acnt[h];
    nbr_count[hb + h] = min(cnt, a.cap);
    if (cnt > a.cap) atomicOr(&st->flags, MDX_FLAG_OVERFLOW);
  }
}

template <bool EXACT64, bool FAST>
__global__ __launch_bounds__(BL_THREADS) void k_build_list_cell(ListCall L) {
  build_list_cell_body<EXACT64, FAST>(L, blockIdx.x);
}

// The common case of the build above as its own kernel: shifted candidates, exact pair mode,
// float32 list radius, window masks present. One wave per home atom, so every ballot, count and
// branch is wave-uniform; exclusions, the atom itself (bit 32) and 1-4 partners are mask bits, the
// window offset is clamped to bit 63 (always clear) instead of range-checked, and the staged
// chunk is padded with far-away candidates up to a multiple of the pass width, so a candidate
// test is distance, two mask bits and the ballot. Home atoms with a partner outside the window
// scan their CSR lists. A group of 64 candidates with none inside the list radius skips the mask
// and store work. Produces the same lists, in the same order, as k_build_list_cell.
#define BLF_PASS (64 * BL_U)
__device__ __forceinline__ void build_list_fast_body(const ListCall& L, const int cell) {
  const ListArgs& a = L.a;
  const float4* __restrict__ xq = L.xq;
  const int* __restrict__ sorted_orig = L.sorted_orig;
  const int* __restrict__ cell_start = L.cell_start;
  const int* __restrict__ cell_count = L.cell_count;
  const int* __restrict__ stencil = L.stencil;
  const int* __restrict__ excl_ptr = L.excl_ptr;
  const int* __restrict__ excl_idx = L.excl_idx;
  const int* __restrict__ s14_ptr = L.s14_ptr;
  const int* __restrict__ s14_idx = L.s14_idx;
  const unsigned long long* __restrict__ ex_win = L.ex_win;
  const unsigned long long* __restrict__ s4_win = L.s4_win;
  const unsigned char* __restrict__ win_bad = L.win_bad;
  int* __restrict__ nbr_count = L.nbr_count;
  unsigned* __restrict__ nbr_list = L.nbr_list;
  DevState* st = L.st;
  const double thr2 = L.thr2;
  static_assert(BL_CH % BLF_PASS == 0, "chunks must hold whole passes");
  if (!need_rebuild(st, thr2)) return;
  __shared__ float4 cxq[BL_CH];
  __shared__ int2 cto[BL_CH];  // (sorted slot, original atom id)
  __shared__ int sb[BL_MAXST], sp[BL_MAXST + 1];
  __shared__ unsigned char shc[BL_MAXST];
  __shared__ int acnt[SORT_MAX];
  __shared__ int exl[BL_THREADS / 64][BL_MAXEX], s4l[BL_THREADS / 64][BL_MAXEX];
  __shared__ float4 hxi[BL_HMAX];
  __shared__ int ho[BL_HMAX];
  __shared__ unsigned long long hexm[BL_HMAX], hs4m[BL_HMAX];
  __shared__ unsigned char hbad[BL_HMAX];
  const GridArgs& gr = a.g;
  const int hb = cell_start[cell], hn = cell_count[cell];
  if (hn > SORT_MAX) {  // oversize cells are flagged where the cells are counted: empty lists
    for (int t = threadIdx.x; t < hn; t += blockDim.x) nbr_count[hb + t] = 0;
    return;
  }
  if (hn == 0) return;
  const int tid = threadIdx.x;
  const int hz = cell % gr.nz, hy = (cell / gr.nz) % gr.ny, hx = cell / (gr.nz * gr.ny);
  for (int k = tid; k < a.nstencil; k += blockDim.x) {
    int p = stencil[k];
    int nx = hx + ((p >> 16) & 0xff) - 128, ny = hy + ((p >> 8) & 0xff) - 128, nz = hz + (p & 0xff) - 128;
    const int cx = nx < 0 ? 0 : (nx >= gr.nx ? 2 : 1);
    const int cy = ny < 0 ? 0 : (ny >= gr.ny ? 2 : 1);
    const int cz = nz < 0 ? 0 : (nz >= gr.nz ? 2 : 1);
    shc[k] = (unsigned char)(cx | (cy << 2) | (cz << 4));
    nx += nx < 0 ? gr.nx : (nx >= gr.nx ? -gr.nx : 0);
    ny += ny < 0 ? gr.ny : (ny >= gr.ny ? -gr.ny : 0);
    nz += nz < 0 ? gr.nz : (nz >= gr.nz ? -gr.nz : 0);
    const int nc = (nx * gr.ny + ny) * gr.nz + nz;
    sb[k] = cell_start[nc];
    sp[k + 1] = cell_count[nc];
  }
  for (int j = tid; j < hn; j += blockDim.x) acnt[j] = 0;
  const bool fit = hn <= BL_HMAX;
  if (fit) {
    for (int h = tid; h < hn; h += blockDim.x) {
      const int o = sorted_orig[hb + h];
      ho[h] = o;
      hxi[h] = xq[hb + h];
      hexm[h] = ex_win[o];
      hs4m[h] = s4_win[o];
      hbad[h] = win_bad[o];
    }
  }
  __syncthreads();
  if (tid == 0) {
    sp[0] = 0;
    for (int k = 0; k < a.nstencil; ++k) sp[k + 1] += sp[k];
  }
  __syncthreads();
  const int total = sp[a.nstencil];
  const float Lx = (float)gr.Lx, Ly = (float)gr.Ly, Lz = (float)gr.Lz;
  float4 pv[BL_SLOTS];
  int2 pt[BL_SLOTS];
  auto fetch = [&](int base) {
    const int nch = min(BL_CH, total - base);
#pragma unroll
    for (int r = 0; r < BL_SLOTS; ++r) {
      const int c = tid + r * BL_THREADS;
      if (c < nch) {
        const int idx = base + c;
        int lo = 0, hi = a.nstencil - 1;  // last stencil cell whose prefix <= idx
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (sp[mid] <= idx) lo = mid;
          else hi = mid - 1;
        }
        const int t = sb[lo] + (idx - sp[lo]);
        float4 v = xq[t];
        const int code = shc[lo];
        v.x += (float)((code & 3) - 1) * Lx;
        v.y += (float)(((code >> 2) & 3) - 1) * Ly;
        v.z += (float)(((code >> 4) & 3) - 1) * Lz;
        pv[r] = v;
        pt[r] = make_int2(t, sorted_orig[t]);
      }
    }
  };
  auto store = [&](int base) {
    const int nch = min(BL_CH, total - base);
    const int npad = min(BL_CH, (nch + BLF_PASS - 1) / BLF_PASS * BLF_PASS);
#pragma unroll
    for (int r = 0; r < BL_SLOTS; ++r) {
      const int c = tid + r * BL_THREADS;
      if (c < nch) {
        cxq[c] = pv[r];
        cto[c] = pt[r];
      } else if (c < npad) {  // far away: fails the distance test
        cxq[c] = make_float4(1e30f, 1e30f, 1e30f, 0.f);
        cto[c] = make_int2(0, 0);
      }
    }
  };
  if (total > 0) fetch(0);
  store(0);
  __syncthreads();
  const int lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: scalar registers
  constexpr int nw = BL_THREADS / 64;
  int* my_ex = exl[w];
  int* my_s4 = s4l[w];
  for (int base = 0; base < total; base += BL_CH) {
    const int nch = min(BL_CH, total - base);
    const bool more = base + BL_CH < total;
    if (more) fetch(base + BL_CH);  // in flight during this chunk's scan
    for (int h = w; h < hn; h += nw) {  // wave-uniform
      const int s = hb + h;
      int o;
      float4 xi;
      unsigned long long exm, s4m;
      bool masked;
      if (fit) {
        o = ho[h];
        xi = hxi[h];
        exm = hexm[h];
        s4m = hs4m[h];
        masked = hbad[h] == 0;
      } else {
        o = sorted_orig[s];
        xi = xq[s];
        exm = ex_win[o];
        s4m = s4_win[o];
        masked = win_bad[o] == 0;
      }
      unsigned* row = nbr_list + (long long)s * a.cap;
      int cnt = acnt[h];
      if (masked) {
        const unsigned o32 = __builtin_amdgcn_readfirstlane((unsigned)(o - 32));
        for (int cb = 0; cb < nch; cb += BLF_PASS) {
          float4 xj[BL_U];
          int2 tj[BL_U];
#pragma unroll
          for (int u = 0; u < BL_U; ++u) {
            xj[u] = cxq[cb + u * 64 + lane];
            tj[u] = cto[cb + u * 64 + lane];
          }
#pragma unroll
          for (int u = 0; u < BL_U; ++u) {
            const float dx = xj[u].x - xi.x, dy = xj[u].y - xi.y, dz = xj[u].z - xi.z;
            const bool near = dx * dx + dy * dy + dz * dz < a.rl2;
            const unsigned long long mn = __ballot(near);
            if (mn == 0) continue;  // wave-uniform: no candidate of these 64 inside the list radius
            const unsigned dw = min((unsigned)tj[u].y - o32, 63u);
            const bool kept = !((exm >> dw) & 1ull);
            const bool acc = near && kept;
            const unsigned long long m = mn & __ballot(kept);  // two compares, no re-materialised mask
            const int slot = cnt + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
            if (acc && slot < a.cap) row[slot] = (unsigned)tj[u].x | ((unsigned)((s4m >> dw) & 1ull) << 31);
            cnt += __popcll(m);
          }
        }
      } else {  // a partner outside the index window: the CSR lists
        const int e0 = excl_ptr ? excl_ptr[o] : 0, t0s = s14_ptr ? s14_ptr[o] : 0;
        const int nex_all = excl_ptr ? excl_ptr[o + 1] - e0 : 0, ns4_all = s14_ptr ? s14_ptr[o + 1] - t0s : 0;
        const int nex = min(nex_all, BL_MAXEX), ns4 = min(ns4_all, BL_MAXEX);
        if (lane < nex) my_ex[lane] = excl_idx[e0 + lane];
        if (lane < ns4) my_s4[lane] = s14_idx[t0s + lane];
        if (lane == 0 && (nex_all > BL_MAXEX || ns4_all > BL_MAXEX)) atomicOr(&st->flags, MDX_FLAG_OVERFLOW);
        wave_lds_sync();
        for (int cb = 0; cb < nch; cb += 64) {
          const float4 xj = cxq[cb + lane];
          const int2 tj = cto[cb + lane];
          const float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
          bool acc = dx * dx + dy * dy + dz * dz < a.rl2 && tj.x != s;
          bool s14 = false;
          for (int r = 0; r < nex; ++r) acc &= (my_ex[r] != tj.y);
          for (int r = 0; r < ns4; ++r) s14 |= (my_s4[r] == tj.y);
          const unsigned long long m = __ballot(acc);
          const int slot = cnt + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
          if (acc && slot < a.cap) row[slot] = (unsigned)tj.x | (s14 ? 0x80000000u : 0u);
          cnt += __popcll(m);
        }
        wave_lds_sync();  // the wave's next home atom overwrites my_ex / my_s4
      }
      if (lane == 0) acnt[h] = cnt;
    }
    __syncthreads();
    if (more) {
      store(base + BL_CH);
      __syncthreads();
    }
  }
  for (int h = tid; h < hn; h += blockDim.x) {
    const int cnt = acnt[h];
    nbr_count[hb + h] = min(cnt, a.cap);
    if (cnt > a.cap) atomicOr(&st->flags, MDX_FLAG_OVERFLOW);
  }
}

__global__ __launch_bounds__(BL_THREADS) void k_build_list_fast(ListCall L) { build_list_fast_body(L, blockIdx.x); }

// every replica's list build in one launch (the fast build, or the cell build of a grid whose
// stencil wraps onto itself)
__global__ __launch_bounds__(BL_THREADS) void k_build_list_fast_batch(BatchMap map, const ListCall* __restrict__ calls) {
  int cell;
  const int r = batch_replica(map, &cell);
  build_list_fast_body(calls[r], cell);
}
__global__ __launch_bounds__(BL_THREADS) void k_build_list_cell_batch(BatchMap map, const ListCall* __restrict__ calls) {
  int cell;
  const int r = batch_replica(map, &cell);
  build_list_cell_body<false, false>(calls[r], cell);
}

__global__ void k_request_rebuild(DevState* st) { st->rebuild = 1; }

int mdx_enqueue_request_rebuild(mdx_ctx* c) {
  k_request_rebuild<<<1, 1, 0, c->stream>>>(c->dstate.p);
  MDX_LAUNCH_CHECK();
  return 0;
}

__global__ void k_clear_rebuild(DevState* st) {
  st->rebuild = 0;
  st->max_disp2_bits = 0;
}

// zero the per-cell counts of the list grid when a rebuild is due (a kernel, not a memset, so
// the step stays safe to capture in a hipGraph) and count rebuilds
__global__ void k_prep_rebuild(DevState* st, double thr2, int* __restrict__ cell_count, int ncell) {
  if (!need_rebuild(st, thr2)) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ncell) cell_count[i] = 0;
  if (i == 0) st->nrebuild += 1;
}


// ============================================================== host side
static void make_grid(mdx_ctx* c, GridArgs* g) {
  g->nx = c->gn[0];
  g->ny = c->gn[1];
  g->nz = c->gn[2];
  g->Lx = c->box[0];
  g->Ly = c->box[1];
  g->Lz = c->box[2];
  g->wx = c->box[0] / c->gn[0];
  g->wy = c->box[1] / c->gn[1];
  g->wz = c->box[2] / c->gn[2];
}

int mdx_nlist_prepare(mdx_ctx* c) {
  double rl = c->rc + c->skin;
  for (int d = 0; d < 3; ++d) {
    int k = (int)std::floor(c->box[d] * c->cells_per_rl / rl);
    if (k < 1) k = 1;
    c->gn[d] = k;
  }
  c->ncell = c->gn[0] * c->gn[1] * c->gn[2];
  // stencil: offsets whose cells can hold an atom within rl of the home cell, each distinct
  // periodic cell once
  c->stencil.clear();
  int lo[3], hi[3];
  double w[3];
  for (int d = 0; d < 3; ++d) {
    w[d] = c->box[d] / c->gn[d];
    int r = (int)std::ceil(rl / w[d]);
    if (2 * r + 1 >= c->gn[d]) {
      lo[d] = -((c->gn[d] - 1) / 2);
      hi[d] = c->gn[d] / 2;
    } else {
      lo[d] = -r;
      hi[d] = r;
    }
    c->reach[d] = r;
  }
  for (int dx = lo[0]; dx <= hi[0]; ++dx)
    for (int dy = lo[1]; dy <= hi[1]; ++dy)
      for (int dz = lo[2]; dz <= hi[2]; ++dz) {
        double mx = std::max(0, std::abs(dx) - 1) * w[0];
        double my = std::max(0, std::abs(dy) - 1) * w[1];
        double mz = std::max(0, std::abs(dz) - 1) * w[2];
        if (mx * mx + my * my + mz * mz > rl * rl) continue;
        c->stencil.push_back(((dx + 128) << 16) | ((dy + 128) << 8) | (dz + 128));
      }
  int rc;
  if ((rc = c->d_stencil.ensure(c->stencil.size()))) return rc;
  MDX_HIP(hipMemcpy(c->d_stencil.p, c->stencil.data(), c->stencil.size() * sizeof(int), hipMemcpyHostToDevice));
  if ((rc = c->cell_count.ensure(c->ncell)) || (rc = c->cell_start.ensure(c->ncell)) ||
      (rc = c->cell_cursor.ensure(c->ncell)) || (rc = c->atom_cell.ensure(c->n)) ||
      (rc = c->sorted_orig.ensure(c->n)) || (rc = c->orig_sorted.ensure(c->n)) || (rc = c->xq.ensure(c->n)) ||
      (rc = c->ljs.ensure(c->n)) || (rc = c->ljs14.ensure(c->n)) || (rc = c->ref_cell_sorted.ensure(c->n)) ||
      (rc = c->pos_at_build.ensure(3 * (size_t)c->n)) || (rc = c->nbr_count.ensure(c->n)) ||
      (rc = c->pos_sorted.ensure(3 * (size_t)c->n)))
    return rc;
  // capacity: expected neighbours in a sphere of radius rl at the system density, with margin
  double vol = c->box[0] * c->box[1] * c->box[2];
  double dens = c->n / vol;
  double expect = 4.0 / 3.0 * M_PI * rl * rl * rl * dens;
  int cap = (int)(expect * 1.5 + 64);
  if (cap > c->n) cap = c->n;
  if (cap < c->nbr_cap) cap = c->nbr_cap;
  c->nbr_cap = cap;
  c->list_stride = cap;
  // slack after the last row: the nonbonded kernel loads the next chunk's indices ahead
  if ((rc = c->nbr_list.ensure((size_t)cap * c->n + MDX_LIST_SLACK))) return rc;
  MDX_HIP(hipMemset(c->pos_at_build.p, 0, 3 * sizeof(double) * c->n));
  int one = 1;
  MDX_HIP(hipMemcpy(&c->dstate.p->rebuild, &one, sizeof(int), hipMemcpyHostToDevice));
  c->list_valid = true;
  return 0;
}

int mdx_nlist_enqueue_pack(mdx_ctx* c) {
  k_pack<<<div_up(c->n, NL_BLOCK), NL_BLOCK, 0, c->stream>>>(c->pos, c->sorted_orig.p,
                                                            c->has_charges ? c->charge_s.p : nullptr, c->n,
                                                            c->box[0], c->box[1], c->box[2], c->xq.p,
                                                            c->pos_sorted.p);
  MDX_LAUNCH_CHECK();
  return 0;
}

// squared displacement that triggers a rebuild; skin == 0 (reference pair set or no buffer):
// rebuild every evaluation
static double rebuild_thr2(const mdx_ctx* c) {
  const double half = 0.5 * c->skin;
  return c->skin > 0 ? half * half : -1.0;
}

// Enqueue: displacement check, conditional rebuild of the list, pack. Leaves the rebuild flag set
// for the consumer (k_clear_rebuild is enqueued by the caller after the last reader).
int mdx_nlist_update(mdx_ctx* c, bool force_rebuild) {
  int rc = mdx_nlist_update_order(c, force_rebuild);
  return rc ? rc : mdx_nlist_update_list(c);
}

// First part: displacement check and, on a rebuild, binning, per-cell sort and fill of the
// sorted-order arrays; then the pack. Everything the PME plane kernels read is final after it.
int mdx_nlist_update_order(mdx_ctx* c, bool force_rebuild) {
  MDX_CHECK(c->nb_configured, MDX_ERR_INVALID, "nonbonded not configured");
  const double thr2 = rebuild_thr2(c);
  if (force_rebuild) {
    int rc0 = mdx_enqueue_request_rebuild(c);
    if (rc0) return rc0;
  }
  GridArgs g;
  make_grid(c, &g);
  int nb = div_up(c->n, NL_BLOCK);
  DevState* st = c->dstate.p;
  const bool pro = c->prologue_done;  // displacement, rebuild decision and packing already done
  if (!pro) {
    k_check_disp<<<nb, NL_BLOCK, 0, c->stream>>>(c->pos, c->pos_at_build.p, c->n, c->box[0], c->box[1], c->box[2], st);
    MDX_LAUNCH_CHECK();
    k_prep_rebuild<<<div_up(c->ncell, NL_BLOCK), NL_BLOCK, 0, c->stream>>>(st, thr2, c->cell_count.p, c->ncell);
    MDX_LAUNCH_CHECK();
  }
  // reference cells are needed for the REFERENCE stencil filter; they depend on the minimum
  // position, so REFERENCE mode runs with skin 0 (rebuild every evaluation) and re-bins here.
  if (c->pair_mode == MDX_PAIRS_REFERENCE) {
    int rc = ref_bin_enqueue(c, c->cell_cutoff, nullptr);
    if (rc) return rc;
  }
  if (!c->prologue_binned) {  // else the step prologue's last workgroup binned the atoms
    k_grid_bin<<<nb, NL_BLOCK, 0, c->stream>>>(c->pos, c->n, g, st, thr2, c->atom_cell.p, c->cell_count.p);
    MDX_LAUNCH_CHECK();
    k_cell_scan<<<1, 1024, 0, c->stream>>>(c->cell_count.p, c->cell_start.p, c->cell_cursor.p, c->ncell, st, st,
                                           thr2);
    MDX_LAUNCH_CHECK();
    k_grid_scatter<<<nb, NL_BLOCK, 0, c->stream>>>(c->atom_cell.p, c->n, c->cell_start.p, c->cell_cursor.p,
                                                   c->sorted_orig.p, st, thr2);
    MDX_LAUNCH_CHECK();
  }
  bool ref = c->pair_mode == MDX_PAIRS_REFERENCE;
  const FillArgs fa = fill_args(c, thr2);
  k_grid_sort_fill<<<c->ncell, 256, 0, c->stream>>>(fa);
  MDX_LAUNCH_CHECK();
  return pro ? 0 : mdx_nlist_enqueue_pack(c);
}

// The list-build launch of a context: its arguments and kernel (0: the fast build; 1-4: the cell
// build <EXACT64, FAST> = <false,false>, <false,true>, <true,false>, <true,true>).
static int list_call(mdx_ctx* c, ListCall* L, int* kind) {
  const bool ref = c->pair_mode == MDX_PAIRS_REFERENCE;
  ListArgs& a = L->a;
  make_grid(c, &a.g);
  a.n = c->n;
  a.cap = c->nbr_cap;
  double rl = c->rc + c->skin;
  a.rl2 = (float)(rl * rl) * (1.0f + 1e-5f);
  a.exact64 = c->skin > 0 ? 0 : 1;
  a.rc = c->rc;
  a.ref_mode = ref ? 1 : 0;
  a.rnx = c->ref_n[0];
  a.rny = c->ref_n[1];
  a.rnz = c->ref_n[2];
  a.nstencil = (int)c->stencil.size();
  MDX_CHECK(a.nstencil <= BL_MAXST, MDX_ERR_INVALID, "stencil of %d cells exceeds %d", a.nstencil, BL_MAXST);
  // each periodic cell appears once in the stencil, on one side only: a fixed shift per stencil
  // cell maps its atoms to the image nearest the home cell
  a.shifted = 1;
  for (int d = 0; d < 3; ++d)
    if (2 * c->reach[d] + 1 >= c->gn[d]) a.shifted = 0;
  const bool win = c->win_ready && c->h_win_bad.size() == (size_t)c->n;
  const bool fastk = a.shifted && !a.ref_mode && win;
  L->xq = c->xq.p;
  L->sorted_orig = c->sorted_orig.p;
  L->cell_start = c->cell_start.p;
  L->cell_count = c->cell_count.p;
  L->stencil = c->d_stencil.p;
  L->excl_ptr = c->has_excl ? c->excl_ptr.p : nullptr;
  L->excl_idx = c->excl_idx.p;
  L->s14_ptr = c->has_s14 ? c->s14_ptr.p : nullptr;
  L->s14_idx = c->s14_idx.p;
  L->ex_win = win ? c->ex_win.p : nullptr;
  L->s4_win = win ? c->s4_win.p : nullptr;
  L->win_bad = win ? c->win_bad.p : nullptr;
  L->ref_cell_sorted = c->ref_cell_sorted.p;
  L->pos = c->pos;
  L->nbr_count = c->nbr_count.p;
  L->nbr_list = c->nbr_list.p;
  L->st = c->dstate.p;
  L->thr2 = rebuild_thr2(c);
  *kind = fastk && !a.exact64 ? 0 : 1 + (a.exact64 ? 2 : 0) + (fastk ? 1 : 0);
  return 0;
}

// Second part: the neighbour-list build (its workgroups return at once unless rebuilding).
int mdx_nlist_update_list(mdx_ctx* c) {
  ListCall L;
  int kind, rc = list_call(c, &L, &kind);
  if (rc) return rc;
  dim3 grid(c->ncell), block(BL_THREADS);
  switch (kind) {
    case 0: k_build_list_fast<<<grid, block, 0, c->stream>>>(L); break;
    case 1: k_build_list_cell<false, false><<<grid, block, 0, c->stream>>>(L); break;
    case 2: k_build_list_cell<false, true><<<grid, block, 0, c->stream>>>(L); break;
    case 3: k_build_list_cell<true, false><<<grid, block, 0, c->stream>>>(L); break;
    default: k_build_list_cell<true, true><<<grid, block, 0, c->stream>>>(L); break;
  }
  MDX_LAUNCH_CHECK();
  return 0;
}

// ---- batched replicas: the sort/fill and list-build launches of every replica
__global__ __launch_bounds__(256) void k_grid_sort_fill_batch(BatchMap map, const FillArgs* __restrict__ calls) {
  int cell;
  const int r = batch_replica(map, &cell);
  const FillArgs& f = calls[r];
  if (!need_rebuild(f.st, f.thr2)) return;
  grid_sort_fill_block(f, cell);
}

int mdx_batch_prepare_list(mdx_ctx* const* cs, int n, BatchKind* fill, BatchKind* fast, BatchKind* cell) {
  std::vector<FillArgs> fa;
  std::vector<ListCall> lf, lc;
  std::vector<int> bfa, bf, bc;
  for (int r = 0; r < n; ++r) {
    mdx_ctx* c = cs[r];
    fa.push_back(fill_args(c, rebuild_thr2(c)));
    bfa.push_back(c->ncell);
    ListCall L;
    int kind, rc = list_call(c, &L, &kind);
    if (rc) return rc;
    MDX_CHECK(kind <= 1, MDX_ERR_INVALID, "replica %d: batched steps need a list radius larger than the pair cutoff", r);
    // every replica holds a place in both builds (zero workgroups in the one it does not use)
    lf.push_back(L);
    lc.push_back(L);
    bf.push_back(kind == 0 ? c->ncell : 0);
    bc.push_back(kind == 1 ? c->ncell : 0);
  }
  int rc;
  if ((rc = batch_kind_upload(fill, fa, bfa, 0)) || (rc = batch_kind_upload(fast, lf, bf, 0)) ||
      (rc = batch_kind_upload(cell, lc, bc, 0)))
    return rc;
  return 0;
}

int mdx_batch_launch_fill(const BatchKind& k, hipStream_t s) {
  if (!k.blocks) return 0;
  k_grid_sort_fill_batch<<<k.blocks, 256, 0, s>>>(k.map, (const FillArgs*)k.calls);
  MDX_LAUNCH_CHECK();
  return 0;
}

int mdx_batch_launch_list(const BatchKind& fast, const BatchKind& cell, hipStream_t s) {
  if (fast.blocks) {
    k_build_list_fast_batch<<<fast.blocks, BL_THREADS, 0, s>>>(fast.map, (const ListCall*)fast.calls);
    MDX_LAUNCH_CHECK();
  }
  if (cell.blocks) {
    k_build_list_cell_batch<<<cell.blocks, BL_THREADS, 0, s>>>(cell.map, (const ListCall*)cell.calls);
    MDX_LAUNCH_CHECK();
  }
  return 0;
}

int mdx_nlist_enqueue_clear(mdx_ctx* c) {
  k_clear_rebuild<<<1, 1, 0, c->stream>>>(c->dstate.p);
  MDX_LAUNCH_CHECK();
  return 0;
}

// ---------------------------------------------------------------- pair enumeration (parity)
__global__ void k_emit_pairs(int n, int cap, const int* __restrict__ nbr_count, const unsigned* __restrict__ nbr_list,
                             const int* __restrict__ sorted_orig, const float4* __restrict__ xq, const double* __restrict__ pos,
                             double Lx, double Ly, double Lz, double rc, int* __restrict__ out,
                             long long out_cap, unsigned long long* __restrict__ counter) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int o = sorted_orig[s];
  int cnt = nbr_count[s];
  double L[3] = {Lx, Ly, Lz};
  for (int k = 0; k < cnt; ++k) {
    unsigned e = nbr_list[(long long)s * cap + k];
    int t = (int)(e & 0x7fffffffu);
    int jo = sorted_orig[t];
    if (jo <= o) continue;  // unique pairs, i < j (original ids)
    double d[3];
    for (int q = 0; q < 3; ++q) {
      double v = pos[3 * jo + q] - pos[3 * o + q];
      double sv = v * (1.0 / L[q]);
      sv -= rint(sv);
      d[q] = sv * L[q];
    }
    double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (!(r <= rc)) continue;
    unsigned long long slot = atomicAdd(counter, 1ull);
    if ((long long)slot < out_cap) {
      out[2 * slot] = o;
      out[2 * slot + 1] = jo;
    }
  }
}

extern "C" int mdx_neighbor_pairs(mdx_ctx* c, int* d_pairs, int64_t capacity, int64_t* h_count) {
  MDX_CHECK(c && c->pos && c->nb_configured, MDX_ERR_INVALID, "mdx_neighbor_pairs: system/nonbonded not set");
  MDX_ENTER();
  c->forces_current = false;
  MDX_HIP(hipMemsetAsync(&c->dstate.p->flags, 0, sizeof(unsigned), c->stream));
  int rc = mdx_nlist_update(c, true);
  if (rc) return rc;
  if ((rc = c->tmp64.ensure(1))) return rc;
  MDX_HIP(hipMemsetAsync(c->tmp64.p, 0, sizeof(long long), c->stream));
  k_emit_pairs<<<div_up(c->n, NL_BLOCK), NL_BLOCK, 0, c->stream>>>(
      c->n, c->nbr_cap, c->nbr_count.p, c->nbr_list.p, c->sorted_orig.p, c->xq.p, c->pos, c->box[0], c->box[1],
      c->box[2], c->rc, d_pairs, capacity, (unsigned long long*)c->tmp64.p);
  MDX_LAUNCH_CHECK();
  rc = mdx_nlist_enqueue_clear(c);
  if (rc) return rc;
  long long cnt = 0;
  unsigned flags = 0;
  MDX_HIP(hipMemcpyAsync(&cnt, c->tmp64.p, sizeof(long long), hipMemcpyDeviceToHost, c->stream));
  MDX_HIP(hipMemcpyAsync(&flags, &c->dstate.p->flags, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
  MDX_HIP(hipStreamSynchronize(c->stream));
  MDX_CHECK(!(flags & MDX_FLAG_OVERFLOW), MDX_ERR_OVERFLOW, "neighbour list or cell capacity exceeded");
  *h_count = cnt;
  MDX_LEAVE_RET(0);
}
