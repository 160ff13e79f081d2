// Reference cell-list binning without any device allocation (no counterpart in the reference:
// for callers that cannot allocate device memory, such as code around stream capture or tight
// serving loops). Same kernels and bit-exact results as mdx_celllist_reference; the reductions
// (per-axis minimum position, largest cell population, flags) and the per-cell counts live in
// one file-scope device array instead of a context.
#include "mdx_internal.h"
#include <cmath>
#include <mutex>

// layout: DevState (the kernels' reduction fields) in the first 256 bytes, then the per-cell
// counts; 16,384 cells = 64 KiB (a reference grid of up to 25 x 25 x 25 cells)
#define NOALLOC_MAX_CELLS 16384
#define NOALLOC_HEAD 256
static_assert(sizeof(DevState) <= NOALLOC_HEAD, "DevState must fit the scratch header");
static __device__ __align__(256) unsigned char scratch[NOALLOC_HEAD + NOALLOC_MAX_CELLS * sizeof(int)];

// per device: the device address of `scratch` (the host-side name is only a stand-in)
#define NOALLOC_MAX_DEVICES 64
static unsigned char* g_scratch_ptr[NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

static int scratch_ptr(unsigned char** out) {
  int dev = 0;
  MDX_HIP(hipGetDevice(&dev));
  MDX_CHECK(dev >= 0 && dev < NOALLOC_MAX_DEVICES, MDX_ERR_INVALID, "device %d beyond %d", dev, NOALLOC_MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    MDX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    MDX_CHECK(p != nullptr, MDX_ERR_HIP, "hipGetSymbolAddress returned no address");
    g_scratch_ptr[dev] = (unsigned char*)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

// kernels shared with mdx_celllist_reference (mdx_nlist.hip)
__global__ void k_ref_reset(DevState* st, int* __restrict__ counts, int nc);
__global__ void k_minpos(const double* __restrict__ pos, int n, DevState* st);
__global__ void k_ref_bin(const double* __restrict__ pos, int n, int nx, int ny, int nz, double ix, double iy,
                          double iz, DevState* st, int* __restrict__ cell_index, int* __restrict__ cell_lin,
                          int* __restrict__ counts);

extern "C" int mdx_celllist_reference_noalloc(void* hip_stream, int n, const double* d_pos, const double* h_box,
                                              double cell_cutoff, int* d_cell_index, int* h_n, int* h_maxpc) {
  MDX_CHECK(n > 0 && d_pos && h_box && d_cell_index && h_n && h_maxpc, MDX_ERR_INVALID,
            "mdx_celllist_reference_noalloc: bad arguments");
  MDX_CHECK(cell_cutoff > 0, MDX_ERR_CELL_CUTOFF, "cutoff radius is poor defined");
  int nn[3];
  double w[3], inv[3];
  for (int d = 0; d < 3; ++d) {
    MDX_CHECK(h_box[d] > 0, MDX_ERR_INVALID, "box edges must be positive");
    MDX_CHECK(std::floor(h_box[d] / cell_cutoff) >= 2, MDX_ERR_CELL_CUTOFF,
              "the cutoff_radius is too large to create cell list");
  }
  mdx_ref_grid(h_box, cell_cutoff, nn, w, inv);
  const long long nc = (long long)nn[0] * nn[1] * nn[2];
  MDX_CHECK(nc <= NOALLOC_MAX_CELLS, MDX_ERR_INVALID,
            "mdx_celllist_reference_noalloc: %lld cells exceed the fixed scratch of %d (use mdx_celllist_reference)",
            nc, NOALLOC_MAX_CELLS);
  unsigned char* base = nullptr;
  int rc = scratch_ptr(&base);
  if (rc) return rc;
  DevState* st = (DevState*)base;
  int* counts = (int*)(base + NOALLOC_HEAD);
  hipStream_t s = (hipStream_t)hip_stream;
  const int B = 256;
  // the reset kernel clears the flags, the minima, the largest population and the counts left
  // by the previous call on this device
  k_ref_reset<<<div_up(nc, B), B, 0, s>>>(st, counts, (int)nc);
  MDX_LAUNCH_CHECK();
  k_minpos<<<div_up(n, B), B, 0, s>>>(d_pos, n, st);
  MDX_LAUNCH_CHECK();
  k_ref_bin<<<div_up(n, B), B, 0, s>>>(d_pos, n, nn[0], nn[1], nn[2], inv[0], inv[1], inv[2], st, d_cell_index,
                                       nullptr, counts);
  MDX_LAUNCH_CHECK();
  int maxpc = 0;
  unsigned flags = 0;
  MDX_HIP(hipMemcpyAsync(&maxpc, &st->maxpc, sizeof(int), hipMemcpyDeviceToHost, s));
  MDX_HIP(hipMemcpyAsync(&flags, &st->flags, sizeof(unsigned), hipMemcpyDeviceToHost, s));
  MDX_HIP(hipStreamSynchronize(s));
  MDX_CHECK(!(flags & MDX_FLAG_OVERFLOW), MDX_ERR_OVERFLOW, "reference binning produced an out-of-range cell");
  for (int d = 0; d < 3; ++d) h_n[d] = nn[d];
  *h_maxpc = maxpc;
  return 0;
}
