// Pair-distance histograms for the RDF and coordination-number analysers
// (mdpy/analyser/rdf_analyser.py:34-85, coordination_number_analyser.py:34-80): for every
// selected atom i of set 1 and every atom j of set 2, the minimum-image distance in float64 with
// the reference's sequence (pbc.py:38-44: s = d/L, s -= rint(s), d = s L; r = sqrt(x^2 + y^2 +
// z^2)), binned exactly as numpy.histogram bins it on uniform edges (index from the scaled
// distance, then moved down or up one bin where the edge comparison disagrees; r outside
// [0, rc] dropped, r == rc in the last bin). Counts are int64 per (i, bin), summed over frames,
// so the result is an integer array that matches the reference bit for bit.
#include "mdx_internal.h"

#define HIST_BLOCK 256
#define HIST_TILE 1024  // set-2 positions staged in LDS per pass

__global__ __launch_bounds__(HIST_BLOCK) void k_pair_hist(const double* __restrict__ pos, int n_atoms,
                                                          const int* __restrict__ sel1, int n1,
                                                          const int* __restrict__ sel2, int n2, double Lx,
                                                          double Ly, double Lz, const double* __restrict__ edges,
                                                          int nbins, unsigned long long* __restrict__ counts) {
  __shared__ double tx[HIST_TILE], ty[HIST_TILE], tz[HIST_TILE];
  const double* x = pos + (long long)blockIdx.y * n_atoms * 3;  // frame
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = a < n1;
  double xi = 0, yi = 0, zi = 0;
  if (active) {
    const int i = sel1[a];
    xi = x[3 * i];
    yi = x[3 * i + 1];
    zi = x[3 * i + 2];
  }
  const double first = edges[0], last = edges[nbins];
  const double norm = (double)nbins / (last - first);
  const double iLx = 1.0 / Lx, iLy = 1.0 / Ly, iLz = 1.0 / Lz;
  for (int t0 = 0; t0 < n2; t0 += HIST_TILE) {
    const int tn = min(HIST_TILE, n2 - t0);
    __syncthreads();
    for (int k = threadIdx.x; k < tn; k += blockDim.x) {
      const int j = sel2[t0 + k];
      tx[k] = x[3 * j];
      ty[k] = x[3 * j + 1];
      tz[k] = x[3 * j + 2];
    }
    __syncthreads();
    if (!active) continue;
    for (int k = 0; k < tn; ++k) {
      // vec = x_i - x_j (rdf_analyser.py:48-52), minimum image
      double dx = xi - tx[k], dy = yi - ty[k], dz = zi - tz[k];
      double s = dx * iLx;
      s -= rint(s);
      dx = s * Lx;
      s = dy * iLy;
      s -= rint(s);
      dy = s * Ly;
      s = dz * iLz;
      s -= rint(s);
      dz = s * Lz;
      // rounded products and sums in numpy's order ((x^2 + y^2) + z^2), never contracted to FMAs
      const double r = sqrt(__dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz)));
      if (!(r >= first && r <= last)) continue;
      const double f = (r - first) * norm;
      int b = f < (double)nbins ? (int)f : nbins;  // f >= 0 here
      if (b == nbins) b = nbins - 1;
      if (r < edges[b]) --b;
      else if (b != nbins - 1 && r >= edges[b + 1]) ++b;
      atomicAdd(&counts[(long long)a * nbins + b], 1ull);
    }
  }
}

extern "C" int mdx_pair_histogram(void* hip_stream, int n_frames, int n_atoms, const double* d_pos,
                                  const double* h_box, const int* d_sel1, int n1, const int* d_sel2, int n2,
                                  const double* d_edges, int nbins, int64_t* d_counts) {
  MDX_CHECK(n_frames > 0 && n_atoms > 0 && d_pos && h_box && d_sel1 && d_sel2 && d_edges && d_counts && n1 > 0 &&
                n2 > 0 && nbins > 0,
            MDX_ERR_INVALID, "mdx_pair_histogram: bad arguments");
  MDX_CHECK(n_frames <= 65535, MDX_ERR_INVALID, "mdx_pair_histogram: at most 65535 frames per call");
  for (int d = 0; d < 3; ++d) MDX_CHECK(h_box[d] > 0, MDX_ERR_INVALID, "box edges must be positive");
  hipStream_t s = (hipStream_t)hip_stream;
  dim3 g(div_up(n1, HIST_BLOCK), n_frames), b(HIST_BLOCK);
  k_pair_hist<<<g, b, 0, s>>>(d_pos, n_atoms, d_sel1, n1, d_sel2, n2, h_box[0], h_box[1], h_box[2], d_edges, nbins,
                              (unsigned long long*)d_counts);
  MDX_LAUNCH_CHECK();
  return 0;
}
