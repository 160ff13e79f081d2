// Smooth particle-mesh Ewald (Essmann et al., J. Chem. Phys. 103, 8577 (1995)), gfx950.
// mdpy has no PME (mdpy/constraint/__init__.py:22 names it; nothing implements it), so this
// is new physics, checked against the float64 oracle in oracle/ewald.py.
//
//   spread   order-4 cardinal B-splines, q_i * sqrt(ke) added to an int64 fixed-point grid
//            (deterministic), one lane per atom
//   FFT      hand-written mixed-radix (2,3,4,5) Stockham passes in LDS, one kernel per axis;
//            the x-axis kernel also applies the influence function B*C, sums the reciprocal
//            energy and runs the inverse x transform on the same LDS tile
//   gather   one lane per atom, 64 grid values, force = -q grad(theta) . conv
//   corrections  float64 excluded-pair erf correction, self energy and neutralising background,
//                launched with the bonded terms (mdx_bonded.hip)
#include "mdx_internal.h"
#include "mdx_bonded_terms.h"
#include "mdx_nonbonded_body.h"
#include <cmath>

#define PME_BLOCK 256
#define FFT_THREADS 256
#define MAX_RADICES 16

// radix of pass p in bits [4p, 4p+4) of `code` (a register-friendly stand-in for an array that
// a dynamically indexed kernel argument would otherwise put in scratch memory)
struct FftPlan {
  int K;
  int nrad;
  unsigned long long code;
};

struct FftLines {
  int nlines;   // number of lines
  int B;        // lines per block
  int inner, inner_stride, outer_stride, estride;
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 mul_i(float2 a, float s) { return make_float2(-s * a.y, s * a.x); }  // (i s) a

// small DFTs, sign s = -1 forward, +1 inverse: U_m = sum_r u_r exp(s 2 pi i r m / R)
__device__ __forceinline__ void dft2(float2* u) {
  float2 a = u[0], b = u[1];
  u[0] = cadd(a, b);
  u[1] = csub(a, b);
}
__device__ __forceinline__ void dft3(float2* u, float s) {
  const float h = 0.86602540378443865f;  // sin(2 pi / 3)
  float2 t1 = cadd(u[1], u[2]);
  float2 t2 = csub(u[0], cscale(t1, 0.5f));
  float2 t3 = mul_i(csub(u[1], u[2]), s * h);
  u[0] = cadd(u[0], t1);
  u[1] = cadd(t2, t3);
  u[2] = csub(t2, t3);
}
__device__ __forceinline__ void dft4(float2* u, float s) {
  float2 a = cadd(u[0], u[2]), b = csub(u[0], u[2]);
  float2 c = cadd(u[1], u[3]), d = mul_i(csub(u[1], u[3]), s);
  u[0] = cadd(a, c);
  u[2] = csub(a, c);
  u[1] = cadd(b, d);
  u[3] = csub(b, d);
}
__device__ __forceinline__ void dft5(float2* u, float s) {
  const float c1 = 0.30901699437494742f, c2 = -0.80901699437494742f;
  const float s1 = 0.95105651629515357f, s2 = 0.58778525229247313f;
  float2 a1 = cadd(u[1], u[4]), b1 = csub(u[1], u[4]);
  float2 a2 = cadd(u[2], u[3]), b2 = csub(u[2], u[3]);
  float2 u0 = u[0];
  float2 p1 = cadd(u0, cadd(cscale(a1, c1), cscale(a2, c2)));
  float2 p2 = cadd(u0, cadd(cscale(a1, c2), cscale(a2, c1)));
  float2 q1 = mul_i(cadd(cscale(b1, s1), cscale(b2, s2)), s);
  float2 q2 = mul_i(csub(cscale(b1, s2), cscale(b2, s1)), s);
  u[0] = cadd(u0, cadd(a1, a2));
  u[1] = cadd(p1, q1);
  u[4] = csub(p1, q1);
  u[2] = cadd(p2, q2);
  u[3] = csub(p2, q2);
}

// one Stockham pass of radix R over B lines (element e of line l at l*ls + e*es); p is the
// product of the radices of the passes already done
template <int DIR, int R>
__device__ __forceinline__ void stockham_pass(const float2* src, float2* dst, int K, int p, int B,
                                              const float2* __restrict__ tw, int ls, int es) {
  const int T = K / R;
  const int tstep = K / (p * R);
  const float s = (float)DIR;
  for (int b = threadIdx.x; b < B * T; b += blockDim.x) {
    int l, i;
    if (es == 1) { l = b / T; i = b - l * T; }
    else { i = b / B; l = b - i * B; }
    const int k = i % p;
    const float2* X = src + l * ls;
    float2* Y = dst + l * ls;
    float2 u[R];
#pragma unroll
    for (int r = 0; r < R; ++r) u[r] = X[(i + r * T) * es];
#pragma unroll
    for (int r = 1; r < R; ++r) {
      float2 w = tw[(r * k * tstep) % K];
      if (DIR > 0) w.y = -w.y;
      u[r] = cmul(u[r], w);
    }
    if constexpr (R == 2) dft2(u);
    else if constexpr (R == 3) dft3(u, s);
    else if constexpr (R == 4) dft4(u, s);
    else dft5(u, s);
    const int j = (i - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) Y[(j + r * p) * es] = u[r];
  }
}

// B lines of length K held in LDS (src); element e of line l sits at l*ls + e*es. Stockham
// passes ping-pong with dst; returns the buffer holding the result. Consecutive threads take
// consecutive elements of a line when es == 1 and consecutive lines otherwise, so LDS accesses
// stay on distinct banks for both contiguous and strided lines.
template <int DIR>
__device__ float2* lds_fft(float2* src, float2* dst, const FftPlan& pl, int B, const float2* __restrict__ tw,
                           int ls = -1, int es = 1) {
  const int K = pl.K;
  if (ls < 0) ls = K;
  int p = 1;
  for (int pass = 0; pass < pl.nrad; ++pass) {
    const int R = (int)((pl.code >> (4 * pass)) & 15ull);
    switch (R) {
      case 2: stockham_pass<DIR, 2>(src, dst, K, p, B, tw, ls, es); break;
      case 3: stockham_pass<DIR, 3>(src, dst, K, p, B, tw, ls, es); break;
      case 4: stockham_pass<DIR, 4>(src, dst, K, p, B, tw, ls, es); break;
      default: stockham_pass<DIR, 5>(src, dst, K, p, B, tw, ls, es); break;
    }
    __syncthreads();
    float2* t = src;
    src = dst;
    dst = t;
    p *= R;
  }
  return src;
}

// ---- 64 x 64 planes: each wave transforms 4 whole lines, so the three radix-4 Stockham passes
// of a 64-point line need only wave-level ordering (no workgroup barrier between passes), and
// every index is a compile-time shift or mask. Planes are held with row stride 65 (PLANE64_LS)
// so that both rows and columns are read across lanes without LDS bank conflicts.
#define PLANE64_LS 65

// one radix-4 pass (p = 4^S) of the 64-point line at X/Y + off, elements es apart; lane i of
// the line's 16 computes butterfly i
template <int DIR, int S>
__device__ __forceinline__ void fft64_pass(const float2* X, float2* Y, int off, int es, int i,
                                           const float2* __restrict__ tw) {
  constexpr int p = 1 << (2 * S);
  constexpr int tstep = 16 >> (2 * S);
  const int k = i & (p - 1);
  float2 u[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) u[r] = X[off + (i + 16 * r) * es];
  if (S > 0) {
#pragma unroll
    for (int r = 1; r < 4; ++r) {
      float2 w = tw[(r * k * tstep) & 63];
      if (DIR > 0) w.y = -w.y;
      u[r] = cmul(u[r], w);
    }
  }
  dft4(u, (float)DIR);
  const int j = ((i - k) << 2) + k;
#pragma unroll
  for (int r = 0; r < 4; ++r) Y[off + (j + r * p) * es] = u[r];
}

// 64-point transforms of the 4 nw rows (rows = true: elements contiguous) or columns of a
// padded plane in a (nw = waves of the block), with b as the ping-pong buffer; the result ends in
// b. Wave w takes lines w, w + nw, w + 2 nw, w + 3 nw.
template <int DIR>
__device__ __forceinline__ void plane64_lines(float2* a, float2* b, bool rows, const float2* __restrict__ tw) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int line = w + nw * (lane >> 4), i = lane & 15;
  const int off = rows ? line * PLANE64_LS : line;
  const int es = rows ? 1 : PLANE64_LS;
  fft64_pass<DIR, 0>(a, b, off, es, i, tw);
  wave_lds_sync();
  fft64_pass<DIR, 1>(b, a, off, es, i, tw);
  wave_lds_sync();
  fft64_pass<DIR, 2>(a, b, off, es, i, tw);
}

__device__ __forceinline__ void load_twiddles(float2* dst, const float2* __restrict__ src, int K) {
  for (int i = threadIdx.x; i < K; i += blockDim.x) dst[i] = src[i];
}

__device__ __forceinline__ long long line_base(const FftLines& ln, int l) {
  return (long long)(l / ln.inner) * ln.outer_stride + (long long)(l % ln.inner) * ln.inner_stride;
}

// MODE 0: complex in/out.  MODE 1: load int64 fixed grid (real), clear it.  MODE 2: store real
// part to a float grid.  MODE 3: forward, multiply by B*C (+ energy), inverse, in place.
// lines [l0, l0 + B) of one axis pass, by one workgroup, in the LDS region lds
template <int DIR, int MODE>
__device__ void fft_axis_body(const FftPlan& pl, const FftLines& ln, const float2* __restrict__ tw,
                              float2* __restrict__ cg, long long* __restrict__ qfix, float* __restrict__ rout,
                              const float* __restrict__ bc, long long* __restrict__ eslot, int l0, float2* lds) {
  const int K = pl.K;
  const int B = ln.B;
  const int nb = min(B, ln.nlines - l0);
  const bool contiguous = ln.estride == 1;
  // lines of strided axes are loaded line-across-lanes: pad the LDS line stride to K + 1 so
  // consecutive lanes fall on different banks
  const int ls = contiguous ? K : K + 1;
  float2* A = lds;
  float2* Bf = lds + B * ls;
  float2* stw = Bf + B * ls;
  load_twiddles(stw, tw, K);
  // load
  for (int idx = threadIdx.x; idx < nb * K; idx += blockDim.x) {
    int l, e;
    if (contiguous) { l = idx / K; e = idx - l * K; }
    else { e = idx / nb; l = idx - e * nb; }
    long long g = line_base(ln, l0 + l) + (long long)e * ln.estride;
    float2 v;
    if (MODE == 1) {
      long long q = qfix[g];
      qfix[g] = 0;
      v = make_float2((float)((double)q * kFixInv), 0.f);
    } else {
      v = cg[g];
    }
    A[l * ls + e] = v;
  }
  __syncthreads();
  // 64-point strided lines, four per wave: wave-level passes (padded stride 65)
  const bool w64 = K == 64 && !contiguous && nb == B && B == 4 * (int)(blockDim.x >> 6);
  float2* res;
  if (w64) {
    plane64_lines<DIR>(A, Bf, true, stw);
    res = Bf;
    __syncthreads();
  } else {
    res = lds_fft<DIR>(A, Bf, pl, nb, stw, ls);
  }
  if (MODE == 3) {
    float2* other = (res == A) ? Bf : A;
    double en = 0.0;
    for (int idx = threadIdx.x; idx < nb * K; idx += blockDim.x) {
      int l = idx / K, e = idx - l * K;
      long long g = line_base(ln, l0 + l) + (long long)e * ln.estride;
      float f = bc[g];
      float2 v = res[l * ls + e];
      en += 0.5 * (double)f * ((double)v.x * v.x + (double)v.y * v.y);
      res[l * ls + e] = cscale(v, f);
    }
    __syncthreads();
    if (w64) {
      plane64_lines<-DIR>(res, other, true, stw);
      res = other;
      __syncthreads();
    } else {
      res = lds_fft<-DIR>(res, other, pl, nb, stw, ls);
    }
    // block energy
    __shared__ double sm[16];
    for (int off = 32; off > 0; off >>= 1) en += __shfl_xor(en, off, 64);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = en;
    __syncthreads();
    if (threadIdx.x == 0) {
      double sacc = 0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) sacc += sm[w];
      atomicAdd((unsigned long long*)eslot, (unsigned long long)to_fix(sacc));
    }
  }
  // store
  for (int idx = threadIdx.x; idx < nb * K; idx += blockDim.x) {
    int l, e;
    if (contiguous) { l = idx / K; e = idx - l * K; }
    else { e = idx / nb; l = idx - e * nb; }
    long long g = line_base(ln, l0 + l) + (long long)e * ln.estride;
    float2 v = res[l * ls + e];
    if (MODE == 2) rout[g] = v.x;
    else cg[g] = v;
  }
}

template <int DIR, int MODE>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_axis(FftPlan pl, FftLines ln, const float2* __restrict__ tw,
                                                        float2* __restrict__ cg, long long* __restrict__ qfix,
                                                        float* __restrict__ rout, const float* __restrict__ bc,
                                                        long long* __restrict__ eslot) {
  extern __shared__ float2 lds[];
  fft_axis_body<DIR, MODE>(pl, ln, tw, cg, qfix, rout, bc, eslot, blockIdx.x * ln.B, lds);
}

// ---------------------------------------------------------------- spread / gather
__device__ __forceinline__ void bspline4(float w, float* th, float* dth) {
  const float w2 = w * w, w3 = w2 * w, om = 1.f - w;
  th[0] = w3 * (1.f / 6.f);
  th[1] = (-3.f * w3 + 3.f * w2 + 3.f * w + 1.f) * (1.f / 6.f);
  th[2] = (3.f * w3 - 6.f * w2 + 4.f) * (1.f / 6.f);
  th[3] = om * om * om * (1.f / 6.f);
  if (dth) {
    dth[0] = 0.5f * w2;
    dth[1] = 0.5f * (-3.f * w2 + 2.f * w + 1.f);
    dth[2] = 0.5f * (3.f * w2 - 4.f * w);
    dth[3] = -0.5f * om * om;
  }
}

struct PmeArgs {
  int n;
  int Kx, Ky, Kz;
  double Lx, Ly, Lz;
  double iLx, iLy, iLz;
};

// fractional grid coordinate of x along an axis with inverse edge iL and K points
__device__ __forceinline__ void frac_coord(double x, double iL, int K, int* base, float* w) {
  double s = x * iL;
  s -= floor(s);
  double u = s * K;
  double f = floor(u);
  // NaN or infinite x: base 0 (range decided before the conversion); u == K after rounding wraps
  int b = f >= 0.0 && f <= (double)K ? (int)f : 0;
  *w = (float)(u - f);
  if (b >= K) b -= K;
  *base = b;
}

// 16 lanes per atom, one per (kx, ky) pencil of the 4x4x4 stencil, each adding the 4 z points
// of its pencil: 16x the parallelism of one lane per atom for the same atomics.
__global__ __launch_bounds__(PME_BLOCK) void k_spread(PmeArgs a, const double* __restrict__ pos,
                                                    const float* __restrict__ qs, long long* __restrict__ qfix) {
  const long long gt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int i = (int)(gt >> 4);
  const int kx = (int)(gt >> 2) & 3, ky = (int)gt & 3;
  if (i >= a.n) return;
  const float q = qs[i];
  if (q == 0.f) return;
  int bx, by, bz;
  float wx, wy, wz;
  frac_coord(pos[3 * i], a.iLx, a.Kx, &bx, &wx);
  frac_coord(pos[3 * i + 1], a.iLy, a.Ky, &by, &wy);
  frac_coord(pos[3 * i + 2], a.iLz, a.Kz, &bz, &wz);
  float tx[4], ty[4], tz[4];
  bspline4(wx, tx, nullptr);
  bspline4(wy, ty, nullptr);
  bspline4(wz, tz, nullptr);
  int gx = bx - kx;
  gx += gx < 0 ? a.Kx : 0;
  int gy = by - ky;
  gy += gy < 0 ? a.Ky : 0;
  const float qxy = q * tx[kx] * ty[ky];
  const long long row = ((long long)gx * a.Ky + gy) * a.Kz;
#pragma unroll
  for (int kz = 0; kz < 4; ++kz) {
    int gz = bz - kz;
    gz += gz < 0 ? a.Kz : 0;
    atomicAdd((unsigned long long*)&qfix[row + gz], (unsigned long long)to_fix_f(qxy * tz[kz]));
  }
}

__global__ __launch_bounds__(PME_BLOCK) void k_gather(PmeArgs a, const double* __restrict__ pos,
                                                    const float* __restrict__ qs, const float* __restrict__ conv,
                                                    long long* __restrict__ ffix) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  float q = qs[i];
  if (q == 0.f) return;
  int bx, by, bz;
  float wx, wy, wz;
  frac_coord(pos[3 * i], a.iLx, a.Kx, &bx, &wx);
  frac_coord(pos[3 * i + 1], a.iLy, a.Ky, &by, &wy);
  frac_coord(pos[3 * i + 2], a.iLz, a.Kz, &bz, &wz);
  float tx[4], ty[4], tz[4], dx[4], dy[4], dz[4];
  bspline4(wx, tx, dx);
  bspline4(wy, ty, dy);
  bspline4(wz, tz, dz);
  float gx = 0.f, gy = 0.f, gz = 0.f;
  for (int kx = 0; kx < 4; ++kx) {
    int ix = bx - kx;
    ix += ix < 0 ? a.Kx : 0;
    for (int ky = 0; ky < 4; ++ky) {
      int iy = by - ky;
      iy += iy < 0 ? a.Ky : 0;
      const float* row = conv + ((long long)ix * a.Ky + iy) * a.Kz;
      for (int kz = 0; kz < 4; ++kz) {
        int iz = bz - kz;
        iz += iz < 0 ? a.Kz : 0;
        float cv = row[iz];
        gx += dx[kx] * ty[ky] * tz[kz] * cv;
        gy += tx[kx] * dy[ky] * tz[kz] * cv;
        gz += tx[kx] * ty[ky] * dz[kz] * cv;
      }
    }
  }
  double fx = -(double)q * gx * (a.Kx / a.Lx);
  double fy = -(double)q * gy * (a.Ky / a.Ly);
  double fz = -(double)q * gz * (a.Kz / a.Lz);
  atomicAdd((unsigned long long*)&ffix[i], (unsigned long long)to_fix(fx));
  atomicAdd((unsigned long long*)&ffix[a.n + i], (unsigned long long)to_fix(fy));
  atomicAdd((unsigned long long*)&ffix[2 * a.n + i], (unsigned long long)to_fix(fz));
}

// ---------------------------------------------------------------- plane kernels
// One workgroup per x-plane x0 of the grid. The atoms that touch the plane (floor(u_x) - kx ==
// x0 for some kx in 0..3) are found in the cell-sorted atom order of the neighbour-list grid:
// cells are x-major, so the x-slices of cells that can hold such atoms form at most two
// contiguous ranges of sorted slots, and no atom is more than skin/2 away from its slice.
// Phase 1 (one lane per candidate) compacts the touching atoms into LDS records; phase 2 gives
// each record 16 lanes, one per (ky, kz) point of its 4x4 stencil in the plane.
#define PLANE_THREADS 1024  // the 64 x 64 wave-level transforms need all 16 waves (4 lines each)

struct PlaneArgs {
  int n;
  int Kx, Ky, Kz;
  double Lx, Ly, Lz;
  double iLx, iLy, iLz;
  int gnx, gny, gnz;  // neighbour-list grid
  double wx;          // its cell width along x
  double margin;      // list skin/2 plus rounding slack
  int W;              // candidates examined per round (bounded by the LDS record space)
  int rec_cap;        // gather records per plane in global memory
};

struct SpreadRec {  // 16 B
  int yz;           // by | bz << 12
  float wy, wz, qx; // fractional offsets, q sqrt(ke) theta_kx(wx)
};
struct GatherRec {  // 24 B (stored in an int buffer, 6 words per record)
  int atom;
  int yzk;          // by | bz << 12 | kx << 24
  float wx, wy, wz, q;
};
static_assert(sizeof(GatherRec) == 6 * sizeof(int), "gather record layout");

// weight k of the order-4 cardinal B-spline at offset w (grid point base - k), and its
// derivative; branch-free so lanes with different k do not diverge
__device__ __forceinline__ float sel4(int k, float a0, float a1, float a2, float a3) {
  return k == 0 ? a0 : (k == 1 ? a1 : (k == 2 ? a2 : a3));
}
__device__ __forceinline__ float bsp4(float w, int k) {
  const float w2 = w * w, w3 = w2 * w, om = 1.f - w;
  return sel4(k, w3, -3.f * w3 + 3.f * w2 + 3.f * w + 1.f, 3.f * w3 - 6.f * w2 + 4.f, om * om * om) * (1.f / 6.f);
}
__device__ __forceinline__ float dbsp4(float w, int k) {
  const float w2 = w * w, om = 1.f - w;
  return sel4(k, w2, -3.f * w2 + 2.f * w + 1.f, 3.f * w2 - 4.f * w, -om * om) * 0.5f;
}

// sorted-slot ranges [b0,e0) and [b1,e1) holding every atom that can touch plane x0
__device__ void plane_ranges(const PlaneArgs& a, const int* __restrict__ cell_start, int x0, int* b0, int* e0,
                             int* b1, int* e1) {
  const double h = a.Lx / a.Kx;
  const int lo = (int)floor((x0 * h - a.margin) / a.wx);
  const int hi = (int)floor(((x0 + 4) * h + a.margin) / a.wx);
  *b1 = *e1 = 0;
  if (hi - lo + 1 >= a.gnx) {
    *b0 = 0;
    *e0 = a.n;
    return;
  }
  const int first = ((lo % a.gnx) + a.gnx) % a.gnx;
  const int last = ((hi % a.gnx) + a.gnx) % a.gnx;
  const int per = a.gny * a.gnz;
  const int s_first = cell_start[first * per];
  const int e_last = last + 1 < a.gnx ? cell_start[(last + 1) * per] : a.n;
  if (first <= last) {
    *b0 = s_first;
    *e0 = e_last;
  } else {
    *b0 = s_first;
    *e0 = a.n;
    *b1 = 0;
    *e1 = e_last;
  }
}

// wave-aggregated slot reservation in an LDS counter
__device__ __forceinline__ int reserve_slot(bool keep, int* counter) {
  const unsigned long long m = __ballot(keep);
  const int lane = threadIdx.x & 63;
  int base = 0;
  const int leader = m ? __ffsll((long long)m) - 1 : 0;
  if (m && lane == leader) base = atomicAdd(counter, __popcll(m));
  base = __shfl(base, leader, 64);
  return base + __popcll(m & ((1ull << lane) - 1ull));
}

// candidates examined per lane per round of phase 1 (independent loads in flight)
#define PLANE_U 4

// spread q*sqrt(ke) of every atom touching plane x0 into an LDS int64 plane (integer sums, so the
// result is independent of the order of the atomics), then forward FFT along z and y; the
// plane goes to cg[x0][y][z].  LDS: twiddles (Kz + Ky), two complex planes; the records
// overlay the second plane and what follows it.
__device__ void plane_fwd_body(const PlaneArgs& a, const FftPlan& plz, const FftPlan& ply,
                               const float2* __restrict__ twz_g, const float2* __restrict__ twy_g,
                               const int* __restrict__ sorted_orig, const int* __restrict__ cell_start,
                               const double* __restrict__ pos_sorted, const float4* __restrict__ xq,
                               float2* __restrict__ cg, GatherRec* __restrict__ grec, int* __restrict__ grec_n,
                               DevState* st, int x0, float2* lds, bool wave64 = true) {
  __shared__ int nrec;
  __shared__ int ngrec;
  const int P = a.Ky * a.Kz;
  // wave-level transforms on a padded plane (they need PLANE_THREADS lanes)
  const bool p64 = wave64 && a.Ky == 64 && a.Kz == 64;
  float2* twz = lds;
  float2* twy = lds + a.Kz;
  float2* A = twy + a.Ky;
  float2* Bf = A + (p64 ? 64 * PLANE64_LS : P);
  long long* plane = (long long*)A;
  SpreadRec* rec = (SpreadRec*)Bf;
  load_twiddles(twz, twz_g, a.Kz);
  load_twiddles(twy, twy_g, a.Ky);
  for (int i = threadIdx.x; i < P; i += blockDim.x) plane[i] = 0;
  if (threadIdx.x == 0) ngrec = 0;
  GatherRec* gout = grec + (long long)x0 * a.rec_cap;
  int b0, e0, b1, e1;
  plane_ranges(a, cell_start, x0, &b0, &e0, &b1, &e1);
  const int len0 = e0 - b0, total = len0 + (e1 - b1);
  for (int w0 = 0; w0 < total; w0 += a.W) {
    if (threadIdx.x == 0) nrec = 0;
    __syncthreads();
    const int wend = min(total, w0 + a.W);
    // uniform trip count so every lane reaches the ballots
    for (int c0 = w0; c0 < wend; c0 += PLANE_U * blockDim.x) {
      int sl[PLANE_U];
      double px[PLANE_U];
#pragma unroll
      for (int u = 0; u < PLANE_U; ++u) {
        const int ci = c0 + u * blockDim.x + threadIdx.x;
        sl[u] = ci < wend ? (ci < len0 ? b0 + ci : b1 + (ci - len0)) : -1;
        px[u] = sl[u] >= 0 ? pos_sorted[3 * sl[u]] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < PLANE_U; ++u) {
        bool keep = false, touch = false;
        SpreadRec r;
        GatherRec g;
        if (sl[u] >= 0) {
          int bx;
          float wx;
          frac_coord(px[u], a.iLx, a.Kx, &bx, &wx);
          int kx = bx - x0;
          kx += kx < 0 ? a.Kx : 0;
          if (kx <= 3) {
            const int s = sl[u];
            const float q = xq[s].w;
            int by, bz;
            frac_coord(pos_sorted[3 * s + 1], a.iLy, a.Ky, &by, &r.wy);
            frac_coord(pos_sorted[3 * s + 2], a.iLz, a.Kz, &bz, &r.wz);
            r.yz = by | (bz << 12);
            r.qx = q * bsp4(wx, kx);
            keep = q != 0.f;
            // the same atom's record for the gather of this plane (zero charges included: the
            // gather writes their slot, so every atom owns exactly four)
            g.atom = sorted_orig[s];
            g.yzk = by | (bz << 12) | (kx << 24);
            g.wx = wx;
            g.wy = r.wy;
            g.wz = r.wz;
            g.q = q;
            touch = true;
          }
        }
        const int slot = reserve_slot(keep, &nrec);
        if (keep) rec[slot] = r;
        const int gslot = reserve_slot(touch, &ngrec);
        if (touch && gslot < a.rec_cap) gout[gslot] = g;
      }
    }
    __syncthreads();
    // four lanes per record, one per ky, each adding its 4 z points
    const int nr = nrec;
    for (int e = threadIdx.x; e < nr * 4; e += blockDim.x) {
      const SpreadRec r = rec[e >> 2];
      const int ky = e & 3;
      int gy = (r.yz & 4095) - ky;
      gy += gy < 0 ? a.Ky : 0;
      const float qxy = r.qx * bsp4(r.wy, ky);
      const float w = r.wz, w2 = w * w, w3 = w2 * w, om = 1.f - w;
      const float tz[4] = {w3 * (1.f / 6.f), (-3.f * w3 + 3.f * w2 + 3.f * w + 1.f) * (1.f / 6.f),
                           (3.f * w3 - 6.f * w2 + 4.f) * (1.f / 6.f), om * om * om * (1.f / 6.f)};
      const int bz = r.yz >> 12;
      long long* row = plane + gy * a.Kz;
#pragma unroll
      for (int kz = 0; kz < 4; ++kz) {
        int gz = bz - kz;
        gz += gz < 0 ? a.Kz : 0;
        atomicAdd((unsigned long long*)&row[gz], (unsigned long long)to_fix_f(qxy * tz[kz]));
      }
    }
    __syncthreads();
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    grec_n[x0] = min(ngrec, a.rec_cap);
    if (ngrec > a.rec_cap) atomicOr(&st->flags, MDX_FLAG_OVERFLOW);
  }
  if (p64) {
    float2* out = cg + (long long)x0 * P;
    for (int i = threadIdx.x; i < P; i += blockDim.x)
      Bf[(i >> 6) * PLANE64_LS + (i & 63)] = make_float2((float)((double)plane[i] * kFixInv), 0.f);
    __syncthreads();
    plane64_lines<-1>(Bf, A, true, twz);   // z lines (rows) -> A
    __syncthreads();
    plane64_lines<-1>(A, Bf, false, twy);  // y lines (columns) -> Bf
    __syncthreads();
    for (int i = threadIdx.x; i < P; i += blockDim.x) out[i] = Bf[(i >> 6) * PLANE64_LS + (i & 63)];
    return;
  }
  // in place: lane i reads int64 element i and writes complex element i (same 8 bytes)
  for (int i = threadIdx.x; i < P; i += blockDim.x) A[i] = make_float2((float)((double)plane[i] * kFixInv), 0.f);
  __syncthreads();
  float2* r1 = lds_fft<-1>(A, Bf, plz, a.Ky, twz);                // z lines: contiguous
  float2* o1 = (r1 == A) ? Bf : A;
  float2* r2 = lds_fft<-1>(r1, o1, ply, a.Kz, twy, 1, a.Kz);      // y lines: stride Kz
  float2* out = cg + (long long)x0 * P;
  for (int i = threadIdx.x; i < P; i += blockDim.x) out[i] = r2[i];
}

__global__ __launch_bounds__(PLANE_THREADS) void k_pme_plane_fwd(
    PlaneArgs a, FftPlan plz, FftPlan ply, const float2* __restrict__ twz_g, const float2* __restrict__ twy_g,
    const int* __restrict__ sorted_orig, const int* __restrict__ cell_start, const double* __restrict__ pos_sorted,
    const float4* __restrict__ xq, float2* __restrict__ cg, GatherRec* __restrict__ grec, int* __restrict__ grec_n,
    DevState* st) {
  extern __shared__ float2 lds[];
  plane_fwd_body(a, plz, ply, twz_g, twy_g, sorted_orig, cell_start, pos_sorted, xq, cg, grec, grec_n, st,
                 blockIdx.x, lds);
}

// inverse FFT along y and z of plane x0, then for every atom touching the plane its share of
// the reciprocal force, -q (dtheta_x theta_y theta_z, ...) . conv over this plane's 4x4 (y,z)
// stencil, written to fpart[kx][3*atom + d] (each slot once per evaluation).
// LDS: twiddles, real part of the convolved plane (floats), two complex planes that the
// records overlay after the transforms.
__device__ void plane_gather_body(const PlaneArgs& a, const FftPlan& plz, const FftPlan& ply,
                                  const float2* __restrict__ twz_g, const float2* __restrict__ twy_g,
                                  const int* __restrict__ sorted_orig, const int* __restrict__ cell_start,
                                  const double* __restrict__ pos_sorted, const float4* __restrict__ xq,
                                  const float2* __restrict__ cg, const GatherRec* __restrict__ grec,
                                  const int* __restrict__ grec_n, double* __restrict__ fpart, DevState* st, int x0,
                                  float2* lds, bool wave64 = true) {
  __shared__ int bcount;
  const int P = a.Ky * a.Kz;
  // wave-level transforms on a padded plane (they need PLANE_THREADS lanes)
  const bool p64 = wave64 && a.Ky == 64 && a.Kz == 64;
  float2* twz = lds;
  float2* twy = lds + a.Kz;
  float* convr = (float*)(twy + a.Ky);
  float2* A = (float2*)(convr + P + (P & 1));
  float2* Bf = A + (p64 ? 64 * PLANE64_LS : P);
  load_twiddles(twz, twz_g, a.Kz);
  load_twiddles(twy, twy_g, a.Ky);
  const float2* in = cg + (long long)x0 * P;
  if (p64) {
    for (int i = threadIdx.x; i < P; i += blockDim.x) A[(i >> 6) * PLANE64_LS + (i & 63)] = in[i];
    __syncthreads();
    plane64_lines<1>(A, Bf, false, twy);  // y lines (columns) -> Bf
    __syncthreads();
    plane64_lines<1>(Bf, A, true, twz);   // z lines (rows) -> A
    __syncthreads();
    for (int i = threadIdx.x; i < P; i += blockDim.x) convr[i] = A[(i >> 6) * PLANE64_LS + (i & 63)].x;
  } else {
    for (int i = threadIdx.x; i < P; i += blockDim.x) A[i] = in[i];
    __syncthreads();
    float2* r1 = lds_fft<1>(A, Bf, ply, a.Kz, twy, 1, a.Kz);        // y lines first (reverse order)
    float2* o1 = (r1 == A) ? Bf : A;
    float2* r2 = lds_fft<1>(r1, o1, plz, a.Ky, twz);                // z lines
    for (int i = threadIdx.x; i < P; i += blockDim.x) convr[i] = r2[i].x;
  }
  const long long n3 = 3LL * a.n;
  // the records of the atoms touching this plane, written by the forward kernel
  const int nr = grec_n[x0];
  const GatherRec* gin = grec + (long long)x0 * a.rec_cap;
  __syncthreads();  // convr complete
  // four lanes per record (one per ky, 4 z points each); nr * 4 and blockDim.x are multiples
  // of 4, so the four lanes of a record stay together through the reduction. Zero charges give
  // zero forces through the same arithmetic (their slot is still written).
  for (int e = threadIdx.x; e < nr * 4; e += blockDim.x) {
    const GatherRec r = gin[e >> 2];
    const int ky = e & 3;
    int gy = (r.yzk & 4095) - ky;
    gy += gy < 0 ? a.Ky : 0;
    const float ty = bsp4(r.wy, ky), dy = dbsp4(r.wy, ky);
    const float w = r.wz, w2 = w * w, w3 = w2 * w, om = 1.f - w;
    const float tz[4] = {w3 * (1.f / 6.f), (-3.f * w3 + 3.f * w2 + 3.f * w + 1.f) * (1.f / 6.f),
                         (3.f * w3 - 6.f * w2 + 4.f) * (1.f / 6.f), om * om * om * (1.f / 6.f)};
    const float dz[4] = {0.5f * w2, 0.5f * (-3.f * w2 + 2.f * w + 1.f), 0.5f * (3.f * w2 - 4.f * w),
                         -0.5f * om * om};
    const int bz = (r.yzk >> 12) & 4095;
    const float* row = convr + gy * a.Kz;
    float c = 0.f, cz = 0.f;
#pragma unroll
    for (int kz = 0; kz < 4; ++kz) {
      int gz = bz - kz;
      gz += gz < 0 ? a.Kz : 0;
      const float cv = row[gz];
      c += tz[kz] * cv;
      cz += dz[kz] * cv;
    }
    float sx = ty * c, sy = dy * c, sz = ty * cz;
#pragma unroll
    for (int off = 2; off > 0; off >>= 1) {
      sx += __shfl_xor(sx, off, 64);
      sy += __shfl_xor(sy, off, 64);
      sz += __shfl_xor(sz, off, 64);
    }
    if (ky == 0) {
      const int kx = r.yzk >> 24;
      const float tx = bsp4(r.wx, kx), dx = dbsp4(r.wx, kx);
      double* dst = fpart + kx * n3 + 3LL * r.atom;
      if (r.q == 0.f) {
        dst[0] = dst[1] = dst[2] = 0.0;
      } else {
        dst[0] = -(double)r.q * (double)(dx * sx) * (a.Kx / a.Lx);
        dst[1] = -(double)r.q * (double)(tx * sy) * (a.Ky / a.Ly);
        dst[2] = -(double)r.q * (double)(tx * sz) * (a.Kz / a.Lz);
      }
    }
  }
  if (threadIdx.x == 0) bcount = nr;
  // every atom owns exactly one slot per kx: the grid-wide count must be 4 n. The last block
  // to finish checks it and re-arms the counters for the next evaluation.
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(&st->pme_slots, bcount);
    __threadfence();
    const unsigned t = atomicAdd(&st->pme_ticket, 1u);
    if (t == (unsigned)a.Kx - 1) {  // the last of the Kx plane workgroups (launches may hold others)
      __threadfence();
      const int tot = atomicAdd(&st->pme_slots, 0);
      if (tot != 4 * a.n) atomicOr(&st->flags, MDX_FLAG_PME_COVERAGE);
      atomicExch(&st->pme_slots, 0);
      atomicExch(&st->pme_ticket, 0u);
    }
  }
}

__global__ __launch_bounds__(PLANE_THREADS) void k_pme_plane_inv_gather(
    PlaneArgs a, FftPlan plz, FftPlan ply, const float2* __restrict__ twz_g, const float2* __restrict__ twy_g,
    const int* __restrict__ sorted_orig, const int* __restrict__ cell_start, const double* __restrict__ pos_sorted,
    const float4* __restrict__ xq, const float2* __restrict__ cg, const GatherRec* __restrict__ grec,
    const int* __restrict__ grec_n, double* __restrict__ fpart, DevState* st) {
  extern __shared__ float2 lds[];
  plane_gather_body(a, plz, ply, twz_g, twy_g, sorted_orig, cell_start, pos_sorted, xq, cg, grec, grec_n, fpart, st,
                    blockIdx.x, lds);
}

// Grid-wide barrier of a kernel whose blocks are all resident at once (at most one block per
// compute unit, grid no larger than the device). The last block to arrive resets the count and
// advances the generation; the others wait for that with a bounded spin, so a launch whose
// blocks cannot all be resident raises MDX_FLAG_SYNC and finishes instead of hanging. One
// thread per block fences on each side (device-scope fences act on whole caches: a fence per
// thread would repeat the same cache write-back a thousand times).
__device__ void grid_barrier(DevState* st) {
  __syncthreads();  // every thread's global stores issued and complete at workgroup scope
  if (threadIdx.x == 0) {
    __threadfence();  // release for the whole block: the cache write-back covers all its stores
    const unsigned gen = atomicAdd(&st->bar_gen, 0u);
    const unsigned arrived = atomicAdd(&st->bar_count, 1u);
    if (arrived == gridDim.x - 1) {
      atomicExch(&st->bar_count, 0u);
      __threadfence();
      atomicAdd(&st->bar_gen, 1u);
    } else {
      for (unsigned spin = 0; atomicAdd(&st->bar_gen, 0u) == gen; ++spin) {
        if (spin > (1u << 18)) {
          atomicOr(&st->flags, MDX_FLAG_SYNC);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    __threadfence();  // acquire: invalidates this compute unit's caches for the whole block
  }
  __syncthreads();
}

// The whole reciprocal-space evaluation in one launch of Kx blocks: block x0 spreads and
// transforms plane x0; after a grid barrier the blocks share the x-axis pass (pencils of the
// Ky x Kz plane, influence function, energy); after a second barrier block x0 transforms plane
// x0 back and gathers its forces. The blocks keep their compute units for the whole
// evaluation, so kernels running concurrently on other streams cannot delay the later phases.
__global__ __launch_bounds__(PLANE_THREADS) void k_pme_fused(
    PlaneArgs a, FftPlan plz, FftPlan ply, FftPlan plx, FftLines lnx, const float2* __restrict__ twz_g,
    const float2* __restrict__ twy_g, const float2* __restrict__ twx_g, const int* __restrict__ sorted_orig,
    const int* __restrict__ cell_start, const double* __restrict__ pos_sorted, const float4* __restrict__ xq,
    float2* __restrict__ cg, const float* __restrict__ bc, long long* __restrict__ eslot,
    GatherRec* __restrict__ grec, int* __restrict__ grec_n, double* __restrict__ fpart, DevState* st) {
  extern __shared__ float2 lds[];
  plane_fwd_body(a, plz, ply, twz_g, twy_g, sorted_orig, cell_start, pos_sorted, xq, cg, grec, grec_n, st,
                 blockIdx.x, lds);
  grid_barrier(st);
  const int nchunks = (lnx.nlines + lnx.B - 1) / lnx.B;
  for (int ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    fft_axis_body<-1, 3>(plx, lnx, twx_g, cg, nullptr, nullptr, bc, eslot, ch * lnx.B, lds);
    __syncthreads();
  }
  grid_barrier(st);
  plane_gather_body(a, plz, ply, twz_g, twy_g, sorted_orig, cell_start, pos_sorted, xq, cg, grec, grec_n, fpart, st,
                    blockIdx.x, lds);
}

// Phase launches of one force evaluation, on one stream: workgroups [0, npme) run a PME
// reciprocal phase (0: spread + 2D forward FFT per x-plane, 1: x-axis pass with the influence
// function, 2: 2D inverse FFT + gather per x-plane), the next nbn run nonbonded workgroups
// nb0 .. nb0 + nbn - 1 (PLANE_THREADS lanes each), the rest the bonded terms. The PME
// workgroups come first, so they take compute units before the nonbonded ones; the nonbonded
// shares are sized to end with their phase. This replaces the side stream of the three-kernel
// path and its fork and join.
struct PhaseArgs {
  PlaneArgs a;
  FftPlan plz, ply, plx;
  FftLines lnx;
  const float2 *twz, *twy, *twx;
  const int* sorted_orig;
  const int* cell_start;
  const double* pos_sorted;
  const float4* xq;
  float2* cg;
  const float* bc;
  long long* eslot;
  GatherRec* grec;
  int* grec_n;
  double* fpart;
  DevState* st;
};

template <bool ENERGY, int G>
__global__ __launch_bounds__(PLANE_THREADS) void k_force_phase(int phase, int npme, int nb0, int nbn, PhaseArgs p,
                                                               NbArgs na, NbPtrs P, BondedArgs ba) {
  extern __shared__ float2 lds[];
  const int b = blockIdx.x;
  if (b < npme) {
    if (phase == 0)
      plane_fwd_body(p.a, p.plz, p.ply, p.twz, p.twy, p.sorted_orig, p.cell_start, p.pos_sorted, p.xq, p.cg, p.grec,
                     p.grec_n, p.st, b, lds);
    else if (phase == 1)
      fft_axis_body<-1, 3>(p.plx, p.lnx, p.twx, p.cg, nullptr, nullptr, p.bc, p.eslot, b * p.lnx.B, lds);
    else
      plane_gather_body(p.a, p.plz, p.ply, p.twz, p.twy, p.sorted_orig, p.cell_start, p.pos_sorted, p.xq, p.cg,
                        p.grec, p.grec_n, p.fpart, p.st, b, lds);
    return;
  }
  const int vb = b - npme;
  if (vb < nbn) {
    nb_block<true, true, ENERGY, G, PLANE_THREADS, true>(na, P, nb0 + vb);
    return;
  }
  bonded_block(ba, vb - nbn);
}

typedef void (*PhaseKernel)(int, int, int, int, PhaseArgs, NbArgs, NbPtrs, BondedArgs);
static PhaseKernel phase_kernel(bool energy, int G) {
  if (G == 16) return energy ? k_force_phase<true, 16> : k_force_phase<false, 16>;
  if (G == 64) return energy ? k_force_phase<true, 64> : k_force_phase<false, 64>;
  return energy ? k_force_phase<true, 32> : k_force_phase<false, 32>;
}

// LDS bytes of the two plane kernels (the forward one for a record window of W candidates)
// (64 x 64 planes are held padded to 64 x PLANE64_LS complex values)
static inline size_t plane_fwd_lds(int Ky, int Kz, int W) {
  const size_t P = (size_t)Ky * Kz;
  const size_t Pp = (Ky == 64 && Kz == 64) ? 64 * PLANE64_LS : P;
  return (size_t)(Ky + Kz) * 8 + std::max(2 * Pp * 8, Pp * 8 + (size_t)W * sizeof(SpreadRec));
}
static inline size_t plane_gather_lds(int Ky, int Kz) {
  const size_t P = (size_t)Ky * Kz;
  const size_t Pp = (Ky == 64 && Kz == 64) ? 64 * PLANE64_LS : P;
  return (size_t)(Ky + Kz) * 8 + (P + (P & 1)) * 4 + 2 * Pp * 8;
}

// ---------------------------------------------------------------- host
static int factorize(int K, FftPlan* pl) {
  pl->K = K;
  pl->nrad = 0;
  pl->code = 0;
  int k = K;
  auto push = [&](unsigned long long r) {
    if (pl->nrad < MAX_RADICES) pl->code |= r << (4 * pl->nrad);
    pl->nrad++;
  };
  while (k % 4 == 0) { push(4); k /= 4; }
  while (k % 2 == 0) { push(2); k /= 2; }
  while (k % 3 == 0) { push(3); k /= 3; }
  while (k % 5 == 0) { push(5); k /= 5; }
  return (k == 1 && pl->nrad <= MAX_RADICES) ? 0 : -1;
}

// |b(m)|^-2 of Essmann eq. 4.4 for order 4: M4(1)=1/6, M4(2)=2/3, M4(3)=1/6
static void bspline_moduli(int K, std::vector<double>& out) {
  const double M[3] = {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0};
  out.assign(K, 0.0);
  std::vector<double> mod(K);
  for (int m = 0; m < K; ++m) {
    double re = 0, im = 0;
    for (int k = 0; k < 3; ++k) {
      double a = 2.0 * M_PI * m * k / K;
      re += M[k] * cos(a);
      im += M[k] * sin(a);
    }
    mod[m] = re * re + im * im;
  }
  for (int m = 0; m < K; ++m)
    if (mod[m] < 1e-10) mod[m] = 0.5 * (mod[(m - 1 + K) % K] + mod[(m + 1) % K]);
  for (int m = 0; m < K; ++m) out[m] = 1.0 / mod[m];
}

extern "C" int mdx_pme_config(mdx_ctx* c, const int* h_grid, int order, double beta) {
  MDX_CHECK(c && c->pos, MDX_ERR_INVALID, "mdx_pme_config: no system bound");
  MDX_CHECK(order == 4, MDX_ERR_INVALID, "only B-spline order 4 is implemented");
  MDX_CHECK(beta > 0, MDX_ERR_INVALID, "beta must be positive");
  FftPlan pl;
  for (int d = 0; d < 3; ++d) {
    MDX_CHECK(h_grid[d] >= 4 && factorize(h_grid[d], &pl) == 0, MDX_ERR_INVALID,
              "PME grid %d must be >= 4 with prime factors 2, 3, 5 only", h_grid[d]);
    c->K[d] = h_grid[d];
  }
  c->order = order;
  c->beta = beta;
  mdx_invalidate_graph(c);
  size_t G = (size_t)c->K[0] * c->K[1] * c->K[2];
  int rc;
  if ((rc = c->qfix.ensure(G)) || (rc = c->cgrid.ensure(G)) || (rc = c->conv.ensure(G)) || (rc = c->bc.ensure(G)) ||
      (rc = c->fpart.ensure(12 * (size_t)c->n)))
    return rc;
  MDX_HIP(hipMemset(c->qfix.p, 0, G * sizeof(long long)));
  // influence function B(m) C(m), C(0) = 0
  std::vector<double> Bm[3];
  for (int d = 0; d < 3; ++d) bspline_moduli(c->K[d], Bm[d]);
  double V = c->box[0] * c->box[1] * c->box[2];
  std::vector<float> h_bc(G);
  for (int x = 0; x < c->K[0]; ++x) {
    double mx = (x <= c->K[0] / 2 ? x : x - c->K[0]) / c->box[0];
    for (int y = 0; y < c->K[1]; ++y) {
      double my = (y <= c->K[1] / 2 ? y : y - c->K[1]) / c->box[1];
      for (int z = 0; z < c->K[2]; ++z) {
        double mz = (z <= c->K[2] / 2 ? z : z - c->K[2]) / c->box[2];
        double m2 = mx * mx + my * my + mz * mz;
        size_t g = ((size_t)x * c->K[1] + y) * c->K[2] + z;
        if (m2 == 0) {
          h_bc[g] = 0.f;
          continue;
        }
        double C = exp(-M_PI * M_PI * m2 / (beta * beta)) / (M_PI * V * m2);
        h_bc[g] = (float)(Bm[0][x] * Bm[1][y] * Bm[2][z] * C);
      }
    }
  }
  MDX_HIP(hipMemcpy(c->bc.p, h_bc.data(), G * sizeof(float), hipMemcpyHostToDevice));
  int Kmax = std::max(c->K[0], std::max(c->K[1], c->K[2]));
  if ((rc = c->twiddle.ensure(3 * (size_t)Kmax))) return rc;
  std::vector<float2> tw(3 * (size_t)Kmax);
  for (int d = 0; d < 3; ++d)
    for (int k = 0; k < c->K[d]; ++k) {
      double a = -2.0 * M_PI * k / c->K[d];
      tw[(size_t)d * Kmax + k] = make_float2((float)cos(a), (float)sin(a));
    }
  MDX_HIP(hipMemcpy(c->twiddle.p, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice));
  // the plane kernels hold twiddles, two complex Ky x Kz planes and a window of atom records
  // in LDS; the window is as large as the device allows (at most 8192 candidates)
  int lds_max = 0;
  MDX_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
  const size_t budget = lds_max > 1024 ? (size_t)lds_max - 1024 : 0;  // static LDS of the kernels
  c->pme_wf = 0;
  for (int W = 8192; W >= PLANE_THREADS; W -= PLANE_THREADS)
    if (!c->pme_wf && plane_fwd_lds(c->K[1], c->K[2], W) <= budget) c->pme_wf = W;
  c->pme_planes = c->pme_wf > 0 && plane_gather_lds(c->K[1], c->K[2]) <= budget && c->K[1] < 4096 &&
                  c->K[2] < 4096;
  // gather records: each atom touches 4 of the Kx planes; room for twice the mean plus slack
  c->pme_rec_cap = (int)std::min<long long>(c->n, 8LL * c->n / c->K[0] + 4096);
  if ((rc = c->pme_grec.ensure((size_t)c->K[0] * c->pme_rec_cap * 6)) || (rc = c->pme_grec_n.ensure(c->K[0])))
    return rc;
  if (c->pme_planes) {  // the whole budget, so contexts with other grids stay launchable
    MDX_HIP(hipFuncSetAttribute((const void*)k_pme_plane_fwd, hipFuncAttributeMaxDynamicSharedMemorySize, (int)budget));
    MDX_HIP(hipFuncSetAttribute((const void*)k_pme_plane_inv_gather, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)budget));
    MDX_HIP(hipFuncSetAttribute((const void*)k_pme_fused, hipFuncAttributeMaxDynamicSharedMemorySize, (int)budget));
  }
  // one-launch kernel: Kx blocks that must all be resident together (one per compute unit);
  // the x-axis pass takes ceil(Ky Kz / Kx) pencils per block if the LDS allows
  int ncu = 0;
  MDX_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
  c->pme_fused_B = 0;
  if (c->pme_planes && c->K[0] <= ncu) {
    const int nl = c->K[1] * c->K[2];
    int B = (nl + c->K[0] - 1) / c->K[0];
    const size_t ls = c->K[0] + 1;  // strided lines: padded LDS line stride
    while (B > 1 && (2 * (size_t)B * ls + c->K[0]) * sizeof(float2) > budget) B = (B + 1) / 2;
    const size_t lds_x = (2 * (size_t)B * ls + c->K[0]) * sizeof(float2);
    c->pme_fused_B = lds_x <= budget ? B : 0;
    c->pme_fused_lds = std::max(lds_x, std::max(plane_fwd_lds(c->K[1], c->K[2], c->pme_wf),
                                                plane_gather_lds(c->K[1], c->K[2])));
  }
  // phase launches: their static LDS (the nonbonded workgroups' table) comes off the budget
  c->pme_wf_phase = 0;
  if (c->pme_planes && c->pme_fused_B > 0) {
    size_t st_max = 0;
    for (int e = 0; e < 2; ++e)
      for (int G = 16; G <= 64; G *= 2) {
        hipFuncAttributes fa;
        MDX_HIP(hipFuncGetAttributes(&fa, (const void*)phase_kernel(e != 0, G)));
        st_max = std::max(st_max, (size_t)fa.sharedSizeBytes);
      }
    const size_t pb = (size_t)lds_max > st_max + 256 ? (size_t)lds_max - st_max - 256 : 0;
    for (int W = 8192; W >= PLANE_THREADS; W -= PLANE_THREADS)
      if (!c->pme_wf_phase && plane_fwd_lds(c->K[1], c->K[2], W) <= pb) c->pme_wf_phase = W;
    const size_t lx = (2 * (size_t)c->pme_fused_B * (c->K[0] + 1) + c->K[0]) * sizeof(float2);
    if (plane_gather_lds(c->K[1], c->K[2]) > pb || lx > pb) c->pme_wf_phase = 0;
    if (c->pme_wf_phase)
      for (int e = 0; e < 2; ++e)
        for (int G = 16; G <= 64; G *= 2)
          MDX_HIP(hipFuncSetAttribute((const void*)phase_kernel(e != 0, G),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)pb));
  }
  if ((rc = mdx_coulomb_table_update(c))) return rc;
  c->pme_configured = true;
  return 0;
}

extern "C" int mdx_pme_set_fused(mdx_ctx* c, int enable) {
  MDX_CHECK(c, MDX_ERR_INVALID, "mdx_pme_set_fused: null context");
  c->pme_fused_on = enable != 0;
  mdx_invalidate_graph(c);
  return 0;
}

// one axis pass of a context's grid: its transform plan, line layout, twiddles and dynamic LDS
static void fft_axis_of(const mdx_ctx* c, int axis, FftPlan* plo, FftLines* lno, const float2** two, size_t* ldso) {
  const int Kx = c->K[0], Ky = c->K[1], Kz = c->K[2];
  FftPlan& pl = *plo;
  FftLines& ln = *lno;
  int Kmax = std::max(Kx, std::max(Ky, Kz));
  if (axis == 2) {
    factorize(Kz, &pl);
    ln.nlines = Kx * Ky;
    ln.inner = ln.nlines;
    ln.inner_stride = Kz;
    ln.outer_stride = 0;
    ln.estride = 1;
    ln.B = std::max(1, 1024 / Kz);
  } else if (axis == 1) {
    factorize(Ky, &pl);
    ln.nlines = Kx * Kz;
    ln.inner = Kz;
    ln.inner_stride = 1;
    ln.outer_stride = Ky * Kz;
    ln.estride = Kz;
    ln.B = std::min(16, Kz);
  } else {
    factorize(Kx, &pl);
    ln.nlines = Ky * Kz;
    ln.inner = ln.nlines;
    ln.inner_stride = 1;
    ln.outer_stride = 0;
    ln.estride = Ky * Kz;
    ln.B = std::min(16, Kz);
  }
  *two = c->twiddle.p + (size_t)axis * Kmax;
  const size_t ls = ln.estride == 1 ? pl.K : pl.K + 1;
  *ldso = (2 * (size_t)ln.B * ls + pl.K) * sizeof(float2);
}

static int fft_pass(mdx_ctx* c, int axis, int dir, int mode) {
  FftPlan pl;
  FftLines ln;
  const float2* tw;
  size_t lds;
  fft_axis_of(c, axis, &pl, &ln, &tw, &lds);
  dim3 grid(div_up(ln.nlines, ln.B));
#define FFT_L(D, M)                                                                                    \
  k_fft_axis<D, M><<<grid, FFT_THREADS, lds, c->stream>>>(pl, ln, tw, c->cgrid.p, c->qfix.p, c->conv.p, c->bc.p, \
                                                          c->efix.p + 2)
  if (mode == 1) FFT_L(-1, 1);
  else if (mode == 3) FFT_L(-1, 3);
  else if (mode == 2) FFT_L(1, 2);
  else if (dir < 0) FFT_L(-1, 0);
  else FFT_L(1, 0);
#undef FFT_L
  MDX_LAUNCH_CHECK();
  return 0;
}

// plane path: spread + 2D forward FFT per x-plane, x-axis pass with the influence function,
// 2D inverse FFT + gather per x-plane. Needs the cell-sorted atom order of the neighbour-list
// grid (current for this evaluation) and two complex planes in LDS.
// parts: bit 0 the plane forward kernel, bit 1 the x-axis transform pass, bit 2 the inverse +
// gather kernel (all three for an evaluation; one at a time for timing)
// the plane kernels' arguments of a context (a.W: the forward kernel's record window)
static void plane_args_of(const mdx_ctx* c, PlaneArgs* out) {
  const int Kx = c->K[0], Ky = c->K[1], Kz = c->K[2];
  PlaneArgs& a = *out;
  a.n = c->n;
  a.Kx = Kx;
  a.Ky = Ky;
  a.Kz = Kz;
  a.Lx = c->box[0];
  a.Ly = c->box[1];
  a.Lz = c->box[2];
  a.iLx = 1.0 / c->box[0];
  a.iLy = 1.0 / c->box[1];
  a.iLz = 1.0 / c->box[2];
  a.gnx = c->gn[0];
  a.gny = c->gn[1];
  a.gnz = c->gn[2];
  a.wx = c->box[0] / c->gn[0];
  a.margin = 0.5 * c->skin + 1e-6 * c->box[0] + 1e-9;
  a.rec_cap = c->pme_rec_cap;
  a.W = c->pme_wf;
}

static int pme_enqueue_planes(mdx_ctx* c, unsigned parts = 7u) {
  const int Kx = c->K[0], Ky = c->K[1], Kz = c->K[2];
  PlaneArgs a;
  plane_args_of(c, &a);
  MDX_CHECK(c->pme_grec.n >= (size_t)Kx * a.rec_cap * 6, MDX_ERR_INVALID, "PME configured for another system size");
  GatherRec* grec = (GatherRec*)c->pme_grec.p;
  FftPlan plz, ply;
  factorize(Kz, &plz);
  factorize(Ky, &ply);
  const int Kmax = std::max(Kx, std::max(Ky, Kz));
  const float2* twy = c->twiddle.p + (size_t)Kmax;
  const float2* twz = c->twiddle.p + 2 * (size_t)Kmax;
  if (parts == 7u && c->pme_fused_B > 0 && c->pme_fused_on && !c->in_batch) {
    FftPlan plx;
    factorize(Kx, &plx);
    FftLines ln;
    ln.nlines = Ky * Kz;
    ln.inner = ln.nlines;
    ln.inner_stride = 1;
    ln.outer_stride = 0;
    ln.estride = Ky * Kz;
    ln.B = c->pme_fused_B;
    a.W = c->pme_wf;
    k_pme_fused<<<Kx, PLANE_THREADS, c->pme_fused_lds, c->stream>>>(
        a, plz, ply, plx, ln, twz, twy, c->twiddle.p, c->sorted_orig.p, c->cell_start.p, c->pos_sorted.p, c->xq.p,
        c->cgrid.p, c->bc.p, c->efix.p + 2, grec, c->pme_grec_n.p, c->fpart.p, c->dstate.p);
    MDX_LAUNCH_CHECK();
    return 0;
  }
  a.W = c->pme_wf;
  if (parts & 1u) {
    k_pme_plane_fwd<<<Kx, PLANE_THREADS, plane_fwd_lds(Ky, Kz, a.W), c->stream>>>(
        a, plz, ply, twz, twy, c->sorted_orig.p, c->cell_start.p, c->pos_sorted.p, c->xq.p, c->cgrid.p, grec,
        c->pme_grec_n.p, c->dstate.p);
    MDX_LAUNCH_CHECK();
  }
  int rc;
  if ((parts & 2u) && (rc = fft_pass(c, 0, -1, 3))) return rc;  // x forward, * B C, energy, x inverse
  if (parts & 4u) {
    k_pme_plane_inv_gather<<<Kx, PLANE_THREADS, plane_gather_lds(Ky, Kz), c->stream>>>(
        a, plz, ply, twz, twy, c->sorted_orig.p, c->cell_start.p, c->pos_sorted.p, c->xq.p, c->cgrid.p, grec,
        c->pme_grec_n.p, c->fpart.p, c->dstate.p);
    MDX_LAUNCH_CHECK();
  }
  return 0;
}

// One of the plane path's three launches alone (part 0, 1 or 2 of pme_enqueue_planes), for
// timing: each reads only what the previous evaluation's launches left in place.
int mdx_pme_enqueue_part(mdx_ctx* c, int part) {
  MDX_CHECK(c->pme_configured && c->pme_planes && c->nb_configured && part >= 0 && part <= 2, MDX_ERR_INVALID,
            "PME part timing needs the plane path");
  return pme_enqueue_planes(c, 1u << part);
}

bool mdx_force_phases_ok(const mdx_ctx* c) {
  return c->phases_on && c->pme_configured && c->pme_planes && c->pme_wf_phase > 0 && c->pme_fused_B > 0 &&
         c->nb_configured && c->has_lj && c->has_charges &&
         (c->nb_lanes == 16 || c->nb_lanes == 32 || c->nb_lanes == 64);
}

int mdx_force_phases_enqueue(mdx_ctx* c, bool energy, unsigned bonded_terms) {
  MDX_CHECK(mdx_force_phases_ok(c), MDX_ERR_INVALID, "phase launches not available for this context");
  MDX_CHECK(c->fpart.p && c->fpart.n >= 12 * (size_t)c->n, MDX_ERR_INVALID, "PME configured for another system size");
  const int Kx = c->K[0], Ky = c->K[1], Kz = c->K[2];
  PhaseArgs p;
  PlaneArgs& a = p.a;
  a.n = c->n;
  a.Kx = Kx;
  a.Ky = Ky;
  a.Kz = Kz;
  a.Lx = c->box[0];
  a.Ly = c->box[1];
  a.Lz = c->box[2];
  a.iLx = 1.0 / c->box[0];
  a.iLy = 1.0 / c->box[1];
  a.iLz = 1.0 / c->box[2];
  a.gnx = c->gn[0];
  a.gny = c->gn[1];
  a.gnz = c->gn[2];
  a.wx = c->box[0] / c->gn[0];
  a.margin = 0.5 * c->skin + 1e-6 * c->box[0] + 1e-9;
  a.rec_cap = c->pme_rec_cap;
  a.W = c->pme_wf_phase;
  MDX_CHECK(c->pme_grec.n >= (size_t)Kx * a.rec_cap * 6, MDX_ERR_INVALID, "PME configured for another system size");
  factorize(Kz, &p.plz);
  factorize(Ky, &p.ply);
  factorize(Kx, &p.plx);
  p.lnx.nlines = Ky * Kz;
  p.lnx.inner = p.lnx.nlines;
  p.lnx.inner_stride = 1;
  p.lnx.outer_stride = 0;
  p.lnx.estride = Ky * Kz;
  p.lnx.B = c->pme_fused_B;
  const int Kmax = std::max(Kx, std::max(Ky, Kz));
  p.twx = c->twiddle.p;
  p.twy = c->twiddle.p + (size_t)Kmax;
  p.twz = c->twiddle.p + 2 * (size_t)Kmax;
  p.sorted_orig = c->sorted_orig.p;
  p.cell_start = c->cell_start.p;
  p.pos_sorted = c->pos_sorted.p;
  p.xq = c->xq.p;
  p.cg = c->cgrid.p;
  p.bc = c->bc.p;
  p.eslot = c->efix.p + 2;
  p.grec = (GatherRec*)c->pme_grec.p;
  p.grec_n = c->pme_grec_n.p;
  p.fpart = c->fpart.p;
  p.st = c->dstate.p;
  const NbArgs na = nb_args(c);
  const NbPtrs P = nb_ptrs(c);
  MDX_CHECK(na.ctab && c->ctab_beta == c->beta && c->ctab_rc == c->rc, MDX_ERR_INVALID,
            "Ewald-correction table not built for the current beta and cutoff");
  BondedArgs ba{};
  const int nbo = bonded_terms ? mdx_bonded_args(c, bonded_terms, &ba, PLANE_THREADS) : 0;
  const int G = c->nb_lanes;
  const int nbt = (int)div_up((long long)c->n * G, PLANE_THREADS);
  const int n0 = std::min(nbt, (int)(c->phase_frac[0] * nbt + 0.5));
  const int n1 = std::min(nbt - n0, (int)(c->phase_frac[1] * nbt + 0.5));
  const int n2 = nbt - n0 - n1;
  const int nx = (p.lnx.nlines + p.lnx.B - 1) / p.lnx.B;  // x-pass workgroups
  const size_t lx = (2 * (size_t)p.lnx.B * (Kx + 1) + Kx) * sizeof(float2);
  PhaseKernel k = phase_kernel(energy, G);
  k<<<Kx + n0, PLANE_THREADS, plane_fwd_lds(Ky, Kz, a.W), c->stream>>>(0, Kx, 0, n0, p, na, P, ba);
  MDX_LAUNCH_CHECK();
  k<<<nx + n1, PLANE_THREADS, lx, c->stream>>>(1, nx, n0, n1, p, na, P, ba);
  MDX_LAUNCH_CHECK();
  k<<<Kx + n2 + nbo, PLANE_THREADS, plane_gather_lds(Ky, Kz), c->stream>>>(2, Kx, n0 + n1, n2, p, na, P, ba);
  MDX_LAUNCH_CHECK();
  c->fpart_live = true;
  return 0;
}

extern "C" int mdx_force_phases_config(mdx_ctx* c, int enable, double frac0, double frac1) {
  MDX_CHECK(c, MDX_ERR_INVALID, "mdx_force_phases_config: no context");
  MDX_CHECK(frac0 >= 0 && frac1 >= 0 && frac0 + frac1 <= 1, MDX_ERR_INVALID,
            "nonbonded shares must be >= 0 and sum to at most 1");
  c->phases_on = enable != 0;
  c->phase_frac[0] = frac0;
  c->phase_frac[1] = frac1;
  mdx_invalidate_graph(c);
  return 0;
}

int mdx_pme_enqueue(mdx_ctx* c, bool energy) {
  (void)energy;
  MDX_CHECK(c->pme_configured, MDX_ERR_INVALID, "PME not configured");
  if (c->pme_planes && c->nb_configured) {
    MDX_CHECK(c->fpart.p && c->fpart.n >= 12 * (size_t)c->n, MDX_ERR_INVALID, "PME configured for another system size");
    int rc;
    if ((rc = pme_enqueue_planes(c))) return rc;
    c->fpart_live = true;
    return 0;
  }
  c->fpart_live = false;
  PmeArgs a;
  a.n = c->n;
  a.Kx = c->K[0];
  a.Ky = c->K[1];
  a.Kz = c->K[2];
  a.Lx = c->box[0];
  a.Ly = c->box[1];
  a.Lz = c->box[2];
  a.iLx = 1.0 / c->box[0];
  a.iLy = 1.0 / c->box[1];
  a.iLz = 1.0 / c->box[2];
  k_spread<<<div_up(16LL * c->n, PME_BLOCK), PME_BLOCK, 0, c->stream>>>(a, c->pos, c->charge_s.p, c->qfix.p);
  MDX_LAUNCH_CHECK();
  int rc;
  if ((rc = fft_pass(c, 2, -1, 1))) return rc;  // z forward, from the fixed-point grid
  if ((rc = fft_pass(c, 1, -1, 0))) return rc;  // y forward
  if ((rc = fft_pass(c, 0, -1, 3))) return rc;  // x forward, * B C, energy, x inverse
  if ((rc = fft_pass(c, 1, 1, 0))) return rc;   // y inverse
  if ((rc = fft_pass(c, 2, 1, 2))) return rc;   // z inverse, real part out
  k_gather<<<div_up(c->n, PME_BLOCK), PME_BLOCK, 0, c->stream>>>(a, c->pos, c->charge_s.p, c->conv.p, c->ffix.p);
  MDX_LAUNCH_CHECK();
  return 0;
}


// ---------------------------------------------------------------- batched replicas
// The three plane-path launches of every replica of a batch: each replica's arguments are the
// record a single-context launch passes by value.
struct PlaneCall {
  PlaneArgs a;
  FftPlan plz, ply;
  const float2* twz;
  const float2* twy;
  const int* sorted_orig;
  const int* cell_start;
  const double* pos_sorted;
  const float4* xq;
  float2* cg;
  GatherRec* grec;
  int* grec_n;
  double* fpart;
  DevState* st;
};
struct FftCall {
  FftPlan pl;
  FftLines ln;
  const float2* tw;
  float2* cg;
  long long* qfix;
  float* rout;
  const float* bc;
  long long* eslot;
};

// Batched plane kernels of BS lanes: PLANE_THREADS, or fewer when no replica has 64 x 64 planes
// (which need all PLANE_THREADS lanes for their wave-level transforms).
template <int BS>
__global__ __launch_bounds__(BS) void k_pme_plane_fwd_batch(BatchMap map, const PlaneCall* __restrict__ calls) {
  extern __shared__ float2 lds[];
  int x0;
  const PlaneCall& p = calls[batch_replica(map, &x0)];
  plane_fwd_body(p.a, p.plz, p.ply, p.twz, p.twy, p.sorted_orig, p.cell_start, p.pos_sorted, p.xq, p.cg, p.grec,
                 p.grec_n, p.st, x0, lds, BS == PLANE_THREADS);
}
__global__ __launch_bounds__(FFT_THREADS) void k_fft_axis_batch(BatchMap map, const FftCall* __restrict__ calls) {
  extern __shared__ float2 lds[];
  int ch;
  const FftCall& f = calls[batch_replica(map, &ch)];
  fft_axis_body<-1, 3>(f.pl, f.ln, f.tw, f.cg, f.qfix, f.rout, f.bc, f.eslot, ch * f.ln.B, lds);
}
template <int BS>
__global__ __launch_bounds__(BS) void k_pme_plane_inv_gather_batch(BatchMap map, const PlaneCall* __restrict__ calls) {
  extern __shared__ float2 lds[];
  int x0;
  const PlaneCall& p = calls[batch_replica(map, &x0)];
  plane_gather_body(p.a, p.plz, p.ply, p.twz, p.twy, p.sorted_orig, p.cell_start, p.pos_sorted, p.xq, p.cg, p.grec,
                    p.grec_n, p.fpart, p.st, x0, lds, BS == PLANE_THREADS);
}
// Without 64 x 64 planes the batched forward kernel runs 256-lane workgroups with a window of
// 1,024 candidates (more planes in flight: 109 vs 133 us over the 64 water replicas); the gather
// keeps PLANE_THREADS lanes (159 us, against 172 with 256 lanes and 180 with 512).
#define PLANE_BATCH_SMALL 256
#define PLANE_BATCH_W 1024

int mdx_batch_prepare_pme(mdx_ctx* const* cs, int n, BatchKind* fwd, BatchKind* fft, BatchKind* gat) {
  std::vector<PlaneCall> pc;
  std::vector<FftCall> fc;
  std::vector<int> bp, bf;
  size_t lf = 0, lx = 0, lg = 0;
  bool any64 = false;
  for (int r = 0; r < n; ++r) any64 |= cs[r]->K[1] == 64 && cs[r]->K[2] == 64;
  const int fthreads = any64 ? PLANE_THREADS : PLANE_BATCH_SMALL;
  for (int r = 0; r < n; ++r) {
    mdx_ctx* c = cs[r];
    MDX_CHECK(c->pme_configured && c->pme_planes && c->nb_configured && c->has_charges, MDX_ERR_INVALID,
              "replica %d: batched steps need the PME plane path", r);
    PlaneCall p;
    plane_args_of(c, &p.a);
    if (!any64) p.a.W = std::min(p.a.W, PLANE_BATCH_W);
    MDX_CHECK(c->pme_grec.n >= (size_t)c->K[0] * p.a.rec_cap * 6 && c->fpart.p && c->fpart.n >= 12 * (size_t)c->n,
              MDX_ERR_INVALID, "replica %d: PME configured for another system size", r);
    factorize(c->K[2], &p.plz);
    factorize(c->K[1], &p.ply);
    const int Kmax = std::max(c->K[0], std::max(c->K[1], c->K[2]));
    p.twy = c->twiddle.p + (size_t)Kmax;
    p.twz = c->twiddle.p + 2 * (size_t)Kmax;
    p.sorted_orig = c->sorted_orig.p;
    p.cell_start = c->cell_start.p;
    p.pos_sorted = c->pos_sorted.p;
    p.xq = c->xq.p;
    p.cg = c->cgrid.p;
    p.grec = (GatherRec*)c->pme_grec.p;
    p.grec_n = c->pme_grec_n.p;
    p.fpart = c->fpart.p;
    p.st = c->dstate.p;
    pc.push_back(p);
    bp.push_back(c->K[0]);
    lf = std::max(lf, plane_fwd_lds(c->K[1], c->K[2], p.a.W));
    lg = std::max(lg, plane_gather_lds(c->K[1], c->K[2]));
    FftCall f;
    size_t l;
    fft_axis_of(c, 0, &f.pl, &f.ln, &f.tw, &l);
    f.cg = c->cgrid.p;
    f.qfix = c->qfix.p;
    f.rout = c->conv.p;
    f.bc = c->bc.p;
    f.eslot = c->efix.p + 2;
    fc.push_back(f);
    bf.push_back(div_up(f.ln.nlines, f.ln.B));
    lx = std::max(lx, l);
  }
  const void* kf = fthreads == PLANE_THREADS ? (const void*)k_pme_plane_fwd_batch<PLANE_THREADS>
                                             : (const void*)k_pme_plane_fwd_batch<PLANE_BATCH_SMALL>;
  MDX_HIP(hipFuncSetAttribute(kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lf));
  MDX_HIP(hipFuncSetAttribute((const void*)k_fft_axis_batch, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lx));
  MDX_HIP(hipFuncSetAttribute((const void*)k_pme_plane_inv_gather_batch<PLANE_THREADS>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lg));
  int rc;
  if ((rc = batch_kind_upload(fwd, pc, bp, lf)) || (rc = batch_kind_upload(fft, fc, bf, lx)) ||
      (rc = batch_kind_upload(gat, pc, bp, lg)))
    return rc;
  fwd->threads = fthreads;
  fft->threads = FFT_THREADS;
  gat->threads = PLANE_THREADS;
  return 0;
}

int mdx_batch_launch_pme(const BatchKind& fwd, const BatchKind& fft, const BatchKind& gat, hipStream_t s,
                         unsigned parts) {
  const bool big = fwd.threads == PLANE_THREADS;
  if (parts & 1u) {
    if (big)
      k_pme_plane_fwd_batch<PLANE_THREADS><<<fwd.blocks, PLANE_THREADS, fwd.lds, s>>>(fwd.map, (const PlaneCall*)fwd.calls);
    else
      k_pme_plane_fwd_batch<PLANE_BATCH_SMALL><<<fwd.blocks, PLANE_BATCH_SMALL, fwd.lds, s>>>(fwd.map,
                                                                                          (const PlaneCall*)fwd.calls);
    MDX_LAUNCH_CHECK();
  }
  if (parts & 2u) {
    k_fft_axis_batch<<<fft.blocks, FFT_THREADS, fft.lds, s>>>(fft.map, (const FftCall*)fft.calls);
    MDX_LAUNCH_CHECK();
  }
  if (parts & 4u) {
    k_pme_plane_inv_gather_batch<PLANE_THREADS><<<gat.blocks, PLANE_THREADS, gat.lds, s>>>(
        gat.map, (const PlaneCall*)gat.calls);
    MDX_LAUNCH_CHECK();
  }
  return 0;
}
