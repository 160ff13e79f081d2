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
//   corrections  float64: excluded-pair erf correction, self energy, neutralising background
#include "mdx_internal.h"
#include <cmath>

#define PME_BLOCK 256
#define FFT_THREADS 256
#define MAX_RADICES 16

struct FftPlan {
  int K;
  int nrad;
  int radix[MAX_RADICES];
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

// One transform of B lines of length K held in LDS (src), Stockham passes ping-ponging with
// dst. Returns the buffer holding the result.
template <int DIR>
__device__ float2* lds_fft(float2* src, float2* dst, const FftPlan& pl, int B, const float2* __restrict__ tw) {
  const int K = pl.K;
  int p = 1;
  const float s = (float)DIR;
  for (int pass = 0; pass < pl.nrad; ++pass) {
    const int R = pl.radix[pass];
    const int T = K / R;
    const int tstep = K / (p * R);
    for (int b = threadIdx.x; b < B * T; b += blockDim.x) {
      const int l = b / T, i = b - l * T;
      const int k = i % p;
      const float2* X = src + l * K;
      float2* Y = dst + l * K;
      float2 u[5];
      for (int r = 0; r < R; ++r) u[r] = X[i + r * T];
      for (int r = 1; r < R; ++r) {
        float2 w = tw[(r * k * tstep) % K];
        if (DIR > 0) w.y = -w.y;
        u[r] = cmul(u[r], w);
      }
      if (R == 2) dft2(u);
      else if (R == 3) dft3(u, s);
      else if (R == 4) dft4(u, s);
      else dft5(u, s);
      const int j = (i - k) * R + k;
      for (int r = 0; r < R; ++r) Y[j + r * p] = u[r];
    }
    __syncthreads();
    float2* t = src;
    src = dst;
    dst = t;
    p *= R;
  }
  return src;
}

__device__ __forceinline__ long long line_base(const FftLines& ln, int l) {
  return (long long)(l / ln.inner) * ln.outer_stride + (long long)(l % ln.inner) * ln.inner_stride;
}

// MODE 0: complex in/out.  MODE 1: load int64 fixed grid (real), clear it.  MODE 2: store real
// part to a float grid.  MODE 3: forward, multiply by B*C (+ energy), inverse, in place.
template <int DIR, int MODE>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_axis(FftPlan pl, FftLines ln, const float2* __restrict__ tw,
                                                        float2* __restrict__ cg, long long* __restrict__ qfix,
                                                        float* __restrict__ rout, const float* __restrict__ bc,
                                                        long long* __restrict__ eslot) {
  extern __shared__ float2 lds[];
  const int K = pl.K;
  const int B = ln.B;
  const int l0 = blockIdx.x * B;
  const int nb = min(B, ln.nlines - l0);
  float2* A = lds;
  float2* Bf = lds + B * K;
  const bool contiguous = ln.estride == 1;
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
    A[l * K + e] = v;
  }
  __syncthreads();
  float2* res = lds_fft<DIR>(A, Bf, pl, nb, tw);
  if (MODE == 3) {
    float2* other = (res == A) ? Bf : A;
    double en = 0.0;
    for (int idx = threadIdx.x; idx < nb * K; idx += blockDim.x) {
      int l = idx / K, e = idx - l * K;
      long long g = line_base(ln, l0 + l) + (long long)e * ln.estride;
      float f = bc[g];
      float2 v = res[idx];
      en += 0.5 * (double)f * ((double)v.x * v.x + (double)v.y * v.y);
      res[idx] = cscale(v, f);
    }
    __syncthreads();
    res = lds_fft<-DIR>(res, other, pl, nb, tw);
    // block energy
    __shared__ double sm[FFT_THREADS / 64];
    for (int off = 32; off > 0; off >>= 1) en += __shfl_xor(en, off, 64);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = en;
    __syncthreads();
    if (threadIdx.x == 0) {
      double sacc = 0;
      for (int w = 0; w < FFT_THREADS / 64; ++w) sacc += sm[w];
      atomicAdd((unsigned long long*)eslot, (unsigned long long)to_fix(sacc));
    }
  }
  // store
  for (int idx = threadIdx.x; idx < nb * K; idx += blockDim.x) {
    int l, e;
    if (contiguous) { l = idx / K; e = idx - l * K; }
    else { e = idx / nb; l = idx - e * nb; }
    long long g = line_base(ln, l0 + l) + (long long)e * ln.estride;
    float2 v = res[l * K + e];
    if (MODE == 2) rout[g] = v.x;
    else cg[g] = v;
  }
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
};

__device__ __forceinline__ void frac_coord(double x, double L, int K, int* base, float* w) {
  double s = x / L;
  s -= floor(s);
  double u = s * K;
  double f = floor(u);
  int b = (int)f;
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
  frac_coord(pos[3 * i], a.Lx, a.Kx, &bx, &wx);
  frac_coord(pos[3 * i + 1], a.Ly, a.Ky, &by, &wy);
  frac_coord(pos[3 * i + 2], a.Lz, a.Kz, &bz, &wz);
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
  frac_coord(pos[3 * i], a.Lx, a.Kx, &bx, &wx);
  frac_coord(pos[3 * i + 1], a.Ly, a.Ky, &by, &wy);
  frac_coord(pos[3 * i + 2], a.Lz, a.Kz, &bz, &wz);
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

// excluded pairs: -ke q_i q_j erf(beta r)/r (float64); block 0 also adds self + background
__global__ void k_pme_excl(int npairs, int n, const int* __restrict__ pairs, const double* __restrict__ pos,
                           const double* __restrict__ q, double ke, double beta, double Lx, double Ly, double Lz,
                           double e_const, long long* __restrict__ ffix, long long* __restrict__ eslot) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0;
  if (t < npairs) {
    int i = pairs[2 * t], j = pairs[2 * t + 1];
    double L[3] = {Lx, Ly, Lz}, d[3];
    for (int k = 0; k < 3; ++k) {
      double v = pos[3 * j + k] - pos[3 * i + k];
      double s = v / L[k];
      s -= rint(s);
      d[k] = s * L[k];
    }
    double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double qq = ke * q[i] * q[j];
    double er = erf(beta * r);
    e = -qq * er / r;
    double dEdr = qq * (er / (r * r) - 2.0 * beta / sqrt(M_PI) * exp(-beta * beta * r * r) / r);
    double g = dEdr / r;  // force on i = +g d, on j = -g d
    atomicAdd((unsigned long long*)&ffix[i], (unsigned long long)to_fix(g * d[0]));
    atomicAdd((unsigned long long*)&ffix[n + i], (unsigned long long)to_fix(g * d[1]));
    atomicAdd((unsigned long long*)&ffix[2 * n + i], (unsigned long long)to_fix(g * d[2]));
    atomicAdd((unsigned long long*)&ffix[j], (unsigned long long)to_fix(-g * d[0]));
    atomicAdd((unsigned long long*)&ffix[n + j], (unsigned long long)to_fix(-g * d[1]));
    atomicAdd((unsigned long long*)&ffix[2 * n + j], (unsigned long long)to_fix(-g * d[2]));
  }
  if (t == 0) e += e_const;
  __shared__ double sm[PME_BLOCK / 64];
  for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off, 64);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
    for (int w = 0; w < PME_BLOCK / 64; ++w) s += sm[w];
    atomicAdd((unsigned long long*)eslot, (unsigned long long)to_fix(s));
  }
}

// ---------------------------------------------------------------- host
static int factorize(int K, FftPlan* pl) {
  pl->K = K;
  pl->nrad = 0;
  int k = K;
  while (k % 4 == 0) { pl->radix[pl->nrad++] = 4; k /= 4; }
  while (k % 2 == 0) { pl->radix[pl->nrad++] = 2; k /= 2; }
  while (k % 3 == 0) { pl->radix[pl->nrad++] = 3; k /= 3; }
  while (k % 5 == 0) { pl->radix[pl->nrad++] = 5; k /= 5; }
  return k == 1 ? 0 : -1;
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
  size_t G = (size_t)c->K[0] * c->K[1] * c->K[2];
  int rc;
  if ((rc = c->qfix.ensure(G)) || (rc = c->cgrid.ensure(G)) || (rc = c->conv.ensure(G)) || (rc = c->bc.ensure(G)))
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
  c->pme_configured = true;
  return 0;
}

static int fft_pass(mdx_ctx* c, int axis, int dir, int mode) {
  const int Kx = c->K[0], Ky = c->K[1], Kz = c->K[2];
  FftPlan pl;
  FftLines ln;
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
  const float2* tw = c->twiddle.p + (size_t)axis * Kmax;
  size_t lds = 2 * (size_t)ln.B * pl.K * sizeof(float2);
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

int mdx_pme_enqueue(mdx_ctx* c, bool energy) {
  (void)energy;
  MDX_CHECK(c->pme_configured, MDX_ERR_INVALID, "PME not configured");
  PmeArgs a;
  a.n = c->n;
  a.Kx = c->K[0];
  a.Ky = c->K[1];
  a.Kz = c->K[2];
  a.Lx = c->box[0];
  a.Ly = c->box[1];
  a.Lz = c->box[2];
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
  return mdx_pme_enqueue_corrections(c);
}

int mdx_pme_enqueue_corrections(mdx_ctx* c) {
  const double ke = 1.0 / (4.0 * M_PI * 0.5727652348434076);
  double V = c->box[0] * c->box[1] * c->box[2];
  double e_self = -ke * c->beta / sqrt(M_PI) * c->q2sum;
  double e_bg = -ke * M_PI * c->qsum * c->qsum / (2.0 * c->beta * c->beta * V);
  int np = std::max(1, c->n_excl_pairs);
  k_pme_excl<<<div_up(np, PME_BLOCK), PME_BLOCK, 0, c->stream>>>(
      c->n_excl_pairs, c->n, c->excl_pairs.p, c->pos, c->charge64.p, ke, c->beta, c->box[0], c->box[1], c->box[2],
      e_self + e_bg, c->ffix.p, c->efix.p + 2);
  MDX_LAUNCH_CHECK();
  return 0;
}
