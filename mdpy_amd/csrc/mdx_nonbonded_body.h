// The nonbonded pair evaluation (CHARMM LJ + erfc-screened Coulomb) as a workgroup-level device
// function, shared by the nonbonded launch and the phase launches that run it beside the PME
// reciprocal kernels.
//
// LJ physics follows mdpy/constraint/charmm_nonbonded_constraint.py:88-107:
//   eps = sqrt(eps_i eps_j), sigma = (sigma_i + sigma_j)/2 ((eps14, sigma14) for 1-4 partners),
//   F_i = -24 eps (2 (s/r)^12 - (s/r)^6) / r^2 * d,  E = 4 eps ((s/r)^12 - (s/r)^6),  r <= rc,
//   with d = minimum-image x_j - x_i.
// The Coulomb part is the PME direct space q_i q_j erfc(beta r)/r over the same pair set.
//
// Layout: a group of G lanes per atom (16, 32 or 64, chosen by mdx_nonbonded_tune) in
// cell-sorted order walks the atom's contiguous
// neighbour list (coalesced index loads), gathers each neighbour's float4 (x, y, z, q*sqrt(ke))
// and float2 (sigma/2, sqrt(24 eps)), and keeps partial forces in registers. Every pair is
// evaluated from both ends (full list), so there are no atomics in the pair loop; the group
// reduces with xor shuffles and adds each atom's force once to the int64 fixed-point
// accumulator (deterministic).
#pragma once
#include "mdx_internal.h"

#ifndef NB_U
#define NB_U 4  // list entries per lane per pass
#endif

struct NbArgs {
  int n, cap;
  float Lx, Ly, Lz, ix, iy, iz;
  float rc2, guard2;
  float beta, twob_sqrtpi;
  float ct_scale;  // Ewald-correction table pieces per A^2
  double Lxd, Lyd, Lzd, rcd;
  const float4* ctab;
};

// Forces-only launches evaluate the direct-space Coulomb force as
//   F_ij / (q_i q_j) = d (1/r^3 - G(r^2)),  G(z) = erf(beta r)/r^3 - (2 beta/sqrt(pi)) exp(-beta^2 z)/z,
// which equals erfc(beta r)/r^3 + (2 beta/sqrt(pi)) exp(-beta^2 r^2)/r^2. G is smooth and finite
// on [0, rc^2] (G(0) = 4 beta^3/(3 sqrt(pi))), so MDX_CTAB_N cubic Hermite pieces in z = r^2,
// built in float64 from G and G' at the knots, reproduce it to ~1e-9 relative; each pair then
// needs one LDS read and three FMAs instead of exp, a reciprocal and the erfc polynomial. The
// subtraction leaves an absolute error of ~1 ulp of 1/r^3 per pair (relative error grows
// towards the cutoff, where the force itself is small). Energy launches keep the erfc form.

// erfc(x) = exp(-x^2) t P(t), t = 1/(1 + p x): the form of Abramowitz & Stegun 7.1.26 refit
// to a 7-term polynomial (least squares at Chebyshev nodes, reweighted towards minimax relative
// error, Lawson's algorithm) with max relative error 5.5e-8 on [0, 3.6] with the coefficients
// rounded to float32 (beta * rc = 3.29 for the 1e-5 Ewald tolerance), so pairs near the cutoff,
// where erfc ~ 1e-5, keep float32 relative accuracy (tests/test_erfc_fit.py).
#define ERFC_P 0.355f
#define ERFC_C1 1.9758294902e-01f
#define ERFC_C2 2.3220149337e-01f
#define ERFC_C3 2.8235364196e-02f
#define ERFC_C4 5.9699421939e-01f
#define ERFC_C5 -5.6119238634e-01f
#define ERFC_C6 6.9343106043e-01f
#define ERFC_C7 -1.8725274521e-01f

// float64 cutoff decision with the reference's minimum-image sequence (pbc.py:40-44)
static __device__ __noinline__ bool within_cutoff64(const double* __restrict__ pos, int oi, int oj, double Lx, double Ly,
                                              double Lz, double rc) {
  double L[3] = {Lx, Ly, Lz};
  double d[3];
  for (int q = 0; q < 3; ++q) {
    double v = pos[3 * oj + q] - pos[3 * oi + q];
    double sv = v * (1.0 / L[q]);
    sv -= rint(sv);
    d[q] = sv * L[q];
  }
  return sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) <= rc;
}

static __device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <bool B>
struct NbTail {
  static constexpr bool value = B;
};

// One workgroup (BS lanes, virtual index vb) of the nonbonded kernel. G lanes share one atom:
// lane g walks list entries g, g + G, ... of the atom's contiguous (row-major) list, so a group's
// index loads and its neighbours' position loads are contiguous; the group reduces its partial
// forces with xor shuffles at the end.
struct NbPtrs {
  const float4* xq;
  const float2* ljs;
  const float2* ljs14;
  const int* nbr_count;
  const unsigned* nbr_list;
  const int* sorted_orig;
  const double* pos;
  long long* ffix;
  long long* efix;
  DevState* st;
};

// S14: the list may hold 1-4 pairs (bit 31), which take the (sigma14, eps14) columns; without
// any, the per-slot column choice is compiled out.
template <bool LJ, bool COUL, bool ENERGY, int G, int BS, bool S14>
__device__ __forceinline__ void nb_block(const NbArgs& a, const NbPtrs& P, int vb) {
  const float4* __restrict__ xq = P.xq;
  const float2* __restrict__ ljs = P.ljs;
  const float2* __restrict__ ljs14 = P.ljs14;
  const int* __restrict__ nbr_count = P.nbr_count;
  const unsigned* __restrict__ nbr_list = P.nbr_list;
  const int* __restrict__ sorted_orig = P.sorted_orig;
  const double* __restrict__ pos = P.pos;
  long long* __restrict__ ffix = P.ffix;
  long long* __restrict__ efix = P.efix;
  DevState* st = P.st;
  constexpr bool TAB = COUL && !ENERGY;
#if NB_TAB_GLOBAL
  const float4* __restrict__ ctab = a.ctab;
#else
  __shared__ float4 ctab[TAB ? MDX_CTAB_N : 1];
  if (TAB) {
    for (int k = threadIdx.x; k < MDX_CTAB_N; k += blockDim.x) ctab[k] = a.ctab[k];
    __syncthreads();
  }
#endif
  const int gt = vb * BS + threadIdx.x;
  const int s = gt / G;
  const int g = gt & (G - 1);
  if (gt == 0) {  // every rebuild kernel has finished reading these
    st->rebuild = 0;
    st->max_disp2_bits = 0;
  }
  float fx = 0.f, fy = 0.f, fz = 0.f;
  double elj = 0.0, ec = 0.0;
  if (s < a.n) {
    const float4 xi = xq[s];
    float2 pi = make_float2(0.f, 0.f), pi14 = pi;
    if (LJ) {
      pi = ljs[s];
      if (S14) pi14 = ljs14[s];
    }
    const int cnt = nbr_count[s];
    const unsigned* lp = nbr_list + (long long)s * a.cap + g;
    // neighbours are addressed by 32-bit byte offsets from the array bases (atom indices stay
    // below 2^28), and the shifts that form them also drop the 1-4 flag in bit 31
    const char* xqb = reinterpret_cast<const char*>(xq);
    const char* ljb = reinterpret_cast<const char*>(ljs);
    const char* lj14b = reinterpret_cast<const char*>(ljs14);
    int oi = -1;
    // one pass over NB_U list entries per lane: all neighbour loads are issued before any is
    // consumed, so each pass waits for memory once, not 2*NB_U times. In the tail pass (tail =
    // NbTail<true>) bit u of `valid` marks the slots inside the row; the others hold the atom
    // itself (zero separation: finite, and zeroed below)
    auto pass = [&](const unsigned* e, unsigned valid, auto tail) {
      constexpr bool TAIL = decltype(tail)::value;
      float4 xj[NB_U];
      float2 pj[NB_U];
#pragma unroll
      for (int u = 0; u < NB_U; ++u) {
        xj[u] = *reinterpret_cast<const float4*>(xqb + (e[u] << 4));
        if (LJ) pj[u] = *reinterpret_cast<const float2*>(((S14 && (e[u] >> 31)) ? lj14b : ljb) + (e[u] << 3));
      }
#pragma unroll
      for (int u = 0; u < NB_U; ++u) {
        float dx = xj[u].x - xi.x, dy = xj[u].y - xi.y, dz = xj[u].z - xi.z;
        dx -= a.Lx * rintf(dx * a.ix);
        dy -= a.Ly * rintf(dy * a.iy);
        dz -= a.Lz * rintf(dz * a.iz);
        const float r2 = dx * dx + dy * dy + dz * dz;
        bool in = r2 <= a.rc2 && (!TAIL || ((valid >> u) & 1u));
        if (fabsf(r2 - a.rc2) < a.guard2) {
          // near the cutoff: decide in float64 like the reference (rare)
          if (oi < 0) oi = sorted_orig[s];
          in = within_cutoff64(pos, oi, sorted_orig[e[u] & 0x7fffffffu], a.Lxd, a.Lyd, a.Lzd, a.rcd);
        }
        // out-of-cutoff and padding slots compute with their own r2 (> rc^2, or 0) and are
        // zeroed below; in-cutoff pairs have r2 >= ~0.5 A^2, never denormal for v_rsq_f32
        const float rinv = __builtin_amdgcn_rsqf(r2);
        const float r2inv = rinv * rinv;
        float fr = 0.f;
        if (LJ) {
          const float2 pp = (S14 && (e[u] >> 31)) ? pi14 : pi;
          const float sig = pp.x + pj[u].x;
          const float eps = pp.y * pj[u].y;
          const float s2 = sig * sig * r2inv;
          const float s6 = s2 * s2 * s2;
          const float s12 = s6 * s6;
          fr -= eps * (2.f * s12 - s6) * r2inv;  // eps holds 24 eps_ij
          if (ENERGY) elj += in ? (double)(eps * (1.f / 6.f) * (s12 - s6)) : 0.0;
        }
        if (TAB) {
          const float zt = r2 * a.ct_scale;
          const float fl = floorf(zt);
          const float t = zt - fl;
          // clamp in float before the conversion: beyond the cutoff the index stays inside
          // the table
          const float4 cf = ctab[(int)fminf(fl, (float)(MDX_CTAB_N - 1))];
          const float gz = fmaf(t, fmaf(t, fmaf(t, cf.w, cf.z), cf.y), cf.x);
          fr -= xi.w * xj[u].w * (rinv * r2inv - gz);
        } else if (COUL) {
          const float qq = xi.w * xj[u].w;
          const float x = a.beta * (r2 * rinv);
          const float ex = __expf(-x * x);
          // 1/(1 + p x): v_rcp_f32 plus one Newton step (the raw 1-ulp reciprocal is biased, and the
          // direct-space energy sum cancels strongly enough to show it)
          const float den = 1.f + ERFC_P * x;
          const float r0 = __builtin_amdgcn_rcpf(den);
          const float tt = fmaf(r0, fmaf(-den, r0, 1.f), r0);
          const float poly =
              ERFC_C1 + tt * (ERFC_C2 + tt * (ERFC_C3 + tt * (ERFC_C4 + tt * (ERFC_C5 + tt * (ERFC_C6 + tt * ERFC_C7)))));
          const float erfcx = tt * poly * ex;
          fr -= qq * (erfcx * rinv + a.twob_sqrtpi * ex) * r2inv;
          if (ENERGY) ec += in ? (double)(qq * erfcx * rinv) : 0.0;
        }
        fr = in ? fr : 0.f;
        fx += fr * dx;
        fy += fr * dy;
        fz += fr * dz;
      }
    };
    // whole chunks of G * NB_U entries: unpredicated index loads, the next chunk's issued before
    // this chunk is evaluated (the rows are followed by slack, so a load past a row's end reads
    // the next row or the slack and is never used)
    int cb = 0;
    if (G * NB_U <= cnt) {
      unsigned e[NB_U];
#pragma unroll
      for (int u = 0; u < NB_U; ++u) e[u] = lp[u * G];
      for (; cb + G * NB_U <= cnt; cb += G * NB_U) {
        unsigned en[NB_U];
#pragma unroll
        for (int u = 0; u < NB_U; ++u) en[u] = lp[cb + G * NB_U + u * G];
        pass(e, 0u, NbTail<false>());
#pragma unroll
        for (int u = 0; u < NB_U; ++u) e[u] = en[u];
      }
    }
    if (cb < cnt) {  // the row's last, partial chunk: slots past the end take the atom itself
      unsigned e[NB_U], valid = 0;
#pragma unroll
      for (int u = 0; u < NB_U; ++u) {
        const bool v = cb + g + u * G < cnt;
        e[u] = v ? lp[cb + u * G] : (unsigned)s;
        valid |= (v ? 1u : 0u) << u;
      }
      pass(e, valid, NbTail<true>());
    }
  }
  for (int off = G / 2; off > 0; off >>= 1) {
    fx += __shfl_xor(fx, off, 64);
    fy += __shfl_xor(fy, off, 64);
    fz += __shfl_xor(fz, off, 64);
  }
  if (s < a.n && g == 0) {
    const int o = sorted_orig[s];
    atomicAdd((unsigned long long*)&ffix[o], (unsigned long long)to_fix_f(fx));
    atomicAdd((unsigned long long*)&ffix[a.n + o], (unsigned long long)to_fix_f(fy));
    atomicAdd((unsigned long long*)&ffix[2 * a.n + o], (unsigned long long)to_fix_f(fz));
  }
  if (ENERGY) {
    __shared__ double sm[2][BS / 64];
    elj = wave_sum(elj);
    ec = wave_sum(ec);
    if ((threadIdx.x & 63) == 0) {
      sm[0][threadIdx.x >> 6] = elj;
      sm[1][threadIdx.x >> 6] = ec;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double a0 = 0, a1 = 0;
      for (int w = 0; w < BS / 64; ++w) {
        a0 += sm[0][w];
        a1 += sm[1][w];
      }
      // every pair was summed from both ends
      if (LJ) atomicAdd((unsigned long long*)&efix[0], (unsigned long long)to_fix(0.5 * a0));
      if (COUL) atomicAdd((unsigned long long*)&efix[1], (unsigned long long)to_fix(0.5 * a1));
    }
  }
}


// kernel arguments of the current configuration (the nonbonded launch and the phase launches)
NbArgs nb_args(const mdx_ctx* c);
NbPtrs nb_ptrs(const mdx_ctx* c);
