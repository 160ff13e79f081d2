// MLP force surrogate, gfx950: descriptors, BF16 MFMA dense layers, analytic forces and
// force-matching training. mdpy has no learned force field; the mathematics is defined (and
// checked) by the float64 oracle oracle/mlp.py:
//
//   descriptors  per atom and neighbour element e in {O, H}, neighbours within r_d = 6 A:
//                32 radial Gaussians sum_j exp(-eta (r - mu_k)^2) fc(r) and 16 angular invariants
//                |T_{e,n,l}|^2, T_l = sum_j g_n(r) rhat^(x l) (4 channels x l = 0..3), i.e.
//                sum_{j,k} g_n g_n (cos theta_jk)^l computed from moments in O(neighbours)
//   network      per central element: (G - mean)/std -> 3 x (256, SiLU) -> 1, BF16 weights and
//                activations, FP32 accumulation on MFMA (v_mfma_f32_32x32x16_bf16)
//   forces       F_j = sum_i dE_j/dr_ji - dE_i/dr_ij, one lane per neighbour pair, int64
//                fixed-point atomics (deterministic)
//   training     force matching; dL/dW by a forward-mode pass along u = dL/d(dE/dG) followed by
//                reverse mode, all on GEMMs + elementwise kernels; Adam on FP32 master weights
//
// Rows of every per-atom matrix are in element order (O atoms, padded to a multiple of 64,
// then H atoms, padded), so each 64-row GEMM tile uses one element's weights.
#include "mdx_internal.h"
#include <cmath>

#define ML_RAD 32
#define ML_CH 4
#define ML_L 4
#define ML_PER 48
#define ML_F 96
#define ML_MOM 20
#define ML_T 160  // 2 elements x 4 channels x 20 moment components
#define ML_H 256
#define ML_TILE 64
// parameter block of one element network (float offsets)
#define P_W1 0
#define P_B1 (P_W1 + ML_F * ML_H)
#define P_W2 (P_B1 + ML_H)
#define P_B2 (P_W2 + ML_H * ML_H)
#define P_W3 (P_B2 + ML_H)
#define P_B3 (P_W3 + ML_H * ML_H)
#define P_W4 (P_B3 + ML_H)
#define P_B4 (P_W4 + ML_H)
#define P_N (P_B4 + 1)

typedef unsigned short bf16_t;
typedef short __attribute__((ext_vector_type(4))) bf16x4;
typedef __bf16 __attribute__((ext_vector_type(8))) bf16x8;  // one gfx950 K=16 MFMA operand per lane
typedef float __attribute__((ext_vector_type(16))) f32x16;

__device__ __forceinline__ bf16_t f2bf(float f) {  // round to nearest even
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float(((unsigned)h) << 16); }

struct MlConst {
  float rd, pi_rd;
  float rad_mu0, rad_dmu, rad_eta;
  float ang_nu[ML_CH];
  float ang_eta;
};

__device__ __forceinline__ void cutoff_fc(const MlConst& k, float r, float* fc, float* dfc) {
  float s, c;
  __sincosf(k.pi_rd * r, &s, &c);
  *fc = 0.5f * (c + 1.f);
  *dfc = -0.5f * k.pi_rd * s;
}

// monomial `comp` (storage order 1 | x y z | xx xy xz yy yz zz | xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz)
__device__ __forceinline__ void mono_exps(int comp, int* a, int* b, int* c) {
  const unsigned char E[20][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0}, {1, 1, 0}, {1, 0, 1},
                                  {0, 2, 0}, {0, 1, 1}, {0, 0, 2}, {3, 0, 0}, {2, 1, 0}, {2, 0, 1}, {1, 2, 0},
                                  {1, 1, 1}, {1, 0, 2}, {0, 3, 0}, {0, 2, 1}, {0, 1, 2}, {0, 0, 3}};
  *a = E[comp][0];
  *b = E[comp][1];
  *c = E[comp][2];
}
__device__ __forceinline__ float pw3(float x, int e) { return e == 0 ? 1.f : (e == 1 ? x : (e == 2 ? x * x : x * x * x)); }
// multiplicity of a symmetric-tensor component (|T|^2 = sum_c w_c T_c^2)
__device__ __forceinline__ float mom_weight(int comp) {
  const float W[20] = {1, 1, 1, 1, 1, 2, 2, 1, 2, 1, 1, 3, 3, 3, 6, 3, 1, 3, 3, 1};
  return W[comp];
}
__device__ __forceinline__ int mom_l(int comp) { return comp == 0 ? 0 : (comp < 4 ? 1 : (comp < 10 ? 2 : 3)); }

// s_l = T_l . u^l and v_l = d(T_l . u^l)/du for one channel's 20 stored components t
__device__ __forceinline__ void contract(const float* t, float ux, float uy, float uz, float s[4], float v[4][3]) {
  const float T2[3][3] = {{t[4], t[5], t[6]}, {t[5], t[7], t[8]}, {t[6], t[8], t[9]}};
  // T3 as a function of sorted index triple; components: 10 xxx 11 xxy 12 xxz 13 xyy 14 xyz 15 xzz 16 yyy 17 yyz 18 yzz 19 zzz
  const float u[3] = {ux, uy, uz};
  s[0] = t[0];
  v[0][0] = v[0][1] = v[0][2] = 0.f;
  s[1] = t[1] * ux + t[2] * uy + t[3] * uz;
  v[1][0] = t[1];
  v[1][1] = t[2];
  v[1][2] = t[3];
  float t2u[3];
  for (int a = 0; a < 3; ++a) t2u[a] = T2[a][0] * ux + T2[a][1] * uy + T2[a][2] * uz;
  s[2] = t2u[0] * ux + t2u[1] * uy + t2u[2] * uz;
  for (int a = 0; a < 3; ++a) v[2][a] = 2.f * t2u[a];
  // M_a = sum_bc T3_abc u_b u_c
  const float xx = ux * ux, yy = uy * uy, zz = uz * uz, xy = ux * uy, xz = ux * uz, yz = uy * uz;
  float M[3];
  M[0] = t[10] * xx + t[13] * yy + t[15] * zz + 2.f * (t[11] * xy + t[12] * xz + t[14] * yz);
  M[1] = t[11] * xx + t[16] * yy + t[18] * zz + 2.f * (t[13] * xy + t[14] * xz + t[17] * yz);
  M[2] = t[12] * xx + t[17] * yy + t[19] * zz + 2.f * (t[14] * xy + t[15] * xz + t[18] * yz);
  s[3] = M[0] * u[0] + M[1] * u[1] + M[2] * u[2];
  for (int a = 0; a < 3; ++a) v[3][a] = 3.f * M[a];
}

struct MlAtoms {
  int n;        // atoms
  int stride;   // force pass: workgroup b takes sorted atom (b * stride) mod n (stride coprime to n)
  int cap;      // neighbour-list row capacity
  float Lx, Ly, Lz, ix, iy, iz;
};

// stage up to 64 neighbours of sorted slot s (lane k = base + lane): unit vector, r, element
struct NbStage {
  float ux[64], uy[64], uz[64], r[64];
  int e[64], t[64];
};

__device__ __forceinline__ void stage_neighbours(const MlAtoms& a, const MlConst& k, const float4& xi,
                                                 const float4* __restrict__ xq, const unsigned* __restrict__ row,
                                                 int cnt, int base, NbStage& S) {
  const int lane = threadIdx.x;
  const int kk = base + lane;
  float ux = 0.f, uy = 0.f, uz = 0.f, r = 1e30f;
  int e = 0, t = 0;
  if (kk < cnt) {
    t = (int)(row[kk] & 0x3fffffffu);
    const float4 xj = xq[t];
    float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
    dx -= a.Lx * rintf(dx * a.ix);
    dy -= a.Ly * rintf(dy * a.iy);
    dz -= a.Lz * rintf(dz * a.iz);
    const float rr = sqrtf(dx * dx + dy * dy + dz * dz);
    if (rr < k.rd && rr > 0.f) {
      r = rr;
      const float ir = __builtin_amdgcn_rcpf(rr);
      ux = dx * ir;
      uy = dy * ir;
      uz = dz * ir;
    }
    e = (int)xj.w;
  }
  S.ux[lane] = ux;
  S.uy[lane] = uy;
  S.uz[lane] = uz;
  S.r[lane] = r;
  S.e[lane] = e;
  S.t[lane] = t;
}

// ---------------------------------------------------------------- descriptors
// One wave per atom. Each chunk of 64 list neighbours is first evaluated with one lane per
// neighbour: the in-range ones (r < r_d) are compacted per element, in list order, into LDS rows
// (element 0 from the front, element 1 from the back) holding r, the cutoff function, the 4
// angular channel weights (times the cutoff function) and the 20 unit-vector monomials. Then, per
// element, the lanes walk that element's rows: the radial Gaussians two rows at a time (lanes
// 0-31 the even row, lanes 32-63 the odd one, Gaussian lane & 31, the halves added at the end),
// the moment components c = lane and 64 + lane (< 80) of the element's tensor one row at a time.
// Per row this is 1 exponential per lane pair and 2 products, against 32 exponentials and 3
// element-selected products when every lane walks every row.
__global__ __launch_bounds__(64) void k_ml_desc(MlAtoms a, MlConst k, const float4* __restrict__ xq,
                                               const int* __restrict__ sorted_orig, const int* __restrict__ nbr_count,
                                               const unsigned* __restrict__ nbr_list, const int* __restrict__ row_of,
                                               const float* __restrict__ mean, const float* __restrict__ rstd,
                                               float* __restrict__ G, float* __restrict__ T, bf16_t* __restrict__ X) {
  __shared__ float rr[64], rfc[64];
  __shared__ float ga[64][ML_CH];
  __shared__ float mo[64][ML_MOM + 1];
  __shared__ float sm[ML_T];
  const int s = blockIdx.x;
  const int lane = threadIdx.x;
  const float4 xi = xq[s];
  const int o = sorted_orig[s];
  const int row = row_of[o];
  const int cnt = nbr_count[s];
  const unsigned* lrow = nbr_list + (long long)s * a.cap;
  const int half = lane >> 5;
  const float mu = k.rad_mu0 + (float)(lane & 31) * k.rad_dmu;  // this lane's Gaussian centre
  // moment components of this lane within an element's 80: c0 = lane, c1 = 64 + lane (lanes < 16)
  const int ch0 = lane / ML_MOM, cp0 = lane % ML_MOM;
  const int c1 = 64 + (lane & 15);
  const int ch1 = c1 / ML_MOM, cp1 = c1 % ML_MOM;
  const bool has1 = lane < 16;
  float rad[2] = {0.f, 0.f};  // per element: Gaussian lane & 31 over this lane's half of the rows
  float m0[2] = {0.f, 0.f}, m1[2] = {0.f, 0.f};
  for (int base = 0; base < cnt; base += 64) {
    // lane = neighbour: geometry, in-range test, per-element compaction
    const int kk = base + lane;
    float ux = 0.f, uy = 0.f, uz = 0.f, r = 1e30f;
    int e = 0;
    if (kk < cnt) {
      const int t = (int)(lrow[kk] & 0x3fffffffu);
      const float4 xj = xq[t];
      float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
      dx -= a.Lx * rintf(dx * a.ix);
      dy -= a.Ly * rintf(dy * a.iy);
      dz -= a.Lz * rintf(dz * a.iz);
      const float rr2 = sqrtf(dx * dx + dy * dy + dz * dz);
      if (rr2 < k.rd && rr2 > 0.f) {
        r = rr2;
        const float ir = __builtin_amdgcn_rcpf(rr2);
        ux = dx * ir;
        uy = dy * ir;
        uz = dz * ir;
      }
      e = (int)xj.w;
    }
    const bool keep = r < k.rd;
    const unsigned long long b0 = __ballot(keep && e == 0), b1 = __ballot(keep && e != 0);
    const unsigned long long below = (1ull << lane) - 1ull;
    const int n0 = __popcll(b0), n1 = __popcll(b1);
    if (keep) {
      const int slot = e == 0 ? __popcll(b0 & below) : 63 - __popcll(b1 & below);
      float fc, dfc;
      cutoff_fc(k, r, &fc, &dfc);
      rr[slot] = r;
      rfc[slot] = fc;
#pragma unroll
      for (int ch = 0; ch < ML_CH; ++ch) {
        const float dn = r - k.ang_nu[ch];
        ga[slot][ch] = __expf(-k.ang_eta * dn * dn) * fc;
      }
      const float xx = ux * ux, yy = uy * uy, zz = uz * uz;
      float* mm = mo[slot];
      mm[0] = 1.f;
      mm[1] = ux;
      mm[2] = uy;
      mm[3] = uz;
      mm[4] = xx;
      mm[5] = ux * uy;
      mm[6] = ux * uz;
      mm[7] = yy;
      mm[8] = uy * uz;
      mm[9] = zz;
      mm[10] = xx * ux;
      mm[11] = xx * uy;
      mm[12] = xx * uz;
      mm[13] = ux * yy;
      mm[14] = ux * uy * uz;
      mm[15] = ux * zz;
      mm[16] = yy * uy;
      mm[17] = yy * uz;
      mm[18] = uy * zz;
      mm[19] = zz * uz;
    }
    __syncthreads();
    // lane = feature: accumulate over each element's rows of the chunk, in list order
#pragma unroll
    for (int el = 0; el < 2; ++el) {
      const int n = el == 0 ? n0 : n1;
      float acc = 0.f;
      for (int jj = half; jj < n; jj += 2) {  // radial: two rows per step
        const int q = el == 0 ? jj : 63 - jj;
        const float d = rr[q] - mu;
        acc += __expf(-k.rad_eta * d * d) * rfc[q];
      }
      rad[el] += acc;
      float a0 = 0.f, a1 = 0.f;
#pragma unroll 4
      for (int jj = 0; jj < n; ++jj) {  // moments: one row per step
        const int q = el == 0 ? jj : 63 - jj;
        a0 += ga[q][ch0] * mo[q][cp0];
        a1 += ga[q][ch1] * mo[q][cp1];
      }
      m0[el] += a0;
      m1[el] += a1;
    }
    __syncthreads();
  }
#pragma unroll
  for (int el = 0; el < 2; ++el) {
    sm[el * 80 + lane] = m0[el];
    T[(long long)row * ML_T + el * 80 + lane] = m0[el];
    if (has1) {
      sm[el * 80 + c1] = m1[el];
      T[(long long)row * ML_T + el * 80 + c1] = m1[el];
    }
  }
  __syncthreads();
  // features: radial f = e*48 + k (lane < 64), angular f = e*48 + 32 + ch*4 + l (lanes 0..31)
  {
    const float g0 = rad[0] + __shfl_xor(rad[0], 32, 64), g1 = rad[1] + __shfl_xor(rad[1], 32, 64);
    const float g = half == 0 ? g0 : g1;
    const int f = half * ML_PER + (lane & 31);
    G[(long long)row * ML_F + f] = g;
    X[(long long)row * ML_F + f] = f2bf((g - mean[f]) * rstd[f]);
  }
  if (lane < 32) {
    const int e = lane >> 4, ch = (lane >> 2) & 3, l = lane & 3;
    const float* t = sm + e * 80 + ch * 20;
    float p = 0.f;
    for (int c = 0; c < ML_MOM; ++c)
      if (mom_l(c) == l) p += mom_weight(c) * t[c] * t[c];
    const int f = e * ML_PER + ML_RAD + ch * ML_L + l;
    G[(long long)row * ML_F + f] = p;
    X[(long long)row * ML_F + f] = f2bf((p - mean[f]) * rstd[f]);
  }
}

// dG_i/dr_ij . w for all 96 features of atom i is assembled by the JVP kernel; the force kernel
// needs the transposed product gG . dG_i/dr_ij for one pair, computed here (gG = dE_i/dG_i,
// tt = the atom's moment tensors, both in LDS).
// Radial terms: only the 2 ML_RWIN + 1 Gaussians centred nearest r. The centres are dmu apart and
// eta dmu^2 = 1/2, so a Gaussian k centres away weighs exp(-k^2/2): the terms left out
// (k >= ML_RWIN + 1 = 9, |r - mu| >= 8.5 dmu) weigh below exp(-36) = 2e-16 of the nearest one,
// far under float32 rounding.
#define ML_RWIN 8
__device__ __forceinline__ void pair_force(const MlConst& k, const float* gG, const float* tt, int e, float ux,
                                           float uy, float uz, float r, float* vx, float* vy, float* vz) {
  float fc, dfc;
  cutoff_fc(k, r, &fc, &dfc);
  const float* gr = gG + e * ML_PER;
  float srad = 0.f;
  const int qc = (int)rintf((r - k.rad_mu0) / k.rad_dmu);
  const int q0 = min(max(qc - ML_RWIN, 0), ML_RAD - (2 * ML_RWIN + 1));
#pragma unroll
  for (int t = 0; t < 2 * ML_RWIN + 1; ++t) {
    const int q = q0 + t;
    const float dr = r - (k.rad_mu0 + q * k.rad_dmu);
    const float ex = __expf(-k.rad_eta * dr * dr);
    srad += gr[q] * ex * (dfc - 2.f * k.rad_eta * dr * fc);
  }
  float ax = srad * ux, ay = srad * uy, az = srad * uz;
  const float ir = 1.f / r;
  // d/dr_ij [g(r) s_l(u)] = g' s_l u + (g/r) (v_l - (v_l . u) u), and v_l . u = l s_l (s_l is a
  // homogeneous polynomial of degree l in u, v_l its gradient): per channel the scalars
  // sum_l cf_l s_l (g' - l g/r) along u and the vector (g/r) sum_l cf_l v_l.
  for (int ch = 0; ch < ML_CH; ++ch) {
    const float dn = r - k.ang_nu[ch];
    const float ex = __expf(-k.ang_eta * dn * dn);
    const float g = ex * fc, dg = ex * (dfc - 2.f * k.ang_eta * dn * fc);
    const float gi = g * ir;
    float s[4], v[4][3];
    contract(tt + e * 80 + ch * 20, ux, uy, uz, s, v);
    float su = 0.f, wx = 0.f, wy = 0.f, wz = 0.f;
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const float cf = 2.f * gr[ML_RAD + ch * ML_L + l];
      su += cf * s[l] * (dg - (float)l * gi);
      wx += cf * v[l][0];
      wy += cf * v[l][1];
      wz += cf * v[l][2];
    }
    ax += su * ux + gi * wx;
    ay += su * uy + gi * wy;
    az += su * uz + gi * wz;
  }
  *vx = ax;
  *vy = ay;
  *vz = az;
}

// ---------------------------------------------------------------- forces
// One wave per atom i. The in-range neighbours (r < r_d) of the list are first compacted into LDS
// in list order, then evaluated one lane per neighbour: v = dE_i/dr_ij, F_j -= v (atomic),
// F_i += v. Compaction keeps every lane busy (about a third of the list is inside r_d).
#define MLF_MAX 256
__global__ __launch_bounds__(64) void k_ml_force(MlAtoms a, MlConst k, const float4* __restrict__ xq,
                                                const int* __restrict__ sorted_orig, const int* __restrict__ nbr_count,
                                                const unsigned* __restrict__ nbr_list, const int* __restrict__ row_of,
                                                const float* __restrict__ GX, const float* __restrict__ rstd,
                                                const float* __restrict__ T, long long* __restrict__ ffix) {
  __shared__ float gG[ML_F];
  __shared__ float tt[ML_T];
  __shared__ float4 nb[MLF_MAX];  // unit vector, r
  __shared__ int nt[MLF_MAX];     // neighbour slot | element << 30
  const int s = (int)(((long long)blockIdx.x * a.stride) % a.n);
  const int lane = threadIdx.x;
  const float4 xi = xq[s];
  const int o = sorted_orig[s];
  const int row = row_of[o];
  for (int f = lane; f < ML_F; f += 64) gG[f] = GX[(long long)row * ML_F + f] * rstd[f];
  for (int c = lane; c < ML_T; c += 64) tt[c] = T[(long long)row * ML_T + c];
  const int cnt = nbr_count[s];
  const unsigned* lrow = nbr_list + (long long)s * a.cap;
  float fx = 0.f, fy = 0.f, fz = 0.f;
  int done = 0;  // list entries consumed
  while (done < cnt) {
    // compact up to MLF_MAX in-range neighbours, in list order
    int m = 0;
    while (done < cnt && m + 64 <= MLF_MAX) {
      const int kk = done + lane;
      bool keep = false;
      float4 u = make_float4(0.f, 0.f, 0.f, 0.f);
      int te = 0;
      if (kk < cnt) {
        const int t = (int)(lrow[kk] & 0x3fffffffu);
        const float4 xj = xq[t];
        float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
        dx -= a.Lx * rintf(dx * a.ix);
        dy -= a.Ly * rintf(dy * a.iy);
        dz -= a.Lz * rintf(dz * a.iz);
        const float r = sqrtf(dx * dx + dy * dy + dz * dz);
        keep = r < k.rd && r > 0.f;
        const float ir = __builtin_amdgcn_rcpf(r);
        u = make_float4(dx * ir, dy * ir, dz * ir, r);
        te = t | ((int)xj.w << 30);
      }
      const unsigned long long ball = __ballot(keep);
      if (keep) {
        const int slot = m + __popcll(ball & ((1ull << lane) - 1ull));
        nb[slot] = u;
        nt[slot] = te;
      }
      m += __popcll(ball);
      done += 64;
    }
    __syncthreads();
    for (int j = lane; j < m; j += 64) {
      const float4 u = nb[j];
      const int te = nt[j];
      const int t = te & 0x3fffffff, e = (int)((unsigned)te >> 30);
      float vx, vy, vz;
      pair_force(k, gG, tt, e, u.x, u.y, u.z, u.w, &vx, &vy, &vz);
      const int oj = sorted_orig[t];
      atomicAdd((unsigned long long*)&ffix[oj], (unsigned long long)to_fix_f(-vx));
      atomicAdd((unsigned long long*)&ffix[a.n + oj], (unsigned long long)to_fix_f(-vy));
      atomicAdd((unsigned long long*)&ffix[2LL * a.n + oj], (unsigned long long)to_fix_f(-vz));
      fx += vx;
      fy += vy;
      fz += vz;
    }
    __syncthreads();
  }
  for (int off = 32; off > 0; off >>= 1) {
    fx += __shfl_xor(fx, off, 64);
    fy += __shfl_xor(fy, off, 64);
    fz += __shfl_xor(fz, off, 64);
  }
  if (lane == 0) {
    atomicAdd((unsigned long long*)&ffix[o], (unsigned long long)to_fix_f(fx));
    atomicAdd((unsigned long long*)&ffix[a.n + o], (unsigned long long)to_fix_f(fy));
    atomicAdd((unsigned long long*)&ffix[2LL * a.n + o], (unsigned long long)to_fix_f(fz));
  }
}

// ---------------------------------------------------------------- descriptor JVP (training)
// U[row][f] = rstd_f * sum_j (w_j - w_i) . dG_if/dr_ij, one wave per atom, lane per feature
// (f = lane and lane + 64 < 96), neighbours staged 64 at a time.
__global__ __launch_bounds__(64) void k_ml_jvp(MlAtoms a, MlConst k, const float4* __restrict__ xq,
                                              const int* __restrict__ sorted_orig, const int* __restrict__ nbr_count,
                                              const unsigned* __restrict__ nbr_list, const int* __restrict__ row_of,
                                              const float* __restrict__ T, const double* __restrict__ w,
                                              const float* __restrict__ rstd, bf16_t* __restrict__ U) {
  __shared__ NbStage S;
  __shared__ float tt[ML_T];
  __shared__ float dw[64][3];
  const int s = blockIdx.x;
  const int lane = threadIdx.x;
  const float4 xi = xq[s];
  const int o = sorted_orig[s];
  const int row = row_of[o];
  for (int c = lane; c < ML_T; c += 64) tt[c] = T[(long long)row * ML_T + c];
  const float wi[3] = {(float)w[3 * o], (float)w[3 * o + 1], (float)w[3 * o + 2]};
  const int cnt = nbr_count[s];
  const unsigned* lrow = nbr_list + (long long)s * a.cap;
  float acc[2] = {0.f, 0.f};
  for (int base = 0; base < cnt; base += 64) {
    stage_neighbours(a, k, xi, xq, lrow, cnt, base, S);
    if (base + lane < cnt) {
      const int oj = sorted_orig[S.t[lane]];
      for (int d = 0; d < 3; ++d) dw[lane][d] = (float)w[3 * oj + d] - wi[d];
    }
    __syncthreads();
    const int m = min(64, cnt - base);
    for (int jj = 0; jj < m; ++jj) {
      const float r = S.r[jj];
      if (r >= k.rd) continue;
      const int e = S.e[jj];
      const float ux = S.ux[jj], uy = S.uy[jj], uz = S.uz[jj];
      const float wx = dw[jj][0], wy = dw[jj][1], wz = dw[jj][2];
      const float uw = ux * wx + uy * wy + uz * wz;
      float fc, dfc;
      cutoff_fc(k, r, &fc, &dfc);
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int f = lane + 64 * q;
        if (f >= ML_F || f / ML_PER != e) continue;
        const int loc = f % ML_PER;
        if (loc < ML_RAD) {
          const float dr = r - (k.rad_mu0 + loc * k.rad_dmu);
          const float ex = __expf(-k.rad_eta * dr * dr);
          acc[q] += ex * (dfc - 2.f * k.rad_eta * dr * fc) * uw;
        } else {
          const int ch = (loc - ML_RAD) >> 2, l = (loc - ML_RAD) & 3;
          const float dn = r - k.ang_nu[ch];
          const float ex = __expf(-k.ang_eta * dn * dn);
          const float g = ex * fc, dg = ex * (dfc - 2.f * k.ang_eta * dn * fc);
          float sl[4], vl[4][3];
          contract(tt + e * 80 + ch * 20, ux, uy, uz, sl, vl);
          const float vw = vl[l][0] * wx + vl[l][1] * wy + vl[l][2] * wz;
          const float vu = vl[l][0] * ux + vl[l][1] * uy + vl[l][2] * uz;
          acc[q] += 2.f * (dg * sl[l] * uw + g * (vw - vu * uw) / r);
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int f = lane + 64 * q;
    if (f < ML_F) U[(long long)row * ML_F + f] = f2bf(acc[q] * rstd[f]);
  }
}

// ---------------------------------------------------------------- BF16 MFMA GEMM
// C[m][n] = sum_k A[m][k] B[k][n] over k in this block's K range, A and B read through
// arbitrary (row, col) strides (transposes are strides). 64x64 tile per 256-lane workgroup,
// each wave a 32x32 sub-tile on gfx950's v_mfma_f32_32x32x16_bf16 (lane l: A row l%32, k
// 8(l/32)..+7; B col l%32, same k; D row 8(i/4) + 4(l/32) + i%4, col l%32 for accumulator i).
// Rows >= row_split use the second operand B1 / bias1 (the second element's weights).
enum { EPI_F32 = 0, EPI_F32_ACC = 1, EPI_BIAS_SILU = 2, EPI_MUL_S = 3 };
struct GemmArgs {
  int M, N, K;
  const bf16_t* A;
  long long a_m, a_k;
  const bf16_t* B0;
  const bf16_t* B1;
  long long b_k, b_n;
  int row_split;
  int kchunk;                 // K per blockIdx.z (split-K)
  float* C;                   // EPI_F32*: partials [z][M][ldc]; EPI_BIAS_SILU: Z (+bias); EPI_MUL_S: raw product
  long long ldc;
  const float* bias0;
  const float* bias1;
  bf16_t* H;                  // bf16 output (SiLU(z) or product * S)
  float* S;                   // EPI_BIAS_SILU: writes SiLU'(z); EPI_MUL_S: reads the multiplier
  float* DD;                  // EPI_BIAS_SILU: writes SiLU''(z) when not null
};
#define GK 32
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// 8 consecutive bf16 of a tile row from global memory (16 B when all 8 are in range and the
// address is 16-byte aligned, element by element with zero fill otherwise)
__device__ __forceinline__ void load8(const bf16_t* __restrict__ base, long long off, long long step, int nvalid,
                                      bf16_t* out) {
  const bf16_t* p = base + off;
  if (nvalid >= 8 && step == 1 && (((unsigned long long)p) & 15ull) == 0) {
    const u32x4 v = *(const u32x4*)p;
    const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      out[2 * i] = (bf16_t)(w[i] & 0xffffu);
      out[2 * i + 1] = (bf16_t)(w[i] >> 16);
    }
  } else {
#pragma unroll
    for (int i = 0; i < 8; ++i) out[i] = i < nvalid ? p[(long long)i * step] : (bf16_t)0;
  }
}

// A_KC: A's k index is contiguous in memory (a_k == 1), else its row index is (a_m == 1);
// B_KC: B's k index is contiguous (b_k == 1), else its column index is (b_n == 1). Each k-step
// every lane moves one 8-element vector of A and one of B into LDS, where both tiles are
// stored k-contiguous per row / column so each MFMA operand is one 8-byte LDS read.
template <bool A_KC, bool B_KC>
__global__ __launch_bounds__(256) void k_gemm_bf16(GemmArgs g, int epi) {
  __shared__ __attribute__((aligned(16))) bf16_t As[ML_TILE][GK + 8];
  __shared__ __attribute__((aligned(16))) bf16_t Bs[ML_TILE][GK + 8];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int row0 = blockIdx.y * ML_TILE, col0 = blockIdx.x * ML_TILE;
  const int kb = blockIdx.z * g.kchunk, ke = min(g.K, kb + g.kchunk);
  const bool second = row0 >= g.row_split;
  const bf16_t* B = second ? g.B1 : g.B0;
  f32x16 acc;
  for (int i = 0; i < 16; ++i) acc[i] = 0.f;
  // this lane's tile vectors: A (row ra / k offset ka, or k ka / rows ra..+7), B likewise
  const int ra = A_KC ? (tid >> 2) : (tid & 7) * 8, ka = A_KC ? (tid & 3) * 8 : (tid >> 3);
  const int rb = B_KC ? (tid >> 2) : (tid & 7) * 8, kbv = B_KC ? (tid & 3) * 8 : (tid >> 3);
  auto fetch = [&](int k0, bf16_t* va, bf16_t* vb) {
    if (A_KC) {
      const int m = row0 + ra, kx = k0 + ka;
      load8(g.A, (long long)m * g.a_m + kx, 1, m < g.M ? max(0, min(8, ke - kx)) : 0, va);
    } else {
      const int m = row0 + ra, kx = k0 + ka;
      load8(g.A, (long long)kx * g.a_k + m, 1, kx < ke ? max(0, min(8, g.M - m)) : 0, va);
    }
    if (B_KC) {
      const int n = col0 + rb, kx = k0 + kbv;
      load8(B, (long long)n * g.b_n + kx, 1, n < g.N ? max(0, min(8, ke - kx)) : 0, vb);
    } else {
      const int n = col0 + rb, kx = k0 + kbv;
      load8(B, (long long)kx * g.b_k + n, 1, kx < ke ? max(0, min(8, g.N - n)) : 0, vb);
    }
  };
  bf16_t va[8], vb[8];
  if (kb < ke) fetch(kb, va, vb);
  for (int k0 = kb; k0 < ke; k0 += GK) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      if (A_KC) As[ra][ka + i] = va[i];
      else As[ra + i][ka] = va[i];
      if (B_KC) Bs[rb][kbv + i] = vb[i];
      else Bs[rb + i][kbv] = vb[i];
    }
    __syncthreads();
    if (k0 + GK < ke) fetch(k0 + GK, va, vb);  // next step's loads in flight during the MFMAs
#pragma unroll
    for (int ks = 0; ks < GK; ks += 16) {
      // v_mfma_f32_32x32x16_bf16: lane l supplies row (or column) l%32 and k 8(l/32)..+7, one
      // 16-byte LDS read per operand (rows are 80 B, k offsets multiples of 8: 16-byte aligned)
      const int ar = wm * 32 + (lane & 31), kq = ks + (lane >> 5) * 8;
      const int bc = wn * 32 + (lane & 31);
      const bf16x8 av = *(const bf16x8*)&As[ar][kq];
      const bf16x8 bv = *(const bf16x8*)&Bs[bc][kq];
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int r = wm * 32 + (i >> 2) * 8 + (lane >> 5) * 4 + (i & 3);
    const int c = wn * 32 + (lane & 31);
    const int m = row0 + r, n = col0 + c;
    if (m >= g.M || n >= g.N) continue;
    const float v = acc[i];
    if (epi == EPI_F32) {
      g.C[((long long)blockIdx.z * g.M + m) * g.ldc + n] = v;
    } else if (epi == EPI_F32_ACC) {
      g.C[((long long)blockIdx.z * g.M + m) * g.ldc + n] += v;
    } else if (epi == EPI_BIAS_SILU) {
      const float z = v + (second ? g.bias1 : g.bias0)[n];
      const float sg = 1.f / (1.f + __expf(-z));
      const long long idx = (long long)m * g.ldc + n;
      if (g.C) g.C[idx] = z;
      g.H[idx] = f2bf(z * sg);
      g.S[idx] = sg * (1.f + z * (1.f - sg));
      if (g.DD) g.DD[idx] = sg * (1.f - sg) * (2.f + z * (1.f - 2.f * sg));
    } else {  // EPI_MUL_S
      const long long idx = (long long)m * g.ldc + n;
      if (g.C) g.C[idx] = v;
      g.H[idx] = f2bf(v * g.S[idx]);
    }
  }
}

// ---------------------------------------------------------------- elementwise / reductions
static_assert(ML_H == 256 && P_W4 % 4 == 0, "head kernel layout: 64 lanes x 4 columns");
// per row: E_i = H3 . w4 + b4, G3 = w4 * S3 (bf16). One wave per row at a time, each lane four
// consecutive columns (8-byte bf16 and 16-byte float accesses); each workgroup covers HEAD_ROWS
// rows and adds its energy with one fixed-point atomic (a per-row atomic on one address
// serialises).
#define HEAD_ROWS 32
__global__ __launch_bounds__(256) void k_ml_head(int rows, int row_split, const bf16_t* __restrict__ H3,
                                                const float* __restrict__ S3, const float* __restrict__ p0,
                                                const float* __restrict__ p1, const int* __restrict__ atom_of_row,
                                                bf16_t* __restrict__ G3, double* __restrict__ Eatom,
                                                long long* __restrict__ eslot) {
  __shared__ double wsum[4];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int c = lane * 4;  // ML_H == 256 == 64 lanes x 4
  double ew = 0.0;
  const int r0 = blockIdx.x * HEAD_ROWS, r1 = min(rows, r0 + HEAD_ROWS);
  for (int row = r0 + w; row < r1; row += 4) {
    const float* p = row >= row_split ? p1 : p0;
    const long long idx = (long long)row * ML_H + c;
    const bf16x4 h = *(const bf16x4*)(H3 + idx);
    const float4 sv = *(const float4*)(S3 + idx);
    const float4 wv = *(const float4*)(p + P_W4 + c);
    float e = bf2f((bf16_t)h[0]) * wv.x + bf2f((bf16_t)h[1]) * wv.y + bf2f((bf16_t)h[2]) * wv.z +
              bf2f((bf16_t)h[3]) * wv.w;
    bf16x4 g;
    g[0] = (short)f2bf(wv.x * sv.x);
    g[1] = (short)f2bf(wv.y * sv.y);
    g[2] = (short)f2bf(wv.z * sv.z);
    g[3] = (short)f2bf(wv.w * sv.w);
    *(bf16x4*)(G3 + idx) = g;
    for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off, 64);
    if (atom_of_row[row] >= 0) {
      const double E = (double)e + (double)p[P_B4];
      if (lane == 0 && Eatom) Eatom[row] = E;
      ew += E;
    }
  }
  if (lane == 0) wsum[w] = ew;
  __syncthreads();
  if (threadIdx.x == 0 && eslot)
    atomicAdd((unsigned long long*)eslot, (unsigned long long)to_fix(wsum[0] + wsum[1] + wsum[2] + wsum[3]));
}

// tangent step of one layer: b = a * S (bf16), c = a * hb * DD (fp32)
__global__ void k_ml_tangent(long long count, const float* __restrict__ a, const float* __restrict__ S,
                             const float* __restrict__ hb, const float* __restrict__ DD, bf16_t* __restrict__ b,
                             float* __restrict__ c) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  b[i] = f2bf(a[i] * S[i]);
  c[i] = a[i] * hb[i] * DD[i];
}
// dz = c + dh * S (bf16; also fp32 for the bias sums)
__global__ void k_ml_dz(long long count, const float* __restrict__ c, const float* __restrict__ dh,
                        const float* __restrict__ S, bf16_t* __restrict__ dz16, float* __restrict__ dz32) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float v = c[i] + (dh ? dh[i] * S[i] : 0.f);
  dz16[i] = f2bf(v);
  dz32[i] = v;
}
// last layer: dw4 terms a3 * S3 (fp32) and dz3 = a3 * w4 * DD3
__global__ void k_ml_top(int rows, int row_split, const float* __restrict__ a3, const float* __restrict__ S3,
                         const float* __restrict__ DD3, const float* __restrict__ p0, const float* __restrict__ p1,
                         float* __restrict__ t4, bf16_t* __restrict__ dz16, float* __restrict__ dz32) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)rows * ML_H) return;
  const int row = (int)(i / ML_H), c = (int)(i % ML_H);
  const float* p = row >= row_split ? p1 : p0;
  t4[i] = a3[i] * S3[i];
  const float v = a3[i] * p[P_W4 + c] * DD3[i];
  dz16[i] = f2bf(v);
  dz32[i] = v;
}
// column sums over rows [r0, r1) of in[row][c] in a fixed order: block b sums its chunk of rows
// (one lane per column, coalesced), then one lane per column adds the chunk sums in order
#define COLSUM_CHUNKS 128
__global__ void k_ml_colsum_part(const float* __restrict__ in, int r0, int r1, int ncol, float* __restrict__ part) {
  const int c = threadIdx.x;
  const int len = (r1 - r0 + COLSUM_CHUNKS - 1) / COLSUM_CHUNKS;
  const int a = r0 + blockIdx.x * len, b = min(r1, a + len);
  if (c >= ncol) return;
  float s = 0.f;
  for (int r = a; r < b; ++r) s += in[(long long)r * ncol + c];
  part[blockIdx.x * ncol + c] = s;
}
__global__ void k_ml_colsum_fin(const float* __restrict__ part, int ncol, float* __restrict__ out) {
  const int c = threadIdx.x;
  if (c >= ncol) return;
  float s = 0.f;
  for (int k = 0; k < COLSUM_CHUNKS; ++k) s += part[k * ncol + c];
  out[c] = s;
}
// sum of split-K partials [z][rows][cols] into out (fixed order)
__global__ void k_ml_reduce_splits(const float* __restrict__ part, int nz, long long count, float* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float s = 0.f;
  for (int z = 0; z < nz; ++z) s += part[(long long)z * count + i];
  out[i] = s;
}
// force error and its loss weight: err = F_hat - F_ref, w = -(2/3N) err, loss += sum err^2/(3N)
__global__ void k_ml_loss(int n, const long long* __restrict__ ffix, const double* __restrict__ fref,
                          double* __restrict__ w, long long* __restrict__ lslot) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double l = 0.0;
  if (i < n) {
    for (int d = 0; d < 3; ++d) {
      const double err = (double)ffix[(long long)d * n + i] * kFixInv - fref[3 * i + d];
      w[3 * i + d] = -2.0 / (3.0 * n) * err;
      l += err * err / (3.0 * n);
    }
  }
  for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off, 64);
  __shared__ double sm[4];
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) s += sm[q];
    atomicAdd((unsigned long long*)lslot, (unsigned long long)to_fix(s));
  }
}
// Adam on the fp32 master parameters, then the bf16 weight copies
__global__ void k_ml_adam(int count, float* __restrict__ p, const float* __restrict__ gr, float* __restrict__ m,
                          float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float gi = gr[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}
__global__ void k_ml_tobf16(int count, const float* __restrict__ src, bf16_t* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) dst[i] = f2bf(src[i]);
}
__global__ void k_ml_fill_f32(long long count, float* __restrict__ p, float v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = v;
}
__global__ void k_ml_zero_i64(long long count, long long* __restrict__ p) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0;
}
// G (rows, element order) -> out (original atom order)
__global__ void k_ml_unpermute(int n, int width, const int* __restrict__ row_of, const float* __restrict__ in,
                               float* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)n * width) return;
  const int o = (int)(i / width), c = (int)(i % width);
  out[i] = in[(long long)row_of[o] * width + c];
}
// ffix (int64 SoA) -> forces (N x 3 f64)
__global__ void k_ml_forces_out(int n, const long long* __restrict__ ffix, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int d = 0; d < 3; ++d) out[3 * i + d] = (double)ffix[(long long)d * n + i] * kFixInv;
}

// ================================================================ host
struct MlNet {
  DevBuf<float> p, grad, m, v;  // P_N floats each
  DevBuf<bf16_t> w1, w2, w3;    // bf16 copies
};

struct mdx_mlp {
  mdx_ctx* host = nullptr;
  mdx_ctx* nl = nullptr;        // internal context: the r_d neighbour list on the host's positions
  int n = 0, nO = 0, nH = 0, rowsO = 0, rows = 0;
  std::vector<int> h_row_of;
  DevBuf<int> row_of, atom_of_row;
  DevBuf<float> mean, rstd;
  MlNet net[2];
  MlConst k;
  // activations (rows x width)
  DevBuf<float> G, T, GX, S1, S2, S3, D1, D2, D3, HB1, HB2, A1, A2, A3, C1, C2, T4, DZ1, DZ2, DZ3, DH1, DH2;
  DevBuf<bf16_t> X, H1, H2, H3, G1, G2, G3, U, B1, B2, DZ16_1, DZ16_2, DZ16_3;
  DevBuf<float> part;           // split-K partials
  DevBuf<double> w;             // loss weights (3N)
  DevBuf<long long> scal;       // [0] energy, [1] loss (fixed point)
  int adam_t = 0;
  float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  bool trained_buffers = false;
};

static MlAtoms ml_atoms(const mdx_mlp* m) {
  MlAtoms a;
  a.n = m->n;
  // a prime stride coprime to n, so the force pass visits every atom once while concurrent
  // workgroups work on distant atoms and update different partners (104 vs 110 us at 23,328 atoms)
  a.stride = 1;
  for (int p : {7919, 7907, 7901, 7883, 7879})
    if (m->n % p != 0) {
      a.stride = p;
      break;
    }
  a.cap = m->nl->nbr_cap;
  a.Lx = (float)m->nl->box[0];
  a.Ly = (float)m->nl->box[1];
  a.Lz = (float)m->nl->box[2];
  a.ix = 1.f / a.Lx;
  a.iy = 1.f / a.Ly;
  a.iz = 1.f / a.Lz;
  return a;
}

static int ml_alloc(mdx_mlp* m, bool training) {
  const size_t R = m->rows;
  int rc;
  if (!m->G.p) {
    if ((rc = m->G.ensure(R * ML_F)) || (rc = m->T.ensure(R * ML_T)) || (rc = m->GX.ensure(R * ML_F)) ||
        (rc = m->X.ensure(R * ML_F)) || (rc = m->S1.ensure(R * ML_H)) || (rc = m->S2.ensure(R * ML_H)) ||
        (rc = m->S3.ensure(R * ML_H)) || (rc = m->H1.ensure(R * ML_H)) || (rc = m->H2.ensure(R * ML_H)) ||
        (rc = m->H3.ensure(R * ML_H)) || (rc = m->G1.ensure(R * ML_H)) || (rc = m->G2.ensure(R * ML_H)) ||
        (rc = m->G3.ensure(R * ML_H)) || (rc = m->scal.ensure(2)))
      return rc;
    // padding rows never receive descriptor writes: keep their inputs zero
    MDX_HIP(hipMemset(m->X.p, 0, R * ML_F * sizeof(bf16_t)));
    MDX_HIP(hipMemset(m->G.p, 0, R * ML_F * sizeof(float)));
    MDX_HIP(hipMemset(m->T.p, 0, R * ML_T * sizeof(float)));
  }
  if (training && !m->trained_buffers) {
    if ((rc = m->D1.ensure(R * ML_H)) || (rc = m->D2.ensure(R * ML_H)) || (rc = m->D3.ensure(R * ML_H)) ||
        (rc = m->HB1.ensure(R * ML_H)) || (rc = m->HB2.ensure(R * ML_H)) || (rc = m->A1.ensure(R * ML_H)) ||
        (rc = m->A2.ensure(R * ML_H)) || (rc = m->A3.ensure(R * ML_H)) || (rc = m->C1.ensure(R * ML_H)) ||
        (rc = m->C2.ensure(R * ML_H)) || (rc = m->T4.ensure(R * ML_H)) || (rc = m->DZ1.ensure(R * ML_H)) ||
        (rc = m->DZ2.ensure(R * ML_H)) || (rc = m->DZ3.ensure(R * ML_H)) || (rc = m->DH1.ensure(R * ML_H)) ||
        (rc = m->DH2.ensure(R * ML_H)) || (rc = m->U.ensure(R * ML_F)) || (rc = m->B1.ensure(R * ML_H)) ||
        (rc = m->B2.ensure(R * ML_H)) || (rc = m->DZ16_1.ensure(R * ML_H)) || (rc = m->DZ16_2.ensure(R * ML_H)) ||
        (rc = m->DZ16_3.ensure(R * ML_H)) || (rc = m->w.ensure(3 * (size_t)m->n)))
      return rc;
    MDX_HIP(hipMemset(m->U.p, 0, R * ML_F * sizeof(bf16_t)));
    const int maxsplit = std::max(1, m->rows / 512 + 1);
    if ((rc = m->part.ensure((size_t)maxsplit * ML_H * ML_H))) return rc;
    for (int e = 0; e < 2; ++e)
      if ((rc = m->net[e].grad.ensure(P_N)) || (rc = m->net[e].m.ensure(P_N)) || (rc = m->net[e].v.ensure(P_N)))
        return rc;
    for (int e = 0; e < 2; ++e) {
      MDX_HIP(hipMemset(m->net[e].m.p, 0, P_N * sizeof(float)));
      MDX_HIP(hipMemset(m->net[e].v.p, 0, P_N * sizeof(float)));
    }
    m->trained_buffers = true;
  }
  return 0;
}

extern "C" int mdx_mlp_create(mdx_ctx* host, const int* h_elem, double rd, double skin, mdx_mlp** out) {
  MDX_CHECK(host && host->n > 0 && h_elem && out, MDX_ERR_INVALID, "mdx_mlp_create: bad arguments");
  MDX_CHECK(rd > 0 && rd <= 0.5 * std::min(host->box[0], std::min(host->box[1], host->box[2])), MDX_ERR_CELL_CUTOFF,
            "descriptor cutoff %.3f exceeds half the box", rd);
  MDX_HIP(hipSetDevice(host->device));
  mdx_mlp* m = new mdx_mlp();
  m->host = host;
  m->n = host->n;
  m->h_row_of.assign(m->n, 0);
  for (int i = 0; i < m->n; ++i) {
    MDX_CHECK(h_elem[i] == 0 || h_elem[i] == 1, MDX_ERR_INVALID, "element of atom %d must be 0 (O) or 1 (H)", i);
    (h_elem[i] == 0 ? m->nO : m->nH)++;
  }
  m->rowsO = div_up(m->nO, ML_TILE) * ML_TILE;
  m->rows = m->rowsO + div_up(m->nH, ML_TILE) * ML_TILE;
  std::vector<int> atom_of(m->rows, -1);
  int ro = 0, rh = m->rowsO;
  for (int i = 0; i < m->n; ++i) {
    const int r = h_elem[i] == 0 ? ro++ : rh++;
    m->h_row_of[i] = r;
    atom_of[r] = i;
  }
  int rc;
  if ((rc = m->row_of.ensure(m->n)) || (rc = m->atom_of_row.ensure(m->rows)) || (rc = m->mean.ensure(ML_F)) ||
      (rc = m->rstd.ensure(ML_F)))
    return rc;
  MDX_HIP(hipMemcpy(m->row_of.p, m->h_row_of.data(), m->n * sizeof(int), hipMemcpyHostToDevice));
  MDX_HIP(hipMemcpy(m->atom_of_row.p, atom_of.data(), m->rows * sizeof(int), hipMemcpyHostToDevice));
  std::vector<float> zero(ML_F, 0.f), one(ML_F, 1.f);
  MDX_HIP(hipMemcpy(m->mean.p, zero.data(), ML_F * sizeof(float), hipMemcpyHostToDevice));
  MDX_HIP(hipMemcpy(m->rstd.p, one.data(), ML_F * sizeof(float), hipMemcpyHostToDevice));
  for (int e = 0; e < 2; ++e) {
    MlNet& N = m->net[e];
    if ((rc = N.p.ensure(P_N)) || (rc = N.w1.ensure(ML_F * ML_H)) || (rc = N.w2.ensure(ML_H * ML_H)) ||
        (rc = N.w3.ensure(ML_H * ML_H)))
      return rc;
    MDX_HIP(hipMemset(N.p.p, 0, P_N * sizeof(float)));
    MDX_HIP(hipMemset(N.w1.p, 0, ML_F * ML_H * sizeof(bf16_t)));
    MDX_HIP(hipMemset(N.w2.p, 0, ML_H * ML_H * sizeof(bf16_t)));
    MDX_HIP(hipMemset(N.w3.p, 0, ML_H * ML_H * sizeof(bf16_t)));
  }
  // descriptor constants (oracle/mlp.py)
  m->k.rd = (float)rd;
  m->k.pi_rd = (float)(M_PI / rd);
  m->k.rad_mu0 = 0.8f;
  m->k.rad_dmu = (float)((rd - 0.8) / (ML_RAD - 1));
  m->k.rad_eta = (float)(1.0 / (2.0 * std::pow((rd - 0.8) / (ML_RAD - 1), 2)));
  const float nu[ML_CH] = {1.0f, 2.5f, 4.0f, 5.5f};
  for (int c = 0; c < ML_CH; ++c) m->k.ang_nu[c] = nu[c];
  m->k.ang_eta = 0.5f;
  // internal list context on the host's positions and stream; "charges" carry the element id
  mdx_ctx* nl = new mdx_ctx();
  nl->device = host->device;
  nl->stream = host->stream;
  nl->owns_stream = false;
  m->nl = nl;
  if ((rc = nl->dstate.ensure(1)) || (rc = nl->efix.ensure(MDX_NTERMS + 1)) || (rc = nl->ffix.ensure(3 * (size_t)m->n)))
    return rc;
  MDX_HIP(hipMemset(nl->dstate.p, 0, sizeof(DevState)));
  MDX_HIP(hipMemset(nl->efix.p, 0, sizeof(long long) * (MDX_NTERMS + 1)));
  int one_i = 1;
  MDX_HIP(hipMemcpy(&nl->dstate.p->rebuild, &one_i, sizeof(int), hipMemcpyHostToDevice));
  nl->n = m->n;
  nl->pos = host->pos;
  nl->vel = host->vel;
  for (int d = 0; d < 3; ++d) nl->box[d] = host->box[d];
  std::vector<float> ef(m->n);
  for (int i = 0; i < m->n; ++i) ef[i] = (float)h_elem[i];
  if ((rc = nl->charge_s.ensure(m->n))) return rc;
  MDX_HIP(hipMemcpy(nl->charge_s.p, ef.data(), m->n * sizeof(float), hipMemcpyHostToDevice));
  nl->has_charges = true;
  nl->rc = rd;
  nl->skin = skin;
  nl->pair_mode = MDX_PAIRS_EXACT;
  nl->cell_cutoff = rd;
  if ((rc = mdx_nlist_prepare(nl))) return rc;
  nl->nb_configured = true;
  if ((rc = ml_alloc(m, false))) return rc;
  *out = m;
  return 0;
}

extern "C" int mdx_mlp_destroy(mdx_mlp* m) {
  if (!m) return 0;
  hipSetDevice(m->host->device);
  hipStreamSynchronize(m->host->stream);
  if (m->host->mlp == m) m->host->mlp = nullptr;
  mdx_ctx_destroy(m->nl);
  for (DevBuf<float>* b : {&m->mean, &m->rstd, &m->G, &m->T, &m->GX, &m->S1, &m->S2, &m->S3,
                           &m->D1, &m->D2, &m->D3, &m->HB1, &m->HB2, &m->A1, &m->A2, &m->A3, &m->C1, &m->C2, &m->T4,
                           &m->DZ1, &m->DZ2, &m->DZ3, &m->DH1, &m->DH2, &m->part})
    b->release();
  for (DevBuf<bf16_t>* b : {&m->X, &m->H1, &m->H2, &m->H3, &m->G1, &m->G2, &m->G3, &m->U, &m->B1, &m->B2, &m->DZ16_1,
                            &m->DZ16_2, &m->DZ16_3})
    b->release();
  for (int e = 0; e < 2; ++e) {
    m->net[e].p.release();
    m->net[e].grad.release();
    m->net[e].m.release();
    m->net[e].v.release();
    m->net[e].w1.release();
    m->net[e].w2.release();
    m->net[e].w3.release();
  }
  m->row_of.release();
  m->atom_of_row.release();
  m->w.release();
  m->scal.release();
  delete m;
  return 0;
}

static void upload_bf16(const float* h, bf16_t* d, size_t count, hipStream_t st) {
  std::vector<bf16_t> tmp(count);
  for (size_t i = 0; i < count; ++i) {
    unsigned u;
    std::memcpy(&u, &h[i], 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    tmp[i] = (bf16_t)(u >> 16);
  }
  (void)hipMemcpy(d, tmp.data(), count * sizeof(bf16_t), hipMemcpyHostToDevice);
  (void)st;
}

extern "C" int mdx_mlp_set_params(mdx_mlp* m, int elem, const float* h_params) {
  MDX_CHECK(m && h_params && (elem == 0 || elem == 1), MDX_ERR_INVALID, "mdx_mlp_set_params: bad arguments");
  MDX_HIP(hipSetDevice(m->host->device));
  MDX_HIP(hipStreamSynchronize(m->host->stream));
  MlNet& N = m->net[elem];
  MDX_HIP(hipMemcpy(N.p.p, h_params, P_N * sizeof(float), hipMemcpyHostToDevice));
  upload_bf16(h_params + P_W1, N.w1.p, ML_F * ML_H, m->host->stream);
  upload_bf16(h_params + P_W2, N.w2.p, ML_H * ML_H, m->host->stream);
  upload_bf16(h_params + P_W3, N.w3.p, ML_H * ML_H, m->host->stream);
  if (N.m.p) {
    MDX_HIP(hipMemset(N.m.p, 0, P_N * sizeof(float)));
    MDX_HIP(hipMemset(N.v.p, 0, P_N * sizeof(float)));
  }
  m->adam_t = 0;
  return 0;
}

extern "C" int mdx_mlp_get_params(mdx_mlp* m, int elem, float* h_params) {
  MDX_CHECK(m && h_params && (elem == 0 || elem == 1), MDX_ERR_INVALID, "mdx_mlp_get_params: bad arguments");
  MDX_HIP(hipStreamSynchronize(m->host->stream));
  MDX_HIP(hipMemcpy(h_params, m->net[elem].p.p, P_N * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int mdx_mlp_set_norm(mdx_mlp* m, const float* h_mean, const float* h_std) {
  MDX_CHECK(m && h_mean && h_std, MDX_ERR_INVALID, "mdx_mlp_set_norm: bad arguments");
  std::vector<float> r(ML_F);
  for (int f = 0; f < ML_F; ++f) {
    MDX_CHECK(h_std[f] > 0, MDX_ERR_INVALID, "feature %d: std must be positive", f);
    r[f] = 1.f / h_std[f];
  }
  MDX_HIP(hipStreamSynchronize(m->host->stream));
  MDX_HIP(hipMemcpy(m->mean.p, h_mean, ML_F * sizeof(float), hipMemcpyHostToDevice));
  MDX_HIP(hipMemcpy(m->rstd.p, r.data(), ML_F * sizeof(float), hipMemcpyHostToDevice));
  return 0;
}

// ---- GEMM launcher
static int gemm(mdx_mlp* m, int epi, int M, int N, int K, const bf16_t* A, long long a_m, long long a_k,
                const bf16_t* B0, const bf16_t* B1, long long b_k, long long b_n, int row_split, float* C,
                long long ldc, int splits = 1, const float* bias0 = nullptr, const float* bias1 = nullptr,
                bf16_t* H = nullptr, float* S = nullptr, float* DD = nullptr) {
  GemmArgs g;
  g.M = M;
  g.N = N;
  g.K = K;
  g.A = A;
  g.a_m = a_m;
  g.a_k = a_k;
  g.B0 = B0;
  g.B1 = B1 ? B1 : B0;
  g.b_k = b_k;
  g.b_n = b_n;
  g.row_split = row_split;
  g.kchunk = div_up(div_up(K, splits), GK) * GK;
  g.C = C;
  g.ldc = ldc;
  g.bias0 = bias0;
  g.bias1 = bias1 ? bias1 : bias0;
  g.H = H;
  g.S = S;
  g.DD = DD;
  dim3 grid(div_up(N, ML_TILE), div_up(M, ML_TILE), splits);
  MDX_CHECK((a_k == 1 || a_m == 1) && (b_k == 1 || b_n == 1), MDX_ERR_INVALID,
            "gemm: one stride of each operand must be 1");
  hipStream_t st = m->host->stream;
  if (a_k == 1 && b_k == 1) k_gemm_bf16<true, true><<<grid, 256, 0, st>>>(g, epi);
  else if (a_k == 1) k_gemm_bf16<true, false><<<grid, 256, 0, st>>>(g, epi);
  else if (b_k == 1) k_gemm_bf16<false, true><<<grid, 256, 0, st>>>(g, epi);
  else k_gemm_bf16<false, false><<<grid, 256, 0, st>>>(g, epi);
  MDX_LAUNCH_CHECK();
  return 0;
}

// descriptors + forward + input gradient (GX); energy into eslot when given
static int ml_forward_backward(mdx_mlp* m, long long* eslot, bool keep_training_state) {
  hipStream_t st = m->host->stream;
  mdx_ctx* nl = m->nl;
  int rc;
  MlAtoms a = ml_atoms(m);
  k_ml_desc<<<m->n, 64, 0, st>>>(a, m->k, nl->xq.p, nl->sorted_orig.p, nl->nbr_count.p, nl->nbr_list.p,
                                m->row_of.p, m->mean.p, m->rstd.p, m->G.p, m->T.p, m->X.p);
  MDX_LAUNCH_CHECK();
  const int R = m->rows, rs = m->rowsO;
  MlNet &N0 = m->net[0], &N1 = m->net[1];
  float* DD1 = keep_training_state ? m->D1.p : nullptr;
  float* DD2 = keep_training_state ? m->D2.p : nullptr;
  float* DD3 = keep_training_state ? m->D3.p : nullptr;
  if ((rc = gemm(m, EPI_BIAS_SILU, R, ML_H, ML_F, m->X.p, ML_F, 1, N0.w1.p, N1.w1.p, ML_H, 1, rs, nullptr, ML_H, 1,
                 N0.p.p + P_B1, N1.p.p + P_B1, m->H1.p, m->S1.p, DD1)) ||
      (rc = gemm(m, EPI_BIAS_SILU, R, ML_H, ML_H, m->H1.p, ML_H, 1, N0.w2.p, N1.w2.p, ML_H, 1, rs, nullptr, ML_H, 1,
                 N0.p.p + P_B2, N1.p.p + P_B2, m->H2.p, m->S2.p, DD2)) ||
      (rc = gemm(m, EPI_BIAS_SILU, R, ML_H, ML_H, m->H2.p, ML_H, 1, N0.w3.p, N1.w3.p, ML_H, 1, rs, nullptr, ML_H, 1,
                 N0.p.p + P_B3, N1.p.p + P_B3, m->H3.p, m->S3.p, DD3)))
    return rc;
  k_ml_head<<<div_up(R, HEAD_ROWS), 256, 0, st>>>(R, rs, m->H3.p, m->S3.p, N0.p.p, N1.p.p, m->atom_of_row.p,
                                                   m->G3.p, nullptr, eslot);
  MDX_LAUNCH_CHECK();
  // input gradient: G2 = (G3 W3^T) * S2, G1 = (G2 W2^T) * S1, GX = G1 W1^T
  if ((rc = gemm(m, EPI_MUL_S, R, ML_H, ML_H, m->G3.p, ML_H, 1, N0.w3.p, N1.w3.p, 1, ML_H, rs,
                 keep_training_state ? m->HB2.p : nullptr, ML_H, 1, nullptr, nullptr, m->G2.p, m->S2.p)) ||
      (rc = gemm(m, EPI_MUL_S, R, ML_H, ML_H, m->G2.p, ML_H, 1, N0.w2.p, N1.w2.p, 1, ML_H, rs,
                 keep_training_state ? m->HB1.p : nullptr, ML_H, 1, nullptr, nullptr, m->G1.p, m->S1.p)) ||
      (rc = gemm(m, EPI_F32, R, ML_F, ML_H, m->G1.p, ML_H, 1, N0.w1.p, N1.w1.p, 1, ML_H, rs, m->GX.p, ML_F)))
    return rc;
  return 0;
}

static int ml_force_kernel(mdx_mlp* m, long long* ffix) {
  mdx_ctx* nl = m->nl;
  MlAtoms a = ml_atoms(m);
  k_ml_force<<<m->n, 64, 0, m->host->stream>>>(a, m->k, nl->xq.p, nl->sorted_orig.p, nl->nbr_count.p,
                                              nl->nbr_list.p, m->row_of.p, m->GX.p, m->rstd.p, m->T.p, ffix);
  MDX_LAUNCH_CHECK();
  return 0;
}

// enqueue the surrogate's forces into the host accumulator (and energy into slot MDX_NTERMS-1)
int mdx_mlp_enqueue(mdx_ctx* c, bool energy) {
  mdx_mlp* m = c->mlp;
  MDX_CHECK(m, MDX_ERR_INVALID, "MLP term requested but no surrogate attached");
  m->nl->pos = c->pos;
  int rc;
  if ((rc = mdx_nlist_update(m->nl, false))) return rc;
  if ((rc = mdx_nlist_enqueue_clear(m->nl))) return rc;
  if ((rc = ml_forward_backward(m, energy ? c->efix.p + MDX_E_MLP : nullptr, false))) return rc;
  return ml_force_kernel(m, c->ffix.p);
}

extern "C" int mdx_ctx_set_mlp(mdx_ctx* c, mdx_mlp* m) {
  MDX_CHECK(c, MDX_ERR_INVALID, "null context");
  MDX_CHECK(!m || m->host == c, MDX_ERR_INVALID, "surrogate belongs to another context");
  c->mlp = m;
  mdx_invalidate_graph(c);
  return 0;
}

// descriptors (N x 96, original atom order) of the host's current positions
extern "C" int mdx_mlp_descriptors(mdx_mlp* m, float* d_out) {
  MDX_CHECK(m && d_out, MDX_ERR_INVALID, "mdx_mlp_descriptors: bad arguments");
  mdx_ctx* c = m->host;
  MDX_ENTER();
  m->nl->pos = c->pos;
  int rc = mdx_nlist_update(m->nl, true);
  if (!rc) rc = mdx_nlist_enqueue_clear(m->nl);
  if (!rc) {
    MlAtoms a = ml_atoms(m);
    k_ml_desc<<<m->n, 64, 0, c->stream>>>(a, m->k, m->nl->xq.p, m->nl->sorted_orig.p, m->nl->nbr_count.p,
                                         m->nl->nbr_list.p, m->row_of.p, m->mean.p, m->rstd.p, m->G.p, m->T.p, m->X.p);
    k_ml_unpermute<<<div_up((long long)m->n * ML_F, 256), 256, 0, c->stream>>>(m->n, ML_F, m->row_of.p, m->G.p, d_out);
    if (hipGetLastError() != hipSuccess) rc = MDX_ERR_HIP;
  }
  MDX_LEAVE_RET(rc);
}

// forces (N x 3 f64, may be null) and total energy of the surrogate at the host's positions
extern "C" int mdx_mlp_eval(mdx_mlp* m, double* d_force_out, double* h_energy) {
  MDX_CHECK(m, MDX_ERR_INVALID, "mdx_mlp_eval: null surrogate");
  mdx_ctx* c = m->host;
  MDX_ENTER();
  m->nl->pos = c->pos;
  long long n3 = 3LL * m->n;
  int rc = mdx_nlist_update(m->nl, true);
  if (!rc) rc = mdx_nlist_enqueue_clear(m->nl);
  if (!rc) {
    k_ml_zero_i64<<<div_up(n3, 256), 256, 0, c->stream>>>(n3, m->nl->ffix.p);
    k_ml_zero_i64<<<1, 64, 0, c->stream>>>(2, m->scal.p);
    rc = ml_forward_backward(m, m->scal.p, false);
  }
  if (!rc) rc = ml_force_kernel(m, m->nl->ffix.p);
  if (!rc && d_force_out) {
    k_ml_forces_out<<<div_up(m->n, 256), 256, 0, c->stream>>>(m->n, m->nl->ffix.p, d_force_out);
    if (hipGetLastError() != hipSuccess) rc = MDX_ERR_HIP;
  }
  if (!rc && h_energy) {
    long long e = 0;
    MDX_HIP(hipMemcpyAsync(&e, m->scal.p, sizeof(long long), hipMemcpyDeviceToHost, c->stream));
    MDX_HIP(hipStreamSynchronize(c->stream));
    *h_energy = (double)e * kFixInv;
  }
  MDX_LEAVE_RET(rc);
}

extern "C" int mdx_mlp_adam_config(mdx_mlp* m, double beta1, double beta2, double eps) {
  MDX_CHECK(m && beta1 >= 0 && beta1 < 1 && beta2 >= 0 && beta2 < 1 && eps > 0, MDX_ERR_INVALID, "bad Adam settings");
  m->beta1 = (float)beta1;
  m->beta2 = (float)beta2;
  m->eps = (float)eps;
  return 0;
}

// weight gradients of one element network from the rows [r0, r1)
static int ml_weight_grads(mdx_mlp* m, int e, int r0, int r1) {
  hipStream_t st = m->host->stream;
  MlNet& N = m->net[e];
  const int K = r1 - r0;
  if (K <= 0) {
    MDX_HIP(hipMemsetAsync(N.grad.p, 0, P_N * sizeof(float), st));
    return 0;
  }
  const int splits = std::max(1, std::min(K / 512, (int)(m->part.n / (ML_H * ML_H))));
  int rc;
  auto wgrad = [&](int M, int Nn, const bf16_t* At, const bf16_t* Bm, const bf16_t* At2, const bf16_t* Bm2,
                   float* out) -> int {
    // out (M x Nn) = At^T Bm + At2^T Bm2, At rows = atoms (row-major [row][M]), Bm [row][Nn]
    int r;
    if ((r = gemm(m, EPI_F32, M, Nn, K, At + (long long)r0 * M, 1, M, Bm + (long long)r0 * Nn, nullptr, Nn, 1,
                  1 << 30, m->part.p, Nn, splits)))
      return r;
    if ((r = gemm(m, EPI_F32_ACC, M, Nn, K, At2 + (long long)r0 * M, 1, M, Bm2 + (long long)r0 * Nn, nullptr, Nn, 1,
                  1 << 30, m->part.p, Nn, splits)))
      return r;
    const long long cnt = (long long)M * Nn;
    k_ml_reduce_splits<<<div_up(cnt, 256), 256, 0, st>>>(m->part.p, splits, cnt, out);
    MDX_LAUNCH_CHECK();
    return 0;
  };
  // dW1 = U^T G1 + X^T dz1 ; dW2 = B1^T G2 + H1^T dz2 ; dW3 = B2^T G3 + H2^T dz3
  if ((rc = wgrad(ML_F, ML_H, m->U.p, m->G1.p, m->X.p, m->DZ16_1.p, N.grad.p + P_W1)) ||
      (rc = wgrad(ML_H, ML_H, m->B1.p, m->G2.p, m->H1.p, m->DZ16_2.p, N.grad.p + P_W2)) ||
      (rc = wgrad(ML_H, ML_H, m->B2.p, m->G3.p, m->H2.p, m->DZ16_3.p, N.grad.p + P_W3)))
    return rc;
  // bias and head gradients: column sums (the split-K partial buffer is free again here)
  const float* cols_in[4] = {m->DZ1.p, m->DZ2.p, m->DZ3.p, m->T4.p};
  const int cols_out[4] = {P_B1, P_B2, P_B3, P_W4};
  for (int q = 0; q < 4; ++q) {
    k_ml_colsum_part<<<COLSUM_CHUNKS, ML_H, 0, st>>>(cols_in[q], r0, r1, ML_H, m->part.p);
    k_ml_colsum_fin<<<1, ML_H, 0, st>>>(m->part.p, ML_H, N.grad.p + cols_out[q]);
  }
  k_ml_fill_f32<<<1, 64, 0, st>>>(1, N.grad.p + P_B4, 0.f);
  MDX_LAUNCH_CHECK();
  return 0;
}

// One force-matching step on one frame: positions d_pos (N x 3 f64, becomes the host's bound
// positions for the call) and reference forces d_fref. lr > 0 applies Adam; lr == 0 only
// computes the gradients (mdx_mlp_get_grads). Loss (mean squared force error per component)
// to *h_loss when not null (synchronises).
extern "C" int mdx_mlp_train_step(mdx_mlp* m, const double* d_pos, const double* d_fref, double lr, double* h_loss) {
  MDX_CHECK(m && d_pos && d_fref && lr >= 0, MDX_ERR_INVALID, "mdx_mlp_train_step: bad arguments");
  mdx_ctx* c = m->host;
  MDX_ENTER();
  hipStream_t st = c->stream;
  int rc = ml_alloc(m, true);
  if (rc) MDX_LEAVE_RET(rc);
  m->nl->pos = const_cast<double*>(d_pos);
  const long long n3 = 3LL * m->n;
  const int R = m->rows, rs = m->rowsO;
  const long long RH = (long long)R * ML_H;
  MlNet &N0 = m->net[0], &N1 = m->net[1];
  if ((rc = mdx_nlist_update(m->nl, true)) || (rc = mdx_nlist_enqueue_clear(m->nl))) MDX_LEAVE_RET(rc);
  k_ml_zero_i64<<<div_up(n3, 256), 256, 0, st>>>(n3, m->nl->ffix.p);
  k_ml_zero_i64<<<1, 64, 0, st>>>(2, m->scal.p);
  if ((rc = ml_forward_backward(m, m->scal.p, true)) || (rc = ml_force_kernel(m, m->nl->ffix.p))) MDX_LEAVE_RET(rc);
  k_ml_loss<<<div_up(m->n, 256), 256, 0, st>>>(m->n, m->nl->ffix.p, d_fref, m->w.p, m->scal.p + 1);
  MDX_LAUNCH_CHECK();
  MlAtoms a = ml_atoms(m);
  k_ml_jvp<<<m->n, 64, 0, st>>>(a, m->k, m->nl->xq.p, m->nl->sorted_orig.p, m->nl->nbr_count.p, m->nl->nbr_list.p,
                               m->row_of.p, m->T.p, m->w.p, m->rstd.p, m->U.p);
  MDX_LAUNCH_CHECK();
  // tangent pass: a1 = U W1 ; b1 = a1 S1, c1 = a1 hb1 DD1 ; a2 = b1 W2 ; ... a3 = b2 W3
  if ((rc = gemm(m, EPI_F32, R, ML_H, ML_F, m->U.p, ML_F, 1, N0.w1.p, N1.w1.p, ML_H, 1, rs, m->A1.p, ML_H)))
    MDX_LEAVE_RET(rc);
  k_ml_tangent<<<div_up(RH, 256), 256, 0, st>>>(RH, m->A1.p, m->S1.p, m->HB1.p, m->D1.p, m->B1.p, m->C1.p);
  if ((rc = gemm(m, EPI_F32, R, ML_H, ML_H, m->B1.p, ML_H, 1, N0.w2.p, N1.w2.p, ML_H, 1, rs, m->A2.p, ML_H)))
    MDX_LEAVE_RET(rc);
  k_ml_tangent<<<div_up(RH, 256), 256, 0, st>>>(RH, m->A2.p, m->S2.p, m->HB2.p, m->D2.p, m->B2.p, m->C2.p);
  if ((rc = gemm(m, EPI_F32, R, ML_H, ML_H, m->B2.p, ML_H, 1, N0.w3.p, N1.w3.p, ML_H, 1, rs, m->A3.p, ML_H)))
    MDX_LEAVE_RET(rc);
  k_ml_top<<<div_up(RH, 256), 256, 0, st>>>(R, rs, m->A3.p, m->S3.p, m->D3.p, N0.p.p, N1.p.p, m->T4.p, m->DZ16_3.p,
                                            m->DZ3.p);
  // reverse through the forward pass: dH2 = dz3 W3^T ; dz2 = c2 + dH2 S2 ; dH1 = dz2 W2^T ; dz1 = c1 + dH1 S1
  if ((rc = gemm(m, EPI_F32, R, ML_H, ML_H, m->DZ16_3.p, ML_H, 1, N0.w3.p, N1.w3.p, 1, ML_H, rs, m->DH2.p, ML_H)))
    MDX_LEAVE_RET(rc);
  k_ml_dz<<<div_up(RH, 256), 256, 0, st>>>(RH, m->C2.p, m->DH2.p, m->S2.p, m->DZ16_2.p, m->DZ2.p);
  if ((rc = gemm(m, EPI_F32, R, ML_H, ML_H, m->DZ16_2.p, ML_H, 1, N0.w2.p, N1.w2.p, 1, ML_H, rs, m->DH1.p, ML_H)))
    MDX_LEAVE_RET(rc);
  k_ml_dz<<<div_up(RH, 256), 256, 0, st>>>(RH, m->C1.p, m->DH1.p, m->S1.p, m->DZ16_1.p, m->DZ1.p);
  MDX_LAUNCH_CHECK();
  if ((rc = ml_weight_grads(m, 0, 0, m->nO)) || (rc = ml_weight_grads(m, 1, rs, rs + m->nH))) MDX_LEAVE_RET(rc);
  if (lr > 0) {
    m->adam_t++;
    const float c1 = 1.f / (1.f - std::pow(m->beta1, (float)m->adam_t));
    const float c2 = 1.f / (1.f - std::pow(m->beta2, (float)m->adam_t));
    for (int e = 0; e < 2; ++e) {
      MlNet& N = m->net[e];
      k_ml_adam<<<div_up(P_N, 256), 256, 0, st>>>(P_N, N.p.p, N.grad.p, N.m.p, N.v.p, (float)lr, m->beta1, m->beta2,
                                                  m->eps, c1, c2);
      k_ml_tobf16<<<div_up(ML_F * ML_H, 256), 256, 0, st>>>(ML_F * ML_H, N.p.p + P_W1, N.w1.p);
      k_ml_tobf16<<<div_up(ML_H * ML_H, 256), 256, 0, st>>>(ML_H * ML_H, N.p.p + P_W2, N.w2.p);
      k_ml_tobf16<<<div_up(ML_H * ML_H, 256), 256, 0, st>>>(ML_H * ML_H, N.p.p + P_W3, N.w3.p);
    }
    MDX_LAUNCH_CHECK();
  }
  m->nl->pos = c->pos;
  if (h_loss) {
    long long l = 0;
    MDX_HIP(hipMemcpyAsync(&l, m->scal.p + 1, sizeof(long long), hipMemcpyDeviceToHost, st));
    MDX_HIP(hipStreamSynchronize(st));
    *h_loss = (double)l * kFixInv;
  }
  MDX_LEAVE_RET(0);
}

extern "C" int mdx_mlp_get_grads(mdx_mlp* m, int elem, float* h_grad) {
  MDX_CHECK(m && h_grad && (elem == 0 || elem == 1) && m->net[elem].grad.p, MDX_ERR_INVALID,
            "mdx_mlp_get_grads: no gradients (run mdx_mlp_train_step first)");
  MDX_HIP(hipStreamSynchronize(m->host->stream));
  MDX_HIP(hipMemcpy(h_grad, m->net[elem].grad.p, P_N * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int mdx_mlp_param_count(void) { return P_N; }

// GEMM check hook: C (M x N f32) = A (M x K bf16 row-major) B (K x N bf16 row-major)
extern "C" int mdx_mlp_gemm_check(mdx_mlp* m, int M, int N, int K, const void* d_A, const void* d_B, float* d_C) {
  MDX_CHECK(m && d_A && d_B && d_C && M > 0 && N > 0 && K > 0, MDX_ERR_INVALID, "mdx_mlp_gemm_check: bad arguments");
  mdx_ctx* c = m->host;
  MDX_ENTER();
  int rc = gemm(m, EPI_F32, M, N, K, (const bf16_t*)d_A, K, 1, (const bf16_t*)d_B, nullptr, N, 1, 1 << 30, d_C, N);
  MDX_LEAVE_RET(rc);
}
