// mag_generate.hip -- seeded synthetic configurations, generated on the GPU.
//
// Streams: Philox4x32-10 (Salmon et al., SC'11) keyed by the 64-bit seed,
// counter = (index, draw, tag, 0); tag 1 = link draws (index = link),
// tag 2 = site gauge draws (index = site).  One Philox block gives two
// uniforms in (0,1) (53 bits each) and, by Box-Muller, two normals.
//   SU(2) link: a ~ N(0,1)^3 (draws 0,1), U = (cos(eps|a|), sin(eps|a|) a_hat)
//   SU(3) link: c ~ N(0,1)^8 (draws 0..3), U = exp(i eps sum c_a lambda_a/2)
//               (Gell-Mann basis of matrix.h:987-1046; Taylor series with
//               scaling and squaring, then Gram-Schmidt)
//   gauge:      SU(2) normalised N(0,1)^4; SU(3) Gram-Schmidt of a complex
//               Gaussian 3x3 times det^(-1/3) (Haar, Mezzadri 2007)
//   U'_x,mu = g(x) U_x,mu g(x+mu)^+ (the form of mag.cpp:1056-1062)
// The gauge draws use their own key, gseed = seed ^ (copy * 0x9E3779B97F4A7C15):
// copy 0 is the plain configuration, copies 1, 2, ... are gauge copies of the
// same physical field (same link draws, different random gauge g).
#include "mag_internal.h"
#include "mag_layout.h"
#include "mag_math.h"

namespace mag {

__device__ __forceinline__ void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                       uint64_t seed, uint32_t* o) {
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
  for (int r = 0; r < 10; r++) {
    uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
    uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  o[0] = c0; o[1] = c1; o[2] = c2; o[3] = c3;
}

__device__ __forceinline__ double u53(uint32_t a, uint32_t b) {
  uint64_t x = (((uint64_t)a << 32) | b) >> 11;
  return ((double)x + 0.5) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ void gauss2(uint64_t seed, uint32_t index, uint32_t draw,
                                       uint32_t tag, double* z) {
  uint32_t o[4];
  philox(index, draw, tag, 0, seed, o);
  double u1 = u53(o[0], o[1]), u2 = u53(o[2], o[3]);
  double r = sqrt(-2.0 * log(u1));
  double s, c;
  sincos(2.0 * 3.14159265358979323846 * u2, &s, &c);
  z[0] = r * c;
  z[1] = r * s;
}

enum { TAG_LINK = 1, TAG_GAUGE = 2 };

// ------------------------------------------------------------------ SU(2)
__device__ __forceinline__ Q<double> su2_gauge(uint64_t seed, uint32_t site) {
  double z[4];
  gauss2(seed, site, 0, TAG_GAUGE, z);
  gauss2(seed, site, 1, TAG_GAUGE, z + 2);
  double n = sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2] + z[3] * z[3]);
  return Q<double>{z[0] / n, z[1] / n, z[2] / n, z[3] / n};
}

// thread per lexicographic site; writes the reference AoS layout
__global__ void k_gen_su2(double* __restrict__ aos, Geom g, double eps, uint64_t seed,
                          uint64_t gseed) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= 2 * g.Vh) return;
  uint32_t c[4];
  uint32_t r = s;
  c[0] = r % (uint32_t)g.L0; r /= (uint32_t)g.L0;
  c[1] = r % (uint32_t)g.L1; r /= (uint32_t)g.L1;
  c[2] = r % (uint32_t)g.L2; c[3] = r / (uint32_t)g.L2;
  const uint32_t L[4] = {(uint32_t)g.L0, (uint32_t)g.L1, (uint32_t)g.L2, (uint32_t)g.L3};
  const uint32_t str[4] = {1u, L[0], L[0] * L[1], L[0] * L[1] * L[2]};
  Q<double> gx = su2_gauge(gseed, s);
  for (int mu = 0; mu < 4; mu++) {
    uint32_t l = 4 * s + mu;
    double z[4];
    gauss2(seed, l, 0, TAG_LINK, z);
    gauss2(seed, l, 1, TAG_LINK, z + 2);
    double an = sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
    Q<double> U{1.0, 0.0, 0.0, 0.0};
    if (an > 0) {
      double sa = sin(eps * an) / an;
      U = Q<double>{cos(eps * an), sa * z[0], sa * z[1], sa * z[2]};
    }
    uint32_t f = (c[mu] + 1 == L[mu]) ? s - (L[mu] - 1) * str[mu] : s + str[mu];
    Q<double> gf = su2_gauge(gseed, f);
    Q<double> A = qmul_dag(qmul(gx, U), gf);
    double* o = aos + (size_t)l * 4;
    o[0] = A.a0; o[1] = A.a1; o[2] = A.a2; o[3] = A.a3;
  }
}

// ------------------------------------------------------------------ SU(3)
typedef C<double> cd;

__device__ __forceinline__ M3<double> m3mul(const M3<double>& A, const M3<double>& B) {
  M3<double> R;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      cd s{0.0, 0.0};
      for (int k = 0; k < 3; k++) s = cadd(s, cmul(A.m[i][k], B.m[k][j]));
      R.m[i][j] = s;
    }
  return R;
}
__device__ __forceinline__ M3<double> m3mul_dag(const M3<double>& A, const M3<double>& B) {
  M3<double> R;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      cd s{0.0, 0.0};
      for (int k = 0; k < 3; k++) s = cadd(s, cmulc(A.m[i][k], B.m[j][k]));
      R.m[i][j] = s;
    }
  return R;
}
__device__ __forceinline__ void reunit3(M3<double>& M) {
  double n0 = 0;
  for (int k = 0; k < 3; k++) n0 += cnorm2(M.m[0][k]);
  n0 = 1.0 / sqrt(n0);
  for (int k = 0; k < 3; k++) M.m[0][k] = {M.m[0][k].re * n0, M.m[0][k].im * n0};
  cd d{0.0, 0.0};
  for (int k = 0; k < 3; k++) d = cadd(d, cmulc(M.m[1][k], M.m[0][k]));
  for (int k = 0; k < 3; k++) M.m[1][k] = csub(M.m[1][k], cmul(d, M.m[0][k]));
  double n1 = 0;
  for (int k = 0; k < 3; k++) n1 += cnorm2(M.m[1][k]);
  n1 = 1.0 / sqrt(n1);
  for (int k = 0; k < 3; k++) M.m[1][k] = {M.m[1][k].re * n1, M.m[1][k].im * n1};
  M2<double> m2;
  for (int i = 0; i < 2; i++)
    for (int k = 0; k < 3; k++) m2.r[i][k] = M.m[i][k];
  row2_of(m2, M.m[2]);
}
__device__ __forceinline__ cd m3det(const M3<double>& A) {
  cd t0 = csub(cmul(A.m[1][1], A.m[2][2]), cmul(A.m[1][2], A.m[2][1]));
  cd t1 = csub(cmul(A.m[1][0], A.m[2][2]), cmul(A.m[1][2], A.m[2][0]));
  cd t2 = csub(cmul(A.m[1][0], A.m[2][1]), cmul(A.m[1][1], A.m[2][0]));
  return cadd(csub(cmul(A.m[0][0], t0), cmul(A.m[0][1], t1)), cmul(A.m[0][2], t2));
}

__device__ M3<double> su3_exp(const double* c, double eps) {
  const double r3 = 1.0 / sqrt(3.0);
  const double e = 0.5 * eps;
  M3<double> H;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) H.m[i][j] = {0.0, 0.0};
  H.m[0][1] = cadd(H.m[0][1], cd{e * c[0], 0});
  H.m[1][0] = cadd(H.m[1][0], cd{e * c[0], 0});
  H.m[0][1] = cadd(H.m[0][1], cd{0, -e * c[1]});
  H.m[1][0] = cadd(H.m[1][0], cd{0, e * c[1]});
  H.m[0][0] = cadd(H.m[0][0], cd{e * c[2], 0});
  H.m[1][1] = cadd(H.m[1][1], cd{-e * c[2], 0});
  H.m[0][2] = cadd(H.m[0][2], cd{e * c[3], 0});
  H.m[2][0] = cadd(H.m[2][0], cd{e * c[3], 0});
  H.m[0][2] = cadd(H.m[0][2], cd{0, -e * c[4]});
  H.m[2][0] = cadd(H.m[2][0], cd{0, e * c[4]});
  H.m[1][2] = cadd(H.m[1][2], cd{e * c[5], 0});
  H.m[2][1] = cadd(H.m[2][1], cd{e * c[5], 0});
  H.m[1][2] = cadd(H.m[1][2], cd{0, -e * c[6]});
  H.m[2][1] = cadd(H.m[2][1], cd{0, e * c[6]});
  H.m[0][0] = cadd(H.m[0][0], cd{e * c[7] * r3, 0});
  H.m[1][1] = cadd(H.m[1][1], cd{e * c[7] * r3, 0});
  H.m[2][2] = cadd(H.m[2][2], cd{-2 * e * c[7] * r3, 0});
  M3<double> A;
  double nrm = 0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      A.m[i][j] = {-H.m[i][j].im, H.m[i][j].re};
      nrm += cnorm2(A.m[i][j]);
    }
  nrm = sqrt(nrm);
  int sq = 0;
  while (nrm > 0.25) {
    nrm *= 0.5;
    sq++;
  }
  double sc = ldexp(1.0, -sq);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) A.m[i][j] = {A.m[i][j].re * sc, A.m[i][j].im * sc};
  M3<double> R, T;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      R.m[i][j] = {i == j ? 1.0 : 0.0, 0.0};
      T.m[i][j] = R.m[i][j];
    }
  for (int k = 1; k <= 20; k++) {
    T = m3mul(T, A);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        T.m[i][j] = {T.m[i][j].re * (1.0 / k), T.m[i][j].im * (1.0 / k)};
        R.m[i][j] = cadd(R.m[i][j], T.m[i][j]);
      }
  }
  for (int k = 0; k < sq; k++) R = m3mul(R, R);
  reunit3(R);
  return R;
}

__device__ M3<double> su3_haar(uint64_t seed, uint32_t site) {
  double z[18];
  for (int k = 0; k < 9; k++) gauss2(seed, site, (uint32_t)k, TAG_GAUGE, z + 2 * k);
  M3<double> Z;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) Z.m[i][j] = {z[2 * (3 * i + j)], z[2 * (3 * i + j) + 1]};
  for (int i = 0; i < 3; i++) {
    for (int p = 0; p < i; p++) {
      cd d{0.0, 0.0};
      for (int k = 0; k < 3; k++) d = cadd(d, cmulc(Z.m[i][k], Z.m[p][k]));
      for (int k = 0; k < 3; k++) Z.m[i][k] = csub(Z.m[i][k], cmul(d, Z.m[p][k]));
    }
    double n = 0;
    for (int k = 0; k < 3; k++) n += cnorm2(Z.m[i][k]);
    n = 1.0 / sqrt(n);
    for (int k = 0; k < 3; k++) Z.m[i][k] = {Z.m[i][k].re * n, Z.m[i][k].im * n};
  }
  cd det = m3det(Z);
  double ph = atan2(det.im, det.re) / 3.0;
  double s, c;
  sincos(ph, &s, &c);
  cd f{c, -s};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) Z.m[i][j] = cmul(Z.m[i][j], f);
  reunit3(Z);
  return Z;
}

// site gauges first (one pass), then links
__global__ void k_gen_su3_gauge(double* __restrict__ gauge, uint32_t V, uint64_t seed) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= V) return;
  M3<double> G = su3_haar(seed, s);
  double* o = gauge + (size_t)s * 18;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      o[2 * (3 * i + j)] = G.m[i][j].re;
      o[2 * (3 * i + j) + 1] = G.m[i][j].im;
    }
}

__device__ __forceinline__ M3<double> ld18(const double* p) {
  M3<double> M;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) M.m[i][j] = {p[2 * (3 * i + j)], p[2 * (3 * i + j) + 1]};
  return M;
}

// thread per link
__global__ void k_gen_su3_links(double* __restrict__ aos, const double* __restrict__ gauge,
                                Geom g, double eps, uint64_t seed) {
  const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= 8 * g.Vh) return;
  const uint32_t s = l >> 2;
  const int mu = (int)(l & 3u);
  uint32_t c[4];
  uint32_t r = s;
  c[0] = r % (uint32_t)g.L0; r /= (uint32_t)g.L0;
  c[1] = r % (uint32_t)g.L1; r /= (uint32_t)g.L1;
  c[2] = r % (uint32_t)g.L2; c[3] = r / (uint32_t)g.L2;
  const uint32_t L[4] = {(uint32_t)g.L0, (uint32_t)g.L1, (uint32_t)g.L2, (uint32_t)g.L3};
  const uint32_t str[4] = {1u, L[0], L[0] * L[1], L[0] * L[1] * L[2]};
  double cc[8];
  for (int k = 0; k < 4; k++) gauss2(seed, l, (uint32_t)k, TAG_LINK, cc + 2 * k);
  M3<double> U = su3_exp(cc, eps);
  uint32_t f = (c[mu] + 1 == L[mu]) ? s - (L[mu] - 1) * str[mu] : s + str[mu];
  M3<double> A = m3mul(ld18(gauge + (size_t)s * 18), U);
  A = m3mul_dag(A, ld18(gauge + (size_t)f * 18));
  reunit3(A);
  double* o = aos + (size_t)l * 18;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      o[2 * (3 * i + j)] = A.m[i][j].re;
      o[2 * (3 * i + j) + 1] = A.m[i][j].im;
    }
}

void launch_gen_su2(double* aos, const Geom& g, double eps, uint64_t seed, uint64_t gseed,
                    hipStream_t st) {
  const uint32_t V = 2 * g.Vh;
  hipLaunchKernelGGL(k_gen_su2, dim3((V + 127) / 128), dim3(128), 0, st, aos, g, eps, seed, gseed);
}
void launch_gen_su3(double* aos, double* gauge_scratch, const Geom& g, double eps,
                    uint64_t seed, uint64_t gseed, hipStream_t st) {
  const uint32_t V = 2 * g.Vh;
  hipLaunchKernelGGL(k_gen_su3_gauge, dim3((V + 127) / 128), dim3(128), 0, st, gauge_scratch,
                     V, gseed);
  const uint32_t nl = 4 * V;
  hipLaunchKernelGGL(k_gen_su3_links, dim3((nl + 127) / 128), dim3(128), 0, st, aos,
                     gauge_scratch, g, eps, seed);
}

// ------------------------------------------------------------------ gauge copies
// Initial spins of gauge copy c (spin form) of the configuration of
// launch_gen_su2(seed, gseed_0): copy c of that configuration is
// g_c U g_c^+ = h U_0 h^+ with h = g_c g_0^+, so it starts from the spin
// of h (row 3 of its rotation); copy 0 starts exactly at e3.  Thread per
// lexicographic site; SoA spins per parity, copy c at S + c*sstride.
__global__ void k_gen_copy_spins(double* __restrict__ S, size_t sstride, int ncopies, Geom g,
                                 uint64_t seed) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= 2 * g.Vh) return;
  uint32_t r = s;
  const uint32_t x = r % (uint32_t)g.L0;
  r /= (uint32_t)g.L0;
  const uint32_t y = r % (uint32_t)g.L1;
  r /= (uint32_t)g.L1;
  const uint32_t z = r % (uint32_t)g.L2, t = r / (uint32_t)g.L2;
  const int p = (int)((x + y + z + t) & 1u);
  const size_t h = s >> 1;
  const Q<double> g0 = su2_gauge(gauge_copy_seed(seed, 0), s);
  for (int c = 0; c < ncopies; c++) {
    double v[3] = {0.0, 0.0, 1.0};
    if (c > 0) {
      const Q<double> q = qmul_dag(su2_gauge(gauge_copy_seed(seed, (uint32_t)c), s), g0);
      v[0] = 2 * (q.a1 * q.a3 + q.a0 * q.a2);
      v[1] = 2 * (q.a2 * q.a3 - q.a0 * q.a1);
      v[2] = q.a0 * q.a0 + q.a3 * q.a3 - q.a1 * q.a1 - q.a2 * q.a2;
    }
    double* b = S + (size_t)c * sstride + (size_t)p * 3 * g.Vh + h;
    b[0] = v[0];
    b[g.Vh] = v[1];
    b[2 * (size_t)g.Vh] = v[2];
  }
}

void launch_gen_copy_spins(double* S, size_t sstride, int ncopies, const Geom& g, uint64_t seed,
                           hipStream_t st) {
  const uint32_t V = 2 * g.Vh;
  hipLaunchKernelGGL(k_gen_copy_spins, dim3((V + 127) / 128), dim3(128), 0, st, S, sstride,
                     ncopies, g, seed);
}

// Starting site gauges of SU(3) gauge copies 0..ncopies-1 (site form, on
// the links of copy 0): h_c(x) = g_c(x) g_0(x)^+ with the Haar gauges of
// launch_gen_su3 (so G_c U_0 G_c^+ is copy c), copy 0 the identity.  Rows
// 0, 1 in the chunked site layout, copy c at G + c*gstride.
__global__ void k_gen_su3_copy_gauges(double* __restrict__ G, size_t gstride, int ncopies,
                                      Geom g, uint64_t seed) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= 2 * g.Vh) return;
  uint32_t r = s;
  const uint32_t x = r % (uint32_t)g.L0;
  r /= (uint32_t)g.L0;
  const uint32_t y = r % (uint32_t)g.L1;
  r /= (uint32_t)g.L1;
  const uint32_t z = r % (uint32_t)g.L2, t = r / (uint32_t)g.L2;
  const int p = (int)((x + y + z + t) & 1u);
  const uint32_t h = s >> 1;
  const M3<double> g0 = su3_haar(gauge_copy_seed(seed, 0), s);
  for (int c = 0; c < ncopies; c++) {
    M2<double> m;
    if (c == 0) {
#pragma unroll
      for (int k = 0; k < 6; k++) m.r[k / 3][k % 3] = {(k == 0 || k == 4) ? 1.0 : 0.0, 0.0};
    } else {
      M3<double> q = m3mul_dag(su3_haar(gauge_copy_seed(seed, (uint32_t)c), s), g0);
      reunit3(q);
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) m.r[i][k] = q.m[i][k];
    }
    stm(site_ptr<12>(G + (size_t)c * gstride, g, p, h), m);
  }
}

void launch_gen_su3_copy_gauges(double* G, size_t gstride, int ncopies, const Geom& g,
                                uint64_t seed, hipStream_t st) {
  const uint32_t V = 2 * g.Vh;
  hipLaunchKernelGGL(k_gen_su3_copy_gauges, dim3((V + 127) / 128), dim3(128), 0, st, G, gstride,
                     ncopies, g, seed);
}

}  // namespace mag
