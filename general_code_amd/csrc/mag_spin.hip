// mag_spin.hip -- spin-form MAG kernels behind the reference's mag.h call
// surface (ilyakudrov/general_code lib/cpu/src/mag.cpp).
//
// State: one unit 3-vector s(x) per site (matrix.h:773-787); links stay
// read-only.  Spins live in HBM as SoA per parity and component:
//   s_c(p, h) = S[(p*3 + c)*Vh + h]
// so consecutive lanes read consecutive doubles of each component array.
// The reference updates spins in place in lexicographic order under an
// OpenMP collapse(4) loop (mag.cpp:1386, 1491), which races; here the sweep
// is a checkerboard: all sites of one parity update at once from the spins
// of the other parity, which is race-free and deterministic.
//
// Per site (contribution_site, mag.cpp:569-626):
//   h(x) = sum_mu [ M(U_x,mu) s(x+mu) + M(U_x-mu,mu)^T s(x-mu) ]
// then one of
//   reflect_fast  s <- 2 (s.h_hat) h_hat - s        (matrix.h:825-840)
//   parallel      s <- h_hat, d = 1 - s_old.h_hat   (matrix.h:842-856)
//   geodesic      s <- rotate s towards h_hat by omega*theta
//   heat bath     s.h_hat ~ exp(|h| x / T) on [-1,1] (mag.cpp:402-433)
#include "mag_internal.h"
#include "mag_layout.h"
#include "mag_math.h"
#include "mag_reduce.h"
#include "mag_spin_math.h"

namespace mag {

// Philox4x32-10 and the uniform of the synthetic generators
__device__ __forceinline__ void philox_s(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                         uint64_t seed, uint32_t* o) {
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
  for (int r = 0; r < 10; r++) {
    const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
    const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
    const uint32_t n0 = hi1 ^ c1 ^ k0;
    const uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  o[0] = c0; o[1] = c1; o[2] = c2; o[3] = c3;
}
__device__ __forceinline__ double u53_s(uint32_t a, uint32_t b) {
  const uint64_t x = (((uint64_t)a << 32) | b) >> 11;
  return ((double)x + 0.5) * (1.0 / 9007199254740992.0);
}
enum { TAG_SPIN = 3, TAG_HEATBATH = 4 };

// heat bath update of one spin (mag.cpp:402-433): p(x) ~ exp(k x) on
// [-1,1], k = |h|/T, drawn by inversion (the reference's rejection loop has
// the same distribution; inversion bounds the work per lane).
__device__ __forceinline__ void heatbath_spin(V3& s, const V3& h, double T, uint64_t seed,
                                              uint32_t site, uint32_t sweep) {
  uint32_t o[4];
  philox_s(site, sweep, TAG_HEATBATH, 0, seed, o);
  const double u = u53_s(o[0], o[1]);
  const double phi = 2 * 3.14159265358979323846 * u53_s(o[2], o[3]);
  const double hn = sqrt(h.a1 * h.a1 + h.a2 * h.a2 + h.a3 * h.a3);
  const double k = hn / T;
  double a3 = (k > 0) ? 1 + log1p(-(1 - u) * (-expm1(-2 * k))) / k : 2 * u - 1;
  a3 = fmin(1.0, fmax(-1.0, a3));
  const double R = sqrt(1 - a3 * a3);
  double sp, cp;
  sincos(phi, &sp, &cp);
  const double a1 = R * cp, a2 = R * sp;
  if (hn == 0) {
    s = V3{a1, a2, a3};
    return;
  }
  const double n3 = h.a3 / hn;
  const double c = sqrt(1 - n3 * n3);
  if (c == 0) {
    s = V3{n3 * a2, -a1, n3 * a3};
    return;
  }
  const double c1 = 1. / (hn * c);
  const double f1 = h.a1 * c1, f2 = h.a2 * c1;
  s = V3{f2 * a1 + f1 * n3 * a2 + f1 * c * a3, -f1 * a1 + f2 * n3 * a2 + f2 * c * a3,
         -c * a2 + n3 * a3};
}

// ------------------------------------------------------------------ sweeps
// mode: SPIN_OR, SPIN_RELAX, SPIN_OMEGA, SPIN_HEATBATH.  Partial triples:
// (0, sum d, max d) with d = 1 - s_old . h_hat.
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_spin_half(const double* __restrict__ U,
                                                      double* __restrict__ S, Geom g, int p,
                                                      double param, uint64_t seed,
                                                      uint32_t sweep, double* __restrict__ part,
                                                      const SpinCtl* __restrict__ ctl) {
  if (ctl->done || (MODE == SPIN_RELAX && ctl->stop_next)) return;
  const uint32_t h = blockIdx.x * kBlock + threadIdx.x;
  double z = 0.0, ds = 0.0, dm = 0.0;
  if (h < g.Vh) {
    const SiteIdx s = decode(g, h, p);
    V3 H{0.0, 0.0, 0.0};
#pragma unroll
    for (int mu = 0; mu < 4; mu++) {
      const uint32_t hf = fwd_h(g, h, s, mu), hb = back_h(g, h, s, mu);
      contribution1(H, ldq(link_ptr<4>(U, g, p, mu, h)), ld_spin(S, g, 1 - p, hf));
      contribution1_conj(H, ldq(link_ptr<4>(U, g, 1 - p, mu, hb)), ld_spin(S, g, 1 - p, hb));
    }
    V3 sp = ld_spin(S, g, p, h);
    const double hn = sqrt(H.a1 * H.a1 + H.a2 * H.a2 + H.a3 * H.a3);
    const double d = hn > 0 ? 1 - (sp.a1 * H.a1 + sp.a2 * H.a2 + sp.a3 * H.a3) / hn : 0.0;
    if (MODE == SPIN_OR) reflect_fast(sp, H);
    else if (MODE == SPIN_RELAX) parallel(sp, H);
    else if (MODE == SPIN_OMEGA) spin_omega(sp, H, param);
    else heatbath_spin(sp, H, param, seed, s.row * (uint32_t)g.L0 + s.x, sweep);
    st_spin(S, g, p, h, sp);
    ds = d;
    dm = d;
  }
  block_reduce3<kWaves>(z, ds, dm);
  write_part(part, z, ds, dm);
}

// F_spin partials (mag.cpp:925-953): sum_mu (M(U_x,mu) s(x+mu)) . s(x)
__global__ __launch_bounds__(kBlock) void k_spin_functional(const double* __restrict__ U,
                                                            const double* __restrict__ S, Geom g,
                                                            double* __restrict__ part,
                                                            const SpinCtl* __restrict__ ctl) {
  if (ctl && ctl->done) return;
  const int nb = (int)((g.Vh + kBlock - 1) / kBlock);
  const int p = blockIdx.x / nb;
  const uint32_t h = (blockIdx.x % nb) * kBlock + threadIdx.x;
  double f = 0.0, z1 = 0.0, z2 = 0.0;
  if (h < g.Vh) {
    const SiteIdx s = decode(g, h, p);
    const V3 sx = ld_spin(S, g, p, h);
#pragma unroll
    for (int mu = 0; mu < 4; mu++) {
      const V3 v = contribution(ld_spin(S, g, 1 - p, fwd_h(g, h, s, mu)),
                                ldq(link_ptr<4>(U, g, p, mu, h)));
      f += v.a1 * sx.a1 + v.a2 * sx.a2 + v.a3 * sx.a3;
    }
  }
  block_reduce3<kWaves>(f, z1, z2);
  write_part(part, f, z1, z2);
}

// GetGaugeMatrix (matrix.h:960-963)
__device__ __forceinline__ Q<double> gauge_matrix(const V3& s) {
  const double sq = sqrt(2. - 2 * s.a3);
  return Q<double>{-s.a1 / sq, 0., -sq / 2., -s.a2 / sq};
}

// gauge_tranformation_spins (mag.cpp:1045-1068): U <- G(s(x)) U G(s(x+mu))^+
__global__ __launch_bounds__(kBlock) void k_spin_gauge_transform(double* __restrict__ U,
                                                                 const double* __restrict__ S,
                                                                 Geom g) {
  const int nb = (int)((g.Vh + kBlock - 1) / kBlock);
  const int p = blockIdx.x / nb;
  const uint32_t h = (blockIdx.x % nb) * kBlock + threadIdx.x;
  if (h >= g.Vh) return;
  const SiteIdx s = decode(g, h, p);
  const Q<double> G = gauge_matrix(ld_spin(S, g, p, h));
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    const Q<double> Gf = gauge_matrix(ld_spin(S, g, 1 - p, fwd_h(g, h, s, mu)));
    const LRef<double> l = link_ptr<4>(U, g, p, mu, h);
    Q<double> A = qmul(G, ldq(l));
    A = qmul(A, Q<double>{Gf.a0, -Gf.a1, -Gf.a2, -Gf.a3});
    stq(l, A);
  }
}

// uniform spins (mag.cpp:166-236 construction, Philox stream keyed by seed,
// counter = (lexicographic site, attempt, 3, 0)); AoS reference layout
__global__ void k_spins_uniform(double* __restrict__ aos, uint32_t V, uint64_t seed) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= V) return;
  double x1, x2;
  uint32_t attempt = 0;
  do {
    uint32_t o[4];
    philox_s(s, attempt++, TAG_SPIN, 0, seed, o);
    x1 = 2 * u53_s(o[0], o[1]) - 1;
    x2 = 2 * u53_s(o[2], o[3]) - 1;
  } while (x1 * x1 + x2 * x2 >= 1 && attempt < 64u);
  const double a = 1 - x1 * x1 - x2 * x2;
  aos[3 * (size_t)s] = 2 * x1 * sqrt(a);
  aos[3 * (size_t)s + 1] = 2 * x2 * sqrt(a);
  aos[3 * (size_t)s + 2] = 1 - 2 * (x1 * x1 + x2 * x2);
}

// spins: reference AoS [V][3] <-> parity-split SoA
__global__ void k_spins_aos_to_soa(const double* __restrict__ aos, double* __restrict__ S, Geom g,
                                   int to_soa) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= 2 * g.Vh) return;
  uint32_t x = s % (uint32_t)g.L0, r = s / (uint32_t)g.L0;
  uint32_t y = r % (uint32_t)g.L1;
  r /= (uint32_t)g.L1;
  uint32_t z = r % (uint32_t)g.L2, t = r / (uint32_t)g.L2;
  const int p = (int)((x + y + z + t) & 1u);
  const uint32_t h = s >> 1;
  double* a = const_cast<double*>(aos) + 3 * (size_t)s;
  if (to_soa) {
    st_spin(S, g, p, h, V3{a[0], a[1], a[2]});
  } else {
    const V3 v = ld_spin(S, g, p, h);
    a[0] = v.a1;
    a[1] = v.a2;
    a[2] = v.a3;
  }
}

// ------------------------------------------------------------------ driver control
// One workgroup: reduce partials and advance the driver state (see
// SpinCtl).  Fixed-order reduction, as for the link sweeps.
__global__ __launch_bounds__(kBlock) void k_spin_reduce(const double* __restrict__ part,
                                                        int nparts, int mode, double invV,
                                                        double invF, SpinCtl* ctl, double* out3) {
  if (mode != SRED_PLAIN && ctl->done) return;
  if ((mode == SRED_RELAX_FINAL || mode == SRED_RELAX_PLAIN) && ctl->stop_next) {
    if (threadIdx.x == 0) ctl->done = 1;
    return;
  }
  double a = 0.0, b = 0.0, c = 0.0;
  for (int i = threadIdx.x; i < nparts; i += kBlock) {
    a += part[3 * i];
    b += part[3 * i + 1];
    c = fmax(c, part[3 * i + 2]);
  }
  block_reduce3<kWaves>(a, b, c);
  if (threadIdx.x != 0) return;
  if (mode == SRED_RELAX_FINAL) {
    // make_maximization_final (mag.cpp:1800-1816)
    const double dmax = c, davg = b * invV;
    ctl->dmax = dmax;
    ctl->davg = davg;
    ctl->iter += 1;
    if (dmax < ctl->tol_max && davg < ctl->tol_avg) {
      ctl->converged = 1;
      ctl->stop_next = 1;
    } else if (ctl->iter >= ctl->max_iter) {
      ctl->stop_next = 1;
    }
  } else if (mode == SRED_RELAX_PLAIN) {
    ctl->dmax = c;
    ctl->davg = b * invV;
  } else if (mode == SRED_FUNC_INIT) {
    ctl->F_prev = a * invF;
  } else if (mode == SRED_FUNC_APPROX) {
    // make_maximization_approximate (mag.cpp:1733-1750), powf kept
    const double F = a * invF, Fo = ctl->F_prev;
    const float scale = ctl->scale;  // the host C library's powf(10, digits)
    const bool eq = (trunc(scale * F) == trunc(scale * Fo)) && (signbit(F) == signbit(Fo));
    ctl->F_prev = F;
    ctl->iter += 1;
    if (eq) {
      ctl->converged = 1;
      ctl->stop_next = 1;
    } else if (ctl->iter >= ctl->max_iter) {
      ctl->stop_next = 1;
    }
  }
  if (out3) {
    out3[0] = a * invF;
    out3[1] = b * invV;
    out3[2] = c;
  }
}

// ------------------------------------------------------------------ launchers
void launch_spin_half(int mode, const double* U, double* S, const Geom& g, int parity,
                      double param, uint64_t seed, uint32_t sweep, double* part,
                      const SpinCtl* ctl, hipStream_t st) {
  const int nb = nblocks(g);
  double* pp = part + (size_t)3 * nb * parity;
#define L_(M)                                                                                  \
  hipLaunchKernelGGL((k_spin_half<M>), dim3(nb), dim3(kBlock), 0, st, U, S, g, parity, param, \
                     seed, sweep, pp, ctl)
  switch (mode) {
    case SPIN_OR: L_(SPIN_OR); break;
    case SPIN_RELAX: L_(SPIN_RELAX); break;
    case SPIN_OMEGA: L_(SPIN_OMEGA); break;
    default: L_(SPIN_HEATBATH); break;
  }
#undef L_
}
void launch_spin_functional(const double* U, const double* S, const Geom& g, double* part,
                            const SpinCtl* ctl, hipStream_t st) {
  hipLaunchKernelGGL(k_spin_functional, dim3(2 * nblocks(g)), dim3(kBlock), 0, st, U, S, g, part,
                     ctl);
}
void launch_spin_reduce(const double* part, int nparts, int mode, double invV, double invF,
                        SpinCtl* ctl, double* out3, hipStream_t st) {
  hipLaunchKernelGGL(k_spin_reduce, dim3(1), dim3(kBlock), 0, st, part, nparts, mode, invV, invF,
                     ctl, out3);
}
void launch_spin_gauge_transform(double* U, const double* S, const Geom& g, hipStream_t st) {
  hipLaunchKernelGGL(k_spin_gauge_transform, dim3(2 * nblocks(g)), dim3(kBlock), 0, st, U, S, g);
}
void launch_spins_uniform(double* aos, const Geom& g, uint64_t seed, hipStream_t st) {
  const uint32_t V = 2 * g.Vh;
  hipLaunchKernelGGL(k_spins_uniform, dim3((V + 255) / 256), dim3(256), 0, st, aos, V, seed);
}
void launch_spins_layout(const double* aos, double* S, const Geom& g, int to_soa,
                         hipStream_t st) {
  const uint32_t V = 2 * g.Vh;
  hipLaunchKernelGGL(k_spins_aos_to_soa, dim3((V + 255) / 256), dim3(256), 0, st, aos, S, g,
                     to_soa);
}

// ------------------------------------------------------------------ gauge copies
__global__ void k_spins_e3(double* __restrict__ S, size_t nreal, uint32_t Vh) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nreal) return;
  // component 2 (a3) of each parity block of each copy is 1
  S[i] = ((i / Vh) % 3 == 2) ? 1.0 : 0.0;
}

// out_x,mu = G(s(x)) U_x,mu G(s(x+mu))^+, G = gauge_of_spin; thread per
// site of either parity (grid 2*nblocks)
__global__ __launch_bounds__(kBlock) void k_spin_links(const double* __restrict__ U,
                                                       const double* __restrict__ S,
                                                       double* __restrict__ out, Geom g) {
  const int nb = (int)((g.Vh + kBlock - 1) / kBlock);
  const int p = blockIdx.x / nb;
  const uint32_t h = (blockIdx.x % nb) * kBlock + threadIdx.x;
  if (h >= g.Vh) return;
  const SiteIdx s = decode(g, h, p);
  const Q<double> G = gauge_of_spin(ld_spin(S, g, p, h));
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    const Q<double> Gf = gauge_of_spin(ld_spin(S, g, 1 - p, fwd_h(g, h, s, mu)));
    const Q<double> A = qmul_dag(qmul(G, ldq(link_ptr<4>(U, g, p, mu, h))), Gf);
    stq(link_ptr<4>(out, g, p, mu, h), A);
  }
}

// the spins of one copy (2 Vh sites) FP64 -> FP32 and FP32 -> FP64, each
// put back on the unit sphere in FP64 (the FP32 phase of the gauge-copy
// spin form)
__global__ void k_spins_to32(const double* __restrict__ a, float* __restrict__ b, Geom g) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * g.Vh) return;
  const int p = (int)(i / g.Vh);
  const uint32_t h = i % g.Vh;
  const V3 v = ld_spin(a, g, p, h);
  const double r = 1.0 / sqrt(v.a1 * v.a1 + v.a2 * v.a2 + v.a3 * v.a3);
  st_spin(b, g, p, h, V3f{(float)(v.a1 * r), (float)(v.a2 * r), (float)(v.a3 * r)});
}
__global__ void k_spins_to64(const float* __restrict__ a, double* __restrict__ b, Geom g) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * g.Vh) return;
  const int p = (int)(i / g.Vh);
  const uint32_t h = i % g.Vh;
  const V3 v = to_d(ld_spin(a, g, p, h));
  const double r = 1.0 / sqrt(v.a1 * v.a1 + v.a2 * v.a2 + v.a3 * v.a3);
  st_spin(b, g, p, h, V3{v.a1 * r, v.a2 * r, v.a3 * r});
}
void launch_spins_to32(const double* a, float* b, const Geom& g, hipStream_t st) {
  hipLaunchKernelGGL(k_spins_to32, dim3((2 * g.Vh + 255) / 256), dim3(256), 0, st, a, b, g);
}
void launch_spins_to64(const float* a, double* b, const Geom& g, hipStream_t st) {
  hipLaunchKernelGGL(k_spins_to64, dim3((2 * g.Vh + 255) / 256), dim3(256), 0, st, a, b, g);
}

void launch_spins_e3(double* S, size_t nreal, const Geom& g, hipStream_t st) {
  if (nreal == 0) return;
  // nreal is a whole number of copies: ncopies * 2 parities * 3 * Vh
  hipLaunchKernelGGL(k_spins_e3, dim3((unsigned)((nreal + 255) / 256)), dim3(256), 0, st, S,
                     nreal, g.Vh);
}
void launch_spin_links(const double* U, const double* S, double* out, const Geom& g,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_spin_links, dim3(2 * nblocks(g)), dim3(kBlock), 0, st, U, S, out, g);
}

// ------------------------------------------------------------------ gauge tracking
// the accumulated gauge of a site, normalised
__device__ __forceinline__ Q<double> acc_gauge(const double* Gacc, const Geom& g, int p,
                                               uint32_t h) {
  const Q<double> q = ldq(site_ptr<4>(Gacc, g, p, h));
  const double n = 1.0 / sqrt(q.a0 * q.a0 + q.a1 * q.a1 + q.a2 * q.a2 + q.a3 * q.a3);
  return Q<double>{q.a0 * n, q.a1 * n, q.a2 * n, q.a3 * n};
}
// s(x) = row 3 of the rotation of G_acc(x): invariant under a diagonal U(1)
// factor on the left, the freedom the MAG functional leaves
__global__ __launch_bounds__(kBlock) void k_gauge_spins(const double* __restrict__ Gacc,
                                                        double* __restrict__ S, Geom g) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= 2 * g.Vh) return;
  const int p = (int)(i / g.Vh);
  const uint32_t h = i % g.Vh;
  st_spin(S, g, p, h, spin_of_gauge(acc_gauge(Gacc, g, p, h)));
}
// u(x) = GetGaugeMatrix(s(x)) G_acc(x)^+ (diagonal up to rounding)
__device__ __forceinline__ Q<double> u1_fix(const double* S, const double* Gacc, const Geom& g,
                                            int p, uint32_t h) {
  return qmul_dag(gauge_matrix(ld_spin(S, g, p, h)), acc_gauge(Gacc, g, p, h));
}
// out_x,mu = u(x) U_x,mu u(x+mu)^+ = G_c(x) U_in G_c(x+mu)^+; thread per site
__global__ __launch_bounds__(kBlock) void k_canonical_links(const double* __restrict__ U,
                                                            const double* __restrict__ S,
                                                            const double* __restrict__ Gacc,
                                                            double* __restrict__ out, Geom g) {
  const int nb = (int)((g.Vh + kBlock - 1) / kBlock);
  const int p = blockIdx.x / nb;
  const uint32_t h = (blockIdx.x % nb) * kBlock + threadIdx.x;
  if (h >= g.Vh) return;
  const SiteIdx s = decode(g, h, p);
  const Q<double> u = u1_fix(S, Gacc, g, p, h);
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    const Q<double> uf = u1_fix(S, Gacc, g, 1 - p, fwd_h(g, h, s, mu));
    stq(link_ptr<4>(out, g, p, mu, h), qmul_dag(qmul(u, ldq(link_ptr<4>(U, g, p, mu, h))), uf));
  }
}
void launch_gauge_spins(const double* Gacc, double* S, const Geom& g, hipStream_t st) {
  hipLaunchKernelGGL(k_gauge_spins, dim3((2 * g.Vh + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                     Gacc, S, g);
}
void launch_canonical_links(const double* U, const double* S, const double* Gacc, double* out,
                            const Geom& g, hipStream_t st) {
  hipLaunchKernelGGL(k_canonical_links, dim3(2 * nblocks(g)), dim3(kBlock), 0, st, U, S, Gacc,
                     out, g);
}

}  // namespace mag
