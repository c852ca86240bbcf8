// mag_kernels.hip -- support kernels of the MAG path (gfx950): functional,
// layout conversion between the reference AoS and the parity-split SoA,
// FP64 rebuild of the working links from the master and the accumulated site
// gauge (mixed precision), reunitarisation, plaquette, and the streaming-copy
// bandwidth reference.  The sweep kernels live in mag_sweep.hip.
#include "mag_internal.h"
#include "mag_gauge.h"
#include "mag_layout.h"
#include "mag_math.h"
#include "mag_reduce.h"

namespace mag {

// ================================================================== functional
// grid = 2*nblocks: block b covers parity b / nblocks
template <typename T>
__global__ __launch_bounds__(kBlock) void k_su2_functional(const T* __restrict__ U, Geom g,
                                                           double* __restrict__ part) {
  const int nb = (int)((g.Vh + kBlock - 1) / kBlock);
  const int p = blockIdx.x / nb;
  const uint32_t h = (blockIdx.x % nb) * kBlock + threadIdx.x;
  double fs = 0.0, z1 = 0.0, z2 = 0.0;
  if (h < g.Vh) {
#pragma unroll
    for (int mu = 0; mu < 4; mu++) fs += link_F(ldq(link_ptr<4>(U, g, p, mu, h)));
  }
  block_reduce3<kWaves>(fs, z1, z2);
  write_part(part, fs, z1, z2);
}
template <typename T>
__global__ __launch_bounds__(kBlock) void k_su3_functional(const T* __restrict__ U, Geom g,
                                                           double* __restrict__ part) {
  const int nb = (int)((g.Vh + kBlock - 1) / kBlock);
  const int p = blockIdx.x / nb;
  const uint32_t h = (blockIdx.x % nb) * kBlock + threadIdx.x;
  double fs = 0.0, z1 = 0.0, z2 = 0.0;
  if (h < g.Vh) {
#pragma unroll
    for (int mu = 0; mu < 4; mu++) fs += su3_link_F(ldm(link_ptr<12>(U, g, p, mu, h)));
  }
  block_reduce3<kWaves>(fs, z1, z2);
  write_part(part, fs, z1, z2);
}
template <typename T>
void launch_su2_functional(const T* U, const Geom& g, double* part, hipStream_t st) {
  hipLaunchKernelGGL((k_su2_functional<T>), dim3(2 * nblocks(g)), dim3(kBlock), 0, st, U, g,
                     part);
}
template <typename T>
void launch_su3_functional(const T* U, const Geom& g, double* part, hipStream_t st) {
  hipLaunchKernelGGL((k_su3_functional<T>), dim3(2 * nblocks(g)), dim3(kBlock), 0, st, U, g,
                     part);
}
template void launch_su2_functional<float>(const float*, const Geom&, double*, hipStream_t);
template void launch_su2_functional<double>(const double*, const Geom&, double*, hipStream_t);
template void launch_su3_functional<float>(const float*, const Geom&, double*, hipStream_t);
template void launch_su3_functional<double>(const double*, const Geom&, double*, hipStream_t);

// ================================================================== layout

template <typename T>
__global__ void k_su2_aos_to_soa(const double* __restrict__ aos, T* __restrict__ U, Geom g) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= 2 * g.Vh) return;
  int p;
  uint32_t h;
  lex_site(g, s, p, h);
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    const double* a = aos + ((size_t)s * 4 + mu) * 4;
    stq(link_ptr<4>(U, g, p, mu, h), Q<T>{(T)a[0], (T)a[1], (T)a[2], (T)a[3]});
  }
}
template <typename T>
__global__ void k_su2_soa_to_aos(const T* __restrict__ U, double* __restrict__ aos, Geom g) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= 2 * g.Vh) return;
  int p;
  uint32_t h;
  lex_site(g, s, p, h);
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    double* a = aos + ((size_t)s * 4 + mu) * 4;
    const Q<T> q = ldq(link_ptr<4>(U, g, p, mu, h));
    a[0] = q.a0;
    a[1] = q.a1;
    a[2] = q.a2;
    a[3] = q.a3;
  }
}
template <typename T>
__global__ void k_su3_aos_to_soa(const double* __restrict__ aos, T* __restrict__ U, Geom g) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= 2 * g.Vh) return;
  int p;
  uint32_t h;
  lex_site(g, s, p, h);
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    const double* a = aos + ((size_t)s * 4 + mu) * 18;
    M2<T> m;
#pragma unroll
    for (int k = 0; k < 6; k++) m.r[k / 3][k % 3] = {(T)a[2 * k], (T)a[2 * k + 1]};
    stm(link_ptr<12>(U, g, p, mu, h), m);
  }
}
template <typename T>
__global__ void k_su3_soa_to_aos(const T* __restrict__ U, double* __restrict__ aos, Geom g) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= 2 * g.Vh) return;
  int p;
  uint32_t h;
  lex_site(g, s, p, h);
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    double* a = aos + ((size_t)s * 4 + mu) * 18;
    const M2<T> o = ldm(link_ptr<12>(U, g, p, mu, h));
    M2<double> m;
#pragma unroll
    for (int k = 0; k < 6; k++) m.r[k / 3][k % 3] = {(double)o.r[k / 3][k % 3].re, (double)o.r[k / 3][k % 3].im};
    C<double> r2[3];
    row2_of(m, r2);
#pragma unroll
    for (int k = 0; k < 6; k++) {
      a[2 * k] = m.r[k / 3][k % 3].re;
      a[2 * k + 1] = m.r[k / 3][k % 3].im;
    }
#pragma unroll
    for (int j = 0; j < 3; j++) {
      a[12 + 2 * j] = r2[j].re;
      a[12 + 2 * j + 1] = r2[j].im;
    }
  }
}

static inline int grid_sites(const Geom& g) { return (int)((2 * g.Vh + 255) / 256); }

template <typename T>
void launch_su2_aos_to_soa(const double* aos, T* U, const Geom& g, hipStream_t st) {
  hipLaunchKernelGGL((k_su2_aos_to_soa<T>), dim3(grid_sites(g)), dim3(256), 0, st, aos, U, g);
}
template <typename T>
void launch_su2_soa_to_aos(const T* U, double* aos, const Geom& g, hipStream_t st) {
  hipLaunchKernelGGL((k_su2_soa_to_aos<T>), dim3(grid_sites(g)), dim3(256), 0, st, U, aos, g);
}
template <typename T>
void launch_su3_aos_to_soa(const double* aos, T* U, const Geom& g, hipStream_t st) {
  hipLaunchKernelGGL((k_su3_aos_to_soa<T>), dim3(grid_sites(g)), dim3(256), 0, st, aos, U, g);
}
template <typename T>
void launch_su3_soa_to_aos(const T* U, double* aos, const Geom& g, hipStream_t st) {
  hipLaunchKernelGGL((k_su3_soa_to_aos<T>), dim3(grid_sites(g)), dim3(256), 0, st, U, aos, g);
}
template void launch_su2_aos_to_soa<float>(const double*, float*, const Geom&, hipStream_t);
template void launch_su2_aos_to_soa<double>(const double*, double*, const Geom&, hipStream_t);
template void launch_su2_soa_to_aos<float>(const float*, double*, const Geom&, hipStream_t);
template void launch_su2_soa_to_aos<double>(const double*, double*, const Geom&, hipStream_t);
template void launch_su3_aos_to_soa<float>(const double*, float*, const Geom&, hipStream_t);
template void launch_su3_aos_to_soa<double>(const double*, double*, const Geom&, hipStream_t);
template void launch_su3_soa_to_aos<float>(const float*, double*, const Geom&, hipStream_t);
template void launch_su3_soa_to_aos<double>(const double*, double*, const Geom&, hipStream_t);

// FP64 links -> FP32 links, link by link (the two chunk layouts differ)
template <int NR>
__global__ void k_links_f64_to_f32(const double* __restrict__ a, float* __restrict__ b, Geom g) {
  const size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= (size_t)8 * g.Vh) return;
  const uint32_t pm = (uint32_t)(l / g.Vh), h = (uint32_t)(l % g.Vh);
  if constexpr (NR == 4) {
    const Q<double> q = ldq(link_ptr_pm<4>(a, g, pm, h));
    stq(link_ptr_pm<4>(b, g, pm, h), Q<float>{(float)q.a0, (float)q.a1, (float)q.a2, (float)q.a3});
  } else {
    const M2<double> m = ldm(link_ptr_pm<12>(a, g, pm, h));
    M2<float> f;
#pragma unroll
    for (int k = 0; k < 6; k++) f.r[k / 3][k % 3] = {(float)m.r[k / 3][k % 3].re, (float)m.r[k / 3][k % 3].im};
    stm(link_ptr_pm<12>(b, g, pm, h), f);
  }
}
void launch_links_f64_to_f32(const double* a, float* b, const Geom& g, int group, hipStream_t st) {
  const unsigned blocks = (unsigned)(((size_t)8 * g.Vh + 255) / 256);
  if (group == 2) hipLaunchKernelGGL(k_links_f64_to_f32<4>, dim3(blocks), dim3(256), 0, st, a, b, g);
  else hipLaunchKernelGGL(k_links_f64_to_f32<12>, dim3(blocks), dim3(256), 0, st, a, b, g);
}

__global__ void k_gauge_identity(float* __restrict__ G, Geom g, int group) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= 2 * g.Vh) return;
  const int p = (int)(s / g.Vh);
  const uint32_t h = s % g.Vh;
  if (group == 2) {
    stq(site_ptr<4>(G, g, p, h), Q<float>{1.f, 0.f, 0.f, 0.f});
  } else {
    M2<float> m;
#pragma unroll
    for (int k = 0; k < 6; k++) m.r[k / 3][k % 3] = {(k == 0 || k == 4) ? 1.f : 0.f, 0.f};
    stm(site_ptr<12>(G, g, p, h), m);
  }
}
void launch_gauge_identity(float* G, const Geom& g, int group, hipStream_t st) {
  const uint32_t n = 2 * g.Vh;
  hipLaunchKernelGGL(k_gauge_identity, dim3((n + 255) / 256), dim3(256), 0, st, G, g, group);
}

// ================================================================== rebuild
// End of the FP32 phase: every link in FP64 from the master, one thread per
// black site x rebuilding its 4 forward and 4 backward links (each link has
// exactly one black end, so out64 may alias U0).  Red gauges G_r(y) are
// normalised in FP64; the black gauge is recovered from W(x,0) (mag_gauge.h).
__global__ __launch_bounds__(256) void k_su2_rebuild_final(const double* U0,
                                                           const float* __restrict__ Gr,
                                                           const float* __restrict__ W,
                                                           double* out64, Geom g,
                                                           double* __restrict__ Gacc) {
  const uint32_t h = blockIdx.x * 256 + threadIdx.x;
  if (h >= g.Vh) return;
  const SiteIdx s = decode(g, h, 1);
  uint32_t hf[4], hb[4];
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    hf[mu] = fwd_h(g, h, s, mu);
    hb[mu] = back_h(g, h, s, mu);
  }
  const Q<double> gx = black_gauge(to_d(ldq(link_ptr<4>(W, g, 1, 0, h))),
                                   norm_q(site_ptr<4>(Gr, g, 0, hf[0])),
                                   ldq(link_ptr<4>(U0, g, 1, 0, h)));
  Q<double> wf[4], wb[4];
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    wf[mu] = qmul_dag(qmul(gx, ldq(link_ptr<4>(U0, g, 1, mu, h))),
                      norm_q(site_ptr<4>(Gr, g, 0, hf[mu])));
    wb[mu] = qmul_dag(qmul(norm_q(site_ptr<4>(Gr, g, 0, hb[mu])),
                           ldq(link_ptr<4>(U0, g, 0, mu, hb[mu]))), gx);
  }
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    stq(link_ptr<4>(out64, g, 1, mu, h), wf[mu]);
    stq(link_ptr<4>(out64, g, 0, mu, hb[mu]), wb[mu]);
  }
  if (Gacc) {  // gauge tracking: the FP32 phase's black gauge joins the total
    const LRef<double> r = site_ptr<4>(Gacc, g, 1, h);
    stq(r, qmul(gx, ldq(r)));
  }
}

// gauge tracking, end of the FP32 phase: red sites' total gauge
// G_acc <- normalise(G_r) G_acc (the black sites' in k_su2_rebuild_final)
__global__ __launch_bounds__(256) void k_su2_compose_red(const float* __restrict__ Gr,
                                                         double* __restrict__ Gacc, Geom g) {
  const uint32_t h = blockIdx.x * 256 + threadIdx.x;
  if (h >= g.Vh) return;
  const LRef<double> r = site_ptr<4>(Gacc, g, 0, h);
  stq(r, qmul(norm_q(site_ptr<4>(Gr, g, 0, h)), ldq(r)));
}

// every site gauge (both parities, FP64, NR = 4) set to the identity
__global__ void k_gauge_identity64(double* __restrict__ G, Geom g) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= 2 * g.Vh) return;
  stq(site_ptr<4>(G, g, (int)(s / g.Vh), s % g.Vh), Q<double>{1.0, 0.0, 0.0, 0.0});
}
void launch_gauge_identity64(double* G, const Geom& g, hipStream_t st) {
  hipLaunchKernelGGL(k_gauge_identity64, dim3((2 * g.Vh + 255) / 256), dim3(256), 0, st, G, g);
}

__global__ __launch_bounds__(256) void k_su3_rebuild_final(const double* U0,
                                                           const float* __restrict__ Gr,
                                                           const float* __restrict__ W,
                                                           double* out64, Geom g) {
  const uint32_t h = blockIdx.x * 256 + threadIdx.x;
  if (h >= g.Vh) return;
  const SiteIdx s = decode(g, h, 1);
  const M3<double> gx = gs3(black_gauge_rows(
      to_d(ldm(link_ptr<12>(W, g, 1, 0, h))), gs3(to_d(ldm(site_ptr<12>(Gr, g, 0, fwd_h(g, h, s, 0))))),
      ldm(link_ptr<12>(U0, g, 1, 0, h))));
  // every read of U0 below precedes the write of the same link (this
  // thread owns the 8 links), so the rebuild may run in place
#pragma unroll 1
  for (int mu = 0; mu < 4; mu++) {
    const uint32_t hf = fwd_h(g, h, s, mu), hb = back_h(g, h, s, mu);
    const M2<double> wf = rebuild_link(gx, ldm(link_ptr<12>(U0, g, 1, mu, h)),
                                       gs3(to_d(ldm(site_ptr<12>(Gr, g, 0, hf)))));
    const M2<double> wb = rebuild_link(gs3(to_d(ldm(site_ptr<12>(Gr, g, 0, hb)))),
                                       ldm(link_ptr<12>(U0, g, 0, mu, hb)), gx);
    stm(link_ptr<12>(out64, g, 1, mu, h), wf);
    stm(link_ptr<12>(out64, g, 0, mu, hb), wb);
  }
}

void launch_su2_rebuild_final(const double* U0, const float* Gr, const float* W, double* out64,
                              const Geom& g, hipStream_t st, double* Gacc) {
  hipLaunchKernelGGL(k_su2_rebuild_final, dim3((g.Vh + 255) / 256), dim3(256), 0, st, U0, Gr, W,
                     out64, g, Gacc);
  if (Gacc)
    hipLaunchKernelGGL(k_su2_compose_red, dim3((g.Vh + 255) / 256), dim3(256), 0, st, Gr, Gacc, g);
}
void launch_su3_rebuild_final(const double* U0, const float* Gr, const float* W, double* out64,
                              const Geom& g, hipStream_t st) {
  hipLaunchKernelGGL(k_su3_rebuild_final, dim3((g.Vh + 255) / 256), dim3(256), 0, st, U0, Gr, W,
                     out64, g);
}

__global__ void k_su2_reunit64(double* __restrict__ U, Geom g) {
  const size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= (size_t)8 * g.Vh) return;
  const LRef<double> r = link_ptr_pm<4>(U, g, (uint32_t)(l / g.Vh), (uint32_t)(l % g.Vh));
  Q<double> q = ldq(r);
  double n = 1.0 / sqrt(q.a0 * q.a0 + q.a1 * q.a1 + q.a2 * q.a2 + q.a3 * q.a3);
  stq(r, Q<double>{q.a0 * n, q.a1 * n, q.a2 * n, q.a3 * n});
}
__global__ void k_su3_reunit64(double* __restrict__ U, Geom g) {
  const size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= (size_t)8 * g.Vh) return;
  const LRef<double> r = link_ptr_pm<12>(U, g, (uint32_t)(l / g.Vh), (uint32_t)(l % g.Vh));
  M3<double> m = gs3(ldm(r));
  M2<double> o;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) o.r[i][k] = m.m[i][k];
  stm(r, o);
}
void launch_su2_reunit64(double* U, const Geom& g, hipStream_t st) {
  size_t n = (size_t)8 * g.Vh;
  hipLaunchKernelGGL(k_su2_reunit64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, U, g);
}
void launch_su3_reunit64(double* U, const Geom& g, hipStream_t st) {
  size_t n = (size_t)8 * g.Vh;
  hipLaunchKernelGGL(k_su3_reunit64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, U, g);
}

// ================================================================== plaquette
// plaket_site (plaket.h:24-42): sum over mu<nu of
// multiply_conj_tr(U_mu(x) U_nu(x+mu) U_mu(x+nu)^+, U_nu(x)), / 6
__global__ __launch_bounds__(kBlock) void k_su2_plaquette(const double* __restrict__ U, Geom g,
                                                          double* __restrict__ part) {
  const int nb = (int)((g.Vh + kBlock - 1) / kBlock);
  const int p = blockIdx.x / nb;
  const uint32_t h = (blockIdx.x % nb) * kBlock + threadIdx.x;
  double acc = 0.0, z1 = 0.0, z2 = 0.0;
  if (h < g.Vh) {
    SiteIdx s = decode(g, h, p);
    uint32_t hf[4];
    for (int mu = 0; mu < 4; mu++) hf[mu] = fwd_h(g, h, s, mu);
    for (int mu = 0; mu < 3; mu++)
      for (int nu = mu + 1; nu < 4; nu++) {
        Q<double> a = ldq(link_ptr<4>(U, g, p, mu, h));
        a = qmul(a, ldq(link_ptr<4>(U, g, 1 - p, nu, hf[mu])));
        a = qmul_dag(a, ldq(link_ptr<4>(U, g, 1 - p, mu, hf[nu])));
        Q<double> b = ldq(link_ptr<4>(U, g, p, nu, h));
        acc += a.a0 * b.a0 + a.a1 * b.a1 + a.a2 * b.a2 + a.a3 * b.a3;
      }
    acc /= 6.0;
  }
  block_reduce3<kWaves>(acc, z1, z2);
  write_part(part, acc, z1, z2);
}

__global__ __launch_bounds__(kBlock) void k_su3_plaquette(const double* __restrict__ U, Geom g,
                                                          double* __restrict__ part) {
  const int nb = (int)((g.Vh + kBlock - 1) / kBlock);
  const int p = blockIdx.x / nb;
  const uint32_t h = (blockIdx.x % nb) * kBlock + threadIdx.x;
  double acc = 0.0, z1 = 0.0, z2 = 0.0;
  if (h < g.Vh) {
    SiteIdx s = decode(g, h, p);
    uint32_t hf[4];
    for (int mu = 0; mu < 4; mu++) hf[mu] = fwd_h(g, h, s, mu);
#pragma unroll 1
    for (int mu = 0; mu < 3; mu++)
#pragma unroll 1
      for (int nu = mu + 1; nu < 4; nu++) {
        M3<double> a = full3(ldm(link_ptr<12>(U, g, p, mu, h)));
        M3<double> b = full3(ldm(link_ptr<12>(U, g, 1 - p, nu, hf[mu])));
        M3<double> c = full3(ldm(link_ptr<12>(U, g, 1 - p, mu, hf[nu])));
        M3<double> d = full3(ldm(link_ptr<12>(U, g, p, nu, h)));
        // Re tr(a b c^+ d^+) / 3
        M3<double> ab, dc;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
          for (int j = 0; j < 3; j++) {
            C<double> x{0.0, 0.0}, y{0.0, 0.0};
#pragma unroll
            for (int k = 0; k < 3; k++) {
              x = cadd(x, cmul(a.m[i][k], b.m[k][j]));
              y = cadd(y, cmul(d.m[i][k], c.m[k][j]));
            }
            ab.m[i][j] = x;
            dc.m[i][j] = y;
          }
        // tr(ab (dc)^+) = sum_ij ab_ij conj(dc_ij)
        double tr = 0.0;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
          for (int j = 0; j < 3; j++) tr += cmulc(ab.m[i][j], dc.m[i][j]).re;
        acc += tr / 3.0;
      }
    acc /= 6.0;
  }
  block_reduce3<kWaves>(acc, z1, z2);
  write_part(part, acc, z1, z2);
}

void launch_su2_plaquette(const double* U, const Geom& g, double* part, hipStream_t st) {
  hipLaunchKernelGGL(k_su2_plaquette, dim3(2 * nblocks(g)), dim3(kBlock), 0, st, U, g, part);
}
void launch_su3_plaquette(const double* U, const Geom& g, double* part, hipStream_t st) {
  hipLaunchKernelGGL(k_su3_plaquette, dim3(2 * nblocks(g)), dim3(kBlock), 0, st, U, g, part);
}

}  // namespace mag

namespace mag {
// streaming copy (16-byte vectors, grid-stride): the bandwidth reference
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ a, float4* __restrict__ b,
                                              size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = a[i];
}
void launch_copy(const void* a, void* b, size_t bytes, hipStream_t st) {
  const size_t n = bytes / 16;
  size_t blocks = (n + 255) / 256;
  if (blocks > 1u << 20) blocks = 1u << 20;
  hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, st, (const float4*)a,
                     (float4*)b, n);
}
// ================================================================== SU(3) site form
// out = G(x) U0 G(x+mu)^+ (FP64) from FP64 site gauges: the links of a copy
// (one thread per site of either parity, its 4 forward links)
__global__ __launch_bounds__(256) void k_su3_site_links(const double* __restrict__ U0,
                                                        const double* __restrict__ G,
                                                        double* __restrict__ out, Geom g) {
  const uint32_t nb = (g.Vh + 255) / 256;
  const int p = (int)(blockIdx.x / nb);
  const uint32_t h = (blockIdx.x % nb) * 256 + threadIdx.x;
  if (h >= g.Vh) return;
  const SiteIdx s = decode(g, h, p);
  const M3<double> gx = gs3(ldm(site_ptr<12>(G, g, p, h)));
#pragma unroll 1
  for (int mu = 0; mu < 4; mu++) {
    const M3<double> gf = gs3(ldm(site_ptr<12>(G, g, 1 - p, fwd_h(g, h, s, mu))));
    const M3<double> u = full3(ldm(link_ptr<12>(U0, g, p, mu, h)));
    // V = u gf^+, then rows 0,1 of gx V
    M3<double> v;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        C<double> acc{0.0, 0.0};
#pragma unroll
        for (int k = 0; k < 3; k++) acc = cadd(acc, cmulc(u.m[i][k], gf.m[j][k]));
        v.m[i][j] = acc;
      }
    M2<double> w;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        C<double> acc{0.0, 0.0};
#pragma unroll
        for (int k = 0; k < 3; k++) acc = cadd(acc, cmul(gx.m[i][k], v.m[k][j]));
        w.r[i][j] = acc;
      }
    stm(link_ptr<12>(out, g, p, mu, h), w);
  }
}
void launch_su3_site_links(const double* U0, const double* G, double* out, const Geom& g,
                           hipStream_t st) {
  hipLaunchKernelGGL(k_su3_site_links, dim3(2 * ((g.Vh + 255) / 256)), dim3(256), 0, st, U0, G,
                     out, g);
}

// site gauges (2 rows, chunked site layout): FP64 -> FP32, and FP32 -> FP64
// with the rows put back on SU(3) in FP64 (Gram-Schmidt, row 2 implied)
__global__ void k_su3_sites_to32(const double* __restrict__ a, float* __restrict__ b, Geom g) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= 2 * g.Vh) return;
  const int p = (int)(s / g.Vh);
  const uint32_t h = s % g.Vh;
  const M2<double> m = ldm(site_ptr<12>(a, g, p, h));
  M2<float> f;
#pragma unroll
  for (int k = 0; k < 6; k++) f.r[k / 3][k % 3] = {(float)m.r[k / 3][k % 3].re, (float)m.r[k / 3][k % 3].im};
  stm(site_ptr<12>(b, g, p, h), f);
}
__global__ void k_su3_sites_to64(const float* __restrict__ a, double* __restrict__ b, Geom g) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= 2 * g.Vh) return;
  const int p = (int)(s / g.Vh);
  const uint32_t h = s % g.Vh;
  const M3<double> m = gs3(to_d(ldm(site_ptr<12>(a, g, p, h))));
  M2<double> r;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) r.r[i][k] = m.m[i][k];
  stm(site_ptr<12>(b, g, p, h), r);
}
__global__ void k_su3_sites_identity(double* __restrict__ G, Geom g) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= 2 * g.Vh) return;
  M2<double> m;
#pragma unroll
  for (int k = 0; k < 6; k++) m.r[k / 3][k % 3] = {(k == 0 || k == 4) ? 1.0 : 0.0, 0.0};
  stm(site_ptr<12>(G, g, (int)(s / g.Vh), s % g.Vh), m);
}
void launch_su3_sites_to32(const double* a, float* b, const Geom& g, hipStream_t st) {
  hipLaunchKernelGGL(k_su3_sites_to32, dim3((2 * g.Vh + 255) / 256), dim3(256), 0, st, a, b, g);
}
void launch_su3_sites_to64(const float* a, double* b, const Geom& g, hipStream_t st) {
  hipLaunchKernelGGL(k_su3_sites_to64, dim3((2 * g.Vh + 255) / 256), dim3(256), 0, st, a, b, g);
}
void launch_su3_sites_identity(double* G, const Geom& g, hipStream_t st) {
  hipLaunchKernelGGL(k_su3_sites_identity, dim3((2 * g.Vh + 255) / 256), dim3(256), 0, st, G, g);
}

}  // namespace mag
