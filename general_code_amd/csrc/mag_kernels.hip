// mag_kernels.hip -- CDNA4 (gfx950) kernels of the MAG gauge-fixing path.
//
// Link layout in HBM ("parity-split SoA"): links of base parity p and
// direction mu form one array of Vh = V/2 links; link (p, mu, h) starts at
//   U + ((p*4 + mu)*Vh + h) * NR          NR = 4 (su2) or 12 (su3 rows 0,1)
// so the 64 lanes of a wave, which own 64 consecutive half indices h of the
// active parity, read and write 64 consecutive links of every (p, mu) array:
// every gather of the 8 links of a site is a contiguous, coalesced access.
// A checkerboard half-sweep reads and writes each link exactly once (every
// link has exactly one endpoint of the active parity), so there is no reuse
// for LDS to capture; the kernels stream links HBM -> VGPR -> HBM.
//
// The site update (build-defined; the reference has only spin-form omega=1
// and omega=2 updates, lib/cpu/src/mag.cpp:1375-1512):
//   h     = sum_mu [ M(U_x,mu) e3 + M(U_x-mu,mu)^T e3 ]      (mag.cpp:569-626
//           with every spin equal to e3, i.e. links already transformed)
//   g     = geodesic omega step of e3 towards h_hat           (mag_math.h)
//   U_x,mu <- g U_x,mu ;  U_x-mu,mu <- U_x-mu,mu g^+
// SU(3): the same update for the Cabibbo-Marinari subgroups (0,1),(1,2),(0,2).
#include "mag_internal.h"
#include "mag_math.h"
#include "mag_reduce.h"

namespace mag {

// ------------------------------------------------------------------ loads
__device__ __forceinline__ Q<float> ldq(const float* p) {
  float4 v = *reinterpret_cast<const float4*>(p);
  return Q<float>{v.x, v.y, v.z, v.w};
}
__device__ __forceinline__ Q<double> ldq(const double* p) {
  double2 a = *reinterpret_cast<const double2*>(p);
  double2 b = *reinterpret_cast<const double2*>(p + 2);
  return Q<double>{a.x, a.y, b.x, b.y};
}
__device__ __forceinline__ void stq(float* p, const Q<float>& q) {
  *reinterpret_cast<float4*>(p) = make_float4(q.a0, q.a1, q.a2, q.a3);
}
__device__ __forceinline__ void stq(double* p, const Q<double>& q) {
  *reinterpret_cast<double2*>(p) = make_double2(q.a0, q.a1);
  *reinterpret_cast<double2*>(p + 2) = make_double2(q.a2, q.a3);
}

template <typename T>
__device__ __forceinline__ M2<T> ldm(const T* p) {
  M2<T> m;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) m.r[i][j] = C<T>{p[6 * i + 2 * j], p[6 * i + 2 * j + 1]};
  return m;
}
template <>
__device__ __forceinline__ M2<float> ldm(const float* p) {
  const float4* v = reinterpret_cast<const float4*>(p);
  float4 a = v[0], b = v[1], c = v[2];
  M2<float> m;
  m.r[0][0] = {a.x, a.y}; m.r[0][1] = {a.z, a.w}; m.r[0][2] = {b.x, b.y};
  m.r[1][0] = {b.z, b.w}; m.r[1][1] = {c.x, c.y}; m.r[1][2] = {c.z, c.w};
  return m;
}
template <>
__device__ __forceinline__ M2<double> ldm(const double* p) {
  const double2* v = reinterpret_cast<const double2*>(p);
  M2<double> m;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double2 d = v[3 * i + j];
      m.r[i][j] = {d.x, d.y};
    }
  return m;
}
__device__ __forceinline__ void stm(float* p, const M2<float>& m) {
  float4* v = reinterpret_cast<float4*>(p);
  v[0] = make_float4(m.r[0][0].re, m.r[0][0].im, m.r[0][1].re, m.r[0][1].im);
  v[1] = make_float4(m.r[0][2].re, m.r[0][2].im, m.r[1][0].re, m.r[1][0].im);
  v[2] = make_float4(m.r[1][1].re, m.r[1][1].im, m.r[1][2].re, m.r[1][2].im);
}
__device__ __forceinline__ void stm(double* p, const M2<double>& m) {
  double2* v = reinterpret_cast<double2*>(p);
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) v[3 * i + j] = make_double2(m.r[i][j].re, m.r[i][j].im);
}

// ------------------------------------------------------------------ geometry
struct SiteIdx {
  uint32_t x, y, z, t, row;
};

__device__ __forceinline__ SiteIdx decode(const Geom& g, uint32_t h, int p) {
  SiteIdx s;
  uint32_t xh = h % g.hx;
  s.row = h / g.hx;
  uint32_t r = s.row;
  s.y = r % (uint32_t)g.L1;
  r /= (uint32_t)g.L1;
  s.z = r % (uint32_t)g.L2;
  s.t = r / (uint32_t)g.L2;
  s.x = 2 * xh + ((uint32_t)(p + s.y + s.z + s.t) & 1u);
  return s;
}

// half index of x - mu (opposite parity)
__device__ __forceinline__ uint32_t back_h(const Geom& g, uint32_t h, const SiteIdx& s, int mu) {
  switch (mu) {
    case 0: {
      uint32_t xb = (s.x == 0) ? (uint32_t)g.L0 - 1 : s.x - 1;
      return s.row * g.hx + (xb >> 1);
    }
    case 1: return (s.y == 0) ? h + ((uint32_t)g.L1 - 1) * g.s1 : h - g.s1;
    case 2: return (s.z == 0) ? h + ((uint32_t)g.L2 - 1) * g.s2 : h - g.s2;
    default: return (s.t == 0) ? h + ((uint32_t)g.L3 - 1) * g.s3 : h - g.s3;
  }
}
// half index of x + mu (opposite parity)
__device__ __forceinline__ uint32_t fwd_h(const Geom& g, uint32_t h, const SiteIdx& s, int mu) {
  switch (mu) {
    case 0: {
      uint32_t xf = (s.x + 1 == (uint32_t)g.L0) ? 0 : s.x + 1;
      return s.row * g.hx + (xf >> 1);
    }
    case 1: return (s.y + 1 == (uint32_t)g.L1) ? h - ((uint32_t)g.L1 - 1) * g.s1 : h + g.s1;
    case 2: return (s.z + 1 == (uint32_t)g.L2) ? h - ((uint32_t)g.L2 - 1) * g.s2 : h + g.s2;
    default: return (s.t + 1 == (uint32_t)g.L3) ? h - ((uint32_t)g.L3 - 1) * g.s3 : h + g.s3;
  }
}

template <int NR, typename T>
__device__ __forceinline__ T* link_ptr(T* U, const Geom& g, int p, int mu, uint32_t h) {
  return U + ((size_t)(p * 4 + mu) * g.Vh + h) * NR;
}
template <int NR, typename T>
__device__ __forceinline__ const T* link_ptr(const T* U, const Geom& g, int p, int mu, uint32_t h) {
  return U + ((size_t)(p * 4 + mu) * g.Vh + h) * NR;
}
template <int NR, typename T>
__device__ __forceinline__ T* site_ptr(T* G, const Geom& g, int p, uint32_t h) {
  return G + ((size_t)p * g.Vh + h) * NR;
}

__device__ __forceinline__ void write_part(double* part, double a, double b, double c) {
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x + 0] = a;
    part[3 * blockIdx.x + 1] = b;
    part[3 * blockIdx.x + 2] = c;
  }
}

// ================================================================== reduction
// Reduce nparts partial triples (F sum, residual sum, residual max) and
// apply the result to the job's control block; called by every thread of
// one workgroup.  Thread t combines partials t, t + kBlock, ... in order,
// then the fixed-order workgroup tree: the result depends only on the job.
__device__ void reduce_partials(const double* part, int nparts, int mode, double invF,
                                double invV, JobCtl* ctl, double* hF, double* hR, double* hA,
                                double* out3) {
  double a = 0.0, b = 0.0, c = 0.0;
  for (int i = threadIdx.x; i < nparts; i += kBlock) {
    a += part[3 * i];
    b += part[3 * i + 1];
    c = fmax(c, part[3 * i + 2]);
  }
  block_reduce3<kWaves>(a, b, c);
  if (threadIdx.x == 0) {
    const double F = a * invF;
    if (mode == RED_INIT) {
      ctl->F_prev = F;
      ctl->F0 = F;
    } else if (mode == RED_RESET) {
      ctl->F_prev = F;
    } else if (mode == RED_SWEEP) {
      const int sw = ctl->sweep;
      hF[sw] = F;
      hR[sw] = c;
      hA[sw] = b * invV;
      const double rel = fabs(F - ctl->F_prev) / fabs(F);
      ctl->F_prev = F;
      ctl->sweep = sw + 1;
      if (rel < ctl->tol) {
        ctl->converged = 1;
        ctl->done = 1;
      }
      if (sw + 1 >= ctl->max_sweeps) ctl->done = 1;
    }
    if (out3) {
      out3[0] = F;
      out3[1] = b * invV;
      out3[2] = c;
    }
  }
}

// End of a black half-sweep with the fused reduction: every workgroup
// publishes its partial (device-scope release) and bumps the arrival
// counter; the last one to arrive (acquire) reduces the sweep.
__device__ __forceinline__ void fused_sweep_epilogue(const SweepRed& r) {
  __shared__ int is_last;
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned prev = atomicAdd(&r.ctl->arrived, 1u);
    is_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (is_last) {
    __threadfence();
    reduce_partials(r.part, r.nparts, RED_SWEEP, r.invF, r.invV, r.ctl, r.hF, r.hR, r.hA,
                    nullptr);
    if (threadIdx.x == 0) r.ctl->arrived = 0u;
  }
}

// ================================================================== SU(2)
// One checkerboard half-sweep over the sites of parity p.  The black
// (p = 1) half also sums the MAG functional of the 8 links each site has
// just written: every link is written exactly once in that half, so the
// sum is the functional after the sweep with no extra pass.
template <typename T, bool GA, bool FUSED>
__global__ __launch_bounds__(kBlock) void k_su2_half(T* __restrict__ U, T* __restrict__ G,
                                                     Geom g, int p, T omega, SweepRed r) {
  if (r.ctl->done) return;
  double* part = r.part + (size_t)3 * gridDim.x * p;
  const uint32_t h = blockIdx.x * kBlock + threadIdx.x;
  double fs = 0.0, rs = 0.0, rm = 0.0;
  if (h < g.Vh) {
    SiteIdx s = decode(g, h, p);
    uint32_t hb[4];
    Q<T> F[4], B[4];
#pragma unroll
    for (int mu = 0; mu < 4; mu++) {
      hb[mu] = back_h(g, h, s, mu);
      F[mu] = ldq(link_ptr<4>(U, g, p, mu, h));
      B[mu] = ldq(link_ptr<4>(U, g, 1 - p, mu, hb[mu]));
    }
    T h0 = T(0), h1 = T(0), h2 = T(0);
#pragma unroll
    for (int mu = 0; mu < 4; mu++) {
      field_fwd(F[mu], h0, h1, h2);
      field_bwd(B[mu], h0, h1, h2);
    }
    T res;
    Q<T> gq = omega_gauge(h0, h1, h2, omega, res);
#pragma unroll
    for (int mu = 0; mu < 4; mu++) {
      F[mu] = gmul(gq, F[mu]);
      B[mu] = mul_gdag(B[mu], gq);
      stq(link_ptr<4>(U, g, p, mu, h), F[mu]);
      stq(link_ptr<4>(U, g, 1 - p, mu, hb[mu]), B[mu]);
    }
    if (GA) {
      T* gp = site_ptr<4>(G, g, p, h);
      stq(gp, gmul(gq, ldq(gp)));
    }
    if (p == 1) {
#pragma unroll
      for (int mu = 0; mu < 4; mu++) fs += link_F(F[mu]) + link_F(B[mu]);
    }
    rs = (double)res;
    rm = (double)res;
  }
  block_reduce3<kWaves>(fs, rs, rm);
  write_part(part, fs, rs, rm);
  if (FUSED && p == 1) fused_sweep_epilogue(r);
}

template <typename T>
void launch_su2_half(T* U, T* G, const Geom& g, int parity, double omega, const SweepRed& r,
                     bool fused, hipStream_t st) {
  const int nb = nblocks(g);
#define L_(GA, FU)                                                                        \
  hipLaunchKernelGGL((k_su2_half<T, GA, FU>), dim3(nb), dim3(kBlock), 0, st, U, G, g, parity, \
                     (T)omega, r)
  if (G) {
    if (fused) L_(true, true); else L_(true, false);
  } else {
    if (fused) L_(false, true); else L_(false, false);
  }
#undef L_
}
template void launch_su2_half<float>(float*, float*, const Geom&, int, double, const SweepRed&,
                                     bool, hipStream_t);
template void launch_su2_half<double>(double*, double*, const Geom&, int, double,
                                      const SweepRed&, bool, hipStream_t);

// ================================================================== SU(3)
// Cabibbo-Marinari subgroup (I,J) update of the 8 links of one site, links
// held as rows 0,1; row 2 is rebuilt where a subgroup touches it.
template <typename T, int I, int J, bool GA>
__device__ __forceinline__ T cm_subgroup(M2<T>* F, M2<T>* B, M2<T>& Gx, T omega) {
  T h0 = T(0), h1 = T(0), h2 = T(0);
  C<T> Fr2[4][3], Br2[4][3];
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    if (J == 2) {
      row2_of(F[mu], Fr2[mu]);
      row2_of(B[mu], Br2[mu]);
    }
    // element (a,b) of the (rebuilt) 3x3 matrix
    auto fe = [&](int a, int b) -> C<T> { return a < 2 ? F[mu].r[a][b] : Fr2[mu][b]; };
    auto be = [&](int a, int b) -> C<T> { return a < 2 ? B[mu].r[a][b] : Br2[mu][b]; };
    blk_fwd(fe(I, I), fe(I, J), fe(J, I), fe(J, J), h0, h1, h2);
    blk_bwd(be(I, I), be(I, J), be(J, I), be(J, J), h0, h1, h2);
  }
  T res;
  Q<T> gq = omega_gauge(h0, h1, h2, omega, res);
  // r = [[g0 + i g3, g2 + i g1], [-g2 + i g1, g0 - i g3]], g3 = 0
  const C<T> r00{gq.a0, T(0)}, r01{gq.a2, gq.a1}, r10{-gq.a2, gq.a1}, r11{gq.a0, T(0)};
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    // forward: rows I,J <- r (rows I,J); only stored rows are written
#pragma unroll
    for (int k = 0; k < 3; k++) {
      C<T> a = F[mu].r[I][k];
      C<T> b = (J < 2) ? F[mu].r[J][k] : Fr2[mu][k];
      F[mu].r[I][k] = cadd(cmul(r00, a), cmul(r01, b));
      if (J < 2) F[mu].r[J][k] = cadd(cmul(r10, a), cmul(r11, b));
    }
    // backward: cols I,J of rows 0,1 <- (cols I,J) r^+
#pragma unroll
    for (int k = 0; k < 2; k++) {
      C<T> a = B[mu].r[k][I], b = B[mu].r[k][J];
      B[mu].r[k][I] = cadd(cmulc(a, r00), cmulc(b, r01));
      B[mu].r[k][J] = cadd(cmulc(a, r10), cmulc(b, r11));
    }
  }
  if (GA) {
    C<T> g2[3];
    if (J == 2) row2_of(Gx, g2);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      C<T> a = Gx.r[I][k];
      C<T> b = (J < 2) ? Gx.r[J][k] : g2[k];
      Gx.r[I][k] = cadd(cmul(r00, a), cmul(r01, b));
      if (J < 2) Gx.r[J][k] = cadd(cmul(r10, a), cmul(r11, b));
    }
  }
  return res;
}

template <typename T>
__device__ __forceinline__ double su3_link_F(const M2<T>& m) {
  // sum_j |U_jj|^2 (mag.cpp:1837-1848); U_22 = conj(U00 U11 - U01 U10)
  C<double> a{(double)m.r[0][0].re, (double)m.r[0][0].im};
  C<double> b{(double)m.r[1][1].re, (double)m.r[1][1].im};
  C<double> c{(double)m.r[0][1].re, (double)m.r[0][1].im};
  C<double> d{(double)m.r[1][0].re, (double)m.r[1][0].im};
  C<double> u22 = csub(cmul(a, b), cmul(c, d));
  return cnorm2(a) + cnorm2(b) + cnorm2(u22);
}

template <typename T, bool GA, bool FUSED>
__global__ __launch_bounds__(kBlock) void k_su3_half(T* __restrict__ U, T* __restrict__ G,
                                                     Geom g, int p, T omega, SweepRed r) {
  if (r.ctl->done) return;
  double* part = r.part + (size_t)3 * gridDim.x * p;
  const uint32_t h = blockIdx.x * kBlock + threadIdx.x;
  double fs = 0.0, rs = 0.0, rm = 0.0;
  if (h < g.Vh) {
    SiteIdx s = decode(g, h, p);
    uint32_t hb[4];
    M2<T> F[4], B[4];
#pragma unroll
    for (int mu = 0; mu < 4; mu++) {
      hb[mu] = back_h(g, h, s, mu);
      F[mu] = ldm<T>(link_ptr<12>(U, g, p, mu, h));
      B[mu] = ldm<T>(link_ptr<12>(U, g, 1 - p, mu, hb[mu]));
    }
    M2<T> Gx;
    if (GA) Gx = ldm<T>(site_ptr<12>(G, g, p, h));
    T res = cm_subgroup<T, 0, 1, GA>(F, B, Gx, omega);
    res += cm_subgroup<T, 1, 2, GA>(F, B, Gx, omega);
    res += cm_subgroup<T, 0, 2, GA>(F, B, Gx, omega);
#pragma unroll
    for (int mu = 0; mu < 4; mu++) {
      stm(link_ptr<12>(U, g, p, mu, h), F[mu]);
      stm(link_ptr<12>(U, g, 1 - p, mu, hb[mu]), B[mu]);
    }
    if (GA) stm(site_ptr<12>(G, g, p, h), Gx);
    if (p == 1) {
#pragma unroll
      for (int mu = 0; mu < 4; mu++) fs += su3_link_F(F[mu]) + su3_link_F(B[mu]);
    }
    rs = (double)res;
    rm = (double)res;
  }
  block_reduce3<kWaves>(fs, rs, rm);
  write_part(part, fs, rs, rm);
  if (FUSED && p == 1) fused_sweep_epilogue(r);
}

template <typename T>
void launch_su3_half(T* U, T* G, const Geom& g, int parity, double omega, const SweepRed& r,
                     bool fused, hipStream_t st) {
  const int nb = nblocks(g);
#define L_(GA, FU)                                                                        \
  hipLaunchKernelGGL((k_su3_half<T, GA, FU>), dim3(nb), dim3(kBlock), 0, st, U, G, g, parity, \
                     (T)omega, r)
  if (G) {
    if (fused) L_(true, true); else L_(true, false);
  } else {
    if (fused) L_(false, true); else L_(false, false);
  }
#undef L_
}
template void launch_su3_half<float>(float*, float*, const Geom&, int, double, const SweepRed&,
                                     bool, hipStream_t);
template void launch_su3_half<double>(double*, double*, const Geom&, int, double,
                                      const SweepRed&, bool, hipStream_t);

// one-workgroup reduction kernel (initial functional, unfused sweeps, checks)
__global__ __launch_bounds__(kBlock) void k_reduce(const double* __restrict__ part, int nparts,
                                                   int mode, double invF, double invV,
                                                   JobCtl* ctl, double* hF, double* hR,
                                                   double* hA, double* out3) {
  if (mode == RED_SWEEP && ctl->done) return;
  reduce_partials(part, nparts, mode, invF, invV, ctl, hF, hR, hA, out3);
}

void launch_reduce(const double* part, int nparts, int mode, double invF, double invV,
                   JobCtl* ctl, double* hF, double* hR, double* hA, double* out3,
                   hipStream_t st) {
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kBlock), 0, st, part, nparts, mode, invF, invV,
                     ctl, hF, hR, hA, out3);
}

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
    for (int mu = 0; mu < 4; mu++) fs += su3_link_F(ldm<T>(link_ptr<12>(U, g, p, mu, h)));
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
// thread per lexicographic site s (reference order, indexing.h:111-123)
__device__ __forceinline__ void lex_site(const Geom& g, uint32_t s, int& p, uint32_t& h) {
  uint32_t x = s % (uint32_t)g.L0, r = s / (uint32_t)g.L0;
  uint32_t y = r % (uint32_t)g.L1;
  r /= (uint32_t)g.L1;
  uint32_t z = r % (uint32_t)g.L2, t = r / (uint32_t)g.L2;
  p = (int)((x + y + z + t) & 1u);
  h = s >> 1;
}

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
    T* o = link_ptr<4>(U, g, p, mu, h);
#pragma unroll
    for (int c = 0; c < 4; c++) o[c] = (T)a[c];
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
    const T* o = link_ptr<4>(U, g, p, mu, h);
#pragma unroll
    for (int c = 0; c < 4; c++) a[c] = (double)o[c];
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
    T* o = link_ptr<12>(U, g, p, mu, h);
#pragma unroll
    for (int c = 0; c < 12; c++) o[c] = (T)a[c];
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
    const T* o = link_ptr<12>(U, g, p, mu, h);
    M2<double> m;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) m.r[i][j] = {(double)o[6 * i + 2 * j], (double)o[6 * i + 2 * j + 1]};
    C<double> r2[3];
    row2_of(m, r2);
#pragma unroll
    for (int c = 0; c < 12; c++) a[c] = (double)o[c];
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

__global__ void k_f64_to_f32(const double* __restrict__ a, float* __restrict__ b, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = (float)a[i];
}
void launch_f64_to_f32(const double* a, float* b, size_t n, hipStream_t st) {
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_f64_to_f32, dim3((unsigned)blocks), dim3(256), 0, st, a, b, n);
}

__global__ void k_gauge_identity(float* __restrict__ G, uint32_t nsites, int nr) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsites) return;
  float* o = G + (size_t)s * nr;
  for (int c = 0; c < nr; c++) o[c] = 0.f;
  if (nr == 4) {
    o[0] = 1.f;
  } else {
    o[0] = 1.f;  // row 0 = (1,0,0)
    o[8] = 1.f;  // row 1 = (0,1,0): element (1,1) real part at 6 + 2
  }
}
void launch_gauge_identity(float* G, const Geom& g, int group, hipStream_t st) {
  const uint32_t n = 2 * g.Vh;
  hipLaunchKernelGGL(k_gauge_identity, dim3((n + 255) / 256), dim3(256), 0, st, G, n,
                     group == 2 ? 4 : 12);
}

// ================================================================== rebuild
__device__ __forceinline__ Q<double> norm_q(const float* p) {
  Q<double> q{(double)p[0], (double)p[1], (double)p[2], (double)p[3]};
  double n = 1.0 / sqrt(q.a0 * q.a0 + q.a1 * q.a1 + q.a2 * q.a2 + q.a3 * q.a3);
  q.a0 *= n; q.a1 *= n; q.a2 *= n; q.a3 *= n;
  return q;
}

// full 3x3 from rows 0,1 after FP64 Gram-Schmidt
__device__ __forceinline__ M3<double> gs3(const M2<double>& in) {
  M2<double> m = in;
  double n0 = 0.0;
#pragma unroll
  for (int k = 0; k < 3; k++) n0 += cnorm2(m.r[0][k]);
  n0 = 1.0 / sqrt(n0);
#pragma unroll
  for (int k = 0; k < 3; k++) m.r[0][k] = {m.r[0][k].re * n0, m.r[0][k].im * n0};
  C<double> d{0.0, 0.0};
#pragma unroll
  for (int k = 0; k < 3; k++) d = cadd(d, cmulc(m.r[1][k], m.r[0][k]));
#pragma unroll
  for (int k = 0; k < 3; k++) m.r[1][k] = csub(m.r[1][k], cmul(d, m.r[0][k]));
  double n1 = 0.0;
#pragma unroll
  for (int k = 0; k < 3; k++) n1 += cnorm2(m.r[1][k]);
  n1 = 1.0 / sqrt(n1);
#pragma unroll
  for (int k = 0; k < 3; k++) m.r[1][k] = {m.r[1][k].re * n1, m.r[1][k].im * n1};
  M3<double> M;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) M.m[i][k] = m.r[i][k];
  row2_of(m, M.m[2]);
  return M;
}
__device__ __forceinline__ M2<double> to_d(const M2<float>& m) {
  M2<double> r;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) r.r[i][k] = {(double)m.r[i][k].re, (double)m.r[i][k].im};
  return r;
}

__global__ __launch_bounds__(256) void k_su2_rebuild(const double* __restrict__ U0,
                                                     const float* __restrict__ G,
                                                     float* __restrict__ out32,
                                                     double* out64, Geom g) {
  const int nb = (int)((g.Vh + 255) / 256);
  const int p = blockIdx.x / nb;
  const uint32_t h = (blockIdx.x % nb) * 256 + threadIdx.x;
  if (h >= g.Vh) return;
  SiteIdx s = decode(g, h, p);
  Q<double> gx = norm_q(G + ((size_t)p * g.Vh + h) * 4);
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    uint32_t hf = fwd_h(g, h, s, mu);
    Q<double> gf = norm_q(G + ((size_t)(1 - p) * g.Vh + hf) * 4);
    Q<double> u = ldq(link_ptr<4>(U0, g, p, mu, h));
    Q<double> w = qmul_dag(qmul(gx, u), gf);
    if (out32) stq(link_ptr<4>(out32, g, p, mu, h), Q<float>{(float)w.a0, (float)w.a1, (float)w.a2, (float)w.a3});
    else stq(link_ptr<4>(out64, g, p, mu, h), w);
  }
}

__device__ __forceinline__ M3<double> full3(const M2<double>& m) {
  M3<double> M;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) M.m[i][k] = m.r[i][k];
  row2_of(m, M.m[2]);
  return M;
}

__global__ __launch_bounds__(256) void k_su3_rebuild(const double* __restrict__ U0,
                                                     const float* __restrict__ G,
                                                     float* __restrict__ out32,
                                                     double* out64, Geom g) {
  const int nb = (int)((g.Vh + 255) / 256);
  const int p = blockIdx.x / nb;
  const uint32_t h = (blockIdx.x % nb) * 256 + threadIdx.x;
  if (h >= g.Vh) return;
  SiteIdx s = decode(g, h, p);
  M3<double> gx = gs3(to_d(ldm<float>(G + ((size_t)p * g.Vh + h) * 12)));
#pragma unroll 1
  for (int mu = 0; mu < 4; mu++) {
    uint32_t hf = fwd_h(g, h, s, mu);
    M3<double> gf = gs3(to_d(ldm<float>(G + ((size_t)(1 - p) * g.Vh + hf) * 12)));
    M3<double> u = full3(ldm<double>(link_ptr<12>(U0, g, p, mu, h)));
    // V = u gf^+
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
    if (out32) {
      M2<float> wf;
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) wf.r[i][j] = {(float)w.r[i][j].re, (float)w.r[i][j].im};
      stm(link_ptr<12>(out32, g, p, mu, h), wf);
    } else {
      stm(link_ptr<12>(out64, g, p, mu, h), w);
    }
  }
}

void launch_su2_rebuild(const double* U0, const float* G, float* out32, double* out64,
                        const Geom& g, hipStream_t st) {
  const int nb = (int)((g.Vh + 255) / 256);
  hipLaunchKernelGGL(k_su2_rebuild, dim3(2 * nb), dim3(256), 0, st, U0, G, out32, out64, g);
}
void launch_su3_rebuild(const double* U0, const float* G, float* out32, double* out64,
                        const Geom& g, hipStream_t st) {
  const int nb = (int)((g.Vh + 255) / 256);
  hipLaunchKernelGGL(k_su3_rebuild, dim3(2 * nb), dim3(256), 0, st, U0, G, out32, out64, g);
}

__global__ void k_gauge_normalize(float* __restrict__ G, uint32_t nsites, int group) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsites) return;
  if (group == 2) {
    float* p = G + (size_t)s * 4;
    Q<double> q = norm_q(p);
    stq(p, Q<float>{(float)q.a0, (float)q.a1, (float)q.a2, (float)q.a3});
  } else {
    float* p = G + (size_t)s * 12;
    M3<double> m = gs3(to_d(ldm<float>(p)));
    M2<float> o;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int k = 0; k < 3; k++) o.r[i][k] = {(float)m.m[i][k].re, (float)m.m[i][k].im};
    stm(p, o);
  }
}
void launch_gauge_normalize(float* G, const Geom& g, int group, hipStream_t st) {
  const uint32_t n = 2 * g.Vh;
  hipLaunchKernelGGL(k_gauge_normalize, dim3((n + 255) / 256), dim3(256), 0, st, G, n, group);
}

__global__ void k_su2_reunit64(double* __restrict__ U, size_t nlinks) {
  size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nlinks) return;
  Q<double> q = ldq(U + l * 4);
  double n = 1.0 / sqrt(q.a0 * q.a0 + q.a1 * q.a1 + q.a2 * q.a2 + q.a3 * q.a3);
  stq(U + l * 4, Q<double>{q.a0 * n, q.a1 * n, q.a2 * n, q.a3 * n});
}
__global__ void k_su3_reunit64(double* __restrict__ U, size_t nlinks) {
  size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nlinks) return;
  M3<double> m = gs3(ldm<double>(U + l * 12));
  M2<double> o;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) o.r[i][k] = m.m[i][k];
  stm(U + l * 12, o);
}
void launch_su2_reunit64(double* U, const Geom& g, hipStream_t st) {
  size_t n = (size_t)8 * g.Vh;
  hipLaunchKernelGGL(k_su2_reunit64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, U, n);
}
void launch_su3_reunit64(double* U, const Geom& g, hipStream_t st) {
  size_t n = (size_t)8 * g.Vh;
  hipLaunchKernelGGL(k_su3_reunit64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, U, n);
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
        M3<double> a = full3(ldm<double>(link_ptr<12>(U, g, p, mu, h)));
        M3<double> b = full3(ldm<double>(link_ptr<12>(U, g, 1 - p, nu, hf[mu])));
        M3<double> c = full3(ldm<double>(link_ptr<12>(U, g, 1 - p, mu, hf[nu])));
        M3<double> d = full3(ldm<double>(link_ptr<12>(U, g, p, nu, h)));
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
}  // namespace mag
