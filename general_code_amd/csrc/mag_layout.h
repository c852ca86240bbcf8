// mag_layout.h -- device-side access to the parity-split SoA link layout.
//
// Link layout in HBM: links of base parity p and direction mu form one array
// of Vh = V/2 links; link (p, mu, h) starts at
//   U + ((p*4 + mu)*Vh + h) * NR          NR = 4 (su2) or 12 (su3 rows 0,1)
// Consecutive lanes own consecutive half indices h, so every gather of the
// 8 links of a site is a contiguous, coalesced access.
#pragma once
#include "mag_internal.h"
#include "mag_math.h"

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

// sum_j |U_jj|^2 of an su3 link (mag.cpp:1837-1848), in FP64;
// U_22 = conj(U00 U11 - U01 U10)
template <typename T>
__device__ __forceinline__ double su3_link_F(const M2<T>& m) {
  C<double> a{(double)m.r[0][0].re, (double)m.r[0][0].im};
  C<double> b{(double)m.r[1][1].re, (double)m.r[1][1].im};
  C<double> c{(double)m.r[0][1].re, (double)m.r[0][1].im};
  C<double> d{(double)m.r[1][0].re, (double)m.r[1][0].im};
  C<double> u22 = csub(cmul(a, b), cmul(c, d));
  return cnorm2(a) + cnorm2(b) + cnorm2(u22);
}

}  // namespace mag
