// mag_math.h -- SU(2) / SU(3) algebra shared by the MAG kernels (device).
//
// Conventions (reference: ilyakudrov/general_code, lib/cpu/include/matrix.h):
//   su2 quaternion {a0,a1,a2,a3} = a0 + i (a1 s1 + a2 s2 + a3 s3)  (matrix.h:8-13)
//     as a 2x2 matrix [[a0 + i a3, a2 + i a1], [-a2 + i a1, a0 - i a3]]
//   su3 rows 0,1 stored, row 2 = conj(row0 x row1)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mag_trig.h"

namespace mag {

// ----------------------------------------------------------------- scalars
__device__ __forceinline__ float m_sqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double m_sqrt(double x) { return sqrt(x); }
// angle functions of the omega rule (y > 0; angles in [0, pi]): FP32 the
// device library's, FP64 the fixed operation sequence of mag_trig.h that
// the host restatement repeats bit for bit
__device__ __forceinline__ float omega_atan2(float y, float x) { return atan2f(y, x); }
__device__ __forceinline__ double omega_atan2(double y, double x) { return trig::atan2_pos(y, x); }
__device__ __forceinline__ void omega_sincos(float a, float* s, float* c) { sincosf(a, s, c); }
__device__ __forceinline__ void omega_sincos(double a, double* s, double* c) {
  trig::sincos_0_2pi(a, s, c);
}
template <typename T> __device__ __forceinline__ T m_pi();
template <> __device__ __forceinline__ float m_pi<float>() { return 3.14159265358979323846f; }
template <> __device__ __forceinline__ double m_pi<double>() { return 3.14159265358979323846; }

// ----------------------------------------------------------------- su2
template <typename T> struct Q { T a0, a1, a2, a3; };

// operator* (matrix.h:93-98)
template <typename T>
__device__ __forceinline__ Q<T> qmul(const Q<T>& A, const Q<T>& B) {
  Q<T> r;
  r.a0 = A.a0 * B.a0 - A.a1 * B.a1 - A.a2 * B.a2 - A.a3 * B.a3;
  r.a1 = A.a0 * B.a1 + B.a0 * A.a1 + A.a3 * B.a2 - A.a2 * B.a3;
  r.a2 = A.a0 * B.a2 + B.a0 * A.a2 + A.a1 * B.a3 - A.a3 * B.a1;
  r.a3 = A.a0 * B.a3 + B.a0 * A.a3 + A.a2 * B.a1 - A.a1 * B.a2;
  return r;
}
// operator^ = A B^+ (matrix.h:99-104)
template <typename T>
__device__ __forceinline__ Q<T> qmul_dag(const Q<T>& A, const Q<T>& B) {
  Q<T> r;
  r.a0 = A.a0 * B.a0 + A.a1 * B.a1 + A.a2 * B.a2 + A.a3 * B.a3;
  r.a1 = -A.a0 * B.a1 + B.a0 * A.a1 - A.a3 * B.a2 + A.a2 * B.a3;
  r.a2 = -A.a0 * B.a2 + B.a0 * A.a2 - A.a1 * B.a3 + A.a3 * B.a1;
  r.a3 = -A.a0 * B.a3 + B.a0 * A.a3 - A.a2 * B.a1 + A.a1 * B.a2;
  return r;
}
// g (with g3 = 0) times U: the site update of a forward link
template <typename T>
__device__ __forceinline__ Q<T> gmul(const Q<T>& g, const Q<T>& B) {
  Q<T> r;
  r.a0 = g.a0 * B.a0 - g.a1 * B.a1 - g.a2 * B.a2;
  r.a1 = g.a0 * B.a1 + B.a0 * g.a1 - g.a2 * B.a3;
  r.a2 = g.a0 * B.a2 + B.a0 * g.a2 + g.a1 * B.a3;
  r.a3 = g.a0 * B.a3 + g.a2 * B.a1 - g.a1 * B.a2;
  return r;
}
// U times g^+ (g3 = 0): the site update of a backward link
template <typename T>
__device__ __forceinline__ Q<T> mul_gdag(const Q<T>& A, const Q<T>& g) {
  Q<T> r;
  r.a0 = A.a0 * g.a0 + A.a1 * g.a1 + A.a2 * g.a2;
  r.a1 = -A.a0 * g.a1 + g.a0 * A.a1 - A.a3 * g.a2;
  r.a2 = -A.a0 * g.a2 + g.a0 * A.a2 + A.a3 * g.a1;
  r.a3 = g.a0 * A.a3 - A.a2 * g.a1 + A.a1 * g.a2;
  return r;
}

// forward field term M(U) e3 (matrix.h:909-926 with b = e3)
template <typename T>
__device__ __forceinline__ void field_fwd(const Q<T>& A, T& h0, T& h1, T& h2) {
  h0 += T(2) * (A.a1 * A.a3 - A.a0 * A.a2);
  h1 += T(2) * (A.a2 * A.a3 + A.a0 * A.a1);
  h2 += A.a0 * A.a0 - A.a1 * A.a1 - A.a2 * A.a2 + A.a3 * A.a3;
}
// backward field term M(U)^T e3 (matrix.h:928-946 with b = e3)
template <typename T>
__device__ __forceinline__ void field_bwd(const Q<T>& A, T& h0, T& h1, T& h2) {
  h0 += T(2) * (A.a1 * A.a3 + A.a0 * A.a2);
  h1 += T(2) * (A.a2 * A.a3 - A.a0 * A.a1);
  h2 += A.a0 * A.a0 - A.a1 * A.a1 - A.a2 * A.a2 + A.a3 * A.a3;
}

// MAG functional of one su2 link (mag.cpp:998): a0^2 + a3^2 - a1^2 - a2^2,
// evaluated in double whatever the storage type.
template <typename T>
__device__ __forceinline__ double link_F(const Q<T>& A) {
  double a0 = A.a0, a1 = A.a1, a2 = A.a2, a3 = A.a3;
  return a0 * a0 - a1 * a1 - a2 * a2 + a3 * a3;
}

// Geodesic omega rule: site gauge g = (cos phi, sin phi * m), g3 = 0,
// phi = omega*theta/2, theta = angle(e3, h), m = (-h1, h0, 0)/|h_perp|.
// omega = 1 aligns e3 with h (spin::parallel), omega = 2 reflects it
// (spin::reflect_fast).  Also returns 1 - cos(theta) (the relaxation
// residual of mag.cpp:1477-1512).
template <typename T>
__device__ __forceinline__ Q<T> omega_gauge(T h0, T h1, T h2, T omega, T& resid) {
  T hp2 = h0 * h0 + h1 * h1;
  T hp = m_sqrt(hp2);
  T hn = m_sqrt(hp2 + h2 * h2);
  Q<T> g;
  g.a3 = T(0);
  if (hp == T(0)) {
    resid = (h2 >= T(0)) ? T(0) : (hn == T(0) ? T(0) : T(2));
    if (h2 >= T(0)) {
      g.a0 = T(1); g.a1 = T(0); g.a2 = T(0);
    } else {
      T s, c;
      omega_sincos(T(0.5) * omega * m_pi<T>(), &s, &c);
      g.a0 = c; g.a1 = s; g.a2 = T(0);
    }
    return g;
  }
  resid = (h2 >= T(0)) ? hp2 / (hn * (hn + h2)) : T(1) - h2 / hn;
  T theta = omega_atan2(hp, h2);
  T s, c;
  omega_sincos(T(0.5) * omega * theta, &s, &c);
  T inv = T(1) / hp;
  g.a0 = c;
  g.a1 = -s * h1 * inv;
  g.a2 = s * h0 * inv;
  return g;
}

// The same rule in FP32 for the FP32 sweeps, with the hardware reciprocal,
// square root and reciprocal square root (about 1 ulp) and the short series
// of mag_trig.h in place of IEEE division and the device library's
// atan2f / sincosf: about a third of the instructions (the rule was some
// 290 of the 2450 VALU instructions per wave of the SU(3) FP32 half-sweep
// per subgroup).  FP32 trajectories are not compared bit for bit with
// anything; the FP64 phases use the exact-IEEE rule above.
template <>
__device__ __forceinline__ Q<float> omega_gauge<float>(float h0, float h1, float h2, float omega,
                                                       float& resid) {
  const float hp2 = h0 * h0 + h1 * h1;
  Q<float> g;
  g.a3 = 0.0f;
  if (hp2 == 0.0f) {  // along e3 or zero: as the FP64 rule
    resid = (h2 >= 0.0f) ? 0.0f : 2.0f;
    if (h2 >= 0.0f) {
      g.a0 = 1.0f; g.a1 = 0.0f; g.a2 = 0.0f;
    } else {
      float s, c;
      trig::sincos_0_pi_f(0.5f * omega * 3.14159265358979323846f, &s, &c);
      g.a0 = c; g.a1 = s; g.a2 = 0.0f;
    }
    return g;
  }
  const float hp = __builtin_amdgcn_sqrtf(hp2);
  const float rhn = __builtin_amdgcn_rsqf(hp2 + h2 * h2);
  resid = (h2 >= 0.0f) ? hp2 * rhn * rhn * __builtin_amdgcn_rcpf(1.0f + h2 * rhn)
                       : 1.0f - h2 * rhn;
  // theta = atan2(hp, h2) in (0, pi): the angle to the e3 axis
  const float ah = fabsf(h2);
  const bool steep = hp > ah;
  const float a = trig::atan01_f((steep ? ah : hp) * __builtin_amdgcn_rcpf(steep ? hp : ah));
  const float th = steep ? 1.57079632679489662f - a : a;
  const float theta = h2 >= 0.0f ? th : 3.14159265358979323846f - th;
  float s, c;
  trig::sincos_0_pi_f(0.5f * omega * theta, &s, &c);
  const float inv = __builtin_amdgcn_rcpf(hp);
  g.a0 = c;
  g.a1 = -s * h1 * inv;
  g.a2 = s * h0 * inv;
  return g;
}

// ----------------------------------------------------------------- complex
template <typename T> struct C { T re, im; };
template <typename T> __device__ __forceinline__ C<T> cmk(T r, T i) { return C<T>{r, i}; }
template <typename T> __device__ __forceinline__ C<T> cadd(C<T> a, C<T> b) { return C<T>{a.re + b.re, a.im + b.im}; }
template <typename T> __device__ __forceinline__ C<T> csub(C<T> a, C<T> b) { return C<T>{a.re - b.re, a.im - b.im}; }
template <typename T> __device__ __forceinline__ C<T> cmul(C<T> a, C<T> b) {
  return C<T>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// a * conj(b)
template <typename T> __device__ __forceinline__ C<T> cmulc(C<T> a, C<T> b) {
  return C<T>{a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}
template <typename T> __device__ __forceinline__ C<T> cconj(C<T> a) { return C<T>{a.re, -a.im}; }
template <typename T> __device__ __forceinline__ T cnorm2(C<T> a) { return a.re * a.re + a.im * a.im; }

// ----------------------------------------------------------------- su3 (2-row)
// Rows 0,1 in registers: r[row][col]; row 2 rebuilt as conj(row0 x row1).
template <typename T> struct M2 { C<T> r[2][3]; };
template <typename T> struct M3 { C<T> m[3][3]; };

template <typename T>
__device__ __forceinline__ void row2_of(const M2<T>& M, C<T>* z) {
  const C<T>* a = M.r[0];
  const C<T>* b = M.r[1];
  z[0] = cconj(csub(cmul(a[1], b[2]), cmul(a[2], b[1])));
  z[1] = cconj(csub(cmul(a[2], b[0]), cmul(a[0], b[2])));
  z[2] = cconj(csub(cmul(a[0], b[1]), cmul(a[1], b[0])));
}

// Block field terms for the SU(2) subgroup acting on rows/cols (i,j):
// B = [[p,q],[s,t]].  Forward: 1/2 tr(s_a B s3 B^+); backward:
// 1/2 tr(s_a B^+ s3 B).  For an su2 link these reduce to field_fwd/bwd.
template <typename T>
__device__ __forceinline__ void blk_fwd(C<T> p, C<T> q, C<T> s, C<T> t, T& h0, T& h1, T& h2) {
  C<T> X01 = csub(cmulc(p, s), cmulc(q, t));
  h0 += X01.re;
  h1 += -X01.im;
  h2 += T(0.5) * (cnorm2(p) - cnorm2(q) - cnorm2(s) + cnorm2(t));
}
template <typename T>
__device__ __forceinline__ void blk_bwd(C<T> p, C<T> q, C<T> s, C<T> t, T& h0, T& h1, T& h2) {
  C<T> Y01 = csub(cmulc(q, p), cmulc(t, s));
  h0 += Y01.re;
  h1 += -Y01.im;
  h2 += T(0.5) * (cnorm2(p) - cnorm2(s) - cnorm2(q) + cnorm2(t));
}

}  // namespace mag
