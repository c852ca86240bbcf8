// mag_spin_math.h -- spin-form site algebra shared by the spin-session
// kernels (mag_spin.hip) and the gauge-copy half-sweep (mag_sweep.hip).
// Spins live in HBM as SoA per parity and component:
//   s_c(p, h) = S[(p*3 + c)*Vh + h]
// No multiply-add contraction in these helpers: every product and sum is
// rounded on its own, as in the oracle (gcc -ffp-contract=off), and a
// copy's arithmetic cannot change with the code around it (the gauge-copy
// kernel instances for different copy counts give identical results).
#pragma once
#include "mag_internal.h"
#include "mag_math.h"

namespace mag {

struct V3 {
  double a1, a2, a3;
};

__device__ __forceinline__ V3 ld_spin(const double* S, const Geom& g, int p, uint32_t h) {
  const double* b = S + (size_t)p * 3 * g.Vh + h;
  return V3{b[0], b[g.Vh], b[2 * (size_t)g.Vh]};
}
__device__ __forceinline__ void st_spin(double* S, const Geom& g, int p, uint32_t h, const V3& s) {
  double* b = S + (size_t)p * 3 * g.Vh + h;
  b[0] = s.a1;
  b[g.Vh] = s.a2;
  b[2 * (size_t)g.Vh] = s.a3;
}

// FP32 spins (the FP32 phase of the gauge-copy spin form)
struct V3f {
  float a1, a2, a3;
};
__device__ __forceinline__ V3f ld_spin(const float* S, const Geom& g, int p, uint32_t h) {
  const float* b = S + (size_t)p * 3 * g.Vh + h;
  return V3f{b[0], b[g.Vh], b[2 * (size_t)g.Vh]};
}
__device__ __forceinline__ void st_spin(float* S, const Geom& g, int p, uint32_t h, const V3f& s) {
  float* b = S + (size_t)p * 3 * g.Vh + h;
  b[0] = s.a1;
  b[g.Vh] = s.a2;
  b[2 * (size_t)g.Vh] = s.a3;
}
__device__ __forceinline__ V3 to_d(const V3f& s) { return V3{s.a1, s.a2, s.a3}; }
// an FP32 spin in FP64, scaled to unit length by the Newton step
// (3 - |s|^2) / 2 (|s|^2 - 1 is of the order of FP32 rounding, so the
// result is a unit vector to about 1e-14): the functional is linear in each
// spin's length, and summed from unscaled FP32 spins it would carry their
// rounding of length (about 6e-8 per spin) at first order
__device__ __forceinline__ V3 to_unit_d(const V3f& s) {
  const V3 v = to_d(s);
  const double c = fma(-0.5, fma(v.a1, v.a1, fma(v.a2, v.a2, v.a3 * v.a3)), 1.5);
  return V3{v.a1 * c, v.a2 * c, v.a3 * c};
}

// h += M(A) b (spin::contribution1, matrix.h:909-926)
__device__ __forceinline__ void contribution1(V3& h, const Q<double>& A, const V3& b) {
#pragma clang fp contract(off)
  const double q1 = A.a1 * A.a1, q2 = A.a2 * A.a2, q3 = A.a3 * A.a3;
  const double q12 = A.a1 * A.a2, q13 = A.a1 * A.a3, q23 = A.a2 * A.a3;
  const double q01 = A.a0 * A.a1, q02 = A.a0 * A.a2, q03 = A.a0 * A.a3;
  const double A5 = A.a0 * A.a0 - q1 - q2 - q3;
  h.a1 += b.a1 * (A5 + 2 * q1) + 2 * b.a2 * (q12 + q03) + 2 * b.a3 * (q13 - q02);
  h.a2 += 2 * b.a1 * (q12 - q03) + b.a2 * (A5 + 2 * q2) + 2 * b.a3 * (q23 + q01);
  h.a3 += 2 * b.a1 * (q13 + q02) + 2 * b.a2 * (q23 - q01) + b.a3 * (A5 + 2 * q3);
}
// h += M(A)^T b (spin::contribution1_conj, matrix.h:928-946)
__device__ __forceinline__ void contribution1_conj(V3& h, const Q<double>& A, const V3& b) {
#pragma clang fp contract(off)
  const double q1 = A.a1 * A.a1, q2 = A.a2 * A.a2, q3 = A.a3 * A.a3;
  const double q12 = A.a1 * A.a2, q13 = A.a1 * A.a3, q23 = A.a2 * A.a3;
  const double q01 = A.a0 * A.a1, q02 = A.a0 * A.a2, q03 = A.a0 * A.a3;
  const double A5 = A.a0 * A.a0 - q1 - q2 - q3;
  h.a1 += b.a1 * (A5 + 2 * q1) + 2 * b.a2 * (q12 - q03) + 2 * b.a3 * (q13 + q02);
  h.a2 += 2 * b.a1 * (q12 + q03) + b.a2 * (A5 + 2 * q2) + 2 * b.a3 * (q23 - q01);
  h.a3 += 2 * b.a1 * (q13 - q02) + 2 * b.a2 * (q23 + q01) + b.a3 * (A5 + 2 * q3);
}
// M(A) s (spin::contribution, matrix.h:870-887)
__device__ __forceinline__ V3 contribution(const V3& s, const Q<double>& A) {
#pragma clang fp contract(off)
  const double q1 = A.a1 * A.a1, q2 = A.a2 * A.a2, q3 = A.a3 * A.a3;
  const double q12 = 2. * A.a1 * A.a2, q13 = 2. * A.a1 * A.a3, q23 = 2. * A.a2 * A.a3;
  const double q01 = 2. * A.a0 * A.a1, q02 = 2. * A.a0 * A.a2, q03 = 2. * A.a0 * A.a3;
  const double A5 = A.a0 * A.a0 - q1 - q2 - q3;
  return V3{s.a1 * (A5 + 2 * q1) + s.a2 * (q12 + q03) + s.a3 * (q13 - q02),
            s.a1 * (q12 - q03) + s.a2 * (A5 + 2 * q2) + s.a3 * (q23 + q01),
            s.a1 * (q13 + q02) + s.a2 * (q23 - q01) + s.a3 * (A5 + 2 * q3)};
}

// reflect_fast (matrix.h:825-840)
__device__ __forceinline__ void reflect_fast(V3& s, const V3& V) {
#pragma clang fp contract(off)
  const double t1 = (V.a1 * V.a1 - V.a2 * V.a2 - V.a3 * V.a3) * s.a1 +
                    2 * V.a1 * (V.a2 * s.a2 + V.a3 * s.a3);
  const double t2 = (-V.a1 * V.a1 + V.a2 * V.a2 - V.a3 * V.a3) * s.a2 +
                    2 * V.a2 * (V.a1 * s.a1 + V.a3 * s.a3);
  const double t3 = (-V.a1 * V.a1 - V.a2 * V.a2 + V.a3 * V.a3) * s.a3 +
                    2 * V.a3 * (V.a1 * s.a1 + V.a2 * s.a2);
  const double n = sqrt(V.a1 * V.a1 + V.a2 * V.a2 + V.a3 * V.a3);
  const double n2 = n * n;
  s = V3{t1 / n2, t2 / n2, t3 / n2};
}
// parallel (matrix.h:842-856)
__device__ __forceinline__ void parallel(V3& s, const V3& V) {
#pragma clang fp contract(off)
  const double vn = sqrt(V.a1 * V.a1 + V.a2 * V.a2 + V.a3 * V.a3);
  s = V3{V.a1 / vn, V.a2 / vn, V.a3 / vn};
}
// geodesic omega step of s towards h_hat
__device__ __forceinline__ void spin_omega(V3& s, const V3& h, double omega) {
#pragma clang fp contract(off)
  const double hn = sqrt(h.a1 * h.a1 + h.a2 * h.a2 + h.a3 * h.a3);
  if (hn == 0.0) return;
  const double ih = 1.0 / hn;  // one division, three products
  const V3 hh{h.a1 * ih, h.a2 * ih, h.a3 * ih};
  const double c = s.a1 * hh.a1 + s.a2 * hh.a2 + s.a3 * hh.a3;
  V3 t{hh.a1 - c * s.a1, hh.a2 - c * s.a2, hh.a3 - c * s.a3};
  double tn = sqrt(t.a1 * t.a1 + t.a2 * t.a2 + t.a3 * t.a3);
  if (tn == 0.0) {
    if (c >= 0) return;
    V3 e{1.0, 0.0, 0.0};
    if (fabs(s.a1) > 0.9) e = V3{0.0, 1.0, 0.0};
    const double d = e.a1 * s.a1 + e.a2 * s.a2 + e.a3 * s.a3;
    t = V3{e.a1 - d * s.a1, e.a2 - d * s.a2, e.a3 - d * s.a3};
    tn = sqrt(t.a1 * t.a1 + t.a2 * t.a2 + t.a3 * t.a3);
  }
  const double theta = atan2(tn, c);
  double sa, ca;
  sincos(omega * theta, &sa, &ca);
  const double st = sa / tn;
  s = V3{ca * s.a1 + st * t.a1, ca * s.a2 + st * t.a2, ca * s.a3 + st * t.a3};
  // back onto the sphere: for omega > 1 this step multiplies a spin's
  // deviation from unit length by about 2 omega - 1 per update, so rounding
  // would grow exponentially over the sweeps without it
  const double r = 1.0 / sqrt(s.a1 * s.a1 + s.a2 * s.a2 + s.a3 * s.a3);
  s = V3{s.a1 * r, s.a2 * r, s.a3 * r};
}


// The geodesic omega step of the FP32 phase: the field h is summed in
// FP64 (from products of FP32 values, exact), its direction and the spin
// step in FP32 with the hardware reciprocal, square root and reciprocal
// square root and the short series of mag_trig.h.  The step direction is
// good to about 1e-7, so the spins settle within about 1e-7 of the
// maximum of the FP32-rounded problem and its functional within about
// 1e-14; the FP64 phase that follows polishes from there.  Returns
// 1 - s.h_hat before the step.
__device__ __forceinline__ float spin_omega_f(V3f& s, const V3& hd, float omega) {
  const float h1 = (float)hd.a1, h2 = (float)hd.a2, h3 = (float)hd.a3;
  const float hn2 = h1 * h1 + h2 * h2 + h3 * h3;
  if (hn2 == 0.0f) return 0.0f;
  const float ih = __builtin_amdgcn_rsqf(hn2);
  const float u1 = h1 * ih, u2 = h2 * ih, u3 = h3 * ih;
  const float c = s.a1 * u1 + s.a2 * u2 + s.a3 * u3;
  float t1 = u1 - c * s.a1, t2 = u2 - c * s.a2, t3 = u3 - c * s.a3;
  float tn2 = t1 * t1 + t2 * t2 + t3 * t3;
  if (tn2 == 0.0f) {
    if (c >= 0.0f) return 1.0f - c;
    // antiparallel: rotate in the plane of s and an axis not along it
    float e1 = 1.0f, e2 = 0.0f;
    if (fabsf(s.a1) > 0.9f) { e1 = 0.0f; e2 = 1.0f; }
    const float d = e1 * s.a1 + e2 * s.a2;
    t1 = e1 - d * s.a1;
    t2 = e2 - d * s.a2;
    t3 = -d * s.a3;
    tn2 = t1 * t1 + t2 * t2 + t3 * t3;
  }
  const float rtn = __builtin_amdgcn_rsqf(tn2);
  const float tn = tn2 * rtn;
  // theta = atan2(tn, c) in (0, pi)
  const float ac = fabsf(c);
  const bool steep = tn > ac;
  const float a = trig::atan01_f((steep ? ac : tn) * __builtin_amdgcn_rcpf(steep ? tn : ac));
  const float th = steep ? 1.57079632679489662f - a : a;
  const float theta = c >= 0.0f ? th : 3.14159265358979323846f - th;
  float sa, ca;
  trig::sincos_0_2pi_f(omega * theta, &sa, &ca);
  const float st = sa * rtn;
  float n1 = ca * s.a1 + st * t1, n2 = ca * s.a2 + st * t2, n3 = ca * s.a3 + st * t3;
  const float r = __builtin_amdgcn_rsqf(n1 * n1 + n2 * n2 + n3 * n3);
  s = V3f{n1 * r, n2 * r, n3 * r};
  return 1.0f - c;
}

// Row 3 of the rotation of the site gauge q: the spin of q, invariant under
// a diagonal U(1) factor on the left (the link functional is invariant
// under the same U(1)).
__device__ __forceinline__ V3 spin_of_gauge(const Q<double>& q) {
#pragma clang fp contract(off)
  return V3{2 * (q.a1 * q.a3 + q.a0 * q.a2), 2 * (q.a2 * q.a3 - q.a0 * q.a1),
            q.a0 * q.a0 + q.a3 * q.a3 - q.a1 * q.a1 - q.a2 * q.a2};
}

// The SU(2) element of smallest rotation with spin s (spin_of_gauge(G) = s):
// G = (c, -s2/(2c), s1/(2c), 0), c^2 = (1 + s3)/2, evaluated as
// (s1^2 + s2^2) / (2 (1 - s3)) for s3 < 0 (no cancellation); at s = -e3,
// G = (0, 1, 0, 0).  Normalised, so a spin off the unit sphere by rounding
// still gives an SU(2) element.
__device__ __forceinline__ Q<double> gauge_of_spin(const V3& s) {
#pragma clang fp contract(off)
  const double t = s.a1 * s.a1 + s.a2 * s.a2;
  const double c2 = s.a3 >= 0 ? 0.5 * (1 + s.a3) : 0.5 * t / (1 - s.a3);
  if (!(c2 > 1e-300)) return Q<double>{0.0, 1.0, 0.0, 0.0};
  const double c = sqrt(c2), r = 0.5 / c;
  const double n = 1.0 / sqrt(c2 + t * r * r);
  return Q<double>{c * n, -s.a2 * r * n, s.a1 * r * n, 0.0};
}

// The adjoint 3x3 matrix M(A) of a link, entries as in contribution1 with
// the off-diagonal factor 2 folded in: computed once per link and applied to
// the neighbour spin of every gauge copy.  add_M / add_MT evaluate the
// terms of contribution1 / contribution1_conj as a chain of explicit fused
// multiply-adds (fma(), not contraction: the same operations in every
// kernel instance, so a copy's result does not depend on the copy count);
// the off-diagonal entries are stored doubled (exact), which gives the
// reference's 2 b_j M_ij terms.
struct Adj {
  double d1, d2, d3;             // A5 + 2 q_i
  double p12, m12, p13, m13, p23, m23;  // 2 (q12 +- q03), 2 (q13 -+ q02), 2 (q23 +- q01)
};
__device__ __forceinline__ Adj adj_of(const Q<double>& A) {
#pragma clang fp contract(off)
  const double q1 = A.a1 * A.a1, q2 = A.a2 * A.a2, q3 = A.a3 * A.a3;
  const double q12 = A.a1 * A.a2, q13 = A.a1 * A.a3, q23 = A.a2 * A.a3;
  const double q01 = A.a0 * A.a1, q02 = A.a0 * A.a2, q03 = A.a0 * A.a3;
  const double A5 = A.a0 * A.a0 - q1 - q2 - q3;
  return Adj{A5 + 2 * q1,       A5 + 2 * q2,       A5 + 2 * q3,
             2 * (q12 + q03),   2 * (q12 - q03),   2 * (q13 - q02),
             2 * (q13 + q02),   2 * (q23 + q01),   2 * (q23 - q01)};
}
// The same entries as fused chains (the FP32 phase: A holds FP32 values,
// so every product is exact and each entry is rounded about once)
__device__ __forceinline__ Adj adj_of_fma(const Q<double>& A) {
  const double b0 = 2 * A.a0, b1 = 2 * A.a1, b2 = 2 * A.a2;
  const double t23 = fma(A.a2, A.a2, A.a3 * A.a3), t13 = fma(A.a1, A.a1, A.a3 * A.a3),
               t12 = fma(A.a1, A.a1, A.a2 * A.a2);
  const double u3 = b0 * A.a3, u2 = b0 * A.a2, u1 = b0 * A.a1;
  return Adj{fma(A.a0, A.a0, fma(A.a1, A.a1, -t23)), fma(A.a0, A.a0, fma(A.a2, A.a2, -t13)),
             fma(A.a0, A.a0, fma(A.a3, A.a3, -t12)), fma(b1, A.a2, u3), fma(b1, A.a2, -u3),
             fma(b1, A.a3, -u2), fma(b1, A.a3, u2), fma(b2, A.a3, u1), fma(b2, A.a3, -u1)};
}
// h += M(A) b
__device__ __forceinline__ void add_M(V3& h, const Adj& M, const V3& b) {
  h.a1 = fma(b.a3, M.p13, fma(b.a2, M.p12, fma(b.a1, M.d1, h.a1)));
  h.a2 = fma(b.a3, M.p23, fma(b.a2, M.d2, fma(b.a1, M.m12, h.a2)));
  h.a3 = fma(b.a3, M.d3, fma(b.a2, M.m23, fma(b.a1, M.m13, h.a3)));
}
// h += M(A)^T b
__device__ __forceinline__ void add_MT(V3& h, const Adj& M, const V3& b) {
  h.a1 = fma(b.a3, M.m13, fma(b.a2, M.m12, fma(b.a1, M.d1, h.a1)));
  h.a2 = fma(b.a3, M.m23, fma(b.a2, M.d2, fma(b.a1, M.p12, h.a2)));
  h.a3 = fma(b.a3, M.d3, fma(b.a2, M.p23, fma(b.a1, M.p13, h.a3)));
}

}  // namespace mag
