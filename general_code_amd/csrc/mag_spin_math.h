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
