// mag_gauge.h -- FP64 site-gauge algebra of the mixed-precision link form:
// normalisation of FP32 gauges (SU(2) q/|q|, SU(3) Gram-Schmidt of rows 0
// and 1, row 2 = conj(row0 x row1)), products of 2-row and full matrices,
// and the rebuild of links from the FP64 master
//   W(x,mu) = G(x) U0(x,mu) G(x+mu)^+.
//
// Only the red (parity 0) sites keep an accumulated gauge G_r.  The gauge
// of a black site x is recovered from one of its links, which the sweeps
// keep at W(x,0) = G(x) U0(x,0) G_r(x+0)^+ up to FP32 rounding:
//   G(x) = normalise( W(x,0) G_r(x+0) U0(x,0)^+ ).
// Any gauge makes the rebuilt links exactly gauge-equivalent to U0; the
// recovered one also keeps them within FP32 rounding of the swept links.
#pragma once
#include "mag_layout.h"
#include "mag_math.h"

namespace mag {

template <typename P>
__device__ __forceinline__ Q<double> norm_q(LRef<P> r) {
  const auto f = ldq(r);
  Q<double> q{(double)f.a0, (double)f.a1, (double)f.a2, (double)f.a3};
  double n = 1.0 / sqrt(q.a0 * q.a0 + q.a1 * q.a1 + q.a2 * q.a2 + q.a3 * q.a3);
  q.a0 *= n; q.a1 *= n; q.a2 *= n; q.a3 *= n;
  return q;
}
__device__ __forceinline__ Q<double> norm_q(const Q<double>& a) {
  const double n = 1.0 / sqrt(a.a0 * a.a0 + a.a1 * a.a1 + a.a2 * a.a2 + a.a3 * a.a3);
  return Q<double>{a.a0 * n, a.a1 * n, a.a2 * n, a.a3 * n};
}
__device__ __forceinline__ Q<double> to_d(const Q<float>& q) {
  return Q<double>{(double)q.a0, (double)q.a1, (double)q.a2, (double)q.a3};
}
__device__ __forceinline__ Q<float> to_f(const Q<double>& q) {
  return Q<float>{(float)q.a0, (float)q.a1, (float)q.a2, (float)q.a3};
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
__device__ __forceinline__ M2<double> to_d(const M2<double>& m) { return m; }
__device__ __forceinline__ M2<double> to_d(const M2<float>& m) {
  M2<double> r;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) r.r[i][k] = {(double)m.r[i][k].re, (double)m.r[i][k].im};
  return r;
}
__device__ __forceinline__ M2<float> to_f(const M2<double>& m) {
  M2<float> r;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) r.r[i][k] = {(float)m.r[i][k].re, (float)m.r[i][k].im};
  return r;
}
__device__ __forceinline__ M2<float> to_f(const M3<double>& m) {
  M2<float> r;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) r.r[i][k] = {(float)m.m[i][k].re, (float)m.m[i][k].im};
  return r;
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
__device__ __forceinline__ M2<double> rows01(const M3<double>& M) {
  M2<double> m;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) m.r[i][k] = M.m[i][k];
  return m;
}
// rows 0,1 of A B and of A B^+ (A given by its rows 0,1)
__device__ __forceinline__ M2<double> mul23(const M2<double>& A, const M3<double>& B) {
  M2<double> r;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      C<double> acc{0.0, 0.0};
#pragma unroll
      for (int k = 0; k < 3; k++) acc = cadd(acc, cmul(A.r[i][k], B.m[k][j]));
      r.r[i][j] = acc;
    }
  return r;
}
__device__ __forceinline__ M2<double> mul23_dag(const M2<double>& A, const M3<double>& B) {
  M2<double> r;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      C<double> acc{0.0, 0.0};
#pragma unroll
      for (int k = 0; k < 3; k++) acc = cadd(acc, cmulc(A.r[i][k], B.m[j][k]));
      r.r[i][j] = acc;
    }
  return r;
}

// G(x) of a black site from its swept link w0 = W(x,0), the normalised red
// gauge gr of x+0 and the master link u0 = U0(x,0)
__device__ __forceinline__ Q<double> black_gauge(const Q<double>& w0, const Q<double>& gr,
                                                 const Q<double>& u0) {
  return norm_q(qmul_dag(qmul(w0, gr), u0));
}
// rows 0,1 of W(x,0) G_r(x+0) U0(x,0)^+ (before normalisation)
__device__ __forceinline__ M2<double> black_gauge_rows(const M2<double>& w0, const M3<double>& gr,
                                                      const M2<double>& u0) {
  return mul23_dag(mul23(w0, gr), full3(u0));
}
// one rebuilt link a U0 b^+ (rows 0,1)
__device__ __forceinline__ M2<double> rebuild_link(const M3<double>& a, const M2<double>& u0,
                                                   const M3<double>& b) {
  return mul23_dag(mul23(rows01(a), full3(u0)), b);
}

}  // namespace mag
