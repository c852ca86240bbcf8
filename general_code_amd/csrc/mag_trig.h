// mag_trig.h -- the FP64 angle functions of the omega rule, written out as
// a fixed sequence of IEEE additions, multiplications and divisions (no
// fused multiply-adds, no library calls), so that the host restatement of
// the same sequence gives bit-identical results.
//
//   atan2_pos(y, x), y > 0:  result in (0, pi).  The octant is folded onto
//     t in [0, 1] (t = y/|x| or |x|/y); t is rounded to the nearest quarter
//     c = k/4 and atan t = atan c + atan u with u = (t - c)/(1 + t c),
//     |u| <= 1/8; atan u is its Taylor series to u^19 (relative truncation
//     error below 4e-20).  atan(k/4) are the correctly rounded constants.
//   sincos_0_2pi(a), a in [0, 2 pi]:  a = q pi/2 + r, q = round(2a/pi),
//     r = (a - q P1) - q P2 with P1 + P2 = pi/2 (Cody-Waite; a - q P1 is
//     exact for q <= 4 by Sterbenz), |r| <= pi/4; sin r to r^17 and cos r to
//     r^18 by their Taylor series (truncation below 1e-18), then the
//     quadrant q.
// Accuracy against the C library: within 2 ulp over their domains.
#pragma once
#include <hip/hip_runtime.h>

namespace mag {
namespace trig {

__device__ __forceinline__ double atan_small(double u) {
#pragma clang fp contract(off)
  const double z = u * u;
  double p = -1.0 / 19.0;
  p = p * z + 1.0 / 17.0;
  p = p * z + -1.0 / 15.0;
  p = p * z + 1.0 / 13.0;
  p = p * z + -1.0 / 11.0;
  p = p * z + 1.0 / 9.0;
  p = p * z + -1.0 / 7.0;
  p = p * z + 1.0 / 5.0;
  p = p * z + -1.0 / 3.0;
  return u + (u * z) * p;
}

// atan t for t in [0, 1]
__device__ __forceinline__ double atan01(double t) {
#pragma clang fp contract(off)
  const int k = (int)(t * 4.0 + 0.5);
  const double c = 0.25 * (double)k;
  const double u = (t - c) / (1.0 + t * c);
  const double base = k == 0   ? 0.0
                      : k == 1 ? 0.24497866312686414
                      : k == 2 ? 0.4636476090008061
                      : k == 3 ? 0.6435011087932844
                               : 0.7853981633974483;
  return base + atan_small(u);
}

__device__ __forceinline__ double atan2_pos(double y, double x) {
#pragma clang fp contract(off)
  const double ax = fabs(x);
  const double pio2 = 1.5707963267948966, pi = 3.141592653589793;
  if (y <= ax) {
    const double a = atan01(y / ax);
    return x >= 0.0 ? a : pi - a;
  }
  const double a = atan01(ax / y);
  return x >= 0.0 ? pio2 - a : pio2 + a;
}

__device__ __forceinline__ void sincos_0_2pi(double a, double* s, double* c) {
#pragma clang fp contract(off)
  const int q = (int)(a * 0.6366197723675814 + 0.5);  // 2/pi
  const double qd = (double)q;
  const double r = (a - qd * 1.5707963267948966) - qd * 6.123233995736766e-17;
  const double z = r * r;
  double ps = 1.0 / 355687428096000.0;  // 1/17!
  ps = ps * z + -1.0 / 1307674368000.0;
  ps = ps * z + 1.0 / 6227020800.0;
  ps = ps * z + -1.0 / 39916800.0;
  ps = ps * z + 1.0 / 362880.0;
  ps = ps * z + -1.0 / 5040.0;
  ps = ps * z + 1.0 / 120.0;
  ps = ps * z + -1.0 / 6.0;
  const double sr = r + (r * z) * ps;
  double pc = -1.0 / 6402373705728000.0;  // -1/18!
  pc = pc * z + 1.0 / 20922789888000.0;
  pc = pc * z + -1.0 / 87178291200.0;
  pc = pc * z + 1.0 / 479001600.0;
  pc = pc * z + -1.0 / 3628800.0;
  pc = pc * z + 1.0 / 40320.0;
  pc = pc * z + -1.0 / 720.0;
  pc = pc * z + 1.0 / 24.0;
  pc = pc * z + -0.5;
  const double cr = 1.0 + z * pc;
  switch (q & 3) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
  }
}

// FP32: the angle functions of the FP32 sweeps, with hardware reciprocal
// and square root (mag_math.h, omega_gauge<float>).  atan on [0, 1] is an
// odd polynomial of degree 15 fitted to atan by weighted least squares
// (Lawson's iteration towards the minimax fit; max error 5e-8 with these
// FP32 coefficients); sin and cos of the angle reduced by pi/2 are their
// Taylor series to r^9 and r^8 (truncation below 3e-8 for |r| <= pi/4).
// The angle of sincos_0_pi_f lies in [0, pi], so q <= 2 and q * (pi/2)_hi
// is exact.
__device__ __forceinline__ float atan01_f(float t) {
  const float z = t * t;
  float p = -0.004054536111652851f;
  p = p * z + 0.021862855181097984f;
  p = p * z + -0.05591219291090965f;
  p = p * z + 0.09642188996076584f;
  p = p * z + -0.1390862762928009f;
  p = p * z + 0.19946564733982086f;
  p = p * z + -0.33329859375953674f;
  p = p * z + 0.9999993443489075f;
  return t * p;
}
__device__ __forceinline__ void sincos_0_pi_f(float a, float* s, float* c) {
  const int q = (int)(a * 0.63661977f + 0.5f);
  const float qd = (float)q;
  const float r = (a - qd * 1.5707963705062866f) - qd * -4.371139006309477e-08f;
  const float z = r * r;
  float ps = 1.0f / 362880.0f;
  ps = ps * z + -1.0f / 5040.0f;
  ps = ps * z + 1.0f / 120.0f;
  ps = ps * z + -1.0f / 6.0f;
  const float sr = r + (r * z) * ps;
  float pc = 1.0f / 40320.0f;
  pc = pc * z + -1.0f / 720.0f;
  pc = pc * z + 1.0f / 24.0f;
  pc = pc * z + -0.5f;
  const float cr = 1.0f + z * pc;
  if (q == 0) {
    *s = sr;
    *c = cr;
  } else if (q == 1) {
    *s = cr;
    *c = -sr;
  } else {
    *s = -sr;
    *c = -cr;
  }
}

// the same for an angle in [0, 2 pi] (q <= 4; q * (pi/2)_hi exact)
__device__ __forceinline__ void sincos_0_2pi_f(float a, float* s, float* c) {
  const int q = (int)(a * 0.63661977f + 0.5f);
  const float qd = (float)q;
  const float r = (a - qd * 1.5707963705062866f) - qd * -4.371139006309477e-08f;
  const float z = r * r;
  float ps = 1.0f / 362880.0f;
  ps = ps * z + -1.0f / 5040.0f;
  ps = ps * z + 1.0f / 120.0f;
  ps = ps * z + -1.0f / 6.0f;
  const float sr = r + (r * z) * ps;
  float pc = 1.0f / 40320.0f;
  pc = pc * z + -1.0f / 720.0f;
  pc = pc * z + 1.0f / 24.0f;
  pc = pc * z + -0.5f;
  const float cr = 1.0f + z * pc;
  switch (q & 3) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
  }
}

}  // namespace trig
}  // namespace mag
