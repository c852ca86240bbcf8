/*
 * mag_oracle.c -- CPU oracle for the MAG gauge-fixing path (FP64).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  The product never links it.
 *
 * Reference citations are relative to ilyakudrov/general_code.
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   su2 algebra ............ 2x2 complex matrix products (numpy), unitarity
 *   spin primitives ........ M(A)_ij = 1/2 tr(s_i A s_j A^+) closed form,
 *                            reflect o reflect = id, parallel = h/|h|,
 *                            row3(R(GetGaugeMatrix(s))) = s
 *   functionals ............ closed forms (identity config -> 1),
 *                            F_spin(U,s) = F(gauge_transform_spins(U,s))
 *   lexicographic sweeps ... exact invariance of F_spin under a sequential
 *                            reflection sweep, monotone relaxation, brute
 *                            force against a pure-Python loop on 2^4
 *   link-form omega sweep .. spin-form/link-form equivalence per sweep,
 *                            omega=1 -> parallel, omega=2 -> reflect_fast,
 *                            F monotone, gauge invariance of plaquette/W
 *   SU(3) CM sweep ......... embedded SU(2) reproduces the SU(2) sweep with
 *                            F3 = (2 + 2 F2)/3 ... (see test), monotone F,
 *                            unitarity, plaquette invariance
 *   Philox4x32-10 .......... Random123 known-answer vectors
 *   SU(3) exp .............. scipy.linalg.expm
 *   plaquette / Wilson ..... abelian constant configs (closed form),
 *                            identity config, gauge invariance
 *   heat bath .............. statistical: distribution of s.h_hat against
 *                            exp(|h| x / T) on [-1,1] (parity unpinned
 *                            stream-wise: the reference RNG is time-seeded)
 *   abelian projection ..... closed forms (exp(i phi s3) -> phi; SU(3)
 *                            diagonal phases summing to 0 mod 2 pi)
 *   monopole currents ...... integer-valued, conserved on the dual lattice,
 *                            invariant under U(1) (SU(3): U(1)^2) gauge
 *                            transformations of the angles, equal to the
 *                            Levi-Civita form -eps_{mu nu rho sigma}
 *                            d_nu P_{rho sigma}(x + mu) / 4 pi (numpy),
 *                            colour sum zero for SU(3)
 *   monopole clusters ...... brute force: connected components of the
 *                            current links by an independent union-find
 *                            (scipy), totals equal to the sums over all
 *                            links, hand-built loops, windings and charges
 *                            with known answers
 */
#include "mag_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* threads                                                              */
/* ------------------------------------------------------------------ */
static int g_threads = 0;
void or_set_threads(int n) {
  g_threads = n;
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#endif
}
int or_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* geometry (indexing.h:71-126)                                        */
/* ------------------------------------------------------------------ */
typedef struct {
  int L[4];
  int64_t V;
  int64_t stride[4];
} geom_t;

static geom_t geom(const int *dims) {
  geom_t g;
  for (int i = 0; i < 4; i++) g.L[i] = dims[i];
  g.stride[0] = 1;
  g.stride[1] = dims[0];
  g.stride[2] = (int64_t)dims[0] * dims[1];
  g.stride[3] = (int64_t)dims[0] * dims[1] * dims[2];
  g.V = g.stride[3] * dims[3];
  return g;
}

static inline void coords(const geom_t *g, int64_t s, int *c) {
  c[0] = (int)(s % g->L[0]);
  s /= g->L[0];
  c[1] = (int)(s % g->L[1]);
  s /= g->L[1];
  c[2] = (int)(s % g->L[2]);
  c[3] = (int)(s / g->L[2]);
}

/* periodic neighbour: DataPatternLexicographical::move_forward/backward */
static inline int64_t nb(const geom_t *g, int64_t s, const int *c, int mu,
                         int dir) {
  int x = c[mu] + dir;
  if (x == g->L[mu]) return s - (int64_t)(g->L[mu] - 1) * g->stride[mu];
  if (x < 0) return s + (int64_t)(g->L[mu] - 1) * g->stride[mu];
  return s + dir * g->stride[mu];
}

static inline int64_t shift_n(const geom_t *g, int64_t s, int mu, int n) {
  int c[4];
  coords(g, s, c);
  int x = ((c[mu] + n) % g->L[mu] + g->L[mu]) % g->L[mu];
  return s + (int64_t)(x - c[mu]) * g->stride[mu];
}

/* ------------------------------------------------------------------ */
/* su2 algebra (matrix.h:9-117)                                        */
/* ------------------------------------------------------------------ */
typedef struct { double a0, a1, a2, a3; } q_t;

static inline q_t qload(const double *p) {
  q_t q = {p[0], p[1], p[2], p[3]};
  return q;
}
static inline void qstore(double *p, q_t q) {
  p[0] = q.a0; p[1] = q.a1; p[2] = q.a2; p[3] = q.a3;
}
/* operator* (matrix.h:93-98) */
static inline q_t qmul(q_t A, q_t B) {
  q_t r;
  r.a0 = A.a0 * B.a0 - A.a1 * B.a1 - A.a2 * B.a2 - A.a3 * B.a3;
  r.a1 = A.a0 * B.a1 + B.a0 * A.a1 + A.a3 * B.a2 - A.a2 * B.a3;
  r.a2 = A.a0 * B.a2 + B.a0 * A.a2 + A.a1 * B.a3 - A.a3 * B.a1;
  r.a3 = A.a0 * B.a3 + B.a0 * A.a3 + A.a2 * B.a1 - A.a1 * B.a2;
  return r;
}
/* operator^ = A * B^+ (matrix.h:99-104) */
static inline q_t qmul_dag(q_t A, q_t B) {
  q_t r;
  r.a0 = A.a0 * B.a0 + A.a1 * B.a1 + A.a2 * B.a2 + A.a3 * B.a3;
  r.a1 = -A.a0 * B.a1 + B.a0 * A.a1 - A.a3 * B.a2 + A.a2 * B.a3;
  r.a2 = -A.a0 * B.a2 + B.a0 * A.a2 - A.a1 * B.a3 + A.a3 * B.a1;
  r.a3 = -A.a0 * B.a3 + B.a0 * A.a3 - A.a2 * B.a1 + A.a1 * B.a2;
  return r;
}
/* operator% = A^+ * B (matrix.h:105-110) */
static inline q_t qdag_mul(q_t A, q_t B) {
  q_t r;
  r.a0 = A.a0 * B.a0 + A.a1 * B.a1 + A.a2 * B.a2 + A.a3 * B.a3;
  r.a1 = A.a0 * B.a1 - B.a0 * A.a1 - A.a3 * B.a2 + A.a2 * B.a3;
  r.a2 = A.a0 * B.a2 - B.a0 * A.a2 - A.a1 * B.a3 + A.a3 * B.a1;
  r.a3 = A.a0 * B.a3 - B.a0 * A.a3 - A.a2 * B.a1 + A.a1 * B.a2;
  return r;
}
static inline q_t qconj(q_t A) { /* matrix.h:73 */
  q_t r = {A.a0, -A.a1, -A.a2, -A.a3};
  return r;
}
/* (sigma3_mult(U) * conj(U)).tr() (mag.cpp:998, matrix.h:78) */
static inline double qfunc(q_t A) {
  q_t s = {A.a0, -A.a1, -A.a2, A.a3};
  return qmul(s, qconj(A)).a0;
}

void or_su2_mul(const double *A, const double *B, double *out) {
  qstore(out, qmul(qload(A), qload(B)));
}
void or_su2_mul_dag(const double *A, const double *B, double *out) {
  qstore(out, qmul_dag(qload(A), qload(B)));
}
void or_su2_dag_mul(const double *A, const double *B, double *out) {
  qstore(out, qdag_mul(qload(A), qload(B)));
}
double or_su2_link_functional(const double *A) { return qfunc(qload(A)); }

/* ------------------------------------------------------------------ */
/* spin primitives (matrix.h:773-985)                                  */
/* ------------------------------------------------------------------ */
/* contribution1: h += M(A) b (matrix.h:909-926) */
void or_spin_contribution1(double *h, const double *Ap, const double *b) {
  q_t A = qload(Ap);
  double q1 = A.a1 * A.a1, q2 = A.a2 * A.a2, q3 = A.a3 * A.a3;
  double q12 = A.a1 * A.a2, q13 = A.a1 * A.a3, q23 = A.a2 * A.a3;
  double q01 = A.a0 * A.a1, q02 = A.a0 * A.a2, q03 = A.a0 * A.a3;
  double A5 = A.a0 * A.a0 - q1 - q2 - q3;
  h[0] += b[0] * (A5 + 2 * q1) + 2 * b[1] * (q12 + q03) + 2 * b[2] * (q13 - q02);
  h[1] += 2 * b[0] * (q12 - q03) + b[1] * (A5 + 2 * q2) + 2 * b[2] * (q23 + q01);
  h[2] += 2 * b[0] * (q13 + q02) + 2 * b[1] * (q23 - q01) + b[2] * (A5 + 2 * q3);
}
/* contribution1_conj: h += M(A)^T b (matrix.h:928-946) */
void or_spin_contribution1_conj(double *h, const double *Ap, const double *b) {
  q_t A = qload(Ap);
  double q1 = A.a1 * A.a1, q2 = A.a2 * A.a2, q3 = A.a3 * A.a3;
  double q12 = A.a1 * A.a2, q13 = A.a1 * A.a3, q23 = A.a2 * A.a3;
  double q01 = A.a0 * A.a1, q02 = A.a0 * A.a2, q03 = A.a0 * A.a3;
  double A5 = A.a0 * A.a0 - q1 - q2 - q3;
  h[0] += b[0] * (A5 + 2 * q1) + 2 * b[1] * (q12 - q03) + 2 * b[2] * (q13 + q02);
  h[1] += 2 * b[0] * (q12 + q03) + b[1] * (A5 + 2 * q2) + 2 * b[2] * (q23 - q01);
  h[2] += 2 * b[0] * (q13 - q02) + 2 * b[1] * (q23 + q01) + b[2] * (A5 + 2 * q3);
}
/* contribution (matrix.h:870-887): returns M(A) this */
static inline void contribution(const double *s, q_t A, double *out) {
  double q1 = A.a1 * A.a1, q2 = A.a2 * A.a2, q3 = A.a3 * A.a3;
  double q12 = 2. * A.a1 * A.a2, q13 = 2. * A.a1 * A.a3, q23 = 2. * A.a2 * A.a3;
  double q01 = 2. * A.a0 * A.a1, q02 = 2. * A.a0 * A.a2, q03 = 2. * A.a0 * A.a3;
  double A5 = A.a0 * A.a0 - q1 - q2 - q3;
  out[0] = s[0] * (A5 + 2 * q1) + s[1] * (q12 + q03) + s[2] * (q13 - q02);
  out[1] = s[0] * (q12 - q03) + s[1] * (A5 + 2 * q2) + s[2] * (q23 + q01);
  out[2] = s[0] * (q13 + q02) + s[1] * (q23 - q01) + s[2] * (A5 + 2 * q3);
}
/* reflect_fast (matrix.h:825-840) */
void or_spin_reflect_fast(double *s, const double *V) {
  double tmp1 = (V[0] * V[0] - V[1] * V[1] - V[2] * V[2]) * s[0] +
                2 * V[0] * (V[1] * s[1] + V[2] * s[2]);
  double tmp2 = (-V[0] * V[0] + V[1] * V[1] - V[2] * V[2]) * s[1] +
                2 * V[1] * (V[0] * s[0] + V[2] * s[2]);
  double tmp3 = (-V[0] * V[0] - V[1] * V[1] + V[2] * V[2]) * s[2] +
                2 * V[2] * (V[0] * s[0] + V[1] * s[1]);
  double n = sqrt(V[0] * V[0] + V[1] * V[1] + V[2] * V[2]);
  double norm2 = n * n;
  s[0] = tmp1 / norm2;
  s[1] = tmp2 / norm2;
  s[2] = tmp3 / norm2;
}
/* parallel (matrix.h:842-856): s <- V/|V|, returns 1 - s_old . V/|V| */
double or_spin_parallel(double *s, const double *V) {
  double vn = sqrt(V[0] * V[0] + V[1] * V[1] + V[2] * V[2]);
  double b1 = V[0] / vn, b2 = V[1] / vn, b3 = V[2] / vn;
  double c = b1 * s[0] + b2 * s[1] + b3 * s[2];
  s[0] = b1;
  s[1] = b2;
  s[2] = b3;
  return 1 - c;
}
/* GetGaugeMatrix (matrix.h:960-963) */
void or_spin_gauge_matrix(const double *s, double *g) {
  double sq = sqrt(2. - 2 * s[2]);
  g[0] = -s[0] / sq;
  g[1] = 0.;
  g[2] = -sq / 2.;
  g[3] = -s[1] / sq;
}

/* ------------------------------------------------------------------ */
/* SU(2) spin-form MAG (lib/cpu/src/mag.cpp)                           */
/* ------------------------------------------------------------------ */
/* contribution_site (mag.cpp:569-626): fwd(mu), bwd(mu) for mu = 0..3 */
static void contribution_site_g(const geom_t *g, const double *links,
                                const double *spins, int64_t s, double *h) {
  int c[4];
  coords(g, s, c);
  h[0] = h[1] = h[2] = 0;
  for (int mu = 0; mu < 4; mu++) {
    int64_t f = nb(g, s, c, mu, +1);
    int64_t b = nb(g, s, c, mu, -1);
    or_spin_contribution1(h, links + 4 * (4 * s + mu), spins + 3 * f);
    or_spin_contribution1_conj(h, links + 4 * (4 * b + mu), spins + 3 * b);
  }
}
void or_su2_contribution_site(const double *links, const double *spins,
                              const int *dims, int64_t site, double *h) {
  geom_t g = geom(dims);
  contribution_site_g(&g, links, spins, site, h);
}

/* MAG_functional_su2 (mag.cpp:993-1001) */
double or_su2_functional(const double *links, const int *dims) {
  geom_t g = geom(dims);
  double result = 0;
  for (int64_t i = 0; i < 4 * g.V; i++) result += qfunc(qload(links + 4 * i));
  return result / (double)(4 * g.V);
}

/* MAG_functional_su2_spin (mag.cpp:925-953) */
double or_su2_functional_spin(const double *links, const double *spins,
                              const int *dims) {
  geom_t g = geom(dims);
  double result = 0;
  for (int64_t s = 0; s < g.V; s++) {
    int c[4];
    coords(&g, s, c);
    for (int mu = 0; mu < 4; mu++) {
      int64_t f = nb(&g, s, c, mu, +1);
      double v[3];
      contribution(spins + 3 * f, qload(links + 4 * (4 * s + mu)), v);
      result += v[0] * spins[3 * s] + v[1] * spins[3 * s + 1] +
                v[2] * spins[3 * s + 2];
    }
  }
  return result / (double)(4 * g.V);
}

/* gauge_tranformation_spins (mag.cpp:1045-1068), in place, serial order */
void or_su2_gauge_transform_spins(double *links, const double *spins,
                                  const int *dims) {
  geom_t g = geom(dims);
  for (int64_t s = 0; s < g.V; s++) {
    int c[4];
    coords(&g, s, c);
    double G[4];
    or_spin_gauge_matrix(spins + 3 * s, G);
    for (int mu = 0; mu < 4; mu++) {
      int64_t f = nb(&g, s, c, mu, +1);
      double Gf[4];
      or_spin_gauge_matrix(spins + 3 * f, Gf);
      q_t A = qmul(qload(G), qload(links + 4 * (4 * s + mu)));
      A = qmul(A, qconj(qload(Gf)));
      qstore(links + 4 * (4 * s + mu), A);
    }
  }
}

/* overrelaxation_update (mag.cpp:1375-1401), single thread = the only
 * deterministic form of the reference's in-place lexicographic loop */
void or_su2_or_sweep_lexi(const double *links, double *spins, const int *dims) {
  geom_t g = geom(dims);
  for (int64_t s = 0; s < g.V; s++) {
    double h[3];
    contribution_site_g(&g, links, spins, s, h);
    or_spin_reflect_fast(spins + 3 * s, h);
  }
}

/* Timing baseline only: the reference's overrelaxation_update
 * (mag.cpp:1375-1401) as it runs there -- an OpenMP loop over all sites in
 * lexicographic order, collapse(4), spins updated in place.  Neighbouring
 * sites updated by different threads race, exactly as in the reference, so
 * the result depends on the thread schedule; it is never used as a check. */
void or_su2_or_sweep_lexi_omp(const double *links, double *spins, const int *dims) {
  geom_t g = geom(dims);
  const int L0 = g.L[0], L1 = g.L[1], L2 = g.L[2], L3 = g.L[3];
#pragma omp parallel for collapse(4) schedule(static)
  for (int t = 0; t < L3; t++)
    for (int z = 0; z < L2; z++)
      for (int y = 0; y < L1; y++)
        for (int x = 0; x < L0; x++) {
          const int64_t s = x + (int64_t)L0 * (y + (int64_t)L1 * (z + (int64_t)L2 * t));
          double h[3];
          contribution_site_g(&g, links, spins, s, h);
          or_spin_reflect_fast(spins + 3 * s, h);
        }
}

/* relaxation_update (mag.cpp:1477-1512) */
void or_su2_relax_sweep_lexi(const double *links, double *spins,
                             const int *dims, double *diff_max,
                             double *diff_avg) {
  geom_t g = geom(dims);
  double dmax = 0, dsum = 0;
  for (int64_t s = 0; s < g.V; s++) {
    double h[3];
    contribution_site_g(&g, links, spins, s, h);
    double d = or_spin_parallel(spins + 3 * s, h);
    if (d > dmax) dmax = d;
    dsum += d;
  }
  *diff_max = dmax;
  *diff_avg = dsum / (double)g.V;
}

/* make_maximization_final (mag.cpp:1800-1816).  max_iter caps the loop,
 * which the reference leaves unbounded; returns iterations or -1. */
int or_su2_maximize_final(const double *links, double *spins, const int *dims,
                          int or_steps, double tol_max, double tol_avg,
                          int max_iter) {
  int it = 0;
  int converged = 0;
  do {
    double dmax, davg;
    or_su2_relax_sweep_lexi(links, spins, dims, &dmax, &davg);
    converged = (dmax < tol_max) && (davg < tol_avg);
    for (int i = 0; i < or_steps; i++) or_su2_or_sweep_lexi(links, spins, dims);
    it++;
  } while (!converged && it < max_iter);
  return converged ? it : -1;
}

/* the stopping test of make_maximization_approximate (mag.cpp:1744-1746):
 * the leading `digits` decimals of F, scaled by the single-precision
 * powf(10, digits) (inexact from 11 digits on) and truncated towards zero,
 * are equal, and so is the sign bit */
int or_approx_same(double functional, double functional_old, int tolerance_digits) {
  return (trunc(powf(10.f, (float)tolerance_digits) * functional) ==
          trunc(powf(10.f, (float)tolerance_digits) * functional_old)) &&
         (signbit(functional) == signbit(functional_old));
}

/* make_maximization_approximate (mag.cpp:1733-1750), powf quirk kept */
int or_su2_maximize_approximate(const double *links, double *spins,
                                const int *dims, int or_steps,
                                int tolerance_digits, int max_iter) {
  double functional = or_su2_functional_spin(links, spins, dims);
  double functional_old = functional;
  int is_equal = 0;
  int it = 0;
  do {
    double dmax, davg;
    or_su2_relax_sweep_lexi(links, spins, dims, &dmax, &davg);
    for (int i = 0; i < or_steps; i++) or_su2_or_sweep_lexi(links, spins, dims);
    functional = or_su2_functional_spin(links, spins, dims);
    is_equal = or_approx_same(functional, functional_old, tolerance_digits);
    functional_old = functional;
    it++;
  } while (!is_equal && it < max_iter);
  return is_equal ? it : -1;
}

/* ------------------------------------------------------------------ */
/* angle functions of the omega rule                                    */
/* ------------------------------------------------------------------ */
/* A fixed sequence of IEEE operations (the library is compiled with
 * -ffp-contract=off, no fma), the same sequence as the device's:
 *   atan2 for y > 0: fold onto t in [0, 1], t = c + ..., c = k/4 nearest,
 *     atan t = atan(k/4) + atan((t - c)/(1 + t c)), Taylor to u^19;
 *   sin / cos on [0, 2 pi]: Cody-Waite reduction by pi/2 (two-part
 *     constant), Taylor to r^17 / r^18, quadrant.
 * Both are checked against the C library's atan2 / sin / cos. */
static double atan_small(double u) {
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

static double atan01(double t) {
  static const double atan_quarter[5] = {0.0, 0.24497866312686414, 0.4636476090008061,
                                         0.6435011087932844, 0.7853981633974483};
  const int k = (int)(t * 4.0 + 0.5);
  const double c = 0.25 * (double)k;
  const double u = (t - c) / (1.0 + t * c);
  return atan_quarter[k] + atan_small(u);
}

double or_atan2_pos(double y, double x) {
  const double ax = fabs(x);
  const double pio2 = 1.5707963267948966, pi = 3.141592653589793;
  if (y <= ax) {
    const double a = atan01(y / ax);
    return x >= 0.0 ? a : pi - a;
  }
  const double a = atan01(ax / y);
  return x >= 0.0 ? pio2 - a : pio2 + a;
}

void or_sincos(double a, double *s, double *c) {
  const int q = (int)(a * 0.6366197723675814 + 0.5);
  const double qd = (double)q;
  const double r = (a - qd * 1.5707963267948966) - qd * 6.123233995736766e-17;
  const double z = r * r;
  double ps = 1.0 / 355687428096000.0;
  ps = ps * z + -1.0 / 1307674368000.0;
  ps = ps * z + 1.0 / 6227020800.0;
  ps = ps * z + -1.0 / 39916800.0;
  ps = ps * z + 1.0 / 362880.0;
  ps = ps * z + -1.0 / 5040.0;
  ps = ps * z + 1.0 / 120.0;
  ps = ps * z + -1.0 / 6.0;
  const double sr = r + (r * z) * ps;
  double pc = -1.0 / 6402373705728000.0;
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

/* ------------------------------------------------------------------ */
/* omega rule (build-defined geodesic rule)                             */
/* ------------------------------------------------------------------ */
/* Site gauge g (quaternion, g3 = 0) that turns n = e3 towards h_hat by
 * omega*theta, theta = angle(e3, h): g = (cos(phi), sin(phi) m),
 * phi = omega*theta/2, m = (-h2, h1, 0)/|h_perp|.  Then g^+ s3 g = n'.s with
 * n' the rotated e3.  omega=1 gives n' = h_hat, omega=2 the reflection. */
void or_omega_gauge(const double *h, double omega, double *g) {
  double hp2 = h[0] * h[0] + h[1] * h[1];
  double hp = sqrt(hp2);
  if (hp == 0.0) {
    if (h[2] >= 0.0) { /* already aligned (or h = 0): identity */
      g[0] = 1; g[1] = 0; g[2] = 0; g[3] = 0;
      return;
    }
    /* anti-aligned: rotate about e1 */
    double phi = 0.5 * omega * M_PI;
    or_sincos(phi, &g[1], &g[0]);
    g[2] = 0; g[3] = 0;
    return;
  }
  double theta = or_atan2_pos(hp, h[2]);
  double phi = 0.5 * omega * theta;
  double sp, cp;
  or_sincos(phi, &sp, &cp);
  double inv = 1.0 / hp;
  g[0] = cp;
  g[1] = -sp * h[1] * inv;
  g[2] = sp * h[0] * inv;
  g[3] = 0;
}

/* 1 - cos(theta) with theta = angle(e3, h), cancellation-free */
static inline double resid_of(const double *h) {
  double hp2 = h[0] * h[0] + h[1] * h[1];
  double hn = sqrt(hp2 + h[2] * h[2]);
  if (hn == 0.0) return 0.0;
  if (h[2] >= 0) return hp2 / (hn * (hn + h[2]));
  return 1.0 - h[2] / hn;
}

/* spin-form omega rule: s' = cos(w th) s + sin(w th) t_hat */
static void spin_omega(double *s, const double *h, double omega) {
  double hn = sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
  if (hn == 0) return;
  double hh[3] = {h[0] / hn, h[1] / hn, h[2] / hn};
  double c = s[0] * hh[0] + s[1] * hh[1] + s[2] * hh[2];
  double t[3] = {hh[0] - c * s[0], hh[1] - c * s[1], hh[2] - c * s[2]};
  double tn = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
  if (tn == 0) {
    if (c >= 0) return;
    /* anti-aligned: pick any perpendicular direction */
    double e[3] = {1, 0, 0};
    if (fabs(s[0]) > 0.9) { e[0] = 0; e[1] = 1; }
    double d = e[0] * s[0] + e[1] * s[1] + e[2] * s[2];
    t[0] = e[0] - d * s[0]; t[1] = e[1] - d * s[1]; t[2] = e[2] - d * s[2];
    tn = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
  }
  double sn = tn; /* sin(theta) */
  double theta = atan2(sn, c);
  double a = omega * theta;
  double ca = cos(a), sa = sin(a);
  for (int k = 0; k < 3; k++) s[k] = ca * s[k] + sa * t[k] / tn;
  /* back onto the sphere: for omega > 1 this step multiplies a spin's
   * deviation from unit length by about 2 omega - 1 per update, so rounding
   * would grow exponentially over the sweeps without it */
  double sn2 = 1.0 / sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
  for (int k = 0; k < 3; k++) s[k] *= sn2;
}

/* ------------------------------------------------------------------ */
/* checkerboard spin-form half-sweeps (build-defined order)            */
/* ------------------------------------------------------------------ */
static inline int parity_of(const int *c) { return (c[0] + c[1] + c[2] + c[3]) & 1; }

void or_su2_spin_halfsweep(const double *links, double *spins, const int *dims,
                           int parity, int mode, double omega,
                           double *diff_max, double *diff_sum) {
  geom_t g = geom(dims);
  int64_t rows = g.V / g.L[0];
  double *rsum = (double *)calloc((size_t)rows, sizeof(double));
  double *rmax = (double *)calloc((size_t)rows, sizeof(double));
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < rows; r++) {
    for (int x = 0; x < g.L[0]; x++) {
      int64_t s = r * g.L[0] + x;
      int c[4];
      coords(&g, s, c);
      if (parity_of(c) != parity) continue;
      double h[3];
      contribution_site_g(&g, links, spins, s, h);
      double *sp = spins + 3 * s;
      double hn = sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
      double d = hn > 0 ? 1 - (sp[0] * h[0] + sp[1] * h[1] + sp[2] * h[2]) / hn : 0;
      if (mode == 0) or_spin_reflect_fast(sp, h);
      else if (mode == 1) or_spin_parallel(sp, h);
      else spin_omega(sp, h, omega);
      rsum[r] += d;
      if (d > rmax[r]) rmax[r] = d;
    }
  }
  double s = 0, m = 0;
  for (int64_t r = 0; r < rows; r++) {
    s += rsum[r];
    if (rmax[r] > m) m = rmax[r];
  }
  free(rsum);
  free(rmax);
  if (diff_max) *diff_max = m;
  if (diff_sum) *diff_sum = s;
}

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)                   */
/* ------------------------------------------------------------------ */
void or_philox4x32_10(const uint32_t *ctr, const uint32_t *key, uint32_t *out) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* stream tags of the synthetic generators */
enum { TAG_LINK = 1, TAG_GAUGE = 2, TAG_SPIN = 3, TAG_HEATBATH = 4 };

/* uniform in (0,1) from 64 random bits (53-bit mantissa, centred) */
static inline double u53(uint32_t a, uint32_t b) {
  uint64_t x = (((uint64_t)a << 32) | b) >> 11;
  return ((double)x + 0.5) * (1.0 / 9007199254740992.0);
}

static inline void philox_draw(uint64_t seed, uint32_t index, uint32_t draw,
                               uint32_t tag, uint32_t *o) {
  uint32_t ctr[4] = {index, draw, tag, 0};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  or_philox4x32_10(ctr, key, o);
}

/* two standard normals from one Philox block (Box-Muller) */
static inline void gauss2(uint64_t seed, uint32_t index, uint32_t draw,
                          uint32_t tag, double *z) {
  uint32_t o[4];
  philox_draw(seed, index, draw, tag, o);
  double u1 = u53(o[0], o[1]);
  double u2 = u53(o[2], o[3]);
  double r = sqrt(-2.0 * log(u1));
  double a = 2.0 * M_PI * u2;
  z[0] = r * cos(a);
  z[1] = r * sin(a);
}

/* ------------------------------------------------------------------ */
/* SU(2) synthetic configuration (build-defined)                       */
/* ------------------------------------------------------------------ */
/* U = exp(i eps a.sigma) = (cos(eps|a|), sin(eps|a|) a_hat), a ~ N(0,1)^3,
 * then U'_xmu = g(x) U_xmu g(x+mu)^+ with Haar g (normalised N(0,1)^4),
 * the transformation form of mag.cpp:1056-1062. */
/* key of gauge copy `copy` of the configuration seeded by `seed`: the same
 * link draws under a different random gauge (copy 0 = the configuration) */
static uint64_t gauge_copy_seed(uint64_t seed, uint32_t copy) {
  return seed ^ ((uint64_t)copy * 0x9E3779B97F4A7C15ull);
}

void or_gen_su2_config(double *links, const int *dims, double eps,
                       uint64_t seed) {
  or_gen_su2_config_copy(links, dims, eps, seed, 0);
}

/* Haar SU(2) site gauge: normalised N(0,1)^4 */
static void su2_site_gauge(uint64_t gseed, uint32_t site, double *q) {
  double z[4];
  gauss2(gseed, site, 0, TAG_GAUGE, z);
  gauss2(gseed, site, 1, TAG_GAUGE, z + 2);
  double n = sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2] + z[3] * z[3]);
  for (int k = 0; k < 4; k++) q[k] = z[k] / n;
}

void or_gen_su2_config_copy(double *links, const int *dims, double eps,
                            uint64_t seed, uint32_t copy) {
  geom_t g = geom(dims);
  const uint64_t gseed = gauge_copy_seed(seed, copy);
  double *gauge = (double *)malloc(sizeof(double) * 4 * (size_t)g.V);
#pragma omp parallel for schedule(static)
  for (int64_t s = 0; s < g.V; s++) su2_site_gauge(gseed, (uint32_t)s, gauge + 4 * s);
#pragma omp parallel for schedule(static)
  for (int64_t s = 0; s < g.V; s++) {
    int c[4];
    coords(&g, s, c);
    for (int mu = 0; mu < 4; mu++) {
      uint32_t l = (uint32_t)(4 * s + mu);
      double z[4];
      gauss2(seed, l, 0, TAG_LINK, z);
      gauss2(seed, l, 1, TAG_LINK, z + 2);
      double an = sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
      q_t U;
      if (an > 0) {
        double sa = sin(eps * an) / an;
        U.a0 = cos(eps * an);
        U.a1 = sa * z[0];
        U.a2 = sa * z[1];
        U.a3 = sa * z[2];
      } else {
        U.a0 = 1; U.a1 = U.a2 = U.a3 = 0;
      }
      int64_t f = nb(&g, s, c, mu, +1);
      q_t A = qmul(qload(gauge + 4 * s), U);
      A = qmul_dag(A, qload(gauge + 4 * f));
      qstore(links + 4 * (4 * s + mu), A);
    }
  }
  free(gauge);
}

/* Starting spins of gauge copy `copy` in the spin form, for the links of
 * copy 0: copy c of the configuration is g_c U g_c^+ = h U_0 h^+ with
 * h = g_c g_0^+ (U_0 = copy 0), so the spin form on U_0 follows it from the
 * spin of h, row 3 of its rotation
 *   (2(a1 a3 + a0 a2), 2(a2 a3 - a0 a1), a0^2 + a3^2 - a1^2 - a2^2);
 * copy 0 starts exactly at e3. */
void or_gen_copy_spins(double *spins, const int *dims, uint64_t seed, uint32_t copy) {
  geom_t g = geom(dims);
#pragma omp parallel for schedule(static)
  for (int64_t s = 0; s < g.V; s++) {
    double *o = spins + 3 * s;
    if (copy == 0) {
      o[0] = 0;
      o[1] = 0;
      o[2] = 1;
      continue;
    }
    double gc[4], g0[4];
    su2_site_gauge(gauge_copy_seed(seed, copy), (uint32_t)s, gc);
    su2_site_gauge(gauge_copy_seed(seed, 0), (uint32_t)s, g0);
    q_t h = qmul_dag(qload(gc), qload(g0));
    o[0] = 2 * (h.a1 * h.a3 + h.a0 * h.a2);
    o[1] = 2 * (h.a2 * h.a3 - h.a0 * h.a1);
    o[2] = h.a0 * h.a0 + h.a3 * h.a3 - h.a1 * h.a1 - h.a2 * h.a2;
  }
}

/* uniform spins on S^2 (mag.cpp:166-236 Marsaglia construction, on a
 * Philox stream instead of the time-seeded ranlux of mag.cpp:143-164) */
void or_gen_spins_uniform(double *spins, const int *dims, uint64_t seed) {
  geom_t g = geom(dims);
#pragma omp parallel for schedule(static)
  for (int64_t s = 0; s < g.V; s++) {
    double x1, x2;
    uint32_t attempt = 0;
    do {
      uint32_t o[4];
      philox_draw(seed, (uint32_t)s, attempt++, TAG_SPIN, o);
      x1 = 2 * u53(o[0], o[1]) - 1;
      x2 = 2 * u53(o[2], o[3]) - 1;
    } while (x1 * x1 + x2 * x2 >= 1 && attempt < 64u); /* bounded, (1-pi/4)^64 ~ 1e-43 */
    double a = 1 - x1 * x1 - x2 * x2;
    spins[3 * s] = 2 * x1 * sqrt(a);
    spins[3 * s + 1] = 2 * x2 * sqrt(a);
    spins[3 * s + 2] = 1 - 2 * (x1 * x1 + x2 * x2);
  }
}

/* heat bath update of one spin (mag.cpp:402-433).  p(x) ~ exp(k x) on
 * [-1,1], k = |h|/T, is drawn by inversion instead of the reference's
 * rejection loop (same distribution, bounded work). */
static void heatbath_spin(double *s, const double *h, double T, uint64_t seed,
                          uint32_t site, uint32_t sweep) {
  uint32_t o[4];
  uint32_t ctr[4] = {site, sweep, TAG_HEATBATH, 0};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  or_philox4x32_10(ctr, key, o);
  double u = u53(o[0], o[1]);
  double phi = 2 * M_PI * u53(o[2], o[3]);
  double hn = sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
  double k = hn / T;
  double a3;
  if (k > 0) a3 = 1 + log1p(-(1 - u) * (-expm1(-2 * k))) / k;
  else a3 = 2 * u - 1;
  if (a3 < -1) a3 = -1;
  if (a3 > 1) a3 = 1;
  double R = sqrt(1 - a3 * a3);
  double a1 = R * cos(phi), a2 = R * sin(phi);
  if (hn == 0) {
    s[0] = a1; s[1] = a2; s[2] = a3;
    return;
  }
  double n3 = h[2] / hn;
  double c = sqrt(1 - n3 * n3);
  if (c == 0) { /* field along +-e3: any frame with third axis h_hat */
    s[0] = n3 * a2; s[1] = -a1; s[2] = n3 * a3;
    return;
  }
  double c1 = 1. / (hn * c);
  double f1 = h[0] * c1, f2 = h[1] * c1;
  s[0] = f2 * a1 + f1 * n3 * a2 + f1 * c * a3;
  s[1] = -f1 * a1 + f2 * n3 * a2 + f2 * c * a3;
  s[2] = -c * a2 + n3 * a3;
}

void or_su2_heatbath_halfsweep(const double *links, double *spins,
                               const int *dims, int parity, double temperature,
                               uint64_t seed, uint32_t sweep) {
  geom_t g = geom(dims);
#pragma omp parallel for schedule(static)
  for (int64_t s = 0; s < g.V; s++) {
    int c[4];
    coords(&g, s, c);
    if (parity_of(c) != parity) continue;
    double h[3];
    contribution_site_g(&g, links, spins, s, h);
    heatbath_spin(spins + 3 * s, h, temperature, seed, (uint32_t)s, sweep);
  }
}

/* ------------------------------------------------------------------ */
/* link-form SU(2) checkerboard omega-OR (build-defined)               */
/* ------------------------------------------------------------------ */
/* forward field term of U: M(U) e3 (= contribution1 with b = e3) */
static inline void field_fwd(q_t A, double *h) {
  h[0] += 2 * (A.a1 * A.a3 - A.a0 * A.a2);
  h[1] += 2 * (A.a2 * A.a3 + A.a0 * A.a1);
  h[2] += A.a0 * A.a0 - A.a1 * A.a1 - A.a2 * A.a2 + A.a3 * A.a3;
}
/* backward field term of U: M(U)^T e3 (= contribution1_conj with b = e3) */
static inline void field_bwd(q_t A, double *h) {
  h[0] += 2 * (A.a1 * A.a3 + A.a0 * A.a2);
  h[1] += 2 * (A.a2 * A.a3 - A.a0 * A.a1);
  h[2] += A.a0 * A.a0 - A.a1 * A.a1 - A.a2 * A.a2 + A.a3 * A.a3;
}

static void su2_link_halfsweep(const geom_t *g, double *links, int parity,
                               double omega, double *rsum, double *rmax) {
  int64_t rows = g->V / g->L[0];
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < rows; r++) {
    for (int x = 0; x < g->L[0]; x++) {
      int64_t s = r * g->L[0] + x;
      int c[4];
      coords(g, s, c);
      if (parity_of(c) != parity) continue;
      int64_t bn[4];
      /* h = [f(0)+b(0)+f(1)+b(1)] + [f(2)+b(2)+f(3)+b(3)] */
      double hp[2][3] = {{0, 0, 0}, {0, 0, 0}};
      for (int mu = 0; mu < 4; mu++) {
        bn[mu] = nb(g, s, c, mu, -1);
        field_fwd(qload(links + 4 * (4 * s + mu)), hp[mu >> 1]);
        field_bwd(qload(links + 4 * (4 * bn[mu] + mu)), hp[mu >> 1]);
      }
      double h[3] = {hp[0][0] + hp[1][0], hp[0][1] + hp[1][1], hp[0][2] + hp[1][2]};
      double d = resid_of(h);
      rsum[r] += d;
      if (d > rmax[r]) rmax[r] = d;
      double gq[4];
      or_omega_gauge(h, omega, gq);
      q_t G = qload(gq);
      for (int mu = 0; mu < 4; mu++) {
        double *pf = links + 4 * (4 * s + mu);
        double *pb = links + 4 * (4 * bn[mu] + mu);
        qstore(pf, qmul(G, qload(pf)));
        qstore(pb, qmul_dag(qload(pb), G));
      }
    }
  }
}

void or_su2_link_sweep(double *links, const int *dims, double omega, double *F,
                       double *rmax, double *ravg) {
  geom_t g = geom(dims);
  int64_t rows = g.V / g.L[0];
  double *rs = (double *)calloc((size_t)rows, sizeof(double));
  double *rm = (double *)calloc((size_t)rows, sizeof(double));
  su2_link_halfsweep(&g, links, 0, omega, rs, rm);
  su2_link_halfsweep(&g, links, 1, omega, rs, rm);
  double s = 0, m = 0;
  for (int64_t r = 0; r < rows; r++) {
    s += rs[r];
    if (rm[r] > m) m = rm[r];
  }
  free(rs);
  free(rm);
  if (rmax) *rmax = m;
  if (ravg) *ravg = s / (double)g.V;
  if (F) *F = or_su2_functional(links, dims);
}

/* ------------------------------------------------------------------ */
/* complex helpers and SU(3)                                            */
/* ------------------------------------------------------------------ */
typedef struct { double re, im; } cx;
static inline cx cmk(double r, double i) { cx z = {r, i}; return z; }
static inline cx cadd(cx a, cx b) { return cmk(a.re + b.re, a.im + b.im); }
static inline cx csub(cx a, cx b) { return cmk(a.re - b.re, a.im - b.im); }
static inline cx cmul(cx a, cx b) {
  return cmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline cx cmulc(cx a, cx b) { /* a * conj(b) */
  return cmk(a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im);
}
static inline cx cconj(cx a) { return cmk(a.re, -a.im); }
static inline cx cscale(cx a, double s) { return cmk(a.re * s, a.im * s); }
static inline double cnorm2(cx a) { return a.re * a.re + a.im * a.im; }

typedef struct { cx m[3][3]; } m3;
static inline m3 m3load(const double *p) {
  m3 M;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) M.m[i][j] = cmk(p[2 * (3 * i + j)], p[2 * (3 * i + j) + 1]);
  return M;
}
static inline void m3store(double *p, const m3 *M) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      p[2 * (3 * i + j)] = M->m[i][j].re;
      p[2 * (3 * i + j) + 1] = M->m[i][j].im;
    }
}
static m3 m3mul(const m3 *A, const m3 *B) {
  m3 C;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      cx s = cmk(0, 0);
      for (int k = 0; k < 3; k++) s = cadd(s, cmul(A->m[i][k], B->m[k][j]));
      C.m[i][j] = s;
    }
  return C;
}
static m3 m3mul_dag(const m3 *A, const m3 *B) { /* A * B^+ */
  m3 C;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      cx s = cmk(0, 0);
      for (int k = 0; k < 3; k++) s = cadd(s, cmulc(A->m[i][k], B->m[j][k]));
      C.m[i][j] = s;
    }
  return C;
}
static m3 m3dag(const m3 *A) {
  m3 C;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) C.m[i][j] = cconj(A->m[j][i]);
  return C;
}
static cx m3det(const m3 *A) {
  const cx(*a)[3] = A->m;
  cx t0 = csub(cmul(a[1][1], a[2][2]), cmul(a[1][2], a[2][1]));
  cx t1 = csub(cmul(a[1][0], a[2][2]), cmul(a[1][2], a[2][0]));
  cx t2 = csub(cmul(a[1][0], a[2][1]), cmul(a[1][1], a[2][0]));
  return cadd(csub(cmul(a[0][0], t0), cmul(a[0][1], t1)), cmul(a[0][2], t2));
}

void or_su3_mul(const double *A, const double *B, double *out) {
  m3 a = m3load(A), b = m3load(B), c = m3mul(&a, &b);
  m3store(out, &c);
}
void or_su3_mul_dag(const double *A, const double *B, double *out) {
  m3 a = m3load(A), b = m3load(B), c = m3mul_dag(&a, &b);
  m3store(out, &c);
}

/* Gram-Schmidt of rows 0,1; row 2 = conj(row0 x row1) (build-defined
 * reunitarisation, the one the 2-row storage implies) */
static void m3_reunit(m3 *M) {
  double n0 = 0;
  for (int k = 0; k < 3; k++) n0 += cnorm2(M->m[0][k]);
  n0 = 1.0 / sqrt(n0);
  for (int k = 0; k < 3; k++) M->m[0][k] = cscale(M->m[0][k], n0);
  cx d = cmk(0, 0); /* <row0, row1> = sum conj(r0) r1 */
  for (int k = 0; k < 3; k++) d = cadd(d, cmulc(M->m[1][k], M->m[0][k]));
  for (int k = 0; k < 3; k++) M->m[1][k] = csub(M->m[1][k], cmul(d, M->m[0][k]));
  double n1 = 0;
  for (int k = 0; k < 3; k++) n1 += cnorm2(M->m[1][k]);
  n1 = 1.0 / sqrt(n1);
  for (int k = 0; k < 3; k++) M->m[1][k] = cscale(M->m[1][k], n1);
  const cx *a = M->m[0], *b = M->m[1];
  M->m[2][0] = cconj(csub(cmul(a[1], b[2]), cmul(a[2], b[1])));
  M->m[2][1] = cconj(csub(cmul(a[2], b[0]), cmul(a[0], b[2])));
  M->m[2][2] = cconj(csub(cmul(a[0], b[1]), cmul(a[1], b[0])));
}
void or_su3_reunitarize(double *m18) {
  m3 M = m3load(m18);
  m3_reunit(&M);
  m3store(m18, &M);
}

/* su3::proj (matrix.h:279-292): 6 Newton polar steps, each followed by a
 * first-order determinant-phase correction */
void or_su3_proj(double *m18) {
  m3 M = m3load(m18);
  m3 A;
  for (int it = 0; it < 6; it++) {
    m3 Md = m3dag(&M);
    m3 t = m3mul(&M, &Md);
    m3 t2 = m3mul(&t, &M);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        A.m[i][j] = csub(cscale(M.m[i][j], 1.5), cscale(t2.m[i][j], 0.5));
    cx det = m3det(&A);
    cx f = csub(cmk(1, 0), cscale(csub(det, cmk(1, 0)), 1.0 / 3));
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) A.m[i][j] = cmul(f, A.m[i][j]);
    M = A;
  }
  m3store(m18, &M);
}

/* mag_functional_su3 (mag.cpp:1837-1862): sum_links sum_j |U_jj|^2 / (12 V),
 * with V taken from dims (the reference's globals are never set, 3.3) */
double or_su3_functional(const double *links18, const int *dims) {
  geom_t g = geom(dims);
  double f = 0;
  for (int64_t l = 0; l < 4 * g.V; l++) {
    const double *p = links18 + 18 * l;
    for (int j = 0; j < 3; j++) {
      double re = p[2 * (3 * j + j)], im = p[2 * (3 * j + j) + 1];
      f += re * re + im * im;
    }
  }
  return f / (double)(g.V * 3 * 4);
}

/* SU(3) exp(i eps sum_a c_a lambda_a / 2), Gell-Mann basis of
 * get_generators_su3 (matrix.h:987-1046); Taylor series with scaling and
 * squaring, then Gram-Schmidt so that the 2-row form is exact */
void or_gen_su3_exp(const double *c, double eps, double *m18) {
  const double r3 = 1.0 / sqrt(3.0);
  double e = 0.5 * eps;
  /* H = e * sum c_a lambda_a (hermitian) */
  m3 H;
  memset(&H, 0, sizeof(H));
  H.m[0][1] = cadd(H.m[0][1], cmk(e * c[0], 0));
  H.m[1][0] = cadd(H.m[1][0], cmk(e * c[0], 0));
  H.m[0][1] = cadd(H.m[0][1], cmk(0, -e * c[1]));
  H.m[1][0] = cadd(H.m[1][0], cmk(0, e * c[1]));
  H.m[0][0] = cadd(H.m[0][0], cmk(e * c[2], 0));
  H.m[1][1] = cadd(H.m[1][1], cmk(-e * c[2], 0));
  H.m[0][2] = cadd(H.m[0][2], cmk(e * c[3], 0));
  H.m[2][0] = cadd(H.m[2][0], cmk(e * c[3], 0));
  H.m[0][2] = cadd(H.m[0][2], cmk(0, -e * c[4]));
  H.m[2][0] = cadd(H.m[2][0], cmk(0, e * c[4]));
  H.m[1][2] = cadd(H.m[1][2], cmk(e * c[5], 0));
  H.m[2][1] = cadd(H.m[2][1], cmk(e * c[5], 0));
  H.m[1][2] = cadd(H.m[1][2], cmk(0, -e * c[6]));
  H.m[2][1] = cadd(H.m[2][1], cmk(0, e * c[6]));
  H.m[0][0] = cadd(H.m[0][0], cmk(e * c[7] * r3, 0));
  H.m[1][1] = cadd(H.m[1][1], cmk(e * c[7] * r3, 0));
  H.m[2][2] = cadd(H.m[2][2], cmk(-2 * e * c[7] * r3, 0));
  /* A = i H */
  m3 A;
  double nrm = 0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      A.m[i][j] = cmk(-H.m[i][j].im, H.m[i][j].re);
      nrm += cnorm2(A.m[i][j]);
    }
  nrm = sqrt(nrm);
  int sq = 0;
  while (nrm > 0.25) { nrm *= 0.5; sq++; }
  double sc = ldexp(1.0, -sq);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) A.m[i][j] = cscale(A.m[i][j], sc);
  m3 R, T;
  memset(&R, 0, sizeof(R));
  memset(&T, 0, sizeof(T));
  for (int i = 0; i < 3; i++) R.m[i][i] = T.m[i][i] = cmk(1, 0);
  for (int k = 1; k <= 20; k++) {
    T = m3mul(&T, &A);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) T.m[i][j] = cscale(T.m[i][j], 1.0 / k);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) R.m[i][j] = cadd(R.m[i][j], T.m[i][j]);
  }
  for (int k = 0; k < sq; k++) R = m3mul(&R, &R);
  m3_reunit(&R);
  m3store(m18, &R);
}

/* Haar SU(3): Gram-Schmidt of a complex Gaussian matrix (Haar on U(3),
 * Mezzadri 2007, R_ii > 0 by construction), times det^(-1/3) */
static void haar_su3(uint64_t seed, uint32_t site, m3 *G) {
  double z[18];
  for (int k = 0; k < 9; k++) gauss2(seed, site, (uint32_t)k, TAG_GAUGE, z + 2 * k);
  m3 Z;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) Z.m[i][j] = cmk(z[2 * (3 * i + j)], z[2 * (3 * i + j) + 1]);
  /* Gram-Schmidt on rows 0,1,2 */
  for (int i = 0; i < 3; i++) {
    for (int p = 0; p < i; p++) {
      cx d = cmk(0, 0);
      for (int k = 0; k < 3; k++) d = cadd(d, cmulc(Z.m[i][k], Z.m[p][k]));
      for (int k = 0; k < 3; k++) Z.m[i][k] = csub(Z.m[i][k], cmul(d, Z.m[p][k]));
    }
    double n = 0;
    for (int k = 0; k < 3; k++) n += cnorm2(Z.m[i][k]);
    n = 1.0 / sqrt(n);
    for (int k = 0; k < 3; k++) Z.m[i][k] = cscale(Z.m[i][k], n);
  }
  cx det = m3det(&Z);
  double ph = atan2(det.im, det.re) / 3.0;
  cx f = cmk(cos(ph), -sin(ph));
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) Z.m[i][j] = cmul(Z.m[i][j], f);
  m3_reunit(&Z);
  *G = Z;
}

void or_gen_su3_config(double *links18, const int *dims, double eps,
                       uint64_t seed) {
  or_gen_su3_config_copy(links18, dims, eps, seed, 0);
}

void or_gen_su3_config_copy(double *links18, const int *dims, double eps,
                            uint64_t seed, uint32_t copy) {
  geom_t g = geom(dims);
  const uint64_t gseed = gauge_copy_seed(seed, copy);
  m3 *gauge = (m3 *)malloc(sizeof(m3) * (size_t)g.V);
#pragma omp parallel for schedule(static)
  for (int64_t s = 0; s < g.V; s++) haar_su3(gseed, (uint32_t)s, gauge + s);
#pragma omp parallel for schedule(static)
  for (int64_t s = 0; s < g.V; s++) {
    int c[4];
    coords(&g, s, c);
    for (int mu = 0; mu < 4; mu++) {
      uint32_t l = (uint32_t)(4 * s + mu);
      double cc[8];
      for (int k = 0; k < 4; k++) gauss2(seed, l, (uint32_t)k, TAG_LINK, cc + 2 * k);
      double u[18];
      or_gen_su3_exp(cc, eps, u);
      m3 U = m3load(u);
      int64_t f = nb(&g, s, c, mu, +1);
      m3 A = m3mul(gauge + s, &U);
      A = m3mul_dag(&A, gauge + f);
      m3_reunit(&A);
      m3store(links18 + 18 * (4 * s + mu), &A);
    }
  }
  free(gauge);
}

/* ------------------------------------------------------------------ */
/* SU(3) link-form checkerboard Cabibbo-Marinari omega-OR (build-defined) */
/* ------------------------------------------------------------------ */
/* Forward field of block B=[[p,q],[s,t]]: x_a = 1/2 tr(s_a B s3 B^+) */
static inline void blk_fwd(cx p, cx q, cx s, cx t, double *h) {
  cx X01 = csub(cmulc(p, s), cmulc(q, t));
  h[0] += X01.re;
  h[1] += -X01.im;
  h[2] += 0.5 * (cnorm2(p) - cnorm2(q) - cnorm2(s) + cnorm2(t));
}
/* Backward field of block B: y_a = 1/2 tr(s_a B^+ s3 B) */
static inline void blk_bwd(cx p, cx q, cx s, cx t, double *h) {
  cx Y01 = csub(cmulc(q, p), cmulc(t, s)); /* p^* q - s^* t */
  h[0] += Y01.re;
  h[1] += -Y01.im;
  h[2] += 0.5 * (cnorm2(p) - cnorm2(s) - cnorm2(q) + cnorm2(t));
}
void or_block_field_fwd(const double *B, double *h) {
  h[0] = h[1] = h[2] = 0;
  blk_fwd(cmk(B[0], B[1]), cmk(B[2], B[3]), cmk(B[4], B[5]), cmk(B[6], B[7]), h);
}
void or_block_field_bwd(const double *B, double *h) {
  h[0] = h[1] = h[2] = 0;
  blk_bwd(cmk(B[0], B[1]), cmk(B[2], B[3]), cmk(B[4], B[5]), cmk(B[6], B[7]), h);
}

/* Cabibbo-Marinari subgroup order */
static const int CM_I[3] = {0, 1, 0};
static const int CM_J[3] = {1, 2, 2};

/* 2-row storage (the build's SU(3) link form): row 2 = conj(row0 x row1),
 * recomputed from rows 0 and 1 wherever it is read */
static void m3_row2(m3 *M) {
  const cx *a = M->m[0], *b = M->m[1];
  M->m[2][0] = cconj(csub(cmul(a[1], b[2]), cmul(a[2], b[1])));
  M->m[2][1] = cconj(csub(cmul(a[2], b[0]), cmul(a[0], b[2])));
  M->m[2][2] = cconj(csub(cmul(a[0], b[1]), cmul(a[1], b[0])));
}

static void su3_link_halfsweep(const geom_t *g, double *links, int parity,
                               double omega, double *rsum, double *rmax) {
  int64_t rows = g->V / g->L[0];
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < rows; r++) {
    for (int x = 0; x < g->L[0]; x++) {
      int64_t s = r * g->L[0] + x;
      int c[4];
      coords(g, s, c);
      if (parity_of(c) != parity) continue;
      m3 F[4], B[4];
      int64_t bn[4];
      for (int mu = 0; mu < 4; mu++) {
        bn[mu] = nb(g, s, c, mu, -1);
        F[mu] = m3load(links + 18 * (4 * s + mu));
        B[mu] = m3load(links + 18 * (4 * bn[mu] + mu));
      }
      double dsite = 0;
      for (int sg = 0; sg < 3; sg++) {
        int i = CM_I[sg], j = CM_J[sg];
        /* rows 0 and 1 are the state; row 2 is rebuilt before a subgroup
         * that reads it (the updates below leave it stale) */
        if (j == 2)
          for (int mu = 0; mu < 4; mu++) {
            m3_row2(&F[mu]);
            m3_row2(&B[mu]);
          }
        /* t_mu = f(mu) + b(mu); h = (t_0 + t_1) + (t_2 + t_3) */
        double t[4][3];
        for (int mu = 0; mu < 4; mu++) {
          t[mu][0] = t[mu][1] = t[mu][2] = 0;
          blk_fwd(F[mu].m[i][i], F[mu].m[i][j], F[mu].m[j][i], F[mu].m[j][j], t[mu]);
          blk_bwd(B[mu].m[i][i], B[mu].m[i][j], B[mu].m[j][i], B[mu].m[j][j], t[mu]);
        }
        double h[3];
        for (int a = 0; a < 3; a++) h[a] = (t[0][a] + t[1][a]) + (t[2][a] + t[3][a]);
        dsite += resid_of(h);
        double gq[4];
        or_omega_gauge(h, omega, gq);
        /* r = [[g0 + i g3, g2 + i g1], [-g2 + i g1, g0 - i g3]] */
        cx r00 = cmk(gq[0], gq[3]), r01 = cmk(gq[2], gq[1]);
        cx r10 = cmk(-gq[2], gq[1]), r11 = cmk(gq[0], -gq[3]);
        for (int mu = 0; mu < 4; mu++) {
          /* forward: rows i,j <- r (rows i,j) */
          for (int k = 0; k < 3; k++) {
            cx a = F[mu].m[i][k], b = F[mu].m[j][k];
            F[mu].m[i][k] = cadd(cmul(r00, a), cmul(r01, b));
            F[mu].m[j][k] = cadd(cmul(r10, a), cmul(r11, b));
          }
          /* backward: cols i,j <- (cols i,j) r^+ */
          for (int k = 0; k < 3; k++) {
            cx a = B[mu].m[k][i], b = B[mu].m[k][j];
            B[mu].m[k][i] = cadd(cmulc(a, r00), cmulc(b, r01));
            B[mu].m[k][j] = cadd(cmulc(a, r10), cmulc(b, r11));
          }
        }
      }
      rsum[r] += dsite;
      if (dsite > rmax[r]) rmax[r] = dsite;
      for (int mu = 0; mu < 4; mu++) {
        m3_row2(&F[mu]);
        m3_row2(&B[mu]);
        m3store(links + 18 * (4 * s + mu), &F[mu]);
        m3store(links + 18 * (4 * bn[mu] + mu), &B[mu]);
      }
    }
  }
}

void or_su3_link_sweep(double *links18, const int *dims, double omega,
                       double *F, double *rmax, double *ravg) {
  geom_t g = geom(dims);
  int64_t rows = g.V / g.L[0];
  double *rs = (double *)calloc((size_t)rows, sizeof(double));
  double *rm = (double *)calloc((size_t)rows, sizeof(double));
  su3_link_halfsweep(&g, links18, 0, omega, rs, rm);
  su3_link_halfsweep(&g, links18, 1, omega, rs, rm);
  double s = 0, m = 0;
  for (int64_t r = 0; r < rows; r++) {
    s += rs[r];
    if (rm[r] > m) m = rm[r];
  }
  free(rs);
  free(rm);
  if (rmax) *rmax = m;
  if (ravg) *ravg = s / (double)g.V;
  if (F) *F = or_su3_functional(links18, dims);
}

/* ------------------------------------------------------------------ */
/* plaquette and Wilson loops                                           */
/* ------------------------------------------------------------------ */
/* plaket_site / plaket (plaket.h:24-42, 84-101), accumulator started at 0 */
double or_su2_plaquette(const double *links, const int *dims) {
  geom_t g = geom(dims);
  double tot = 0;
  for (int64_t s = 0; s < g.V; s++) {
    int c[4];
    coords(&g, s, c);
    double p = 0;
    for (int mu = 0; mu < 3; mu++)
      for (int nu = mu + 1; nu < 4; nu++) {
        int64_t smu = nb(&g, s, c, mu, +1), snu = nb(&g, s, c, nu, +1);
        q_t A = qload(links + 4 * (4 * s + mu));
        A = qmul(A, qload(links + 4 * (4 * smu + nu)));
        A = qmul_dag(A, qload(links + 4 * (4 * snu + mu)));
        q_t Bn = qload(links + 4 * (4 * s + nu));
        p += A.a0 * Bn.a0 + A.a1 * Bn.a1 + A.a2 * Bn.a2 + A.a3 * Bn.a3;
      }
    tot += p / 6;
  }
  return tot / (double)g.V;
}

double or_su3_plaquette(const double *links18, const int *dims) {
  geom_t g = geom(dims);
  double tot = 0;
  for (int64_t s = 0; s < g.V; s++) {
    int c[4];
    coords(&g, s, c);
    double p = 0;
    for (int mu = 0; mu < 3; mu++)
      for (int nu = mu + 1; nu < 4; nu++) {
        int64_t smu = nb(&g, s, c, mu, +1), snu = nb(&g, s, c, nu, +1);
        m3 A = m3load(links18 + 18 * (4 * s + mu));
        m3 X = m3load(links18 + 18 * (4 * smu + nu));
        A = m3mul(&A, &X);
        X = m3load(links18 + 18 * (4 * snu + mu));
        A = m3mul_dag(&A, &X);
        X = m3load(links18 + 18 * (4 * s + nu));
        A = m3mul_dag(&A, &X);
        p += (A.m[0][0].re + A.m[1][1].re + A.m[2][2].re) / 3;
      }
    tot += p / 6;
  }
  return tot / (double)g.V;
}

/* straight Wilson line of n links from s along mu */
static q_t su2_line(const geom_t *g, const double *links, int64_t s, int mu,
                    int n) {
  q_t L = {1, 0, 0, 0};
  for (int k = 0; k < n; k++) {
    L = qmul(L, qload(links + 4 * (4 * s + mu)));
    s = shift_n(g, s, mu, 1);
  }
  return L;
}
static m3 su3_line(const geom_t *g, const double *links, int64_t s, int mu,
                   int n) {
  m3 L;
  memset(&L, 0, sizeof(L));
  for (int i = 0; i < 3; i++) L.m[i][i] = cmk(1, 0);
  for (int k = 0; k < n; k++) {
    m3 U = m3load(links + 18 * (4 * s + mu));
    L = m3mul(&L, &U);
    s = shift_n(g, s, mu, 1);
  }
  return L;
}

/* wilson_plane_indexed_single_rxt (wilson_loops.h:728-754) at one (R,T):
 * mean over sites and the 3 (space, time) planes of
 * multiply_conj_tr(L_t(x,T) L_nu(x+T t) L_t(x+R nu)^+, L_nu(x,R)) */
double or_su2_wilson_loop(const double *links, const int *dims, int R, int T) {
  geom_t g = geom(dims);
  double tot = 0;
  for (int64_t s = 0; s < g.V; s++) {
    q_t Lt = su2_line(&g, links, s, 3, T);
    int64_t sT = shift_n(&g, s, 3, T);
    for (int nu = 0; nu < 3; nu++) {
      q_t W = qmul(Lt, su2_line(&g, links, sT, nu, R));
      W = qmul_dag(W, su2_line(&g, links, shift_n(&g, s, nu, R), 3, T));
      q_t Ls = su2_line(&g, links, s, nu, R);
      tot += W.a0 * Ls.a0 + W.a1 * Ls.a1 + W.a2 * Ls.a2 + W.a3 * Ls.a3;
    }
  }
  return tot / (double)(g.V * 3);
}

double or_su3_wilson_loop(const double *links18, const int *dims, int R,
                          int T) {
  geom_t g = geom(dims);
  double tot = 0;
  for (int64_t s = 0; s < g.V; s++) {
    m3 Lt = su3_line(&g, links18, s, 3, T);
    int64_t sT = shift_n(&g, s, 3, T);
    for (int nu = 0; nu < 3; nu++) {
      m3 X = su3_line(&g, links18, sT, nu, R);
      m3 W = m3mul(&Lt, &X);
      X = su3_line(&g, links18, shift_n(&g, s, nu, R), 3, T);
      W = m3mul_dag(&W, &X);
      X = su3_line(&g, links18, s, nu, R);
      W = m3mul_dag(&W, &X);
      tot += (W.m[0][0].re + W.m[1][1].re + W.m[2][2].re) / 3;
    }
  }
  return tot / (double)(g.V * 3);
}

/* ------------------------------------------------------------------ */
/* abelian projection and monopole currents                            */
/* ------------------------------------------------------------------ */
/* plane index of (mu, nu), mu < nu: (0,1)=0 (0,2)=1 (0,3)=2 (1,2)=3 (1,3)=4
 * (2,3)=5, the order of the mu/nu loops of monopoles.cpp:415-457 */
static int plane_of(int mu, int nu) {
  static const int P[4][4] = {{-1, 0, 1, 2}, {0, -1, 3, 4}, {1, 3, -1, 5}, {2, 4, 5, -1}};
  return P[mu][nu];
}

/* MatrixConverter<su2, abelian>: phi = atan2(a3, a0) (matrix.h:1059-1062) */
void or_su2_abelian_angles(const double *links, const int *dims, double *angles) {
  geom_t g = geom(dims);
  for (int64_t l = 0; l < 4 * g.V; l++) angles[l] = atan2(links[4 * l + 3], links[4 * l]);
}

/* MatrixConverter<su3, su3_angles> (matrix.h:1082-1100): the phases of the
 * diagonal, less a third of their sum taken into [-pi, pi).  Output [4V][3]
 * (the su3_angles element order of LatticeData) */
void or_su3_abelian_angles(const double *links18, const int *dims, double *angles3) {
  geom_t g = geom(dims);
  for (int64_t l = 0; l < 4 * g.V; l++) {
    const double *m = links18 + 18 * l;
    double a[3], sum = 0;
    for (int c = 0; c < 3; c++) {
      a[c] = atan2(m[2 * (4 * c) + 1], m[2 * (4 * c)]);
      sum += a[c];
    }
    while (sum >= M_PI) sum -= 2 * M_PI;
    while (sum < -M_PI) sum += 2 * M_PI;
    for (int c = 0; c < 3; c++) angles3[3 * l + c] = a[c] - sum / 3;
  }
}

/* plaquette angle of plane (mu,nu) at every site, taken into [-pi, pi]:
 * phi_mu(x) + phi_nu(x+mu) - phi_mu(x+nu) - phi_nu(x), in that order
 * (monopoles.cpp:415-457, the abelian operator* and ^ of matrix.h:194-199);
 * angles[stride * (4 * site + mu) + off], out[6][V] */
static void u1_plaquettes(const geom_t *g, const double *angles, int stride, int off,
                          double *out) {
  for (int mu = 0; mu < 4; mu++)
    for (int nu = mu + 1; nu < 4; nu++) {
      double *o = out + (int64_t)plane_of(mu, nu) * g->V;
      for (int64_t s = 0; s < g->V; s++) {
        int c[4];
        coords(g, s, c);
        const int64_t smu = nb(g, s, c, mu, +1), snu = nb(g, s, c, nu, +1);
        double p = angles[stride * (4 * s + mu) + off];
        p = p + angles[stride * (4 * smu + nu) + off];
        p = p - angles[stride * (4 * snu + mu) + off];
        p = p - angles[stride * (4 * s + nu) + off];
        while (p > M_PI) p -= 2 * M_PI;
        while (p < -M_PI) p += 2 * M_PI;
        o[s] = p;
      }
    }
}

void or_u1_monopole_plaquettes(const double *angles, const int *dims, double *out) {
  geom_t g = geom(dims);
  u1_plaquettes(&g, angles, 1, 0, out);
}

/* SU(3): per colour as above (monopoles.cpp:368-413), then the sum of the
 * three colours brought into (-pi, pi) by moving 2 pi off the largest
 * (sum >= pi) or onto the smallest (sum <= -pi) colour, later colours
 * winning ties (monopoles.cpp:844-884).  out[3][6][V] */
void or_su3_monopole_plaquettes(const double *angles3, const int *dims, double *out) {
  geom_t g = geom(dims);
  const int64_t n = 6 * g.V;
  for (int c = 0; c < 3; c++) u1_plaquettes(&g, angles3, 3, c, out + c * n);
  for (int64_t k = 0; k < n; k++) {
    double *p0 = out + k, *p1 = out + n + k, *p2 = out + 2 * n + k;
    const double sum = *p0 + *p1 + *p2;
    if (sum >= M_PI) {
      double *ext = p0;
      if (*p1 >= *ext) ext = p1;
      if (*p2 >= *ext) ext = p2;
      *ext -= 2 * M_PI;
    }
    if (sum <= -M_PI) {
      double *ext = p0;
      if (*p1 <= *ext) ext = p1;
      if (*p2 <= *ext) ext = p2;
      *ext += 2 * M_PI;
    }
  }
}

/* monopole current on the link (x, mu) of the dual lattice from the six
 * plaquette fields P[6][V] (get_current, monopoles.cpp:44-122): forward
 * differences of the three planes orthogonal to mu, taken at x + mu, over
 * 2 pi; J[4V] */
void or_monopole_currents(const double *P, const int *dims, double *J) {
  geom_t g = geom(dims);
  const int64_t V = g.V;
  for (int64_t s = 0; s < V; s++) {
    int64_t b[4], bx[4][4];
    for (int mu = 0; mu < 4; mu++) {
      b[mu] = shift_n(&g, s, mu, 1);
      for (int nu = 0; nu < 4; nu++) bx[mu][nu] = shift_n(&g, b[mu], nu, 1);
    }
#define PL(pl, site) P[(int64_t)(pl) * V + (site)]
    /* t: planes yz (3), xz (1), xy (0) */
    double j3 = PL(3, bx[3][0]) - PL(3, b[3]) - (PL(1, bx[3][1]) - PL(1, b[3])) +
                PL(0, bx[3][2]) - PL(0, b[3]);
    /* x: planes yz (3), zt (5), yt (4) */
    double j0 = -(PL(3, bx[0][3]) - PL(3, b[0])) - (PL(5, bx[0][1]) - PL(5, b[0])) +
                PL(4, bx[0][2]) - PL(4, b[0]);
    /* y: planes xz (1), zt (5), xt (2) */
    double j1 = PL(1, bx[1][3]) - PL(1, b[1]) + PL(5, bx[1][0]) - PL(5, b[1]) -
                (PL(2, bx[1][2]) - PL(2, b[1]));
    /* z: planes xy (0), yt (4), xt (2) */
    double j2 = -(PL(0, bx[2][3]) - PL(0, b[2])) - (PL(4, bx[2][0]) - PL(4, b[2])) +
                PL(2, bx[2][1]) - PL(2, b[2]);
#undef PL
    J[4 * s + 0] = j0 / 2 / M_PI;
    J[4 * s + 1] = j1 / 2 / M_PI;
    J[4 * s + 2] = j2 / 2 / M_PI;
    J[4 * s + 3] = j3 / 2 / M_PI;
  }
}

/* ------------------------------------------------------------------ */
/* monopole clusters (monopoles.cpp:1120-1303, 1343-1449)               */
/* ------------------------------------------------------------------ */
/* The reference's walk, step for step: sites in lexicographic order; a
 * site starts a cluster when one of its 8 links carries |J| > 0.3
 * (find_current, 1120-1138); from the cluster's node list it follows, for
 * every node in turn, the forward links (x, mu) with J > 0.3, then the
 * backward links (x - mu, mu) with J < -0.3, mu = 0..3, zeroing each link
 * it takes and adding the far site as a new node with charge lround(J)
 * (find_paths, 1185-1228), until a pass adds no node (find_cluster,
 * 1241-1250).  Per cluster it keeps what the observables sum over the
 * tree's edges: cluster_length = sum |charge| (1343-1354), length_mu[mu] =
 * sum charge over edges in direction mu (1380-1398), currents_directions =
 * (sum |charge| over mu < 3, over mu = 3) (1423-1447), and the number of
 * distinct sites of the nodes (site_number, 1510-1531).
 * rec[9 n + .] = {first site, length, length_mu[4], space, time, sites};
 * returns the number of clusters (records beyond cap are not written). */
int64_t or_monopole_clusters(const double *J_in, const int *dims, int32_t *rec, int64_t cap) {
  geom_t g = geom(dims);
  const int64_t V = g.V;
  double *J = (double *)malloc(sizeof(double) * 4 * V);
  int64_t *cur = (int64_t *)malloc(sizeof(int64_t) * 8 * V + 8);
  int64_t *nxt = (int64_t *)malloc(sizeof(int64_t) * 8 * V + 8);
  int64_t *mark = (int64_t *)malloc(sizeof(int64_t) * V);
  memcpy(J, J_in, sizeof(double) * 4 * V);
  for (int64_t s = 0; s < V; s++) mark[s] = -1;
  int64_t n = 0;
  for (int64_t s = 0; s < V; s++) {
    int start = 0;
    for (int mu = 0; mu < 4 && !start; mu++) {
      if (fabs(J[4 * s + mu]) > 0.3) start = 1;
      if (fabs(J[4 * shift_n(&g, s, mu, -1) + mu]) > 0.3) start = 1;
    }
    if (!start) continue;
    int32_t length = 0, lmu[4] = {0, 0, 0, 0}, space = 0, time = 0, sites = 1;
    mark[s] = n;
    int64_t ncur = 1, nnext;
    cur[0] = s;
    do {
      nnext = 0;
      for (int64_t i = 0; i < ncur; i++) {
        const int64_t x = cur[i];
        for (int dir = 0; dir < 8; dir++) {
          const int mu = dir & 3;
          const int64_t l = dir < 4 ? 4 * x + mu : 4 * shift_n(&g, x, mu, -1) + mu;
          const double j = J[l];
          if (dir < 4 ? !(j > 0.3) : !(j < -0.3)) continue;
          const int32_t q = (int32_t)lround(j);
          J[l] = 0.;
          const int64_t y = shift_n(&g, x, mu, dir < 4 ? 1 : -1);
          nxt[nnext++] = y;
          length += abs(q);
          lmu[mu] += q;
          if (mu == 3) time += abs(q);
          else space += abs(q);
          if (mark[y] != n) {
            mark[y] = n;
            sites++;
          }
        }
      }
      int64_t *t = cur;
      cur = nxt;
      nxt = t;
      ncur = nnext;
    } while (ncur > 0);
    if (n < cap) {
      int32_t *r = rec + 9 * n;
      r[0] = (int32_t)s;
      r[1] = length;
      for (int mu = 0; mu < 4; mu++) r[2 + mu] = lmu[mu];
      r[6] = space;
      r[7] = time;
      r[8] = sites;
    }
    n++;
  }
  free(J);
  free(cur);
  free(nxt);
  free(mark);
  return n;
}
