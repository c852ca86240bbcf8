/*
 * mag_oracle.h -- CPU oracle for the MAG gauge-fixing path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in general_code_amd/: only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py may load it.  The product never links
 * it and never falls back to it.
 *
 * It is a plain-C restatement of the reference's CPU algorithm
 * (ilyakudrov/general_code, lib/cpu), plus the build-defined checkerboard
 * omega-overrelaxation sweeps that the GPU kernels run (the reference has
 * no checkerboard, no omega and no SU(3) fixing).  Every function names the
 * reference file:line it follows, or says "build-defined".
 *
 * Layouts (reference layout, lib/cpu/include/indexing.h:71-126):
 *   site  = x + Lx*(y + Ly*(z + Lz*t))
 *   link  = 4*site + mu
 *   su2   = 4 doubles {a0,a1,a2,a3} meaning a0 + i*(a1 s1 + a2 s2 + a3 s3)
 *           (lib/cpu/include/matrix.h:8-13)
 *   su3   = 18 doubles, row-major (re,im): element (i,j) at 2*(3*i+j)
 *           (lib/cpu/include/matrix.h:236-254)
 *   spin  = 3 doubles {a1,a2,a3} per site (matrix.h:773-787)
 *
 * Parity status of each function (how it is pinned) is listed in the
 * oracle header comment of mag_oracle.c.
 */
#ifndef MAG_ORACLE_H
#define MAG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- su2 algebra (matrix.h:9-117) ---------------- */
void or_su2_mul(const double *A, const double *B, double *out);     /* A*B   matrix.h:93-98   */
void or_su2_mul_dag(const double *A, const double *B, double *out); /* A*B^+ matrix.h:99-104  */
void or_su2_dag_mul(const double *A, const double *B, double *out); /* A^+*B matrix.h:105-110 */
double or_su2_link_functional(const double *A);                      /* mag.cpp:998            */

/* ---------------- spin primitives (matrix.h:773-985) ---------------- */
void or_spin_contribution1(double *h, const double *A, const double *b);      /* matrix.h:909-926 */
void or_spin_contribution1_conj(double *h, const double *A, const double *b); /* matrix.h:928-946 */
void or_spin_reflect_fast(double *s, const double *V);                        /* matrix.h:825-840 */
double or_spin_parallel(double *s, const double *V);                          /* matrix.h:842-856 */
void or_spin_gauge_matrix(const double *s, double *g);                        /* matrix.h:960-963 */

/* ---------------- SU(2) spin-form MAG (lib/cpu/src/mag.cpp) ---------------- */
void or_su2_contribution_site(const double *links, const double *spins,
                              const int *dims, int64_t site, double *h); /* mag.cpp:569-626 */
double or_su2_functional(const double *links, const int *dims);          /* mag.cpp:993-1001 */
double or_su2_functional_spin(const double *links, const double *spins,
                              const int *dims);                          /* mag.cpp:925-953 */
void or_su2_gauge_transform_spins(double *links, const double *spins,
                                  const int *dims);                      /* mag.cpp:1045-1068 */
void or_su2_or_sweep_lexi(const double *links, double *spins, const int *dims);   /* mag.cpp:1375-1401 */
/* timing baseline only: the reference's racy in-place OpenMP sweep */
void or_su2_or_sweep_lexi_omp(const double *links, double *spins, const int *dims);
void or_su2_relax_sweep_lexi(const double *links, double *spins, const int *dims,
                             double *diff_max, double *diff_avg);                 /* mag.cpp:1477-1512 */
int or_su2_maximize_final(const double *links, double *spins, const int *dims,
                          int or_steps, double tol_max, double tol_avg,
                          int max_iter);                                 /* mag.cpp:1800-1816 */
/* the omega rule's angle functions (the device's operation sequence):
 * atan2 for y > 0, sin and cos on [0, 2 pi] */
double or_atan2_pos(double y, double x);
void or_sincos(double a, double *s, double *c);
int or_approx_same(double functional, double functional_old, int tolerance_digits);
int or_su2_maximize_approximate(const double *links, double *spins,
                                const int *dims, int or_steps,
                                int tolerance_digits, int max_iter);     /* mag.cpp:1733-1750 */

/* Build-defined checkerboard spin-form half-sweeps (the order the GPU spin
 * kernels use).  mode 0 = reflection (reflect_fast), 1 = relaxation
 * (parallel), 2 = geodesic omega rule.  diff = 1 - s_old . h_hat per site,
 * summed / maxed over the updated sites. */
void or_su2_spin_halfsweep(const double *links, double *spins, const int *dims,
                           int parity, int mode, double omega,
                           double *diff_max, double *diff_sum);
/* Checkerboard heat bath half-sweep (mag.cpp:402-433 update, Philox stream
 * keyed by (seed, sweep) instead of a time-seeded ranlux). */
void or_su2_heatbath_halfsweep(const double *links, double *spins,
                               const int *dims, int parity, double temperature,
                               uint64_t seed, uint32_t sweep);

/* ---------------- link-form checkerboard omega-OR (build-defined) ------ */
/* One full sweep = parity-0 half-sweep then parity-1 half-sweep.
 * On return: *F = MAG functional after the sweep (mag.cpp:993-1001 for
 * SU(2), mag.cpp:1837-1848 for SU(3)), *rmax / *ravg = max / mean over all
 * V site updates of 1 - cos(theta) measured before each update. */
void or_su2_link_sweep(double *links, const int *dims, double omega,
                       double *F, double *rmax, double *ravg);
void or_su3_link_sweep(double *links18, const int *dims, double omega,
                       double *F, double *rmax, double *ravg);
/* Omega rule: site gauge quaternion g (g3 = 0) for local field h. */
void or_omega_gauge(const double *h, double omega, double *g);
/* Local-field terms for one SU(2)-subgroup block B = [[p,q],[s,t]]
 * (complex, 8 reals re/im p,q,s,t): forward term of B s3 B^+ and
 * backward term of B^+ s3 B. */
void or_block_field_fwd(const double *B, double *h);
void or_block_field_bwd(const double *B, double *h);

/* ---------------- SU(3) ---------------- */
double or_su3_functional(const double *links18, const int *dims);   /* mag.cpp:1837-1862 */
void or_su3_reunitarize(double *m18);       /* build-defined Gram-Schmidt of rows 0,1 */
void or_su3_proj(double *m18);              /* matrix.h:279-292 (Newton polar + det fix) */
void or_su3_mul(const double *A, const double *B, double *out);
void or_su3_mul_dag(const double *A, const double *B, double *out);

/* ---------------- gauge-invariant checks ---------------- */
double or_su2_plaquette(const double *links, const int *dims);               /* plaket.h:24-42,84-101 */
double or_su3_plaquette(const double *links18, const int *dims);
double or_su2_wilson_loop(const double *links, const int *dims, int R, int T); /* wilson_loops.h:666-754,1049-1073 */
double or_su3_wilson_loop(const double *links18, const int *dims, int R, int T);

/* ---------------- abelian projection, monopoles (monopoles.cpp) ---------- */
void or_su2_abelian_angles(const double *links, const int *dims, double *angles);  /* matrix.h:1059-1062 */
void or_su3_abelian_angles(const double *links18, const int *dims, double *angles3); /* matrix.h:1082-1100, [4V][3] */
void or_u1_monopole_plaquettes(const double *angles, const int *dims, double *out);  /* monopoles.cpp:415-457, [6][V] */
void or_su3_monopole_plaquettes(const double *angles3, const int *dims, double *out); /* monopoles.cpp:368-413,844-884, [3][6][V] */
void or_monopole_currents(const double *P, const int *dims, double *J);              /* monopoles.cpp:44-122, [4V] */
/* clusters of the currents J[4V]: rec[9n] = {first site, length, length_mu[4], space, time, sites},
 * in the reference's order; returns the count (monopoles.cpp:1120-1303, 1343-1449, 1510-1531) */
int64_t or_monopole_clusters(const double *J, const int *dims, int32_t *rec, int64_t cap);

/* ---------------- synthetic configurations (build-defined) ------------- */
void or_philox4x32_10(const uint32_t *ctr, const uint32_t *key, uint32_t *out);
void or_gen_su2_config(double *links, const int *dims, double eps, uint64_t seed);
void or_gen_su3_config(double *links18, const int *dims, double eps, uint64_t seed);
/* gauge copy `copy`: same link draws, gauge key seed ^ copy * 0x9E3779B97F4A7C15 */
/* starting spins of gauge copy `copy` (spin form, links of copy 0) */
void or_gen_copy_spins(double *spins, const int *dims, uint64_t seed, uint32_t copy);
void or_gen_su2_config_copy(double *links, const int *dims, double eps, uint64_t seed,
                            uint32_t copy);
void or_gen_su3_config_copy(double *links18, const int *dims, double eps, uint64_t seed,
                            uint32_t copy);
void or_gen_su3_exp(const double *c8, double eps, double *m18);
void or_gen_spins_uniform(double *spins, const int *dims, uint64_t seed);   /* mag.cpp:143-164,166-236 */

/* OpenMP thread count used by the sweeps (0 = leave the runtime default). */
void or_set_threads(int n);
int or_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
