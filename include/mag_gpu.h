/*
 * mag_gpu.h -- C ABI of the MI355X Maximal Abelian gauge (MAG) fixing path.
 *
 * Plain pointers and sizes only; no exceptions cross the boundary; every
 * function returns an int status (MAG_OK = 0, negative on error, message in
 * mag_last_error()).  Host buffers are copied; device memory is owned by the
 * job.  Jobs are independent: calls on different jobs may come from
 * different host threads.
 *
 * Link arrays at this boundary use the reference's layout
 * (ilyakudrov/general_code lib/cpu/include/indexing.h:71-126, data.h:78-237):
 *   link index = 4*site + mu, site = x + Lx*(y + Ly*(z + Lz*t))
 *   SU(2): 4 doubles per link {a0,a1,a2,a3} = a0 + i a.sigma (matrix.h:8-13)
 *   SU(3): 18 doubles per link, row-major (re,im) (matrix.h:236-254)
 *   spins: 3 doubles per site (matrix.h:773-787)
 * dims = {Lx, Ly, Lz, Lt}; every extent must be even (checkerboard).
 *
 * Reference interfaces replaced (lib/cpu/include/mag.h unless noted):
 *   MAG_functional_su2 ............. mag.h:69-70      -> mag_su2_functional
 *   MAG_functional_su2_spin ........ mag.h:58-60      -> mag_su2_functional_spin
 *   mag_functional_su3 ............. mag.h:177        -> mag_su3_functional
 *   gauge_tranformation_spins ...... mag.h:80-82      -> mag_su2_gauge_transform_spins
 *   make_maximization_final ........ mag.h:166-169    -> mag_su2_maximize_final
 *   make_maximization_approximate .. mag.h:151-153    -> mag_su2_maximize_approximate
 *   make_simulated_annealing ....... mag.h:119-122    -> mag_su2_simulated_annealing
 *   generate_spins_uniform ......... mag.h:179-199    -> mag_su2_generate_spins_uniform
 *   plaket ......................... plaket.h:84-101  -> mag_plaquette
 *   calculate_current (abelian) .... monopoles.cpp:669-687 -> mag_u1_monopole_currents,
 *                                    mag_job_monopole_currents
 *   calculate_current (su3_angles) . monopoles.cpp:646-667 -> mag_su3_monopole_currents,
 *                                    mag_job_monopole_currents
 *   calculate_clusters + cluster_length / length_mu / currents_directions
 *                                    monopoles.cpp:1277-1303, 1343-1449
 *                                    -> mag_current_clusters, mag_job_monopole_clusters
 *   MatrixConverter<su2, abelian>,
 *   MatrixConverter<su3, su3_angles> matrix.h:1059-1062, 1082-1100 -> mag_job_abelian_angles
 *   (none: the reference has no link-form checkerboard MAG, no SU(3) MAG
 *    fixing, no mixed precision, batching or serving)  -> mag_job_* / mag_batch_*
 */
#ifndef MAG_GPU_H
#define MAG_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status codes */
#define MAG_OK 0
#define MAG_ERR_ARG (-1)           /* bad argument (null pointer, range)       */
#define MAG_ERR_DIMS (-2)          /* odd / non-positive / too large extents   */
#define MAG_ERR_NOMEM (-3)         /* device allocation failed                 */
#define MAG_ERR_NONFINITE (-4)     /* input holds NaN or Inf                   */
#define MAG_ERR_NOT_CONVERGED (-5) /* sweep cap reached before the tolerance   */
#define MAG_ERR_HIP (-6)           /* HIP runtime error                        */
#define MAG_ERR_STATE (-7)         /* call not valid in the job's state        */
#define MAG_ERR_CANCELLED (-8)     /* served request cancelled                 */

/* ------------------------------------------------------------ parameters */
#define MAG_SU2 2
#define MAG_SU3 3
#define MAG_PREC_F64 0   /* FP64 links, FP64 sweeps                            */
#define MAG_PREC_MIXED 1 /* FP32 sweeps, FP64 functional, FP64 master links,  */
                         /* FP64 reunitarisation every reunit_every sweeps     */

typedef struct mag_params {
  int group;          /* MAG_SU2 or MAG_SU3                                    */
  int dims[4];        /* Lx, Ly, Lz, Lt                                        */
  int precision;      /* MAG_PREC_F64 or MAG_PREC_MIXED                        */
  int device;         /* HIP device ordinal                                    */
  double omega;       /* overrelaxation parameter, 0 < omega <= 2              */
  double tol;         /* stop when |F_k - F_{k-1}| < tol |F_k|                 */
  int max_sweeps;     /* sweep cap (all phases together)                       */
  int reunit_every;   /* mixed: rebuild FP32 links from the FP64 master      */
                      /*        every reunit_every sweeps                      */
  double tol_mixed;   /* mixed: FP32 phase ends when the mean relative     */
                      /*        change per sweep drops below it; link form:   */
                      /*        measured on the FP64 links of two successive  */
                      /*        rebuilds (free of FP32 rounding), site form:  */
                      /*        per sweep.  FP64 sweeps then run down to tol  */
  int form;           /* MAG_FORM_LINK or MAG_FORM_SPIN                        */
  int copies;         /* gauge copies of the configuration swept together    */
                      /* (spin form, 1..MAG_MAX_COPIES; 1 in the link form)  */
  int fp64_warmup;    /* mixed: FP64 sweeps before the FP32 phase (a multiple */
                      /*        of 20); the early sweeps pick the maximum     */
                      /*        the run converges to                          */
  int track_gauge;    /* SU(2) link form: also accumulate the site gauge G(x) */
                      /*        of the fixing (current links = G U_in G^+),   */
                      /*        for mag_job_download_links_canonical and      */
                      /*        mag_job_download_spins; not in batches        */
} mag_params;

/* Link form: the links carry the gauge transformation and are rewritten by
 * every half-sweep.  Spin form (SU(2)): the links stay read-only and each
 * gauge copy is a field of unit 3-vectors s(x) (the reference's spin form,
 * mag.cpp:569-626, 1375-1512); SU(3): each copy is a site gauge field G(x)
 * (the site form).  All copies of one configuration are swept together,
 * so the links are read once per half-sweep for all of them.  Both forms
 * run the same checkerboard omega step and give the same functional per
 * sweep to rounding.  In mixed precision the spin form's FP32 phase sweeps
 * FP32 spins over FP32-rounded links (field and functional summed in
 * FP64); the FP64 spins follow at the end of the phase. */
#define MAG_FORM_LINK 0
#define MAG_FORM_SPIN 1
#define MAG_MAX_COPIES 8

/* Fill defaults: SU(2), omega 1.7, tol 1e-12, 1000 sweeps, FP64,
 * reunit_every 10, tol_mixed 1e-9, link form, 1 copy, no FP64 warm-up,
 * device 0. */
void mag_params_default(mag_params *p);

const char *mag_last_error(void);
int mag_device_count(int *n);

/* ------------------------------------------------------------ jobs
 * A job owns one configuration in HBM (parity-split SoA) and runs the
 * checkerboard omega-overrelaxation (SU(3): Cabibbo-Marinari over the
 * subgroups (0,1),(1,2),(0,2)) until the functional's relative change drops
 * below tol.  All calls are asynchronous on the job's stream unless they
 * return data to the host. */
typedef struct mag_job mag_job;

int mag_job_create(const mag_params *p, mag_job **out);
int mag_job_destroy(mag_job *job);
/* hipStream_t of the job (created by the job) */
int mag_job_stream(mag_job *job, void **stream);
/* bytes of HBM the job holds */
int mag_job_footprint(const mag_params *p, uint64_t *bytes);

/* Upload links in the reference layout (host memory) and reset the job. */
int mag_job_upload_links(mag_job *job, const double *links);
/* Same from a device buffer in the reference layout (no host copy). */
int mag_job_upload_links_device(mag_job *job, const double *dev_links);
/* Fill the job with a seeded synthetic configuration generated on the GPU:
 * SU(2) U = exp(i eps a.sigma), a ~ N(0,1)^3; SU(3) U = exp(i eps
 * sum_a c_a lambda_a / 2), c ~ N(0,1)^8; then a Haar random gauge
 * transformation, all from a Philox4x32-10 stream keyed by seed. */
int mag_job_generate(mag_job *job, double eps, uint64_t seed);
/* Gauge copy `copy` of that configuration: the same link draws under a
 * different random gauge (gauge key seed ^ copy * 0x9E3779B97F4A7C15), so
 * copies differ only by a gauge transformation; copy 0 = mag_job_generate.
 * Stands for the reference's gauge-copy loop, where each copy starts the
 * maximisation from new random spins (apps/mag_su2/mag_fixation.cpp:164-172). */
int mag_job_generate_copy(mag_job *job, double eps, uint64_t seed, uint32_t copy);

/* Enqueue exactly n sweeps (fewer take effect once the job is done). */
int mag_job_sweeps(mag_job *job, int n);
/* Run to convergence or the sweep cap (blocking).  Mixed precision: FP32
 * phase, final FP64 rebuild, FP64 polish.  Returns MAG_OK or
 * MAG_ERR_NOT_CONVERGED. */
int mag_job_run(mag_job *job);
/* Finish a mixed-precision job's FP32 phase now (FP64 rebuild of the links
 * from the master and the accumulated gauge); no-op for FP64 jobs. */
int mag_job_finalize(mag_job *job);

/* sweeps done, converged flag, current functional (blocking) */
int mag_job_status(mag_job *job, int *sweeps, int *converged, double *F);
/* initial functional F0 (blocking) */
int mag_job_initial_functional(mag_job *job, double *F0);
/* per-sweep history: functional, max and mean of 1 - cos(theta) (the
 * relaxation residual of mag.cpp:1477-1512); n = entries written. */
int mag_job_history(mag_job *job, double *F, double *rmax, double *ravg,
                    int cap, int *n);
/* sweep index at which a mixed job switched to FP64 (-1: not yet) */
int mag_job_fp64_start(mag_job *job, int *sweep);
/* Download the current links (reference layout, FP64).  A mixed job that is
 * still in its FP32 phase returns its FP32 working links. */
int mag_job_download_links(mag_job *job, double *links);
/* SU(2) link form with track_gauge, after the run (FP64 phase): the fixed
 * links in the convention of the reference's spin form, G_c(x) U_in
 * G_c(x+mu)^+ with G_c = GetGaugeMatrix(s(x)) (matrix.h:960-963) and s(x)
 * the spin of the accumulated gauge (row 3 of its rotation, invariant
 * under the diagonal U(1) that the MAG functional leaves free): exactly
 * what gauge_tranformation_spins (mag.cpp:1045-1068) writes for the spins
 * that mag_job_download_spins returns (V x 3 doubles, lexicographic).  The
 * plain download returns the same links up to a diagonal U(1) gauge.
 * MAG_ERR_NONFINITE if a spin is exactly (0,0,1), the pole of
 * GetGaugeMatrix. */
int mag_job_download_links_canonical(mag_job *job, double *links);
int mag_job_download_spins(mag_job *job, double *spins);
/* MAG functional / plaquette of the device-resident links (blocking) */
int mag_job_functional(mag_job *job, double *F);
int mag_job_plaquette(mag_job *job, double *P);

/* ---- launch shapes
 * The link-form half-sweep kernels run with 64, 128, 256 or 512 threads per
 * workgroup and 1 or 2 tiles per thread; their partial sums are per wave,
 * so every shape gives bit-identical results.  By default the first sweep
 * on a new (group, precision, lattice) times every shape on a scratch copy
 * of the job's links and keeps the fastest for the process (MAG_TUNE_AUTO);
 * MAG_TUNE_FIXED uses (nt, spt) everywhere, MAG_TUNE_DEFAULT uses (256, 1)
 * without tuning.  A batch class's half-sweep is tuned on its own batched
 * launch (every shape over scratch copies of the class's jobs, cached per
 * group, precision, device and log2 of the class's lanes) under the same
 * modes.  No counterpart in the reference. */
#define MAG_TUNE_AUTO 0
#define MAG_TUNE_FIXED 1
#define MAG_TUNE_DEFAULT 2
int mag_tune_set(int mode, int nt, int spt);
/* forget the tuned shapes */
int mag_tune_clear(void);
/* shape the job's FP64 (fp32 = 0) or FP32 (fp32 = 1) link-form kernel uses;
 * 0, 0 before its first sweep */
int mag_job_shape(mag_job *job, int fp32, int *nt, int *spt);

/* ---- gauge copies (spin form)
 * A spin-form job holds one SU(2) configuration and `copies` spin fields.
 * mag_job_upload_links starts every copy at s = e3 (the identity gauge);
 * mag_job_generate(eps, seed) uploads copy 0 of the synthetic configuration
 * and starts copy k at the spins of the gauge g_k g_0^+ (g_k: the Haar
 * gauge of mag_job_generate_copy(eps, seed, k)), so copy k follows the same
 * trajectory as a link-form job of mag_job_generate_copy(eps, seed, k).
 * Each copy converges on its own (its own history and status); the
 * link-form calls above (status, history, functional, download) report
 * copy 0, mag_job_status reports converged only when every copy has.
 * Stands for the reference's gauge-copy loop with best-copy selection
 * (apps/mag_su2/mag_fixation.cpp:164-249). */
/* start copy `copy` from the spins s (reference layout [V][3]) */
int mag_job_set_copy_spins(mag_job *job, int copy, const double *spins);
int mag_job_get_copy_spins(mag_job *job, int copy, double *spins);
int mag_job_copy_status(mag_job *job, int copy, int *sweeps, int *converged, double *F);
int mag_job_copy_history(mag_job *job, int copy, double *F, double *rmax, double *ravg,
                         int cap, int *n);
/* copy with the largest functional (the reference keeps the spins of the
 * higher functional, apps/mag_su2/mag_fixation.cpp:223-240) */
int mag_job_best_copy(mag_job *job, int *copy, double *F);
/* Links of copy `copy` gauge-transformed by its spins:
 * U'_x,mu = G(s(x)) U_x,mu G(s(x+mu))^+ with G(s) the SU(2) element of
 * smallest rotation that takes s to e3 (regular everywhere except s = -e3,
 * where any G with row 3 of its rotation = -e3 is used).  Equal to the
 * reference's gauge_tranformation_spins (mag.cpp:1045-1068) up to a
 * diagonal U(1) gauge, which leaves the MAG functional and every
 * a0^2 + a3^2 unchanged. */
int mag_job_download_copy_links(mag_job *job, int copy, double *links);

/* Measurement: n sweeps bracketed by HIP events on the job's stream
 * (blocking); total_ms = elapsed time. */
int mag_job_sweeps_timed(mag_job *job, int n, double *total_ms);
/* Measurement of the dominant kernel: n half-sweep launches back to back
 * (parities alternating) between two events, avg_ms = elapsed / n.  The
 * sweeps are real and move the links, but are not entered in the job's
 * sweep history (no per-sweep reduction runs). */
int mag_job_halfsweep_timed(mag_job *job, int n, double *avg_ms);

/* ------------------------------------------------------------ batches
 * No counterpart in the reference.  A batch sweeps many jobs (mixed SU(2) /
 * SU(3), mixed sizes, FP64 or mixed precision) together: each half-sweep is
 * one launch per (group, storage type) class whose grid concatenates the
 * member jobs' grids, each sweep ends with one reduction launch (a workgroup
 * per job), and the sequence of `mag_batch_replay_sweeps` sweeps is captured
 * once per class in a hipGraph and replayed.  Classes run as independent
 * chains on streams of their own: mag_batch_step / mag_batch_run relaunch a
 * class as soon as its previous replay is done, whatever the other classes
 * are doing.  Every job runs exactly the code and the partition of its own
 * single launch, so its results are bit-identical alone and in any batch.
 * Jobs that converge stop consuming work (their workgroups return at once);
 * a mixed-precision job whose FP32 phase ends is rebuilt in FP64 when its
 * class's replay has been examined and continues in the FP64 class.
 * Member jobs must not be swept through the mag_job_* calls meanwhile. */
typedef struct mag_batch mag_batch;

int mag_batch_create(int device, mag_batch **out);
int mag_batch_destroy(mag_batch *b);
/* Add a loaded job (same device; fresh, or at a multiple of its
 * reunit_every sweeps).  All members share reunit_every. */
int mag_batch_add(mag_batch *b, mag_job *job);
int mag_batch_remove(mag_batch *b, mag_job *job);
int mag_batch_size(mag_batch *b, int *n);
/* sweeps per graph replay (default 20; a multiple of reunit_every) */
int mag_batch_replay_sweeps(mag_batch *b, int *n);
/* Enqueue n replays (n * replay_sweeps sweeps of every active member),
 * joined on the batch stream. */
int mag_batch_replays(mag_batch *b, int n);
/* Blocking: wait for every class, move mixed jobs whose FP32 phase is
 * done to FP64, report members still running and members finished. */
int mag_batch_poll(mag_batch *b, int *n_running, int *n_finished);
/* Launch a replay of every class that is idle, wait until at least one
 * class's replay has finished, examine its members (phase changes, finished
 * jobs) and relaunch the idle classes; returns with replays in flight.
 * Finished members may be removed (and destroyed) right after it returns. */
int mag_batch_step(mag_batch *b, int *n_running, int *n_finished);
/* Blocking: step until every member is finished. */
int mag_batch_run(mag_batch *b);
/* Measurement: n replays between HIP events on the batch stream. */
int mag_batch_replays_timed(mag_batch *b, int n, double *ms);
int mag_batch_stream(mag_batch *b, void **stream);

/* ------------------------------------------------- abelian projection, monopoles
 * What the reference's monopole apps compute from MAG-fixed links
 * (apps/monopoles_su2/monopoles_su2.cpp:110, apps/monopoles_su3/
 * monopoles_su3.cpp:103-120), on the device.  Layouts are the reference's:
 *   SU(2) angles  [4V]     phi of link 4*site+mu
 *   SU(3) angles  [4V][3]  the su3_angles of link 4*site+mu
 *   SU(2) J       [4V]     J of dual link 4*site+mu
 *   SU(3) J       [3][4V]  per colour
 * The currents are integers up to rounding (J = n / 1 with |J - n| ~ 1e-15).
 *
 * Angles of the job's links (copy `copy` of a site-form job; 0 for the link
 * form): SU(2) atan2(a3, a0) (MatrixConverter<su2, abelian>,
 * matrix.h:1059-1062); SU(3) the diagonal phases less a third of their sum
 * taken into [-pi, pi) (MatrixConverter<su3, su3_angles>, matrix.h:1082-1100). */
int mag_job_abelian_angles(mag_job *job, int copy, double *angles);
/* Monopole currents of the job's links: calculate_current of the abelian
 * (monopoles.cpp:669-687) / su3_angles (monopoles.cpp:646-667, with
 * make_monopole_plakets 844-884) projection, links kept on the device. */
int mag_job_monopole_currents(mag_job *job, int copy, double *J);
/* Measurement: mean device time of angles + plaquettes + currents over reps
 * runs (copy 0), no host transfer. */
int mag_job_monopole_timed(mag_job *job, int reps, double *ms);
/* Currents from host angles (the monopole apps read abelian configurations
 * from files): calculate_current(LatticeData<abelian>) (monopoles.cpp:669-687)
 * and calculate_current(LatticeData<su3_angles>) (monopoles.cpp:646-667).
 * Extents >= 2 (odd allowed); every angle finite (else MAG_ERR_NONFINITE)
 * with |angle| <= 1000 (else MAG_ERR_ARG). */
int mag_u1_monopole_currents(const int *dims, const double *angles, double *J);
int mag_su3_monopole_currents(const int *dims, const double *angles3, double *J3);

/* Monopole clusters: the connected clusters of current links
 * (calculate_clusters(J, data_pattern), monopoles.cpp:1277-1303) and, per
 * cluster, what the monopole apps read from its tree (cluster_length,
 * length_mu, currents_directions, monopoles.cpp:1343-1449; site_number,
 * 1510-1531).  A link carries a current when |J| > 0.3, with charge
 * lround(J) (monopoles.cpp:1105-1228).  One record of MAG_CLUSTER_INTS
 * int32 per cluster:
 *   {first site, length = sum |charge|, length_mu[4] = sum of charges per
 *    direction, space = sum |charge| over mu < 3, time = over mu = 3,
 *    number of sites}
 * in the reference's order (by first site in lexicographic order; the
 * first site is the cluster's smallest site index).  Records of colour k
 * at records + k * cap * MAG_CLUSTER_INTS, n[k] = number of clusters of
 * colour k; more than cap clusters: n is still set, cap records are
 * written, MAG_ERR_ARG.  The charges must balance at every site (true of
 * every current from the functions above) with no current of charge 0;
 * otherwise MAG_ERR_ARG.
 * mag_job_monopole_clusters: angles -> currents -> clusters of a job's
 *   copy on the device (SU(2) one colour, SU(3) three), only the records
 *   leave the GPU.
 * mag_current_clusters: from host currents J ([ncol][4V], ncol 1 or 3),
 *   as the apps call calculate_clusters (extents >= 2; J finite, else
 *   MAG_ERR_NONFINITE).
 * mag_job_clusters_timed: mean device time of the cluster kernels over
 *   reps runs (copy 0's currents, all colours), no host transfer. */
#define MAG_CLUSTER_INTS 9
int mag_job_monopole_clusters(mag_job *job, int copy, int32_t *records, int64_t cap, int64_t *n);
int mag_current_clusters(const int *dims, int ncol, const double *J, int32_t *records, int64_t cap,
                         int64_t *n);
int mag_job_clusters_timed(mag_job *job, int reps, double *ms);

/* ------------------------------------------------------------ serving
 * No counterpart in the reference (one configuration per process,
 * apps/mag_su2/mag_fixation.cpp).  A server owns one device, a scheduler
 * thread and a batch.  Requests are submitted from any thread and admitted
 * in arrival order while the HBM footprint of the admitted jobs
 * (mag_job_footprint) stays within the budget and their count within
 * max_active (0: no limit); an admitted job joins the batch at the next
 * replay (continuous batching), keeps its links in HBM until it has
 * converged (or hit its sweep cap), and leaves the batch at once, freeing
 * its memory.  Each result is the one the job gives when run alone.
 * Mixed-precision requests served together must share reunit_every. */
typedef struct mag_server mag_server;
typedef struct mag_request {
  mag_params params;
  const double *links; /* links in the reference layout, copied at submit, */
                       /* or NULL: synthetic input (mag_job_generate_copy;  */
                       /* the spin form generates its copies 0..copies-1)   */
  double eps;
  uint64_t seed;
  uint32_t copy;
} mag_request;
typedef struct mag_result {
  uint64_t id;
  int status;       /* MAG_OK, MAG_ERR_NOT_CONVERGED, or the job's error     */
  int sweeps;       /* sweeps of the reported copy                          */
  int converged;    /* every copy converged                                 */
  int best_copy;    /* copy with the largest functional (spin form)         */
  int fp64_start;   /* mixed: sweep of the switch to the FP64 polish        */
  double F;         /* final functional of best_copy                        */
  double t_submit;  /* seconds on the server clock: submitted, admitted,    */
  double t_admit;   /* finished                                             */
  double t_done;
} mag_result;
typedef struct mag_server_stats {
  uint64_t budget;     /* HBM budget in bytes                               */
  uint64_t peak_hbm;   /* largest footprint admitted at once                */
  uint64_t submitted;
  uint64_t finished;
  uint64_t admitted;
  uint64_t replays;    /* batch replays launched                            */
  int peak_active;     /* most jobs on the GPU at once                      */
} mag_server_stats;

/* hbm_budget 0: 80% of the device's free memory */
int mag_server_create(int device, uint64_t hbm_budget, int max_active, mag_server **out);
/* finishes every submitted request, then stops the scheduler */
int mag_server_destroy(mag_server *s);
/* non-blocking; a request larger than the whole budget is refused */
int mag_server_submit(mag_server *s, const mag_request *r, uint64_t *id);
/* blocks until request id is finished */
int mag_server_wait(mag_server *s, uint64_t id, mag_result *out);
/* blocks until every request submitted so far is finished */
int mag_server_drain(mag_server *s);
/* Cancel request id: a queued request is dropped at once, an admitted one
 * leaves the batch at the scheduler's next step (after the replay that
 * holds it has finished) and its memory is freed; either way its result
 * has status MAG_ERR_CANCELLED (with the sweeps and F reached, if it was
 * running).  MAG_ERR_STATE if it has already finished.  The other requests
 * are not affected: their results stay those of each job alone. */
int mag_server_cancel(mag_server *s, uint64_t id);
int mag_server_clock(mag_server *s, double *seconds);
int mag_server_get_stats(mag_server *s, mag_server_stats *st);

/* ------------------------------------------------------------ spin form
 * The reference's own algorithm (lib/cpu/src/mag.cpp): one unit spin s(x)
 * per site, links read-only, fixed links produced at the end by
 * gauge_tranformation_spins.  On the GPU the sweeps run in checkerboard
 * order (the reference's in-place lexicographic OpenMP loop races and is
 * only deterministic on one thread).  Spins: V*3 doubles, site order of the
 * links.  A session keeps links and spins resident between calls. */
typedef struct mag_spin_session mag_spin_session;

int mag_spin_session_create(const int *dims, int device, mag_spin_session **out);
int mag_spin_session_destroy(mag_spin_session *s);
int mag_spin_session_upload_links(mag_spin_session *s, const double *links);
int mag_spin_session_download_links(mag_spin_session *s, double *links);
int mag_spin_session_set_spins(mag_spin_session *s, const double *spins);
int mag_spin_session_get_spins(mag_spin_session *s, double *spins);
/* generate_spins_uniform (mag.h:179-199) on a Philox stream keyed by seed */
int mag_spin_session_generate_spins(mag_spin_session *s, uint64_t seed);
/* MAG_functional_su2_spin (mag.cpp:925-953) */
int mag_spin_session_functional_spin(mag_spin_session *s, double *F);
/* MAG_functional_su2 (mag.cpp:993-1001) of the resident links */
int mag_spin_session_functional(mag_spin_session *s, double *F);
/* One checkerboard sweep: mode 0 reflection (overrelaxation_update,
 * mag.cpp:1375-1401), 1 relaxation (relaxation_update, mag.cpp:1477-1512),
 * 2 geodesic omega rule (param = omega), 3 heat bath (heat_bath_update_fast,
 * mag.cpp:785-814; param = T).  dmax/davg: max / mean of 1 - s_old.h_hat. */
int mag_spin_session_sweep(mag_spin_session *s, int mode, double param, uint64_t seed,
                           double *dmax, double *davg);
/* make_simulated_annealing (mag.cpp:1585-1603), heat bath on a Philox
 * stream keyed by seed */
int mag_spin_session_simulated_annealing(mag_spin_session *s, double T_init, double T_final,
                                         double T_step, int OR_steps,
                                         int thermalization_steps, uint64_t seed);
/* make_maximization_final (mag.cpp:1800-1816), at most max_iter iterations */
int mag_spin_session_maximize_final(mag_spin_session *s, int OR_steps, double tol_max,
                                    double tol_avg, int max_iter, int *iters);
/* make_maximization_approximate (mag.cpp:1733-1750) */
int mag_spin_session_maximize_approximate(mag_spin_session *s, int OR_steps,
                                          int tolerance_digits, int max_iter, int *iters);
/* gauge_tranformation_spins (mag.cpp:1045-1068) on the resident links */
int mag_spin_session_gauge_transform(mag_spin_session *s);

/* One-shot forms on host arrays (current HIP device), mirroring mag.h */
int mag_su2_generate_spins_uniform(const int *dims, uint64_t seed, double *spins);
int mag_su2_functional_spin(const int *dims, const double *links, const double *spins,
                            double *F);
int mag_su2_gauge_transform_spins(const int *dims, double *links, const double *spins);
int mag_su2_maximize_final(const int *dims, const double *links, double *spins, int OR_steps,
                           double tol_max, double tol_avg, int max_iter, int *iters);
int mag_su2_maximize_approximate(const int *dims, const double *links, double *spins,
                                 int OR_steps, int tolerance_digits, int max_iter, int *iters);
int mag_su2_simulated_annealing(const int *dims, const double *links, double *spins,
                                double T_init, double T_final, double T_step, int OR_steps,
                                int thermalization_steps, uint64_t seed);

/* Measured HBM bandwidth of a plain streaming copy kernel (read bytes +
 * write bytes per second / 1e9) over buffers of the given size: the
 * practical ceiling the sweep kernels are compared with.  No counterpart in
 * the reference (measurement aid). */
int mag_hbm_copy_bandwidth(int device, uint64_t bytes, int iters, double *gbs);

/* ------------------------------------------------------------ one-shot
 * functionals and checks on host arrays (reference layout) */
int mag_su2_functional(const int *dims, const double *links, double *F);
int mag_su3_functional(const int *dims, const double *links18, double *F);
int mag_plaquette(int group, const int *dims, const double *links, double *P);

/* The functional of mag_su2_functional / mag_su3_functional (group
 * MAG_SU2 / MAG_SU3) from links already in device memory (reference
 * layout, 4*V*4 or 4*V*18 doubles), written to dev_F (one double in device
 * memory), enqueued on `stream` (a hipStream_t, NULL: the default stream)
 * of the current device; returns without waiting.  It allocates nothing:
 * its partial sums live in a fixed 64 KiB device array of the library,
 * one per device, so calls that overlap on one device race.  At most
 * 16 777 216 sites (MAG_ERR_DIMS beyond); extents even and >= 2; the links
 * are not checked for NaN / Inf (a non-finite link gives a non-finite F).
 * No counterpart in the reference: for callers that cannot allocate (code
 * under HIP graph capture, serving loops that forbid hipMalloc). */
int mag_functional_noalloc(int group, const int *dims, const double *dev_links, double *dev_F,
                           void *stream);

#ifdef __cplusplus
}
#endif
#endif /* MAG_GPU_H */
