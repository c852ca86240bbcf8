// mag_internal.h -- kernel launchers and device-side job structures shared
// by the C-ABI layer (mag_capi.hip) and the kernels (mag_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace mag {

constexpr int kBlock = 256;  // threads per workgroup for the sweep kernels
constexpr int kWaves = kBlock / 64;

// Lattice geometry in the parity-split layout.  Site (x,y,z,t) with
// lexicographic index i = x + Lx(y + Ly(z + Lz t)) (indexing.h:111-123) has
// parity p = (x+y+z+t)&1 and half index h = i>>1; every dimension is even.
// Pitch padding of the chunk fields, in half-sites: consecutive fields of a
// link or site array start NC*(16*Vh + 4096) bytes apart (Vh rounded up to
// whole tiles).  With a power-of-two pitch the many streams of one
// half-sweep fall on the same DRAM channels; measured on a 48-stream copy,
// 4 KB of padding raised 4.3 to 5.3 TB/s.
constexpr uint32_t kChunkPad = 256;

// MAG_TILE_LAYOUT=1 (default): the chunks of 64 consecutive half-sites are
// stored together (tile-interleaved); 0: one array per chunk index.  The
// pitch is rounded up to whole tiles of 64 half-sites.
#ifndef MAG_TILE_LAYOUT
#define MAG_TILE_LAYOUT 1
#endif
constexpr uint32_t kTile = 64;

struct Geom {
  int L0, L1, L2, L3;
  uint32_t hx;  // L0/2: half sites per x-row
  uint32_t Vh;  // V/2
  uint32_t s1;  // half-index stride of y  (= hx)
  uint32_t s2;  // half-index stride of z  (= hx*L1)
  uint32_t s3;  // half-index stride of t  (= hx*L1*L2)
  uint32_t hp;  // pitch of a chunk field in half-sites (Vh in whole tiles + kChunkPad)
};

inline Geom make_geom(const int* d) {
  Geom g;
  g.L0 = d[0]; g.L1 = d[1]; g.L2 = d[2]; g.L3 = d[3];
  g.hx = (uint32_t)(d[0] / 2);
  g.s1 = g.hx;
  g.s2 = g.hx * (uint32_t)d[1];
  g.s3 = g.s2 * (uint32_t)d[2];
  g.Vh = g.s3 * (uint32_t)d[3];
  g.hp = (MAG_TILE_LAYOUT ? (g.Vh + kTile - 1) / kTile * kTile : g.Vh) + kChunkPad;
  return g;
}

// Per-job control block, device resident.  Written only by the one-block
// reduction kernel; read by the sweep kernels to skip work once done.
struct JobCtl {
  int done;        // 1: converged or sweep cap reached (sweep kernels exit)
  int sweep;       // completed sweeps
  int max_sweeps;  // cap
  int converged;   // 1: |F - F_prev| < tol |F|
  double tol;      // relative functional change for convergence
  double F_prev;   // functional after the last sweep (initial value first)
  double F0;       // functional of the uploaded configuration
  int window;      // > 1: convergence on the mean change per sweep over the
                   // last `window` sweeps; -1: no test per sweep (tested
                   // at the FP64 rebuilds instead); else on the last sweep
  int rb_every;    // sweeps between two FP64 rebuilds (mixed link form)
  double F_ref;    // F of the links of the previous rebuild (0: none yet)
  int arrive;      // workgroups of the running per-sweep reduction that have
                   // folded their share (the last one resets it to 0)
  int pad;
};

// reduction kernel modes
//   RED_INIT : F0 = F_prev = F      RED_SWEEP: append a sweep to the history
//   RED_PLAIN: out3 only            RED_RESET: F_prev = F (phase change)
enum { RED_INIT = 0, RED_SWEEP = 1, RED_PLAIN = 2, RED_RESET = 3 };

// ---------------------------------------------------------------- launchers
// Partial buffers hold triples of doubles (F sum, residual sum, residual
// max).  The half-sweep kernels write one triple per wave of 64 lanes
// (sweep_waves below): a sweep uses 2*nw triples, parity p at 3*nw*p.  In
// the link form the black half folds the red triple of each wave into its
// own, and the sweep's reductions read the nw black triples at 3*nw.
// nblocks(g) is the workgroup count of a kBlock-thread launch over Vh
// sites (the functional, spin-form and copy launches).
inline int nblocks(const Geom& g) { return (int)((g.Vh + kBlock - 1) / kBlock); }
// lanes per site of the link-form SU(3) half-sweep: 2 (lane q holds the
// links of mu = 2q, 2q+1) or 4 (lane q holds mu = q)
#ifndef MAG_SU3_LANES
#define MAG_SU3_LANES 2
#endif
constexpr int kSu3Lanes = MAG_SU3_LANES;
static_assert(kSu3Lanes == 2 || kSu3Lanes == 4, "SU(3) lanes per site: 2 or 4");
// lanes of a link-form half-sweep: SU(2) one per site, SU(3) kSu3Lanes;
// the SU(3) site form runs one site per lane (site3_lanes)
inline size_t sweep_lanes(const Geom& g, int group) {
  return (size_t)g.Vh * (group == 2 ? 1 : kSu3Lanes);
}
inline size_t site3_lanes(const Geom& g) { return (size_t)g.Vh; }
// workgroups of a half-sweep at the default shape
inline int sweep_blocks(const Geom& g, int group) {
  return (int)((sweep_lanes(g, group) + kBlock - 1) / kBlock);
}
// partial triples of a half-sweep: one per wave of 64 lanes
inline int sweep_waves(const Geom& g, int group) {
  return (int)((sweep_lanes(g, group) + 63) / 64);
}
// Launch shape of a half-sweep kernel: nt threads per workgroup (64, 128,
// 256 or 512) and spt tiles of nt lanes per thread (1 or 2).  Partials are
// per wave, so the shape never changes a result.
struct Shape {
  int nt = 256;
  int spt = 1;
};

// Mixed-precision link form, FP32 phase: sweeps run in periods of
// reunit_every.  Only red (parity 0) sites accumulate a site gauge G; the
// red half-sweep that opens a period first normalises its G in FP64
// (HALF_RENORM); the black half-sweep that closes it rebuilds every link
// from the FP64 master U0 with the red gauges and the black gauges
// recovered from the swept links (HALF_REBUILD, mag_gauge.h) and writes
// the functional partials of the rebuilt links to part_rb.  The per-sweep
// partials of the FP32 phase are kept in kSlots slots (slot s at
// part + s * part_stride(nw)) and reduced together at the end of a period
// or when the slots are full (launch_reduce_slots / launch_reduce_period);
// the rebuild partials sit after the slots, at part + kSlots *
// part_stride(nw).
enum { HALF_PLAIN = 0, HALF_RENORM = 1, HALF_REBUILD = 2 };
constexpr int kSlots = 20;

struct HalfOpt {
  int slot = 0;               // sweep slot of the partials (0 outside the FP32 phase)
  int mode = HALF_PLAIN;
  const double* U0 = nullptr; // HALF_REBUILD: FP64 master
  double* part_rb = nullptr;  // HALF_REBUILD: partials of the rebuilt links' F
};

// One half-sweep of parity `parity`; partial triples go to
// part + slot*part_stride(nw) + 3*nw*parity, nw = sweep_waves(g, group); G is the
// red site gauge (mixed FP32 phase: accumulated by red PLAIN / RENORM
// half-sweeps, read by HALF_REBUILD) or nullptr.  The kernels return at
// once when ctl->done is set or ctl->sweep + slot reaches ctl->max_sweeps.
template <typename T>
void launch_su2_half(T* U, T* G, const Geom& g, int parity, double omega,
                     double* part, const JobCtl* ctl, Shape sh, hipStream_t st,
                     const HalfOpt& o = HalfOpt{});
template <typename T>
void launch_su3_half(T* U, T* G, const Geom& g, int parity, double omega,
                     double* part, const JobCtl* ctl, Shape sh, hipStream_t st,
                     const HalfOpt& o = HalfOpt{});
// FP32 phase: reduce slots 0..n-1 into the history (entries ctl->sweep + k,
// below max_sweeps), then advance ctl->sweep by n (capped at max_sweeps,
// done at the cap); with rebuild: the rebuilt links' F from part_rb (nrb
// triples) decides convergence: the mean change per sweep since the
// previous rebuild below tol
void launch_reduce_slots(const double* part, size_t pstride, int nparts, int n, double invF,
                         double invV, JobCtl* ctl, double* hF, double* hR, double* hA,
                         hipStream_t st);
void launch_reduce_period(const double* part_rb, int nrb, int n, int rebuild, double invF,
                          JobCtl* ctl, double* hF, hipStream_t st);

// Gauge-copy (spin form, SU(2), FP64) half-sweep over all copies of one
// configuration: the links U (read-only) are loaded once per site and
// every copy not yet done updates its spins, copy c at S + c*sstride; its
// partial triples (one per wave) go to part + c*pstride + 3*nw*parity
// (nw = sweep_waves(g, 2)) and it is skipped once ctl[c].done.
// TL = float: the FP32 phase of the mixed form, FP32-rounded links and
// FP32 spins (field and functional summed in FP64, FP32 omega step).
template <typename TL>
void launch_su2_copies_half(const TL* U, TL* S, size_t sstride, int ncopies,
                            const Geom& g, int parity, double omega, double* part,
                            size_t pstride, const JobCtl* ctl, hipStream_t st);
// SU(3) site form: links U read-only, site gauge of copy c at G + c*gstride
// (2 rows, chunked site layout, the links' type); partials and control as
// for the spin form, one triple per wave of the site3_lanes(g) lanes.
// The FP32 phase (T = float) keeps the partials of its sweeps in slots:
// sweep slot s of copy c at part + (s*ncopies + c)*pstride, a sweep whose
// slot lies beyond the cap (ctl[c].sweep + s >= max_sweeps) is skipped;
// the slots are reduced together at the end of a period of reunit_every
// sweeps (launch_reduce_slots_copies / launch_reduce_period_copies), as in
// the mixed link form.  Outside the FP32 phase slot = 0.
template <typename T>
void launch_su3_site_copies(const T* U, T* G, size_t gstride, int ncopies, const Geom& g,
                            int parity, double omega, double* part, size_t pstride,
                            const JobCtl* ctl, hipStream_t st, int slot = 0);
// SU(3) site form, FP32 phase, at the end of every period of reunit_every
// sweeps: F of every copy's links in FP64 from the FP32 state (gauges put
// on SU(3) in FP64), partials of copy c at part + c*pstride (part: slot
// kSlots of the buffer); the period reduction then tests it (apply_rebuild:
// the change per sweep since the previous evaluation below tol)
void launch_su3_site_F(const float* U, const float* G, size_t gstride, int ncopies,
                       const Geom& g, double* part, size_t pstride, const JobCtl* ctl,
                       hipStream_t st);
// the slots of one copy set: n slots of each of the ncopies copies into the
// copies' histories (hF + c*hstride ...), then per copy the period
// (reduce_period: sweep count, F_prev, cap, and with check = 1 the
// stopping test on the F partials at part_check + c*pstride, nw of them)
void launch_reduce_slots_copies(const double* part, size_t pstride, int nparts, int ncopies, int n,
                                double invF, double invV, const JobCtl* ctl, double* hF,
                                double* hR, double* hA, size_t hstride, hipStream_t st);
void launch_reduce_period_copies(const double* part_check, size_t pstride, int nw, int ncopies,
                                 int n, int check, double invF, JobCtl* ctl, const double* hF,
                                 size_t hstride, hipStream_t st);
// end-of-sweep reduction of every copy (one workgroup per copy): copy c
// reduces part + c*pstride into ctl[c] and the history at hF + c*hstride
void launch_reduce_copies(const double* part, size_t pstride, int nparts, int ncopies,
                          double invF, double invV, JobCtl* ctl, double* hF, double* hR,
                          double* hA, size_t hstride, hipStream_t st);

// per-sweep (or initial) reduction of the partials into the job history
constexpr int kRedBlock = 1024;  // threads of the one-workgroup reduction
// The per-sweep reductions (RED_SWEEP) run on two levels in one launch:
// workgroup b folds partial triples [b kFold, (b+1) kFold) into fold triple
// b (fixed order), and the workgroup that arrives last (ctl->arrive) sums
// the fold triples in index order and updates the control block; the fold
// triples of a job (or copy) sit right after its 2*nw sweep partials.
// One workgroup reading every partial took up to a fifth of a sweep (2 MB
// of partials at 48^4); the device-scope fences of the two levels cost a
// few microseconds, more than they save below kFoldMin fold blocks, where
// one workgroup reduces everything (red_blocks = 1).  Either way the
// summation order depends only on nparts.
constexpr int kFold = kRedBlock;
constexpr int kFoldMin = 16;
__host__ __device__ inline int fold_blocks(int nparts) { return (nparts + kFold - 1) / kFold; }
__host__ __device__ inline int red_blocks(int nparts) {
  return fold_blocks(nparts) > kFoldMin ? fold_blocks(nparts) : 1;
}
// doubles per job (or copy, or FP32-phase sweep slot) of a partial buffer:
// the 2*nw sweep triples, then their fold triples
__host__ __device__ inline size_t part_stride(int nw) {
  return (size_t)3 * (2 * nw + fold_blocks(2 * nw));
}
void launch_reduce(const double* part, int nparts, int mode, double invF,
                   double invV, JobCtl* ctl, double* hist_F, double* hist_rmax,
                   double* hist_ravg, double* out3, hipStream_t st);

// ---------------------------------------------------------------- batches
// One job of a batched launch.  A batched half-sweep grid is the
// concatenation of the member jobs' own grids (block0 = first workgroup of
// the job); each workgroup runs exactly the code and the partition of the
// job's single launch, so a job's results do not depend on its batch.
struct BatchJob {
  void* U;        // links of the sweep (FP32 working or FP64)
  void* G;        // accumulated site gauge (FP32 phase) or nullptr
  const double* U0;  // FP64 master (mixed) for the rebuild
  Geom g;
  double* part;
  JobCtl* ctl;
  double* hF;
  double* hR;
  double* hA;
  double omega;
  double invF;
  double invV;
  int nb;      // workgroups of one half-sweep of this job (default shape)
  int nw;      // partial triples of one half-sweep (one per wave)
  int nparts;  // partial triples a sweep's reduction reads, from part + roff: the
               // nw black triples (link form: they hold the red residuals too),
               // or 2*nw (site forms); their fold triples follow them
  int roff;    // offset in doubles of those triples from part (3*nw or 0)
  int block0;  // first workgroup of this job in a batched half-sweep grid
  int red0;    // first workgroup of this entry in a batched per-sweep reduction
  // spin form (gauge copies): spins of copy c at S + c*sstride, partials
  // at part + c*pstride, control block ctl + c; ncopies = 0 in the link form
  void* S;  // spins (double) / site gauges (the link type)
  size_t sstride;
  size_t pstride;
  int ncopies;
};

// batched half-sweep of one class (group, FP32 phase or FP64) table with
// launch shape sh (block0 / nblocks_total in workgroups of that shape);
// slot and mode as in HalfOpt (the FP32 class: G of the red sites, U0 and
// the rebuild partials of each job from its table entry)
void launch_half_batch(int group, bool fp32, Shape sh, const BatchJob* jobs, int njobs,
                       int nblocks_total, int parity, hipStream_t st, int slot = 0,
                       int mode = HALF_PLAIN);
// batched launch_reduce_slots / launch_reduce_period over a table (link-form
// jobs, or the copy entries of FP32 SU(3) site-form sets: slot s of copy c
// at the set's part + (s*ncopies + c)*pstride, check partials in slot kSlots)
void launch_reduce_slots_batch(const BatchJob* jobs, int njobs, int n, hipStream_t st);
void launch_reduce_period_batch(const BatchJob* jobs, int njobs, int n, int rebuild,
                                hipStream_t st);
// batched spin-form (gauge-copy) half-sweep of a table of copy sets that
// all hold ncopies copies
void launch_copies_half_batch(int ncopies, bool fp32, const BatchJob* jobs, int njobs,
                              int nblocks_total, int parity, hipStream_t st);
// batched SU(3) site-form half-sweep: the FP64 class (any copy counts) or
// the FP32 class of sets of ncopies copies
void launch_su3_site_copies_batch(int ncopies, bool fp32, const BatchJob* jobs, int njobs,
                                  int nblocks_total, int parity, hipStream_t st, int slot = 0);
// launch_su3_site_F for an FP32 site-form class (half-sweep table, grid
// nblocks_total): F partials of each set's copies in its slot kSlots; the
// period reduction of the class's copy entries tests them
void launch_su3_site_F_batch(const BatchJob* jobs, int njobs, int nblocks_total, hipStream_t st);
// batched per-sweep reduction (RED_SWEEP) of every entry of a table:
// red_blocks(nparts) workgroups per entry (red0 ascending, nblocks_total
// in all)
void launch_reduce_batch(const BatchJob* jobs, int njobs, int nblocks_total, hipStream_t st);
// every entry's control block into out (host-visible memory, n entries)
void launch_gather_ctl(const BatchJob* jobs, int n, JobCtl* out, hipStream_t st);

// functional partials of a whole configuration (both parities): 2*nblocks
template <typename T>
void launch_su2_functional(const T* U, const Geom& g, double* part, hipStream_t st);
template <typename T>
void launch_su3_functional(const T* U, const Geom& g, double* part, hipStream_t st);

// layout conversion: reference AoS doubles <-> parity-split SoA
template <typename T>
void launch_su2_aos_to_soa(const double* aos, T* U, const Geom& g, hipStream_t st);
template <typename T>
void launch_su2_soa_to_aos(const T* U, double* aos, const Geom& g, hipStream_t st);
template <typename T>
void launch_su3_aos_to_soa(const double* aos, T* U, const Geom& g, hipStream_t st);
template <typename T>
void launch_su3_soa_to_aos(const T* U, double* aos, const Geom& g, hipStream_t st);

// FP64 site gauges (SU(2), both parities) set to the identity
void launch_gauge_identity64(double* G, const Geom& g, hipStream_t st);
// gauge tracking (SU(2) link form): spins s(x) = row 3 of the rotation of
// the accumulated gauge, in the spin SoA layout (S[(p*3+c)*Vh + h]); the
// links in the reference's convention G_c(x) U_in G_c(x+mu)^+ computed from
// the current links as u(x) U u(x+mu)^+, u = G_c G_acc^+, G_c =
// GetGaugeMatrix(s) (chunked FP64 links into out)
void launch_gauge_spins(const double* Gacc, double* S, const Geom& g, hipStream_t st);
void launch_canonical_links(const double* U, const double* S, const double* Gacc, double* out,
                            const Geom& g, hipStream_t st);
// FP64 links -> FP32 links (both in the chunked SoA layout)
void launch_links_f64_to_f32(const double* a, float* b, const Geom& g, int group, hipStream_t st);
// site gauges set to the identity (NR reals per site, both parities)
void launch_gauge_identity(float* G, const Geom& g, int group, hipStream_t st);

// The FP64 links at the end of the FP32 phase: out64 (may alias U0) =
// G(x) U0 G(x+mu)^+, one thread per black site rebuilding its 8 links,
// with the red gauges G_r normalised in FP64 and the black gauges recovered
// from the working links W (mag_gauge.h).
// Gacc (gauge tracking, SU(2)): the accumulated FP64 site gauges, which
// take the FP32 phase's gauge (G_acc <- G G_acc)
void launch_su2_rebuild_final(const double* U0, const float* Gr, const float* W, double* out64,
                              const Geom& g, hipStream_t st, double* Gacc = nullptr);
void launch_su3_rebuild_final(const double* U0, const float* Gr, const float* W, double* out64,
                              const Geom& g, hipStream_t st);
// FP64 reunitarisation of links in place (SU(2) q/|q|, SU(3) Gram-Schmidt)
void launch_su2_reunit64(double* U, const Geom& g, hipStream_t st);
void launch_su3_reunit64(double* U, const Geom& g, hipStream_t st);

// seeded synthetic configurations in the reference AoS layout (FP64);
// SU(3) needs V*18 doubles of scratch for the site gauges
void launch_gen_su2(double* aos, const Geom& g, double eps, uint64_t seed, uint64_t gseed,
                    hipStream_t st);
void launch_gen_su3(double* aos, double* gauge_scratch, const Geom& g, double eps,
                    uint64_t seed, uint64_t gseed, hipStream_t st);
// key of gauge copy `copy` of the configuration seeded by `seed` (copy 0 = seed)
__host__ __device__ inline uint64_t gauge_copy_seed(uint64_t seed, uint32_t copy) {
  return seed ^ ((uint64_t)copy * 0x9E3779B97F4A7C15ull);
}

// plaquette partials (sum over sites of the 6-plane mean), 2*nblocks triples
void launch_su2_plaquette(const double* U, const Geom& g, double* part, hipStream_t st);
void launch_su3_plaquette(const double* U, const Geom& g, double* part, hipStream_t st);

}  // namespace mag

namespace mag {
// 16-byte streaming copy of bytes (a multiple of 16)
void launch_copy(const void* a, void* b, size_t bytes, hipStream_t st);
}  // namespace mag

namespace mag {
// ---------------------------------------------------------------- spin form
// Driver state of the spin-form maximisation, device resident.  The relax
// half-sweeps return at once when stop_next or done is set; the reduction
// that follows a relax sweep turns stop_next into done, so the reflection
// sweeps of the iteration that met the criterion still run, as in
// make_maximization_final (mag.cpp:1800-1816).
struct SpinCtl {
  int done;
  int stop_next;
  int iter;
  int max_iter;
  int converged;
  int digits;      // make_maximization_approximate tolerance digits
  float scale;     // powf(10, digits), evaluated on the host by the C library
                   // as the reference's criterion does (mag.cpp:1745-1746)
  double tol_max;  // make_maximization_final criteria
  double tol_avg;
  double dmax;     // last relaxation residuals
  double davg;
  double F_prev;   // last spin functional (approximate criterion)
};

enum { SPIN_OR = 0, SPIN_RELAX = 1, SPIN_OMEGA = 2, SPIN_HEATBATH = 3 };
enum { SRED_RELAX_FINAL = 0, SRED_RELAX_PLAIN = 1, SRED_FUNC_INIT = 2, SRED_FUNC_APPROX = 3,
       SRED_PLAIN = 4 };

// checkerboard spin half-sweep; partial triples at part[3*nblocks*parity]
void launch_spin_half(int mode, const double* U, double* S, const Geom& g, int parity,
                      double param, uint64_t seed, uint32_t sweep, double* part,
                      const SpinCtl* ctl, hipStream_t st);
// spin functional partials, 2*nblocks triples (ctl may be nullptr)
void launch_spin_functional(const double* U, const double* S, const Geom& g, double* part,
                            const SpinCtl* ctl, hipStream_t st);
void launch_spin_reduce(const double* part, int nparts, int mode, double invV, double invF,
                        SpinCtl* ctl, double* out3, hipStream_t st);
void launch_spin_gauge_transform(double* U, const double* S, const Geom& g, hipStream_t st);
void launch_spins_uniform(double* aos, const Geom& g, uint64_t seed, hipStream_t st);
// to_soa = 1: AoS -> SoA, 0: SoA -> AoS
void launch_spins_layout(const double* aos, double* S, const Geom& g, int to_soa,
                         hipStream_t st);
// every spin of n copies (nreal = n*6*Vh reals) set to e3
void launch_spins_e3(double* S, size_t nreal, const Geom& g, hipStream_t st);
// spins of one copy FP64 -> FP32 / FP32 -> FP64, normalised in FP64
void launch_spins_to32(const double* a, float* b, const Geom& g, hipStream_t st);
void launch_spins_to64(const float* a, double* b, const Geom& g, hipStream_t st);

// initial spins of gauge copies 0..ncopies-1 of a synthetic configuration
// (copy c at S + c*sstride): spin of g_c(x) g_0(x)^+, copy 0 exactly e3
void launch_gen_copy_spins(double* S, size_t sstride, int ncopies, const Geom& g,
                           uint64_t seed, hipStream_t st);
// SU(3) site form: links of a copy out = G(x) U0 G(x+mu)^+ (FP64); site
// gauge arrays FP64 <-> FP32 (to FP64 with Gram-Schmidt); identity gauges;
// starting gauges of the synthetic gauge copies (copy c at G + c*gstride)
void launch_su3_site_links(const double* U0, const double* G, double* out, const Geom& g,
                           hipStream_t st);
void launch_su3_sites_to32(const double* a, float* b, const Geom& g, hipStream_t st);
void launch_su3_sites_to64(const float* a, double* b, const Geom& g, hipStream_t st);
void launch_su3_sites_identity(double* G, const Geom& g, hipStream_t st);
void launch_gen_su3_copy_gauges(double* G, size_t gstride, int ncopies, const Geom& g,
                                uint64_t seed, hipStream_t st);
// out = G(s(x)) U G(s(x+mu))^+ with G = gauge_of_spin (chunked FP64 links)
void launch_spin_links(const double* U, const double* S, double* out, const Geom& g,
                       hipStream_t st);
}  // namespace mag
