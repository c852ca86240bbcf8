// mag_sweep.hip -- per-sweep reductions, the FP32 instantiations of the
// link-form half-sweeps (mag_sweep_link.h; the FP64 ones are in
// mag_sweep64.hip), and the site-form gauge-copy half-sweeps (SU(2) spins,
// SU(3) site gauges) of the MAG path (gfx950).
#include "mag_internal.h"
#include "mag_layout.h"
#include "mag_math.h"
#include "mag_reduce.h"
#include "mag_spin_math.h"
#include "mag_sweep_link.h"

namespace mag {

// ================================================================== reduction
// Reduce nparts partial triples (F sum, residual sum, residual max) and
// apply the result to the job's control block; one workgroup of kRedBlock
// threads.  Thread t combines partials t, t + kRedBlock, ... in order, then
// the fixed-order workgroup tree, so the result depends only on the job's
// partials (alone or in a batch).
// (sum a, sum b, max c) of nparts triples in a fixed order; valid in thread 0
__device__ void reduce_triples(const double* __restrict__ part, int nparts, double& a, double& b,
                               double& c) {
  a = 0.0;
  b = 0.0;
  c = 0.0;
  int i = threadIdx.x;
  // kRedUnroll triples in flight per thread, combined in index order
  constexpr int kRedUnroll = 8;
  for (; i + (kRedUnroll - 1) * kRedBlock < nparts; i += kRedUnroll * kRedBlock) {
    double x[kRedUnroll][3];
#pragma unroll
    for (int k = 0; k < kRedUnroll; k++)
#pragma unroll
      for (int j = 0; j < 3; j++) x[k][j] = part[3 * (i + k * kRedBlock) + j];
#pragma unroll
    for (int k = 0; k < kRedUnroll; k++) {
      a += x[k][0];
      b += x[k][1];
      c = fmax(c, x[k][2]);
    }
  }
  for (; i < nparts; i += kRedBlock) {
    a += part[3 * i];
    b += part[3 * i + 1];
    c = fmax(c, part[3 * i + 2]);
  }
  block_reduce3<kRedBlock / 64>(a, b, c);
}

// F of freshly rebuilt FP64 links (mixed FP32 phase): converged when the
// mean change per sweep since the previous rebuild is below tol (thread 0)
__device__ __forceinline__ void apply_rebuild(double F, JobCtl* ctl) {
  if (ctl->F_ref != 0.0) {
    const double rel = fabs(F - ctl->F_ref) / (ctl->rb_every * fabs(F));
    if (rel < ctl->tol) {
      ctl->converged = 1;
      ctl->done = 1;
    }
  }
  ctl->F_ref = F;
}

__device__ void reduce_job(const double* __restrict__ part, int nparts, int mode, double invF,
                           double invV, JobCtl* ctl, double* hF, double* hR, double* hA,
                           double* out3) {
  double a, b, c;
  reduce_triples(part, nparts, a, b, c);
  if (threadIdx.x == 0) {
    const double F = a * invF;
    if (mode == RED_INIT) {
      ctl->F_prev = F;
      ctl->F0 = F;
    } else if (mode == RED_RESET) {
      ctl->F_prev = F;
    } else if (mode == RED_SWEEP) {
      const int sw = ctl->sweep;
      hF[sw] = F;
      hR[sw] = c;
      hA[sw] = b * invV;
      const int w = ctl->window;
      double rel = fabs(F - ctl->F_prev) / fabs(F);
      if (w > 1) rel = sw >= w ? fabs(F - hF[sw - w]) / (w * fabs(F)) : INFINITY;
      if (w < 0) rel = INFINITY;
      ctl->F_prev = F;
      ctl->sweep = sw + 1;
      if (rel < ctl->tol) {
        ctl->converged = 1;
        ctl->done = 1;
      }
      if (sw + 1 >= ctl->max_sweeps) ctl->done = 1;
    }
    if (out3) {
      out3[0] = F;
      out3[1] = b * invV;
      out3[2] = c;
    }
  }
}

__global__ __launch_bounds__(kRedBlock) void k_reduce(const double* __restrict__ part,
                                                      int nparts, int mode, double invF,
                                                      double invV, JobCtl* ctl, double* hF,
                                                      double* hR, double* hA, double* out3) {
  if (mode == RED_SWEEP && ctl->done) return;
  reduce_job(part, nparts, mode, invF, invV, ctl, hF, hR, hA, out3);
}

// one workgroup per job of the table
__global__ __launch_bounds__(kRedBlock) void k_reduce_batch(const BatchJob* __restrict__ jobs,
                                                            int mode) {
  const BatchJob& j = jobs[blockIdx.x];
  if (j.ctl->done) return;
  reduce_job(j.part, j.nparts, mode, j.invF, j.invV, j.ctl,
             j.hF, j.hR, j.hA, nullptr);
}

void launch_reduce(const double* part, int nparts, int mode, double invF, double invV,
                   JobCtl* ctl, double* hF, double* hR, double* hA, double* out3,
                   hipStream_t st) {
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kRedBlock), 0, st, part, nparts, mode, invF, invV,
                     ctl, hF, hR, hA, out3);
}
void launch_reduce_batch(const BatchJob* jobs, int njobs, int mode, hipStream_t st) {
  if (njobs == 0) return;
  hipLaunchKernelGGL(k_reduce_batch, dim3(njobs), dim3(kRedBlock), 0, st, jobs, mode);
}

// ------------------------------------------------------ FP32-phase slots
// Block k: the history entry of slot k (sweep ctl->sweep + k, below the cap),
// reduced exactly as RED_SWEEP reduces one sweep.  Nothing in the control
// block changes here, so every block sees the same ctl.
__device__ __forceinline__ void reduce_slot(const double* __restrict__ part, size_t pstride,
                                            int nparts, double invF, double invV,
                                            const JobCtl* ctl, double* hF, double* hR,
                                            double* hA, int k) {
  if (ctl->done) return;
  const int sw = ctl->sweep + k;
  if (sw >= ctl->max_sweeps) return;
  double a, b, c;
  reduce_triples(part + (size_t)k * pstride, nparts, a, b, c);
  if (threadIdx.x == 0) {
    hF[sw] = a * invF;
    hR[sw] = c;
    hA[sw] = b * invV;
  }
}
__global__ __launch_bounds__(kRedBlock) void k_reduce_slots(const double* __restrict__ part,
                                                            size_t pstride, int nparts,
                                                            double invF, double invV,
                                                            const JobCtl* ctl, double* hF,
                                                            double* hR, double* hA) {
  reduce_slot(part, pstride, nparts, invF, invV, ctl, hF, hR, hA, blockIdx.x);
}
// After the slots: the sweep count advances by n (up to the cap), F_prev is
// the last recorded F, and a period that ended with a rebuild (its
// rebuilding sweep ran) tests convergence on the rebuilt links' F.
__device__ __forceinline__ void reduce_period(const double* __restrict__ part_rb, int nrb, int n,
                                              int rebuild, double invF, JobCtl* ctl,
                                              const double* hF) {
  if (ctl->done) return;
  const int s0 = ctl->sweep, mx = ctl->max_sweeps;
  const bool rb = rebuild && s0 + n <= mx;
  double a = 0.0, b = 0.0, c = 0.0;
  if (rb) reduce_triples(part_rb, nrb, a, b, c);
  if (threadIdx.x == 0) {
    const int sw = s0 + n < mx ? s0 + n : mx;
    if (sw > s0) ctl->F_prev = hF[sw - 1];
    ctl->sweep = sw;
    if (sw >= mx) ctl->done = 1;
    if (rb) apply_rebuild(a * invF, ctl);
  }
}
__global__ __launch_bounds__(kRedBlock) void k_reduce_period(const double* __restrict__ part_rb,
                                                             int nrb, int n, int rebuild,
                                                             double invF, JobCtl* ctl,
                                                             const double* hF) {
  reduce_period(part_rb, nrb, n, rebuild, invF, ctl, hF);
}
void launch_reduce_slots(const double* part, size_t pstride, int nparts, int n, double invF,
                         double invV, JobCtl* ctl, double* hF, double* hR, double* hA,
                         hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_reduce_slots, dim3(n), dim3(kRedBlock), 0, st, part, pstride, nparts,
                     invF, invV, ctl, hF, hR, hA);
}
void launch_reduce_period(const double* part_rb, int nrb, int n, int rebuild, double invF,
                          JobCtl* ctl, double* hF, hipStream_t st) {
  hipLaunchKernelGGL(k_reduce_period, dim3(1), dim3(kRedBlock), 0, st, part_rb, nrb, n, rebuild,
                     invF, ctl, hF);
}
// batched: slots of job b / n, slot b % n; the period: one block per job
__global__ __launch_bounds__(kRedBlock) void k_reduce_slots_batch(const BatchJob* __restrict__ jobs,
                                                                  int n) {
  const BatchJob& j = jobs[blockIdx.x / n];
  reduce_slot(j.part, j.pstride, j.nparts, j.invF, j.invV, j.ctl, j.hF, j.hR, j.hA,
              (int)(blockIdx.x % n));
}
__global__ __launch_bounds__(kRedBlock) void k_reduce_period_batch(
    const BatchJob* __restrict__ jobs, int n, int rebuild) {
  const BatchJob& j = jobs[blockIdx.x];
  reduce_period(j.part + (size_t)kSlots * j.pstride, j.nw, n, rebuild, j.invF, j.ctl, j.hF);
}
void launch_reduce_slots_batch(const BatchJob* jobs, int njobs, int n, hipStream_t st) {
  if (njobs == 0 || n <= 0) return;
  hipLaunchKernelGGL(k_reduce_slots_batch, dim3(njobs * n), dim3(kRedBlock), 0, st, jobs, n);
}
void launch_reduce_period_batch(const BatchJob* jobs, int njobs, int n, int rebuild,
                                hipStream_t st) {
  if (njobs == 0) return;
  hipLaunchKernelGGL(k_reduce_period_batch, dim3(njobs), dim3(kRedBlock), 0, st, jobs, n, rebuild);
}

// control block of every entry of a table into out[i] (pinned host memory):
// the host reads a finished replay's state without a copy call per job
__global__ __launch_bounds__(64) void k_gather_ctl(const BatchJob* __restrict__ jobs, int n,
                                                   JobCtl* __restrict__ out) {
  constexpr int W = sizeof(JobCtl) / 8;
  const int i = blockIdx.x * 64 + (int)threadIdx.x;
  if (i >= n * W) return;
  const int e = i / W, w = i % W;
  reinterpret_cast<uint64_t*>(out + e)[w] = reinterpret_cast<const uint64_t*>(jobs[e].ctl)[w];
}
void launch_gather_ctl(const BatchJob* jobs, int n, JobCtl* out, hipStream_t st) {
  constexpr int W = sizeof(JobCtl) / 8;
  if (n == 0) return;
  hipLaunchKernelGGL(k_gather_ctl, dim3((n * W + 63) / 64), dim3(64), 0, st, jobs, n, out);
}

// ================================================================== link form, FP32
template void launch_su2_half<float>(float*, float*, const Geom&, int, double, double*,
                                     const JobCtl*, Shape, hipStream_t,
                                     const HalfOpt&);
template void launch_su3_half<float>(float*, float*, const Geom&, int, double, double*,
                                     const JobCtl*, Shape, hipStream_t,
                                     const HalfOpt&);

// FP64 tables: mag_sweep64.hip
void launch_half_batch_f64(int group, Shape sh, const BatchJob* jobs, int njobs,
                           int nblocks_total, int parity, hipStream_t st);

void launch_half_batch(int group, bool fp32, Shape sh, const BatchJob* jobs, int njobs,
                       int nblocks_total, int parity, hipStream_t st, int slot, int mode) {
  if (fp32) launch_half_batch_t<float>(group, sh, jobs, njobs, nblocks_total, parity, st, slot, mode);
  else launch_half_batch_f64(group, sh, jobs, njobs, nblocks_total, parity, st);
}

// ================================================================== gauge copies
// Spin-form half-sweep over the gauge copies of one SU(2) configuration,
// one site per lane (the reference's spin form, mag.cpp:569-626, with the
// omega step of the link form).  K (the set's copy count) is a template
// parameter so the K field accumulators live in registers.  Each of the 8
// links of the site is loaded once, its adjoint matrix M formed once, and
// applied to the neighbour spin of every copy that is not done:
//   h = sum_mu [ M(U_x,mu) s(x+mu) + M(U_x-mu,mu)^T s(x-mu) ]
// (in the order f0, b0, f1, b1, f2, b2, f3, b3 of contribution_site); then
// each copy rotates s(x) towards h_hat by omega*theta.  Per copy and workgroup the
// partial triple is (sum s'.h on the black half, sum d, max d) with
// d = 1 - s.h_hat before the step: every link has exactly one black end,
// so the black half's sum of s'.h is 4V times the spin functional
// (mag.cpp:925-953) after the sweep.  A copy's arithmetic and partition do
// not depend on the other copies, so its trajectory is the same in a set
// of any size.
template <typename TL>
__device__ __forceinline__ Q<double> ld_link64(const TL* U, const Geom& g, int p, int mu,
                                               uint32_t h) {
  const Q<TL> q = ldq(link_ptr<4>(U, g, p, mu, h));
  return Q<double>{(double)q.a0, (double)q.a1, (double)q.a2, (double)q.a3};
}

template <int K, typename TL>
__device__ __forceinline__ void su2_copies_block(const TL* __restrict__ U,
                                                 double* __restrict__ S, size_t sstride,
                                                 const Geom& g, int p, double omega,
                                                 double* __restrict__ part, size_t pstride,
                                                 const JobCtl* __restrict__ ctl, uint32_t lb) {
#pragma clang fp contract(off)
  bool act[K];
  bool live = false;
#pragma unroll
  for (int k = 0; k < K; k++) {
    act[k] = !ctl[k].done;  // uniform over the workgroup
    live |= act[k];
  }
  if (!live) return;
  const uint32_t h = lb * kBlock + threadIdx.x;
  double fs[K], ds[K], dm[K];
#pragma unroll
  for (int k = 0; k < K; k++) fs[k] = ds[k] = dm[k] = 0.0;
  if (h < g.Vh) {
    const SiteIdx s = decode(g, h, p);
    V3 H[K];
#pragma unroll
    for (int k = 0; k < K; k++) H[k] = V3{0.0, 0.0, 0.0};
#pragma unroll
    for (int mu = 0; mu < 4; mu++) {
      const uint32_t hf = fwd_h(g, h, s, mu);
      const Adj Mf = adj_of(ld_link64(U, g, p, mu, h));
#pragma unroll
      for (int k = 0; k < K; k++)
        if (act[k]) add_M(H[k], Mf, ld_spin(S + k * sstride, g, 1 - p, hf));
      const uint32_t hb = back_h(g, h, s, mu);
      const Adj Mb = adj_of(ld_link64(U, g, 1 - p, mu, hb));
#pragma unroll
      for (int k = 0; k < K; k++)
        if (act[k]) add_MT(H[k], Mb, ld_spin(S + k * sstride, g, 1 - p, hb));
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
      if (!act[k]) continue;
      double* Sk = S + k * sstride;
      V3 sp = ld_spin(Sk, g, p, h);
      const V3& Hk = H[k];
      const double hn = sqrt(Hk.a1 * Hk.a1 + Hk.a2 * Hk.a2 + Hk.a3 * Hk.a3);
      const double d = hn > 0 ? 1 - (sp.a1 * Hk.a1 + sp.a2 * Hk.a2 + sp.a3 * Hk.a3) / hn : 0.0;
      spin_omega(sp, Hk, omega);
      st_spin(Sk, g, p, h, sp);
      if (p == 1) fs[k] = sp.a1 * Hk.a1 + sp.a2 * Hk.a2 + sp.a3 * Hk.a3;
      ds[k] = d;
      dm[k] = d;
    }
  }
#pragma unroll
  for (int k = 0; k < K; k++)
    if (act[k])  // uniform
      wave_partial(part + k * pstride, lb * kBlock + (threadIdx.x & ~63u), g.Vh, fs[k], ds[k],
                   dm[k]);
}

template <int K, typename TL>
__global__ __launch_bounds__(kBlock) void k_su2_copies_half(const TL* __restrict__ U,
                                                            double* __restrict__ S,
                                                            size_t sstride, Geom g, int p,
                                                            double omega,
                                                            double* __restrict__ part,
                                                            size_t pstride,
                                                            const JobCtl* __restrict__ ctl) {
  su2_copies_block<K, TL>(U, S, sstride, g, p, omega, part, pstride, ctl,
                          xcd_block(blockIdx.x, gridDim.x));
}

// every copy set of the table has K copies and links of type TL
template <int K, typename TL>
__global__ __launch_bounds__(kBlock) void k_copies_half_batch(const BatchJob* __restrict__ jobs,
                                                              int njobs, int p) {
  const uint32_t lb = xcd_block(blockIdx.x, gridDim.x);
  const BatchJob& j = jobs[batch_job(jobs, njobs, (int)lb)];
  su2_copies_block<K, TL>((const TL*)j.U, (double*)j.S, j.sstride, j.g, p, j.omega,
                          j.part + (size_t)3 * j.nw * p, j.pstride, j.ctl,
                          lb - j.block0);
}

// one workgroup per copy
__global__ __launch_bounds__(kRedBlock) void k_reduce_copies(const double* __restrict__ part,
                                                             size_t pstride, int nparts,
                                                             double invF, double invV,
                                                             JobCtl* ctl, double* hF, double* hR,
                                                             double* hA, size_t hstride) {
  const int c = blockIdx.x;
  if (ctl[c].done) return;
  reduce_job(part + (size_t)c * pstride, nparts, RED_SWEEP, invF, invV, ctl + c,
             hF + (size_t)c * hstride, hR + (size_t)c * hstride, hA + (size_t)c * hstride,
             nullptr);
}

template <typename TL>
void launch_su2_copies_half(const TL* U, double* S, size_t sstride, int ncopies, const Geom& g,
                            int parity, double omega, double* part, size_t pstride,
                            const JobCtl* ctl, hipStream_t st) {
  const int nb = nblocks(g);
  double* pp = part + (size_t)3 * sweep_waves(g, 2) * parity;
#define L_(K)                                                                                   \
  hipLaunchKernelGGL((k_su2_copies_half<K, TL>), dim3(nb), dim3(kBlock), 0, st, U, S, sstride, \
                     g, parity, omega, pp, pstride, ctl)
  switch (ncopies) {
    case 1: L_(1); break;
    case 2: L_(2); break;
    case 3: L_(3); break;
    case 4: L_(4); break;
    case 5: L_(5); break;
    case 6: L_(6); break;
    case 7: L_(7); break;
    default: L_(8); break;
  }
#undef L_
}
template void launch_su2_copies_half<double>(const double*, double*, size_t, int, const Geom&,
                                             int, double, double*, size_t, const JobCtl*,
                                             hipStream_t);
template void launch_su2_copies_half<float>(const float*, double*, size_t, int, const Geom&, int,
                                            double, double*, size_t, const JobCtl*,
                                            hipStream_t);
void launch_reduce_copies(const double* part, size_t pstride, int nparts, int ncopies,
                          double invF, double invV, JobCtl* ctl, double* hF, double* hR,
                          double* hA, size_t hstride, hipStream_t st) {
  hipLaunchKernelGGL(k_reduce_copies, dim3(ncopies), dim3(kRedBlock), 0, st, part, pstride,
                     nparts, invF, invV, ctl, hF, hR, hA, hstride);
}
void launch_copies_half_batch(int ncopies, bool fp32, const BatchJob* jobs, int njobs,
                              int nblocks_total, int parity, hipStream_t st) {
  if (njobs == 0 || nblocks_total == 0) return;
#define L_(K)                                                                                 \
  do {                                                                                        \
    if (fp32)                                                                                 \
      hipLaunchKernelGGL((k_copies_half_batch<K, float>), dim3(nblocks_total), dim3(kBlock), \
                         0, st, jobs, njobs, parity);                                         \
    else                                                                                      \
      hipLaunchKernelGGL((k_copies_half_batch<K, double>), dim3(nblocks_total), dim3(kBlock),\
                         0, st, jobs, njobs, parity);                                         \
  } while (0)
  switch (ncopies) {
    case 1: L_(1); break;
    case 2: L_(2); break;
    case 3: L_(3); break;
    case 4: L_(4); break;
    case 5: L_(5); break;
    case 6: L_(6); break;
    case 7: L_(7); break;
    default: L_(8); break;
  }
#undef L_
}

// ================================================================== SU(3) gauge copies
// Site form of SU(3) (the SU(3) counterpart of the spin form): the links U
// stay read-only and each gauge copy c holds a site gauge field G_c.  The
// site's effective links are
//   forward  W_f = G X,  X = U_x,mu G_c(x+mu)^+      backward  W_b = Y G^+,  Y = G_c(x-mu) U_x-mu,mu
// with G = G_c(x).  Their diagonals are bilinear in the rows g_i of G:
// (W_f)_ii = g_i . x_i (x_i = column i of X) and (W_b)_ii = y_i . conj(g_i)
// (y_i = row i of Y), so every quantity the Cabibbo-Marinari update of G
// needs is a Hermitian form of three 3x3 matrices that do not depend on G:
//   H^(i) = sum_fwd x_i x_i^+ + sum_bwd conj(y_i) y_i^T.
// For the subgroup (I,J), with D = H^(I) - H^(J), the field of the link
// form's block sums (blk_fwd / blk_bwd) is
//   h0 - i h1 = g_I D g_J^+,   h2 = (g_I D g_I^+ - g_J D g_J^+) / 2,
// and the update rotates rows I, J of G.  A site builds H once from its 8
// links (X and Y each from two products of rows and the unitary row-2 rule)
// and then updates only G; only G is written.  Copy c follows the
// link-form trajectory of the links G_c U G_c^+; all copies of a
// configuration share the link reads.  Two lanes per site as in the link
// form: each builds H from two directions, the pair adds them.
template <typename T>
struct Herm3 {
  T d0, d1, d2;       // diagonal
  C<T> o01, o02, o12;  // upper off-diagonal (H_ba = conj(H_ab))
};
// H += v v^+  (H_ab += v_a conj(v_b))
template <typename T>
__device__ __forceinline__ void herm_add_col(Herm3<T>& H, C<T> v0, C<T> v1, C<T> v2) {
  H.d0 += cnorm2(v0);
  H.d1 += cnorm2(v1);
  H.d2 += cnorm2(v2);
  H.o01 = cadd(H.o01, cmulc(v0, v1));
  H.o02 = cadd(H.o02, cmulc(v0, v2));
  H.o12 = cadd(H.o12, cmulc(v1, v2));
}
// H += conj(y) y^T  (H_ab += conj(y_a) y_b)
template <typename T>
__device__ __forceinline__ void herm_add_row(Herm3<T>& H, C<T> y0, C<T> y1, C<T> y2) {
  H.d0 += cnorm2(y0);
  H.d1 += cnorm2(y1);
  H.d2 += cnorm2(y2);
  H.o01 = cadd(H.o01, cmulc(y1, y0));
  H.o02 = cadd(H.o02, cmulc(y2, y0));
  H.o12 = cadd(H.o12, cmulc(y2, y1));
}
// r = D w^+ (a column: r_a = sum_b D_ab conj(w_b))
template <typename T>
__device__ __forceinline__ void herm_mul_dag(const Herm3<T>& D, const C<T>* w, C<T>* r) {
  const C<T> c0 = cconj(w[0]), c1 = cconj(w[1]), c2 = cconj(w[2]);
  r[0] = cadd(cadd(C<T>{D.d0 * c0.re, D.d0 * c0.im}, cmul(D.o01, c1)), cmul(D.o02, c2));
  r[1] = cadd(cadd(cmul(cconj(D.o01), c0), C<T>{D.d1 * c1.re, D.d1 * c1.im}), cmul(D.o12, c2));
  r[2] = cadd(cadd(cmul(cconj(D.o02), c0), cmul(cconj(D.o12), c1)), C<T>{D.d2 * c2.re, D.d2 * c2.im});
}
template <typename T>
__device__ __forceinline__ C<T> dot3(const C<T>* u, const C<T>* r) {
  return cadd(cadd(cmul(u[0], r[0]), cmul(u[1], r[1])), cmul(u[2], r[2]));
}
template <typename T>
__device__ __forceinline__ void herm_pair_sum(Herm3<T>& H) {
  H.d0 += pair_xchg(H.d0);
  H.d1 += pair_xchg(H.d1);
  H.d2 += pair_xchg(H.d2);
  H.o01.re += pair_xchg(H.o01.re);
  H.o01.im += pair_xchg(H.o01.im);
  H.o02.re += pair_xchg(H.o02.re);
  H.o02.im += pair_xchg(H.o02.im);
  H.o12.re += pair_xchg(H.o12.re);
  H.o12.im += pair_xchg(H.o12.im);
}

// subgroup (I,J) update of the site gauge G (3 rows); returns 1 - cos(theta)
template <typename T, int I, int J>
__device__ __forceinline__ T site_subgroup(const Herm3<T>* H, C<T> (&G)[3][3], T omega) {
  Herm3<T> D;
  D.d0 = H[I].d0 - H[J].d0;
  D.d1 = H[I].d1 - H[J].d1;
  D.d2 = H[I].d2 - H[J].d2;
  D.o01 = csub(H[I].o01, H[J].o01);
  D.o02 = csub(H[I].o02, H[J].o02);
  D.o12 = csub(H[I].o12, H[J].o12);
  C<T> rJ[3], rI[3];
  herm_mul_dag(D, G[J], rJ);
  herm_mul_dag(D, G[I], rI);
  const C<T> z = dot3(G[I], rJ);
  const T h0 = z.re, h1 = -z.im;
  const T h2 = T(0.5) * (dot3(G[I], rI).re - dot3(G[J], rJ).re);
  T res;
  const Q<T> gq = omega_gauge(h0, h1, h2, omega, res);
  // r = [[g0 + i g3, g2 + i g1], [-g2 + i g1, g0 - i g3]], g3 = 0 (as in cm_subgroup)
  const C<T> r00{gq.a0, T(0)}, r01{gq.a2, gq.a1}, r10{-gq.a2, gq.a1}, r11{gq.a0, T(0)};
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const C<T> a = G[I][c], b = G[J][c];
    G[I][c] = cadd(cmul(r00, a), cmul(r01, b));
    G[J][c] = cadd(cmul(r10, a), cmul(r11, b));
  }
  return res;
}

// the contributions of one direction to H^(0..2): the forward link U_f
// with the site gauge G_f at x + mu, the backward link U_b with G_b at x - mu
template <typename T>
__device__ __forceinline__ void site_fwd_field(Herm3<T>* H, const M2<T>& Uf, const M2<T>& Gf) {
  {
    // X = U_f G_f^+: rows 0, 1 from the stored rows of U_f and all rows of
    // G_f, row 2 = conj(row0 x row1)
    C<T> gf2[3];
    row2_of(Gf, gf2);
    M2<T> X;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 3; b++) {
        const C<T>* gb = b < 2 ? Gf.r[b] : gf2;
        X.r[a][b] = cadd(cadd(cmulc(Uf.r[a][0], gb[0]), cmulc(Uf.r[a][1], gb[1])),
                         cmulc(Uf.r[a][2], gb[2]));
      }
    C<T> x2[3];
    row2_of(X, x2);
#pragma unroll
    for (int i = 0; i < 3; i++) herm_add_col(H[i], X.r[0][i], X.r[1][i], x2[i]);
  }
}
template <typename T>
__device__ __forceinline__ void site_bwd_field(Herm3<T>* H, const M2<T>& Ub, const M2<T>& Gb) {
  {
    // Y = G_b U_b: rows 0, 1 from the stored rows of G_b and all rows of
    // U_b, row 2 = conj(row0 x row1)
    C<T> ub2[3];
    row2_of(Ub, ub2);
    M2<T> Y;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 3; b++)
        Y.r[a][b] = cadd(cadd(cmul(Gb.r[a][0], Ub.r[0][b]), cmul(Gb.r[a][1], Ub.r[1][b])),
                         cmul(Gb.r[a][2], ub2[b]));
    C<T> y2[3];
    row2_of(Y, y2);
    herm_add_row(H[0], Y.r[0][0], Y.r[0][1], Y.r[0][2]);
    herm_add_row(H[1], Y.r[1][0], Y.r[1][1], Y.r[1][2]);
    herm_add_row(H[2], y2[0], y2[1], y2[2]);
  }
}

template <typename T>
__device__ __forceinline__ void su3_site_copies_block(const T* __restrict__ U,
                                                      T* __restrict__ G, size_t gstride,
                                                      int ncopies, const Geom& g, int p,
                                                      T omega, double* __restrict__ part,
                                                      size_t pstride,
                                                      const JobCtl* __restrict__ ctl,
                                                      uint32_t lb) {
  bool live = false;
  for (int c = 0; c < ncopies; c++) live |= !ctl[c].done;
  if (!live) return;  // uniform over the workgroup
  const uint32_t tid = lb * kBlock + threadIdx.x;
  const uint32_t h = tid >> 1;
  const int q = (int)(tid & 1u);
  const bool in = h < g.Vh;
  SiteIdx s{};
  uint32_t hb[4] = {0, 0, 0, 0};
  if (in) {
    s = decode(g, h, p);
    back_all(g, h, s, hb);
  }
  for (int c = 0; c < ncopies; c++) {
    if (ctl[c].done) continue;  // uniform over the workgroup
    T* Gc = G + (size_t)c * gstride;
    double fs = 0.0, rs = 0.0, rm = 0.0;
    if (in) {
      Herm3<T> H[3];
#pragma unroll
      for (int i = 0; i < 3; i++) H[i] = Herm3<T>{T(0), T(0), T(0), {T(0), T(0)}, {T(0), T(0)}, {T(0), T(0)}};
      // FP32: the eight loads of the two directions first, then the
      // arithmetic (16^3x32, one copy: 24.3 -> 11.3 us per half-sweep);
      // FP64 loads each direction's links where it uses them (its
      // registers are full: 250 VGPRs)
      if constexpr (sizeof(T) == 4) {
        M2<T> Lf[2], Lgf[2], Lb[2], Lgb[2];
        auto load_dir = [&](int k) {
          const int mu = 2 * q + k;
          const uint32_t hfm = fwd_h(g, h, s, mu);
          const uint32_t hbm = q ? hb[2 + k] : hb[k];
          Lf[k] = ldm(link_ptr<12>(U, g, p, mu, h));
          Lgf[k] = ldm(site_ptr<12>(Gc, g, 1 - p, hfm));
          Lb[k] = ldm(link_ptr<12>(U, g, 1 - p, mu, hbm));
          Lgb[k] = ldm(site_ptr<12>(Gc, g, 1 - p, hbm));
        };
        load_dir(0);
        load_dir(1);
#pragma unroll
        for (int k = 0; k < 2; k++) {
          site_fwd_field(H, Lf[k], Lgf[k]);
          site_bwd_field(H, Lb[k], Lgb[k]);
        }
      } else {
#pragma unroll
        for (int k = 0; k < 2; k++) {
          const int mu = 2 * q + k;
          const uint32_t hbm = q ? hb[2 + k] : hb[k];
          site_fwd_field(H, ldm(link_ptr<12>(U, g, p, mu, h)),
                         ldm(site_ptr<12>(Gc, g, 1 - p, fwd_h(g, h, s, mu))));
          site_bwd_field(H, ldm(link_ptr<12>(U, g, 1 - p, mu, hbm)), ldm(site_ptr<12>(Gc, g, 1 - p, hbm)));
        }
      }
#pragma unroll
      for (int i = 0; i < 3; i++) herm_pair_sum(H[i]);
      const LRef<T> gp = site_ptr<12>(Gc, g, p, h);
      M2<T> Gx2 = ldm(gp);
      C<T> Gx[3][3];
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) Gx[a][b] = Gx2.r[a][b];
      row2_of(Gx2, Gx[2]);
      T res = site_subgroup<T, 0, 1>(H, Gx, omega);
      res += site_subgroup<T, 1, 2>(H, Gx, omega);
      res += site_subgroup<T, 0, 2>(H, Gx, omega);
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) Gx2.r[a][b] = Gx[a][b];
      if constexpr (sizeof(T) == 4) renorm_rows(Gx2);
      if (q == 0) {
        stm(gp, Gx2);
        rs = (double)res;
        rm = (double)res;
        if (p == 1) {
          // sum_i g_i H^(i) g_i^+ = sum over the site's 8 links of
          // sum_i |W_ii|^2, in double
#pragma unroll
          for (int i = 0; i < 3; i++) {
            Herm3<double> Hd{(double)H[i].d0, (double)H[i].d1, (double)H[i].d2,
                             {(double)H[i].o01.re, (double)H[i].o01.im},
                             {(double)H[i].o02.re, (double)H[i].o02.im},
                             {(double)H[i].o12.re, (double)H[i].o12.im}};
            C<double> gi[3], r[3];
#pragma unroll
            for (int b = 0; b < 3; b++) gi[b] = C<double>{(double)Gx[i][b].re, (double)Gx[i][b].im};
            herm_mul_dag(Hd, gi, r);
            fs += dot3(gi, r).re;
          }
        }
      }
    }
    wave_partial(part + (size_t)c * pstride, lb * kBlock + (threadIdx.x & ~63u), 2 * g.Vh, fs,
                 rs, rm);
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_su3_site_copies(const T* __restrict__ U,
                                                            T* __restrict__ G, size_t gstride,
                                                            int ncopies, Geom g, int p,
                                                            T omega, double* __restrict__ part,
                                                            size_t pstride,
                                                            const JobCtl* __restrict__ ctl) {
  su3_site_copies_block<T>(U, G, gstride, ncopies, g, p, omega, part, pstride, ctl,
                           xcd_block(blockIdx.x, gridDim.x));
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_su3_site_copies_batch(
    const BatchJob* __restrict__ jobs, int njobs, int p) {
  const uint32_t lb = xcd_block(blockIdx.x, gridDim.x);
  const BatchJob& j = jobs[batch_job(jobs, njobs, (int)lb)];
  su3_site_copies_block<T>((const T*)j.U, (T*)j.S, j.sstride, j.ncopies, j.g, p, (T)j.omega,
                           j.part + (size_t)3 * j.nw * p, j.pstride, j.ctl,
                           lb - j.block0);
}

template <typename T>
void launch_su3_site_copies(const T* U, T* G, size_t gstride, int ncopies, const Geom& g,
                            int parity, double omega, double* part, size_t pstride,
                            const JobCtl* ctl, hipStream_t st) {
  const int nb = (int)((site3_lanes(g) + kBlock - 1) / kBlock);
  const size_t nw = (site3_lanes(g) + 63) / 64;
  hipLaunchKernelGGL((k_su3_site_copies<T>), dim3(nb), dim3(kBlock), 0, st, U, G, gstride,
                     ncopies, g, parity, (T)omega, part + 3 * nw * parity, pstride, ctl);
}
template void launch_su3_site_copies<float>(const float*, float*, size_t, int, const Geom&, int,
                                            double, double*, size_t, const JobCtl*,
                                            hipStream_t);
template void launch_su3_site_copies<double>(const double*, double*, size_t, int, const Geom&,
                                             int, double, double*, size_t, const JobCtl*,
                                             hipStream_t);

void launch_su3_site_copies_batch(bool fp32, const BatchJob* jobs, int njobs, int nblocks_total,
                                  int parity, hipStream_t st) {
  if (njobs == 0 || nblocks_total == 0) return;
  if (fp32)
    hipLaunchKernelGGL((k_su3_site_copies_batch<float>), dim3(nblocks_total), dim3(kBlock), 0,
                       st, jobs, njobs, parity);
  else
    hipLaunchKernelGGL((k_su3_site_copies_batch<double>), dim3(nblocks_total), dim3(kBlock), 0,
                       st, jobs, njobs, parity);
}

}  // namespace mag
