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
  reduce_job(part, nparts, mode, invF, invV, ctl, hF, hR, hA, out3);
}

// Per-sweep reduction, workgroup b of the nb = red_blocks(nparts) of one
// job (or copy); with nb = 1 the one workgroup reduces every partial.
// Else: fold partial triples [b kFold, (b+1) kFold) into fold
// triple b at part + 3 nparts (one triple per thread, the fixed-order
// workgroup tree); the workgroup that arrives last on ctl->arrive (after a
// device-scope release of its fold triple) acquires the others', reduces
// the nb fold triples in index order (reduce_job, RED_SWEEP) and resets
// the counter.  The result depends only on the partials.  Every workgroup
// reads ctl->done before the last one can write it.
__device__ void reduce_sweep2(const double* __restrict__ part, int nparts, double invF,
                              double invV, JobCtl* ctl, double* hF, double* hR, double* hA,
                              int b) {
  if (ctl->done) return;  // uniform over the launch
  const int nb = red_blocks(nparts);
  if (nb == 1) {  // one workgroup: every partial in index order
    reduce_job(part, nparts, RED_SWEEP, invF, invV, ctl, hF, hR, hA, nullptr);
    return;
  }
  double* fold = const_cast<double*>(part) + 3 * (size_t)nparts;
  double a = 0.0, r = 0.0, m = 0.0;
  const int i = b * kFold + (int)threadIdx.x;
  if (i < nparts) {
    a = part[3 * i];
    r = part[3 * i + 1];
    m = part[3 * i + 2];
  }
  block_reduce3<kRedBlock / 64>(a, r, m);
  __shared__ int last;
  if (threadIdx.x == 0) {
    fold[3 * b] = a;
    fold[3 * b + 1] = r;
    fold[3 * b + 2] = m;
    __threadfence();
    last = atomicAdd(&ctl->arrive, 1) == nb - 1;
  }
  __syncthreads();
  if (!last) return;  // uniform over the workgroup
  __threadfence();
  reduce_job(fold, nb, RED_SWEEP, invF, invV, ctl, hF, hR, hA, nullptr);
  if (threadIdx.x == 0) ctl->arrive = 0;
}
__global__ __launch_bounds__(kRedBlock) void k_reduce_sweep(const double* __restrict__ part,
                                                            int nparts, double invF, double invV,
                                                            JobCtl* ctl, double* hF, double* hR,
                                                            double* hA) {
  reduce_sweep2(part, nparts, invF, invV, ctl, hF, hR, hA, (int)blockIdx.x);
}

// per-sweep reduction of a table: red_blocks(nparts) workgroups per entry
__device__ __forceinline__ int red_entry(const BatchJob* __restrict__ jobs, int n, int b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].red0 <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}
__global__ __launch_bounds__(kRedBlock) void k_reduce_batch(const BatchJob* __restrict__ jobs,
                                                            int njobs) {
  const BatchJob& j = jobs[red_entry(jobs, njobs, (int)blockIdx.x)];
  reduce_sweep2(j.part + j.roff, j.nparts, j.invF, j.invV, j.ctl, j.hF, j.hR, j.hA,
                (int)blockIdx.x - j.red0);
}

void launch_reduce(const double* part, int nparts, int mode, double invF, double invV,
                   JobCtl* ctl, double* hF, double* hR, double* hA, double* out3,
                   hipStream_t st) {
  if (mode == RED_SWEEP)
    hipLaunchKernelGGL(k_reduce_sweep, dim3(red_blocks(nparts)), dim3(kRedBlock), 0, st, part,
                       nparts, invF, invV, ctl, hF, hR, hA);
  else
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kRedBlock), 0, st, part, nparts, mode, invF, invV,
                       ctl, hF, hR, hA, out3);
}
void launch_reduce_batch(const BatchJob* jobs, int njobs, int nblocks_total, hipStream_t st) {
  if (njobs == 0 || nblocks_total == 0) return;
  hipLaunchKernelGGL(k_reduce_batch, dim3(nblocks_total), dim3(kRedBlock), 0, st, jobs, njobs);
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
// rebuilding sweep ran) tests convergence on the F of its end state: the
// rebuilt links' F in the mixed link form, the F evaluated in FP64 from the
// FP32 state in the SU(3) site form (part_rb: the partials of that F).
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
// batched: slots of entry b / n, slot b % n; the period: one block per
// entry.  A link-form entry's slot s is at part + s * pstride; a copy-set
// entry (copy c of an SU(3) site-form set of K copies, part = the set's
// buffer + c * pstride) has its slot s at part + s * K * pstride, and its
// check partials (the F of the stopping test) in slot kSlots.
__device__ __forceinline__ size_t slot_stride(const BatchJob& j) {
  return (size_t)(j.ncopies > 0 ? j.ncopies : 1) * j.pstride;
}
__global__ __launch_bounds__(kRedBlock) void k_reduce_slots_batch(const BatchJob* __restrict__ jobs,
                                                                  int n) {
  const BatchJob& j = jobs[blockIdx.x / n];
  reduce_slot(j.part + j.roff, slot_stride(j), j.nparts, j.invF, j.invV, j.ctl, j.hF, j.hR, j.hA,
              (int)(blockIdx.x % n));
}
__global__ __launch_bounds__(kRedBlock) void k_reduce_period_batch(
    const BatchJob* __restrict__ jobs, int n, int rebuild) {
  const BatchJob& j = jobs[blockIdx.x];
  reduce_period(j.part + (size_t)kSlots * slot_stride(j), j.nw, n, rebuild, j.invF, j.ctl, j.hF);
}
// one job's copy set: block b reduces slot b / ncopies of copy b % ncopies
// (history of copy c at hF + c * hstride); the period: one block per copy
__global__ __launch_bounds__(kRedBlock) void k_reduce_slots_copies(
    const double* __restrict__ part, size_t pstride, int nparts, int ncopies, double invF,
    double invV, const JobCtl* ctl, double* hF, double* hR, double* hA, size_t hstride) {
  const int c = (int)blockIdx.x % ncopies, k = (int)blockIdx.x / ncopies;
  const size_t hs = (size_t)c * hstride;
  reduce_slot(part + (size_t)c * pstride, (size_t)ncopies * pstride, nparts, invF, invV, ctl + c,
              hF + hs, hR + hs, hA + hs, k);
}
__global__ __launch_bounds__(kRedBlock) void k_reduce_period_copies(
    const double* __restrict__ part_check, size_t pstride, int nw, int n, int check,
    double invF, JobCtl* ctl, const double* hF, size_t hstride) {
  const int c = (int)blockIdx.x;
  reduce_period(part_check + (size_t)c * pstride, nw, n, check, invF, ctl + c,
                hF + (size_t)c * hstride);
}
void launch_reduce_slots_copies(const double* part, size_t pstride, int nparts, int ncopies, int n,
                                double invF, double invV, const JobCtl* ctl, double* hF,
                                double* hR, double* hA, size_t hstride, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_reduce_slots_copies, dim3(n * ncopies), dim3(kRedBlock), 0, st, part,
                     pstride, nparts, ncopies, invF, invV, ctl, hF, hR, hA, hstride);
}
void launch_reduce_period_copies(const double* part_check, size_t pstride, int nw, int ncopies,
                                 int n, int check, double invF, JobCtl* ctl, const double* hF,
                                 size_t hstride, hipStream_t st) {
  hipLaunchKernelGGL(k_reduce_period_copies, dim3(ncopies), dim3(kRedBlock), 0, st, part_check,
                     pstride, nw, n, check, invF, ctl, hF, hstride);
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
// TL = double: FP64 links and spins, the FP64 omega step.  TL = float (the
// FP32 phase): FP32-rounded links and FP32 spins; M is formed in FP64 from
// the FP32 link and h summed in FP64 from the spins scaled to unit length
// in FP64 (to_unit_d), so h and s'.h are functions of the FP32 state good
// to about 1e-14 and the functional resolves changes far below FP32
// rounding; the omega step is FP32 (spin_omega_f).
template <typename TL>
__device__ __forceinline__ Q<double> ld_link64(const TL* U, const Geom& g, int p, int mu,
                                               uint32_t h, bool nt) {
  const Q<TL> q = nt ? ldq_nt(link_ptr<4>(U, g, p, mu, h)) : ldq(link_ptr<4>(U, g, p, mu, h));
  return Q<double>{(double)q.a0, (double)q.a1, (double)q.a2, (double)q.a3};
}
__device__ __forceinline__ V3 ld_spin64(const double* S, const Geom& g, int p, uint32_t h) {
  return ld_spin(S, g, p, h);
}
__device__ __forceinline__ V3 ld_spin64(const float* S, const Geom& g, int p, uint32_t h) {
  return to_unit_d(ld_spin(S, g, p, h));
}

template <typename TL>
__device__ __forceinline__ Adj adj_of_t(const Q<double>& A) {
  if constexpr (sizeof(TL) == 4) return adj_of_fma(A);
  else return adj_of(A);
}

template <int K, typename TL, bool NT>
__device__ __forceinline__ void su2_copies_block_nt(const TL* __restrict__ U,
                                                    TL* __restrict__ S, size_t sstride,
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
      const Adj Mf = adj_of_t<TL>(ld_link64(U, g, p, mu, h, NT));
#pragma unroll
      for (int k = 0; k < K; k++)
        if (act[k]) add_M(H[k], Mf, ld_spin64(S + k * sstride, g, 1 - p, hf));
      const uint32_t hb = back_h(g, h, s, mu);
      const Adj Mb = adj_of_t<TL>(ld_link64(U, g, 1 - p, mu, hb, NT));
#pragma unroll
      for (int k = 0; k < K; k++)
        if (act[k]) add_MT(H[k], Mb, ld_spin64(S + k * sstride, g, 1 - p, hb));
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
      if (!act[k]) continue;
      TL* Sk = S + k * sstride;
      const V3& Hk = H[k];
      double d;
      V3 sn;
      if constexpr (sizeof(TL) == 4) {
        V3f sp = ld_spin(Sk, g, p, h);
        d = (double)spin_omega_f(sp, Hk, (float)omega);
        st_spin(Sk, g, p, h, sp);
        sn = to_unit_d(sp);
      } else {
        V3 sp = ld_spin(Sk, g, p, h);
        const double hn = sqrt(Hk.a1 * Hk.a1 + Hk.a2 * Hk.a2 + Hk.a3 * Hk.a3);
        d = hn > 0 ? 1 - (sp.a1 * Hk.a1 + sp.a2 * Hk.a2 + sp.a3 * Hk.a3) / hn : 0.0;
        spin_omega(sp, Hk, omega);
        st_spin(Sk, g, p, h, sp);
        sn = sp;
      }
      if (p == 1) fs[k] = sn.a1 * Hk.a1 + sn.a2 * Hk.a2 + sn.a3 * Hk.a3;
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

// links with the nontemporal hint where they exceed the MALL (uniform)
template <int K, typename TL>
__device__ __forceinline__ void su2_copies_block(const TL* __restrict__ U, TL* __restrict__ S,
                                                 size_t sstride, const Geom& g, int p,
                                                 double omega, double* __restrict__ part,
                                                 size_t pstride, const JobCtl* __restrict__ ctl,
                                                 uint32_t lb) {
  if (stream_links(g, 4 * sizeof(TL)))
    su2_copies_block_nt<K, TL, true>(U, S, sstride, g, p, omega, part, pstride, ctl, lb);
  else
    su2_copies_block_nt<K, TL, false>(U, S, sstride, g, p, omega, part, pstride, ctl, lb);
}

template <int K, typename TL>
__global__ __launch_bounds__(kBlock) void k_su2_copies_half(const TL* __restrict__ U,
                                                            TL* __restrict__ S,
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
  su2_copies_block<K, TL>((const TL*)j.U, (TL*)j.S, j.sstride, j.g, p, j.omega,
                          j.part + (size_t)3 * j.nw * p, j.pstride, j.ctl,
                          lb - j.block0);
}

// red_blocks(nparts) workgroups per copy
__global__ __launch_bounds__(kRedBlock) void k_reduce_copies(const double* __restrict__ part,
                                                             size_t pstride, int nparts,
                                                             double invF, double invV,
                                                             JobCtl* ctl, double* hF, double* hR,
                                                             double* hA, size_t hstride) {
  const int nb = red_blocks(nparts);
  const int c = (int)blockIdx.x / nb;
  reduce_sweep2(part + (size_t)c * pstride, nparts, invF, invV, ctl + c,
                hF + (size_t)c * hstride, hR + (size_t)c * hstride, hA + (size_t)c * hstride,
                (int)blockIdx.x % nb);
}

template <typename TL>
void launch_su2_copies_half(const TL* U, TL* S, size_t sstride, int ncopies, const Geom& g,
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
template void launch_su2_copies_half<float>(const float*, float*, size_t, int, const Geom&, int,
                                            double, double*, size_t, const JobCtl*,
                                            hipStream_t);
void launch_reduce_copies(const double* part, size_t pstride, int nparts, int ncopies,
                          double invF, double invV, JobCtl* ctl, double* hF, double* hR,
                          double* hA, size_t hstride, hipStream_t st) {
  hipLaunchKernelGGL(k_reduce_copies, dim3(ncopies * red_blocks(nparts)), dim3(kRedBlock), 0, st,
                     part, pstride, nparts, invF, invV, ctl, hF, hR, hA, hstride);
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
// and the update rotates rows I, J of G.  X and Y are unitary, so
// sum_i x_i x_i^+ = X X^+ = 1 and sum_i conj(y_i) y_i^T = conj(Y^+ Y) = 1
// for each of the 8 links: H^(2) = 8 - H^(0) - H^(1).  A link therefore
// contributes only x_0, x_1 (x_i = U conj(g'_i), g'_i the stored row i of
// the neighbour's gauge) or y_0, y_1 (y_i = g'_i U), and
//   D01 = H0 - H1,   D12 = H0 + 2 H1 - 8,   D02 = 2 H0 + H1 - 8.
// One site per lane: a lane builds H0 and H1 from its 8 links and then
// updates only G; only G is written.  (The first version ran a site on a
// lane pair, each lane summing four links, and both lanes repeated the
// three subgroup updates: about 4500 VALU instructions per site update in
// FP32 against about 1900 here.)  Copy c follows the link-form trajectory
// of the links G_c U G_c^+; all copies of a configuration share the link
// reads.
template <typename T>
struct Herm3 {
  T d0, d1, d2;       // diagonal
  C<T> o01, o02, o12;  // upper off-diagonal (H_ba = conj(H_ab))
};
template <typename T>
__device__ __forceinline__ Herm3<T> herm_zero() {
  return Herm3<T>{T(0), T(0), T(0), {T(0), T(0)}, {T(0), T(0)}, {T(0), T(0)}};
}
// sum_k a_k conj(b_k) over k = 0..2, one fused chain per component
template <typename T>
__device__ __forceinline__ C<T> cdot3c(const C<T>* a, const C<T>* b) {
  T re = a[0].re * b[0].re, im = a[0].im * b[0].re;
  re = fma(a[0].im, b[0].im, re);
  im = fma(-a[0].re, b[0].im, im);
#pragma unroll
  for (int k = 1; k < 3; k++) {
    re = fma(a[k].re, b[k].re, re);
    re = fma(a[k].im, b[k].im, re);
    im = fma(a[k].im, b[k].re, im);
    im = fma(-a[k].re, b[k].im, im);
  }
  return C<T>{re, im};
}
// a0 b0 + a1 b1 + a2 b2 (complex), one fused chain per component
template <typename T>
__device__ __forceinline__ C<T> cdot3(C<T> a0, C<T> b0, C<T> a1, C<T> b1, C<T> a2, C<T> b2) {
  T re = a0.re * b0.re, im = a0.re * b0.im;
  re = fma(-a0.im, b0.im, re);
  im = fma(a0.im, b0.re, im);
  re = fma(a1.re, b1.re, fma(-a1.im, b1.im, re));
  im = fma(a1.re, b1.im, fma(a1.im, b1.re, im));
  re = fma(a2.re, b2.re, fma(-a2.im, b2.im, re));
  im = fma(a2.re, b2.im, fma(a2.im, b2.re, im));
  return C<T>{re, im};
}
// H += v v^+  (H_ab += v_a conj(v_b)), as fused multiply-add chains
template <typename T>
__device__ __forceinline__ void herm_add_col(Herm3<T>& H, C<T> v0, C<T> v1, C<T> v2) {
  H.d0 = fma(v0.re, v0.re, fma(v0.im, v0.im, H.d0));
  H.d1 = fma(v1.re, v1.re, fma(v1.im, v1.im, H.d1));
  H.d2 = fma(v2.re, v2.re, fma(v2.im, v2.im, H.d2));
  H.o01 = C<T>{fma(v0.re, v1.re, fma(v0.im, v1.im, H.o01.re)),
               fma(v0.im, v1.re, fma(-v0.re, v1.im, H.o01.im))};
  H.o02 = C<T>{fma(v0.re, v2.re, fma(v0.im, v2.im, H.o02.re)),
               fma(v0.im, v2.re, fma(-v0.re, v2.im, H.o02.im))};
  H.o12 = C<T>{fma(v1.re, v2.re, fma(v1.im, v2.im, H.o12.re)),
               fma(v1.im, v2.re, fma(-v1.re, v2.im, H.o12.im))};
}
// H += conj(y) y^T  (H_ab += conj(y_a) y_b)
template <typename T>
__device__ __forceinline__ void herm_add_row(Herm3<T>& H, C<T> y0, C<T> y1, C<T> y2) {
  herm_add_col(H, cconj(y0), cconj(y1), cconj(y2));
}
// a A + b B + c 1
template <typename T>
__device__ __forceinline__ Herm3<T> herm_comb(const Herm3<T>& A, T a, const Herm3<T>& B, T b,
                                              T c) {
  Herm3<T> D;
  D.d0 = a * A.d0 + b * B.d0 + c;
  D.d1 = a * A.d1 + b * B.d1 + c;
  D.d2 = a * A.d2 + b * B.d2 + c;
  D.o01 = C<T>{a * A.o01.re + b * B.o01.re, a * A.o01.im + b * B.o01.im};
  D.o02 = C<T>{a * A.o02.re + b * B.o02.re, a * A.o02.im + b * B.o02.im};
  D.o12 = C<T>{a * A.o12.re + b * B.o12.re, a * A.o12.im + b * B.o12.im};
  return D;
}
// a + d conj(w): (x + iy)(u - iv) = (xu + yv) + i(yu - xv), fused
template <typename T>
__device__ __forceinline__ C<T> cacc_c(C<T> a, C<T> d, C<T> w) {
  return C<T>{fma(d.re, w.re, fma(d.im, w.im, a.re)), fma(d.im, w.re, fma(-d.re, w.im, a.im))};
}
// r = D w^+ (a column: r_a = sum_b D_ab conj(w_b)), fused chains
template <typename T>
__device__ __forceinline__ void herm_mul_dag(const Herm3<T>& D, const C<T>* w, C<T>* r) {
  r[0] = cacc_c(cacc_c(C<T>{D.d0 * w[0].re, -D.d0 * w[0].im}, D.o01, w[1]), D.o02, w[2]);
  r[1] = cacc_c(cacc_c(C<T>{D.d1 * w[1].re, -D.d1 * w[1].im}, cconj(D.o01), w[0]), D.o12, w[2]);
  r[2] = cacc_c(cacc_c(C<T>{D.d2 * w[2].re, -D.d2 * w[2].im}, cconj(D.o02), w[0]), cconj(D.o12),
                w[1]);
}
// sum_c u_c r_c, one fused chain per component
template <typename T>
__device__ __forceinline__ C<T> dot3(const C<T>* u, const C<T>* r) {
  return cdot3(u[0], r[0], u[1], r[1], u[2], r[2]);
}
// g H g^+ of a Hermitian H, in double
__device__ __forceinline__ double herm_form(const Herm3<double>& H, const C<double>* g) {
  const double d = H.d0 * cnorm2(g[0]) + H.d1 * cnorm2(g[1]) + H.d2 * cnorm2(g[2]);
  const double o = cmul(g[0], cmulc(H.o01, g[1])).re + cmul(g[0], cmulc(H.o02, g[2])).re +
                   cmul(g[1], cmulc(H.o12, g[2])).re;
  return d + 2.0 * o;
}
template <typename T>
__device__ __forceinline__ Herm3<double> herm_d(const Herm3<T>& H) {
  return Herm3<double>{(double)H.d0, (double)H.d1, (double)H.d2,
                       {(double)H.o01.re, (double)H.o01.im},
                       {(double)H.o02.re, (double)H.o02.im},
                       {(double)H.o12.re, (double)H.o12.im}};
}

// subgroup (I,J) update of the site gauge G (3 rows) in the field of
// D = H^(I) - H^(J); returns 1 - cos(theta)
template <typename T, int I, int J>
__device__ __forceinline__ T site_subgroup(const Herm3<T>& D, C<T> (&G)[3][3], T omega) {
  C<T> rJ[3], rI[3];
  herm_mul_dag(D, G[J], rJ);
  herm_mul_dag(D, G[I], rI);
  const C<T> z = dot3(G[I], rJ);
  const T h0 = z.re, h1 = -z.im;
  const T h2 = T(0.5) * (dot3(G[I], rI).re - dot3(G[J], rJ).re);
  T res;
  const Q<T> gq = omega_gauge(h0, h1, h2, omega, res);
  // r = [[g0 + i g3, g2 + i g1], [-g2 + i g1, g0 - i g3]], g3 = 0 (as in cm_subgroup)
  // row I <- g0 a + (g2 + i g1) b,  row J <- (-g2 + i g1) a + g0 b
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const C<T> a = G[I][c], b = G[J][c];
    G[I][c] = C<T>{fma(gq.a0, a.re, fma(gq.a2, b.re, -gq.a1 * b.im)),
                   fma(gq.a0, a.im, fma(gq.a2, b.im, gq.a1 * b.re))};
    G[J][c] = C<T>{fma(gq.a0, b.re, fma(-gq.a2, a.re, -gq.a1 * a.im)),
                   fma(gq.a0, b.im, fma(-gq.a2, a.im, gq.a1 * a.re))};
  }
  return res;
}

// row 2 = conj(row0 x row1) of an SU(3) matrix, fused chains
template <typename T>
__device__ __forceinline__ void row2_fma(const M2<T>& M, C<T>* z) {
  const C<T>* a = M.r[0];
  const C<T>* b = M.r[1];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int i = (k + 1) % 3, j = (k + 2) % 3;
    // conj(a_i b_j - a_j b_i)
    T re = a[i].re * b[j].re, nim = -a[i].re * b[j].im;
    re = fma(-a[i].im, b[j].im, re);
    nim = fma(-a[i].im, b[j].re, nim);
    re = fma(-a[j].re, b[i].re, re);
    nim = fma(a[j].re, b[i].im, nim);
    re = fma(a[j].im, b[i].im, re);
    nim = fma(a[j].im, b[i].re, nim);
    z[k] = C<T>{re, nim};
  }
}

// x_i = U conj(g_i), i = 0, 1 (g_i the stored rows of the forward
// neighbour's gauge; u2 the row 2 of U): H_i += x_i x_i^+
template <typename T>
__device__ __forceinline__ void site_fwd2(Herm3<T>& H0, Herm3<T>& H1, const M2<T>& U,
                                          const C<T>* u2, const M2<T>& Gf) {
#pragma unroll
  for (int i = 0; i < 2; i++) {
    const C<T>* gi = Gf.r[i];
    herm_add_col(i == 0 ? H0 : H1, cdot3c(U.r[0], gi), cdot3c(U.r[1], gi), cdot3c(u2, gi));
  }
}
// y_i = g_i U, i = 0, 1 (g_i the stored rows of the backward neighbour's
// gauge; u2 the row 2 of U): H_i += conj(y_i) y_i^T
template <typename T>
__device__ __forceinline__ void site_bwd2(Herm3<T>& H0, Herm3<T>& H1, const M2<T>& U,
                                          const C<T>* u2, const M2<T>& Gb) {
#pragma unroll
  for (int i = 0; i < 2; i++) {
    const C<T>* gi = Gb.r[i];
    C<T> y[3];
#pragma unroll
    for (int b = 0; b < 3; b++) y[b] = cdot3(gi[0], U.r[0][b], gi[1], U.r[1][b], gi[2], u2[b]);
    herm_add_row(i == 0 ? H0 : H1, y[0], y[1], y[2]);
  }
}

// rows 0, 1 back to orthonormal: the FP32 site gauges are never rebuilt in
// FP64 during their FP32 phase, and every update leaves about 1e-7 in row
// norm and in the overlap of the rows, which would otherwise accumulate
// over the sweeps.  Row 0 is scaled by (3 - |row|^2) / 2 (the Newton step
// of 1/|row| at 1); row 1 loses its component along row 0 and is scaled
// the same way.
template <typename T>
__device__ __forceinline__ void scale_row(C<T>* r) {
  T n = T(0);
#pragma unroll
  for (int c = 0; c < 3; c++) n = fma(r[c].re, r[c].re, fma(r[c].im, r[c].im, n));
  const T f = T(1.5) - T(0.5) * n;
#pragma unroll
  for (int c = 0; c < 3; c++) r[c] = C<T>{r[c].re * f, r[c].im * f};
}
template <typename T>
__device__ __forceinline__ void orthonorm_rows(M2<T>& m) {
  scale_row(m.r[0]);
  const C<T> o = cdot3c(m.r[1], m.r[0]);  // <row 0, row 1> = sum r1_c conj(r0_c)
#pragma unroll
  for (int c = 0; c < 3; c++) m.r[1][c] = csub(m.r[1][c], cmul(o, m.r[0][c]));
  scale_row(m.r[1]);
}

// The end of a site's update for one copy: its gauge G (at gp) rotated in
// the three subgroups by the fields of H0, H1 and stored; residual and (on
// the black half) the site's functional sum_i g_i H^(i) g_i^+ (= sum over
// its 8 links of sum_i |W_ii|^2, H^(2) = 8 - H0 - H1), in double
template <typename T>
__device__ __forceinline__ void site_finish(const Herm3<T>& H0, const Herm3<T>& H1,
                                            LRef<T> gp, int p, T omega, double& fs, double& rs,
                                            double& rm) {
  M2<T> G2 = ldm(gp);
  C<T> Gx[3][3];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) Gx[a][b] = G2.r[a][b];
  row2_of(G2, Gx[2]);
  T res = site_subgroup<T, 0, 1>(herm_comb(H0, T(1), H1, T(-1), T(0)), Gx, omega);
  res += site_subgroup<T, 1, 2>(herm_comb(H0, T(1), H1, T(2), T(-8)), Gx, omega);
  res += site_subgroup<T, 0, 2>(herm_comb(H0, T(2), H1, T(1), T(-8)), Gx, omega);
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) G2.r[a][b] = Gx[a][b];
  if constexpr (sizeof(T) == 4) orthonorm_rows(G2);
  stm(gp, G2);
  rs = (double)res;
  rm = (double)res;
  fs = 0.0;
  if (p == 1) {
    const Herm3<double> A = herm_d(H0), B = herm_d(H1);
    C<double> gd[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++) gd[a][b] = C<double>{(double)Gx[a][b].re, (double)Gx[a][b].im};
    fs = herm_form(A, gd[0]) + herm_form(B, gd[1]) +
         (8.0 * (cnorm2(gd[2][0]) + cnorm2(gd[2][1]) + cnorm2(gd[2][2])) -
          herm_form(A, gd[2]) - herm_form(B, gd[2]));
  }
}

// FP64: the copies one after the other, each reading the links again (the
// FP64 field matrices of every copy would not fit in registers); the FP64
// phases are the short warm-up and polish.
template <typename T, bool NT>
__device__ __forceinline__ void su3_site_copies_block_nt(const T* __restrict__ U,
                                                         T* __restrict__ G, size_t gstride,
                                                         int ncopies, const Geom& g, int p,
                                                         T omega, double* __restrict__ part,
                                                         size_t pstride,
                                                         const JobCtl* __restrict__ ctl,
                                                         uint32_t lb) {
  bool live = false;
  for (int c = 0; c < ncopies; c++) live |= !ctl[c].done;
  if (!live) return;  // uniform over the workgroup
  const uint32_t h = lb * kBlock + threadIdx.x;
  const bool in = h < g.Vh;
  const SiteIdx s = in ? decode(g, h, p) : SiteIdx{};
  for (int c = 0; c < ncopies; c++) {
    if (ctl[c].done) continue;  // uniform over the workgroup
    T* Gc = G + (size_t)c * gstride;
    double fs = 0.0, rs = 0.0, rm = 0.0;
    if (in) {
      Herm3<T> H0 = herm_zero<T>(), H1 = herm_zero<T>();
#pragma unroll 1
      for (int mu = 0; mu < 4; mu++) {
        const uint32_t hbm = back_h(g, h, s, mu);
        const M2<T> Lf = NT ? ldm_nt(link_ptr<12>(U, g, p, mu, h)) : ldm(link_ptr<12>(U, g, p, mu, h));
        const M2<T> Gf = ldm(site_ptr<12>(Gc, g, 1 - p, fwd_h(g, h, s, mu)));
        const M2<T> Lb =
            NT ? ldm_nt(link_ptr<12>(U, g, 1 - p, mu, hbm)) : ldm(link_ptr<12>(U, g, 1 - p, mu, hbm));
        const M2<T> Gb = ldm(site_ptr<12>(Gc, g, 1 - p, hbm));
        C<T> u2[3];
        row2_fma(Lf, u2);
        site_fwd2(H0, H1, Lf, u2, Gf);
        row2_fma(Lb, u2);
        site_bwd2(H0, H1, Lb, u2, Gb);
      }
      site_finish(H0, H1, site_ptr<12>(Gc, g, p, h), p, omega, fs, rs, rm);
    }
    wave_partial(part + (size_t)c * pstride, lb * kBlock + (threadIdx.x & ~63u), g.Vh, fs, rs,
                 rm);
  }
}

// links with the nontemporal hint where they exceed the MALL (uniform)
template <typename T>
__device__ __forceinline__ void su3_site_copies_block(const T* __restrict__ U, T* __restrict__ G,
                                                      size_t gstride, int ncopies, const Geom& g,
                                                      int p, T omega, double* __restrict__ part,
                                                      size_t pstride,
                                                      const JobCtl* __restrict__ ctl,
                                                      uint32_t lb) {
  if (stream_links(g, 12 * sizeof(T)))
    su3_site_copies_block_nt<T, true>(U, G, gstride, ncopies, g, p, omega, part, pstride, ctl, lb);
  else
    su3_site_copies_block_nt<T, false>(U, G, gstride, ncopies, g, p, omega, part, pstride, ctl, lb);
}

// FP32 with the K copies of the set inside the link loop: each link is
// loaded, and its row 2 formed, once for all copies, which then add their
// terms with their own neighbour gauges (with the copies outside, four
// copies at 32^4 read 4.1 times the links of one: the 4 MB L2 of an XCD
// does not keep them from one copy to the next).  K = 1 loads two
// directions per pass, more copies one.  A
// copy's operations and their order are the same for every K.  Sets of
// two or more copies load the links with the nontemporal hint (32^4, four
// copies: 140 -> 126 us per half-sweep; the copies' neighbour gauges, not
// the links, are then what the caches hold); one copy loads them plainly
// where the links fit the MALL (the hint cost 10% at 32^4, where the next
// half-sweep reads the links again from the MALL), with the hint beyond it.
template <int K, bool NT>
__device__ __forceinline__ void su3_site_set_block_nt(const float* __restrict__ U,
                                                   float* __restrict__ G, size_t gstride,
                                                   const Geom& g, int p, float omega,
                                                   double* __restrict__ part, size_t pstride,
                                                   const JobCtl* __restrict__ ctl, int slot,
                                                   uint32_t lb) {
  bool act[K];
  bool live = false;
#pragma unroll
  for (int c = 0; c < K; c++) {
    // uniform over the workgroup; a slot's sweep beyond the cap is skipped
    act[c] = !ctl[c].done && ctl[c].sweep + slot < ctl[c].max_sweeps;
    live |= act[c];
  }
  if (!live) return;
  const uint32_t h = lb * kBlock + threadIdx.x;
  const bool in = h < g.Vh;
  double fs[K], rs[K], rm[K];
#pragma unroll
  for (int c = 0; c < K; c++) fs[c] = rs[c] = rm[c] = 0.0;
  if (in) {
    const SiteIdx s = decode(g, h, p);
    Herm3<float> H0[K], H1[K];
#pragma unroll
    for (int c = 0; c < K; c++) H0[c] = H1[c] = herm_zero<float>();
    constexpr int ND = K == 1 ? 2 : 1;
#pragma unroll 1
    for (int m0 = 0; m0 < 4; m0 += ND) {
      M2<float> Lf[ND], Lb[ND];
      uint32_t hf[ND], hb[ND];
#pragma unroll
      for (int k = 0; k < ND; k++) {
        const int mu = m0 + k;
        hf[k] = fwd_h(g, h, s, mu);
        hb[k] = back_h(g, h, s, mu);
        if constexpr (!NT) {
          Lf[k] = ldm(link_ptr<12>(U, g, p, mu, h));
          Lb[k] = ldm(link_ptr<12>(U, g, 1 - p, mu, hb[k]));
        } else {
          Lf[k] = ldm_nt(link_ptr<12>(U, g, p, mu, h));
          Lb[k] = ldm_nt(link_ptr<12>(U, g, 1 - p, mu, hb[k]));
        }
      }
#pragma unroll
      for (int c = 0; c < K; c++) {
        if (!act[c]) continue;
        const float* Gc = G + (size_t)c * gstride;
        M2<float> Gf[ND], Gb[ND];
#pragma unroll
        for (int k = 0; k < ND; k++) {
          Gf[k] = ldm(site_ptr<12>(Gc, g, 1 - p, hf[k]));
          Gb[k] = ldm(site_ptr<12>(Gc, g, 1 - p, hb[k]));
        }
#pragma unroll
        for (int k = 0; k < ND; k++) {
          C<float> u2[3];
          row2_fma(Lf[k], u2);
          site_fwd2(H0[c], H1[c], Lf[k], u2, Gf[k]);
          row2_fma(Lb[k], u2);
          site_bwd2(H0[c], H1[c], Lb[k], u2, Gb[k]);
        }
      }
    }
#pragma unroll
    for (int c = 0; c < K; c++)
      if (act[c])
        site_finish(H0[c], H1[c], site_ptr<12>(G + (size_t)c * gstride, g, p, h), p, omega, fs[c],
                    rs[c], rm[c]);
  }
#pragma unroll
  for (int c = 0; c < K; c++)
    if (act[c])  // uniform
      wave_partial(part + (size_t)c * pstride, lb * kBlock + (threadIdx.x & ~63u), g.Vh, fs[c],
                   rs[c], rm[c]);
}

template <int K>
__device__ __forceinline__ void su3_site_set_block(const float* __restrict__ U,
                                                   float* __restrict__ G, size_t gstride,
                                                   const Geom& g, int p, float omega,
                                                   double* __restrict__ part, size_t pstride,
                                                   const JobCtl* __restrict__ ctl, int slot,
                                                   uint32_t lb) {
  // both variants in one kernel (K = 1: 165 VGPRs, 3 waves per SIMD) ran
  // 4% faster at 32^4 than the plain-load variant alone (100 VGPRs)
  if (K > 1 || stream_links(g, 48))  // uniform
    su3_site_set_block_nt<K, true>(U, G, gstride, g, p, omega, part, pstride, ctl, slot, lb);
  else
    su3_site_set_block_nt<K, false>(U, G, gstride, g, p, omega, part, pstride, ctl, slot, lb);
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
template <int K>
__global__ __launch_bounds__(kBlock) void k_su3_site_set(const float* __restrict__ U,
                                                         float* __restrict__ G, size_t gstride,
                                                         Geom g, int p, float omega,
                                                         double* __restrict__ part,
                                                         size_t pstride,
                                                         const JobCtl* __restrict__ ctl,
                                                         int slot) {
  su3_site_set_block<K>(U, G, gstride, g, p, omega, part, pstride, ctl, slot,
                        xcd_block(blockIdx.x, gridDim.x));
}

// FP64 table: any copy counts; FP32 table: sets of K copies
__global__ __launch_bounds__(kBlock) void k_su3_site_copies_batch(
    const BatchJob* __restrict__ jobs, int njobs, int p) {
  const uint32_t lb = xcd_block(blockIdx.x, gridDim.x);
  const BatchJob& j = jobs[batch_job(jobs, njobs, (int)lb)];
  su3_site_copies_block<double>((const double*)j.U, (double*)j.S, j.sstride, j.ncopies, j.g, p,
                                j.omega, j.part + (size_t)3 * j.nw * p, j.pstride, j.ctl,
                                lb - j.block0);
}
template <int K>
__global__ __launch_bounds__(kBlock) void k_su3_site_set_batch(const BatchJob* __restrict__ jobs,
                                                               int njobs, int p, int slot) {
  const uint32_t lb = xcd_block(blockIdx.x, gridDim.x);
  const BatchJob& j = jobs[batch_job(jobs, njobs, (int)lb)];
  su3_site_set_block<K>((const float*)j.U, (float*)j.S, j.sstride, j.g, p, (float)j.omega,
                        j.part + (size_t)slot * K * j.pstride + (size_t)3 * j.nw * p, j.pstride,
                        j.ctl, slot, lb - j.block0);
}

// The stopping test of the SU(3) site form's FP32 phase.  The functional
// that the FP32 half-sweeps sum comes from FP32 field matrices, whose
// rounding moves it by ~1e-7 / sqrt(V) (relative) from one sweep to the
// next -- more than the tolerance -- so it cannot decide convergence.
// Every reunit_every sweeps this kernel evaluates, for every copy not done,
// F of the links G_c U G_c^+ in FP64 from the FP32 state, with the FP32
// links taken exactly and every gauge put on SU(3) in FP64 (gs3): a
// function of the state free of FP32 rounding.  Lane h covers the sites
// (p, h), p = 0, 1, and their 4 forward links (every link once); one
// partial triple (F, 0, 0) per wave at part + c * pstride.
__device__ __forceinline__ void su3_site_F_block(const float* __restrict__ U,
                                                 const float* __restrict__ G, size_t gstride,
                                                 int ncopies, const Geom& g,
                                                 double* __restrict__ part, size_t pstride,
                                                 const JobCtl* __restrict__ ctl, uint32_t lb) {
  bool live = false;
  for (int c = 0; c < ncopies; c++) live |= !ctl[c].done;
  if (!live) return;  // uniform over the workgroup
  const uint32_t h = lb * kBlock + threadIdx.x;
  const bool in = h < g.Vh;
  for (int c = 0; c < ncopies; c++) {
    if (ctl[c].done) continue;  // uniform over the workgroup
    const float* Gc = G + (size_t)c * gstride;
    double fs = 0.0;
    if (in) {
#pragma unroll 1
      for (int p = 0; p < 2; p++) {
        const SiteIdx s = decode(g, h, p);
        const M3<double> gx = gs3(to_d(ldm(site_ptr<12>(Gc, g, p, h))));
#pragma unroll 1
        for (int mu = 0; mu < 4; mu++) {
          const M3<double> gy = gs3(to_d(ldm(site_ptr<12>(Gc, g, 1 - p, fwd_h(g, h, s, mu)))));
          const M3<double> u = full3(to_d(ldm(link_ptr<12>(U, g, p, mu, h))));
          // W_ii = g_i U g'_i^+
#pragma unroll
          for (int i = 0; i < 3; i++) {
            C<double> v[3];
#pragma unroll
            for (int a = 0; a < 3; a++) v[a] = cdot3c(u.m[a], gy.m[i]);
            fs += cnorm2(cdot3(gx.m[i][0], v[0], gx.m[i][1], v[1], gx.m[i][2], v[2]));
          }
        }
      }
    }
    wave_partial(part + (size_t)c * pstride, lb * kBlock + (threadIdx.x & ~63u), g.Vh, fs, 0.0,
                 0.0);
  }
}
__global__ __launch_bounds__(kBlock) void k_su3_site_F(const float* __restrict__ U,
                                                       const float* __restrict__ G,
                                                       size_t gstride, int ncopies, Geom g,
                                                       double* __restrict__ part, size_t pstride,
                                                       const JobCtl* __restrict__ ctl) {
  su3_site_F_block(U, G, gstride, ncopies, g, part, pstride, ctl, blockIdx.x);
}
__global__ __launch_bounds__(kBlock) void k_su3_site_F_batch(const BatchJob* __restrict__ jobs,
                                                             int njobs) {
  const BatchJob& j = jobs[batch_job(jobs, njobs, (int)blockIdx.x)];
  su3_site_F_block((const float*)j.U, (const float*)j.S, j.sstride, j.ncopies, j.g,
                   j.part + (size_t)kSlots * j.ncopies * j.pstride, j.pstride, j.ctl,
                   blockIdx.x - j.block0);
}
// the F of that evaluation (nparts triples at part) tested against the
// previous one: converged when their difference per sweep is below tol
// (apply_rebuild); the sweep count does not change
// F partials of every copy (check region of the partial buffer); the
// period reduction (launch_reduce_period_copies / _batch) applies the test
void launch_su3_site_F(const float* U, const float* G, size_t gstride, int ncopies,
                       const Geom& g, double* part, size_t pstride, const JobCtl* ctl,
                       hipStream_t st) {
  const int nb = (int)((site3_lanes(g) + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_su3_site_F, dim3(nb), dim3(kBlock), 0, st, U, G, gstride, ncopies, g, part,
                     pstride, ctl);
}
void launch_su3_site_F_batch(const BatchJob* jobs, int njobs, int nblocks_total, hipStream_t st) {
  if (njobs == 0 || nblocks_total == 0) return;
  hipLaunchKernelGGL(k_su3_site_F_batch, dim3(nblocks_total), dim3(kBlock), 0, st, jobs, njobs);
}

template <typename T>
void launch_su3_site_copies(const T* U, T* G, size_t gstride, int ncopies, const Geom& g,
                            int parity, double omega, double* part, size_t pstride,
                            const JobCtl* ctl, hipStream_t st, int slot) {
  const int nb = (int)((site3_lanes(g) + kBlock - 1) / kBlock);
  const size_t nw = (site3_lanes(g) + 63) / 64;
  double* pp = part + (size_t)slot * ncopies * pstride + 3 * nw * parity;
  if constexpr (sizeof(T) == 4) {
#define L_(K)                                                                                  \
  hipLaunchKernelGGL((k_su3_site_set<K>), dim3(nb), dim3(kBlock), 0, st, U, G, gstride, g,     \
                     parity, (float)omega, pp, pstride, ctl, slot)
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
  } else {
    hipLaunchKernelGGL((k_su3_site_copies<T>), dim3(nb), dim3(kBlock), 0, st, U, G, gstride,
                       ncopies, g, parity, (T)omega, pp, pstride, ctl);
  }
}
template void launch_su3_site_copies<float>(const float*, float*, size_t, int, const Geom&, int,
                                            double, double*, size_t, const JobCtl*,
                                            hipStream_t, int);
template void launch_su3_site_copies<double>(const double*, double*, size_t, int, const Geom&,
                                             int, double, double*, size_t, const JobCtl*,
                                             hipStream_t, int);

void launch_su3_site_copies_batch(int ncopies, bool fp32, const BatchJob* jobs, int njobs,
                                  int nblocks_total, int parity, hipStream_t st, int slot) {
  if (njobs == 0 || nblocks_total == 0) return;
  if (!fp32) {
    hipLaunchKernelGGL(k_su3_site_copies_batch, dim3(nblocks_total), dim3(kBlock), 0, st, jobs,
                       njobs, parity);
    return;
  }
#define L_(K)                                                                                  \
  hipLaunchKernelGGL((k_su3_site_set_batch<K>), dim3(nblocks_total), dim3(kBlock), 0, st, jobs, \
                     njobs, parity, slot)
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

}  // namespace mag
