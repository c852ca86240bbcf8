// mag_sweep_link.h -- the link-form half-sweep kernels (SU(2) one site
// per lane, SU(3) Cabibbo-Marinari on lane pairs) and their launchers, as
// templates over the storage type.  Instantiated for float in
// mag_sweep.hip and for double in mag_sweep64.hip; the latter is compiled
// without floating-point contraction, so that every FP64 site update is
// the same sequence of IEEE operations as the host restatement's and the
// two agree bit for bit.
//
// The site update (build-defined; the reference has only the spin-form
// omega=1 and omega=2 updates, lib/cpu/src/mag.cpp:1375-1512):
//   h      = sum_mu [ M(U_x,mu) e3 + M(U_x-mu,mu)^T e3 ]     (mag.cpp:569-626
//            with every spin e3: the links carry the gauge transformation)
//   g      = geodesic omega step of e3 towards h_hat          (mag_math.h)
//   U_x,mu <- g U_x,mu ;  U_x-mu,mu <- U_x-mu,mu g^+
// SU(3): the same update for the Cabibbo-Marinari subgroups (0,1), (1,2),
// (0,2) in turn, with the block field terms of mag_math.h.
//
// Work split: SU(2) runs one site per lane.  SU(3) runs a site on a pair of
// adjacent lanes: lane q of the pair holds the 4 links of directions
// mu = 2q, 2q+1 (forward and backward), sums their field terms, and the
// pair exchanges the two partial fields with one DPP quad_perm; both lanes
// then derive the same g and update their own links.  Control flow is
// identical in both lanes and the register footprint per lane is halved
// (4 links instead of 8): FP64 SU(3) goes from 1 to 3 waves per SIMD.
//
// Field summation order (matched by the oracle):
//   h = [f(0) + b(0) + f(1) + b(1)] + [f(2) + b(2) + f(3) + b(3)]
#pragma once
#include "mag_gauge.h"
#include "mag_internal.h"
#include "mag_layout.h"
#include "mag_math.h"
#include "mag_reduce.h"

namespace mag {

// ================================================================== helpers
// value of the partner lane of the pair (lane ^ 1), and of lane ^ 2
__device__ __forceinline__ double pair_xchg(double v) { return dpp_f64<DPP_QUAD_XOR1>(v); }
__device__ __forceinline__ double quad_xchg2(double v) { return dpp_f64<DPP_QUAD_XOR2>(v); }
__device__ __forceinline__ float quad_xchg2(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), DPP_QUAD_XOR2, 0xF, 0xF, true));
}
__device__ __forceinline__ float pair_xchg(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), DPP_QUAD_XOR1, 0xF, 0xF, true));
}

// half indices of x - mu for all mu
__device__ __forceinline__ void back_all(const Geom& g, uint32_t h, const SiteIdx& s,
                                         uint32_t* hb) {
#pragma unroll
  for (int mu = 0; mu < 4; mu++) hb[mu] = back_h(g, h, s, mu);
}

// job of workgroup b in a batch table (block0 ascending)
__device__ __forceinline__ int batch_job(const BatchJob* __restrict__ jobs, int n, int b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].block0 <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// XCD-aware workgroup order (xcd_block, mag_layout.h): the site-form
// kernels reuse neighbour spins / gauges (read by up to 8 sites).

// ================================================================== launch shapes
// The partial triples of a half-sweep are per wave: wave w of a job covers
// its lanes 64w .. 64w+63 (64 SU(2) sites, 32 SU(3) sites) and its triple
// goes to part[3w].  A launch shape -- NT threads per workgroup, SPT
// 64-lane-aligned tiles of NT lanes per thread (workgroup lb, tile k covers
// lanes (lb*SPT + k)*NT ...) -- only decides which thread computes which
// wave, never the arithmetic or the order of a partial, so every shape
// gives bit-identical results and the shape is free for the autotuner.
__device__ __forceinline__ void wave_partial(double* __restrict__ part, uint32_t lane0,
                                             uint32_t nlanes, double f, double r, double m) {
  f = wave_sum(f);
  r = wave_sum(r);
  m = wave_max(m);
  if ((threadIdx.x & 63u) == 0 && lane0 < nlanes) {
    const uint32_t w = lane0 >> 6;
    part[3 * w] = f;
    part[3 * w + 1] = r;
    part[3 * w + 2] = m;
  }
}
// Link-form half-sweeps: the black half (p = 1) folds the residuals of the
// red half's triple of the same wave (written by the previous launch, at
// part - 3 nw, nw = the half-sweep's waves) into its own; the red half
// sums no functional.  A sweep's reduction then reads the nw black triples
// alone, half the partials (the functional's sum is unchanged: the red
// triples' F is exactly zero).
__device__ __forceinline__ void wave_partial_sweep(double* __restrict__ part, int p,
                                                   uint32_t lane0, uint32_t nlanes, double f,
                                                   double r, double m) {
  f = wave_sum(f);
  r = wave_sum(r);
  m = wave_max(m);
  if ((threadIdx.x & 63u) == 0 && lane0 < nlanes) {
    const uint32_t w = lane0 >> 6;
    if (p == 1) {
      const double* red = part - 3 * (size_t)((nlanes + 63) / 64);
      r += red[3 * w + 1];
      m = fmax(m, red[3 * w + 2]);
    }
    part[3 * w] = f;
    part[3 * w + 1] = r;
    part[3 * w + 2] = m;
  }
}

// ================================================================== SU(2)
// Plain half-sweeps (every shape): SPT tiles per thread with every tile's
// links loaded before the first update (no two sites of one parity share a
// link, so a later tile's loads never alias an earlier tile's stores), so
// the loads of the second site are in flight while the first is updated.
// Measured at 32^4 against the same update tile by tile: FP64 256x2 40.0 us
// per half-sweep (best one-tile shape 41.0; 181 VGPRs, 2 waves per SIMD),
// FP32 128x2 21.8 us (23.2; 96 VGPRs).  All shapes of a plain half-sweep run
// this code: the FP32 kernels are compiled with contraction, and the
// one-tile code's fused multiply-adds came out differently from these.
template <typename T, bool GA, int NT, int SPT, bool ST>
__device__ __forceinline__ void su2_half_tiles(T* __restrict__ U, T* __restrict__ G,
                                               const Geom& g, int p, T omega,
                                               double* __restrict__ part, uint32_t lb) {
  Q<T> F[SPT][4], B[SPT][4];
  uint32_t hb[SPT][4];
#pragma unroll
  for (int k = 0; k < SPT; k++) {
    const uint32_t h = (lb * SPT + k) * NT + threadIdx.x;
    if (h < g.Vh) {
      const SiteIdx s = decode(g, h, p);
      back_all(g, h, s, hb[k]);
#pragma unroll
      for (int mu = 0; mu < 4; mu++) {
        if constexpr (ST) {
          F[k][mu] = ldq_nt(link_ptr<4>((const T*)U, g, p, mu, h));
          B[k][mu] = ldq_nt(link_ptr<4>((const T*)U, g, 1 - p, mu, hb[k][mu]));
        } else {
          F[k][mu] = ldq(link_ptr<4>(U, g, p, mu, h));
          B[k][mu] = ldq(link_ptr<4>(U, g, 1 - p, mu, hb[k][mu]));
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < SPT; k++) {
    const uint32_t base = (lb * SPT + k) * NT;
    const uint32_t h = base + threadIdx.x;
    double fs = 0.0, rs = 0.0, rm = 0.0;
    if (h < g.Vh) {
      T a0 = T(0), a1 = T(0), a2 = T(0), b0 = T(0), b1 = T(0), b2 = T(0);
#pragma unroll
      for (int mu = 0; mu < 2; mu++) {
        field_fwd(F[k][mu], a0, a1, a2);
        field_bwd(B[k][mu], a0, a1, a2);
        field_fwd(F[k][mu + 2], b0, b1, b2);
        field_bwd(B[k][mu + 2], b0, b1, b2);
      }
      T res;
      const Q<T> gq = omega_gauge(a0 + b0, a1 + b1, a2 + b2, omega, res);
#pragma unroll
      for (int mu = 0; mu < 4; mu++) {
        F[k][mu] = gmul(gq, F[k][mu]);
        B[k][mu] = mul_gdag(B[k][mu], gq);
      }
      if (p == 1) {
#pragma unroll
        for (int mu = 0; mu < 4; mu++) fs += link_F(F[k][mu]) + link_F(B[k][mu]);
      }
#pragma unroll
      for (int mu = 0; mu < 4; mu++) {
        if constexpr (ST) {
          stq_nt(link_ptr<4>(U, g, p, mu, h), F[k][mu]);
          stq_nt(link_ptr<4>(U, g, 1 - p, mu, hb[k][mu]), B[k][mu]);
        } else {
          stq(link_ptr<4>(U, g, p, mu, h), F[k][mu]);
          stq(link_ptr<4>(U, g, 1 - p, mu, hb[k][mu]), B[k][mu]);
        }
      }
      if constexpr (GA) {
        const LRef<T> gp = site_ptr<4>(G, g, p, h);
        stq(gp, gmul(gq, ldq(gp)));
      }
      rs = (double)res;
      rm = (double)res;
    }
    wave_partial_sweep(part, p, base + (threadIdx.x & ~63u), g.Vh, fs, rs, rm);
  }
}

// Tile k of workgroup lb of a half-sweep over the sites of parity p, one
// site per lane (one site per lane measured faster than a lane pair).  The
// black (p = 1) half also sums the functional of the links just written.
// GA: the red half-sweep of the mixed FP32 phase accumulates the site gauge
// (MODE HALF_RENORM: after normalising it in FP64); MODE HALF_REBUILD
// (black half closing an FP32 period): the swept links are replaced by
// links rebuilt from the master, and the rebuilt links' F goes to part_rb.
// Plain half-sweeps go to su2_half_tiles; this loop runs the other modes.
template <typename T, bool GA, int NT, int SPT, int MODE, bool ST>
__device__ __forceinline__ void su2_half_block_st(T* __restrict__ U, T* __restrict__ G,
                                                  const Geom& g, int p, T omega,
                                                  double* __restrict__ part, uint32_t lb,
                                                  const double* __restrict__ U0,
                                                  double* __restrict__ part_rb) {
  if constexpr (MODE == HALF_PLAIN) {
    su2_half_tiles<T, GA, NT, SPT, ST>(U, G, g, p, omega, part, lb);
    return;
  }
#pragma unroll
  for (int k = 0; k < SPT; k++) {
  const uint32_t base = (lb * SPT + k) * NT;
  const uint32_t h = base + threadIdx.x;
  double fs = 0.0, rs = 0.0, rm = 0.0, fr = 0.0;
  if (h < g.Vh) {
    const SiteIdx s = decode(g, h, p);
    uint32_t hb[4];
    back_all(g, h, s, hb);
    Q<T> F[4], B[4];
#pragma unroll
    for (int mu = 0; mu < 4; mu++) {
      if constexpr (ST) {
        F[mu] = ldq_nt(link_ptr<4>((const T*)U, g, p, mu, h));
        B[mu] = ldq_nt(link_ptr<4>((const T*)U, g, 1 - p, mu, hb[mu]));
      } else {
        F[mu] = ldq(link_ptr<4>(U, g, p, mu, h));
        B[mu] = ldq(link_ptr<4>(U, g, 1 - p, mu, hb[mu]));
      }
    }
    T a0 = T(0), a1 = T(0), a2 = T(0), b0 = T(0), b1 = T(0), b2 = T(0);
#pragma unroll
    for (int mu = 0; mu < 2; mu++) {
      field_fwd(F[mu], a0, a1, a2);
      field_bwd(B[mu], a0, a1, a2);
      field_fwd(F[mu + 2], b0, b1, b2);
      field_bwd(B[mu + 2], b0, b1, b2);
    }
    T res;
    const Q<T> gq = omega_gauge(a0 + b0, a1 + b1, a2 + b2, omega, res);
#pragma unroll
    for (int mu = 0; mu < 4; mu++) {
      F[mu] = gmul(gq, F[mu]);
      B[mu] = mul_gdag(B[mu], gq);
    }
    if (p == 1) {
#pragma unroll
      for (int mu = 0; mu < 4; mu++) fs += link_F(F[mu]) + link_F(B[mu]);
    }
    if constexpr (MODE == HALF_REBUILD) {
      // the black gauge from the swept W(x,0), then all 8 links in FP64
      const Q<double> gx = black_gauge(to_d(F[0]), norm_q(site_ptr<4>(G, g, 0, fwd_h(g, h, s, 0))),
                                       ldq(link_ptr<4>(U0, g, 1, 0, h)));
#pragma unroll
      for (int mu = 0; mu < 4; mu++) {
        const Q<double> wf = qmul_dag(qmul(gx, ldq(link_ptr<4>(U0, g, 1, mu, h))),
                                      norm_q(site_ptr<4>(G, g, 0, fwd_h(g, h, s, mu))));
        const Q<double> wb = qmul_dag(qmul(norm_q(site_ptr<4>(G, g, 0, hb[mu])),
                                           ldq(link_ptr<4>(U0, g, 0, mu, hb[mu]))), gx);
        fr += link_F(wf) + link_F(wb);
        F[mu] = to_f(wf);
        B[mu] = to_f(wb);
      }
    }
#pragma unroll
    for (int mu = 0; mu < 4; mu++) {
      if constexpr (ST) {
        stq_nt(link_ptr<4>(U, g, p, mu, h), F[mu]);
        stq_nt(link_ptr<4>(U, g, 1 - p, mu, hb[mu]), B[mu]);
      } else {
        stq(link_ptr<4>(U, g, p, mu, h), F[mu]);
        stq(link_ptr<4>(U, g, 1 - p, mu, hb[mu]), B[mu]);
      }
    }
    if constexpr (GA) {
      const LRef<T> gp = site_ptr<4>(G, g, p, h);
      Q<T> gc = ldq(gp);
      if constexpr (MODE == HALF_RENORM) gc = to_f(norm_q(to_d(gc)));
      stq(gp, gmul(gq, gc));
    }
    rs = (double)res;
    rm = (double)res;
  }
  wave_partial_sweep(part, p, base + (threadIdx.x & ~63u), g.Vh, fs, rs, rm);
  if constexpr (MODE == HALF_REBUILD) wave_partial(part_rb, base + (threadIdx.x & ~63u), g.Vh, fr, 0.0, 0.0);
  }
}

// Links past the MALL (8 Vh links of 4 T, 48^4 SU(2) in either precision)
// are loaded and stored with the nontemporal hint: each is read and written
// once per half-sweep and the next half-sweep reads it from HBM whatever the
// hint (48^4 per half-sweep FP64 267 -> 251 us, mixed 133 -> 125 us); below,
// the plain accesses let the next half-sweep hit the MALL.
template <typename T, bool GA, int NT, int SPT, int MODE>
__device__ __forceinline__ void su2_half_block(T* __restrict__ U, T* __restrict__ G,
                                               const Geom& g, int p, T omega,
                                               double* __restrict__ part, uint32_t lb,
                                               const double* __restrict__ U0,
                                               double* __restrict__ part_rb) {
  if (stream_links(g, 4 * sizeof(T)))  // uniform
    su2_half_block_st<T, GA, NT, SPT, MODE, true>(U, G, g, p, omega, part, lb, U0, part_rb);
  else
    su2_half_block_st<T, GA, NT, SPT, MODE, false>(U, G, g, p, omega, part, lb, U0, part_rb);
}

// the sweep of a slot is beyond the cap (or the job is done): return
__device__ __forceinline__ bool half_skip(const JobCtl* __restrict__ ctl, int slot) {
  return ctl->done || ctl->sweep + slot >= ctl->max_sweeps;
}

// one job: the launcher picks the plain or the streaming kernel by lattice
template <typename T, bool GA, int NT, int SPT, int MODE>
__global__ __launch_bounds__(NT) void k_su2_half(T* __restrict__ U, T* __restrict__ G, Geom g,
                                                 int p, T omega, double* __restrict__ part,
                                                 const JobCtl* __restrict__ ctl, int slot,
                                                 const double* __restrict__ U0,
                                                 double* __restrict__ part_rb) {
  if (half_skip(ctl, slot)) return;
  su2_half_block_st<T, GA, NT, SPT, MODE, false>(U, G, g, p, omega, part, blockIdx.x, U0,
                                                 part_rb);
}
template <typename T, bool GA, int NT, int SPT, int MODE>
__global__ __launch_bounds__(NT) void k_su2_half_st(T* __restrict__ U, T* __restrict__ G, Geom g,
                                                    int p, T omega, double* __restrict__ part,
                                                    const JobCtl* __restrict__ ctl, int slot,
                                                    const double* __restrict__ U0,
                                                    double* __restrict__ part_rb) {
  if (half_skip(ctl, slot)) return;
  su2_half_block_st<T, GA, NT, SPT, MODE, true>(U, G, g, p, omega, part, blockIdx.x, U0,
                                                part_rb);
}

// batched: workgroup b runs block b - block0 of its job with the class's
// launch shape (NT, SPT); the job's partials are per wave, so the shape
// leaves its results unchanged; plain or streaming accesses per job
template <typename T, bool GA, int NT, int SPT, int MODE>
__global__ __launch_bounds__(NT) void k_su2_half_batch(const BatchJob* __restrict__ jobs,
                                                       int njobs, int p, int slot) {
  const BatchJob& j = jobs[batch_job(jobs, njobs, blockIdx.x)];
  if (half_skip(j.ctl, slot)) return;
  su2_half_block<T, GA, NT, SPT, MODE>((T*)j.U, (T*)j.G, j.g, p, (T)j.omega,
                                       j.part + (size_t)slot * j.pstride + (size_t)3 * j.nw * p,
                                       blockIdx.x - j.block0, j.U0, j.part + (size_t)kSlots * j.pstride);
}

// launch of a shape (NT, SPT) of the supported set
#define MAG_SHAPE_SWITCH(sh, L_)                  \
  switch (sh.nt * 4 + sh.spt) {                   \
    case 64 * 4 + 1: L_(64, 1); break;            \
    case 64 * 4 + 2: L_(64, 2); break;            \
    case 128 * 4 + 1: L_(128, 1); break;          \
    case 128 * 4 + 2: L_(128, 2); break;          \
    case 256 * 4 + 2: L_(256, 2); break;          \
    case 512 * 4 + 1: L_(512, 1); break;          \
    case 512 * 4 + 2: L_(512, 2); break;          \
    default: L_(256, 1); break;                   \
  }

// kernel variant of a launch: 0 plain, 1 plain + site gauge (red FP32
// half), 2 renormalise + site gauge, 3 rebuild (black half, G = red gauges)
template <typename T>
int half_kind(const void* G, int mode) {
  if (!G) return 0;
  if (sizeof(T) == 8) return 1;
  return mode == HALF_REBUILD ? 3 : mode == HALF_RENORM ? 2 : 1;
}

// one launch of kernel template K<T, GA, NT, SPT, MODE> for variant kind;
// FP64 launches of DGA kernels (SU(2)) may accumulate the site gauge too
// (kind 1: gauge tracking of an FP64 job)
#define MAG_HALF_VARIANTS(K, DGA, NT, SPT, grid, st, ...)                                     \
  do {                                                                                         \
    if constexpr (sizeof(T) == 4) {                                                            \
      switch (kind) {                                                                          \
        case 1: hipLaunchKernelGGL((K<T, true, NT, SPT, HALF_PLAIN>), grid, dim3(NT), 0, st,   \
                                   __VA_ARGS__); break;                                        \
        case 2: hipLaunchKernelGGL((K<T, true, NT, SPT, HALF_RENORM>), grid, dim3(NT), 0, st,  \
                                   __VA_ARGS__); break;                                        \
        case 3: hipLaunchKernelGGL((K<T, false, NT, SPT, HALF_REBUILD>), grid, dim3(NT), 0,    \
                                   st, __VA_ARGS__); break;                                    \
        default: hipLaunchKernelGGL((K<T, false, NT, SPT, HALF_PLAIN>), grid, dim3(NT), 0, st, \
                                    __VA_ARGS__); break;                                       \
      }                                                                                        \
    } else if constexpr (DGA) {                                                                \
      if (kind == 1)                                                                           \
        hipLaunchKernelGGL((K<T, true, NT, SPT, HALF_PLAIN>), grid, dim3(NT), 0, st,           \
                           __VA_ARGS__);                                                       \
      else                                                                                     \
        hipLaunchKernelGGL((K<T, false, NT, SPT, HALF_PLAIN>), grid, dim3(NT), 0, st,          \
                           __VA_ARGS__);                                                       \
    } else {                                                                                   \
      hipLaunchKernelGGL((K<T, false, NT, SPT, HALF_PLAIN>), grid, dim3(NT), 0, st,            \
                         __VA_ARGS__);                                                         \
    }                                                                                          \
  } while (0)

template <typename T>
void launch_su2_half(T* U, T* G, const Geom& g, int parity, double omega, double* part,
                     const JobCtl* ctl, Shape sh, hipStream_t st, const HalfOpt& o) {
  const size_t nw = (size_t)sweep_waves(g, 2);
  double* pp = part + (size_t)o.slot * part_stride((int)nw) + 3 * nw * parity;
  const uint32_t lanes = g.Vh;
  const int kind = half_kind<T>(G, o.mode);
#define L_(NT, SPT)                                                                        \
  do {                                                                                     \
    const dim3 grid((lanes + NT * SPT - 1) / (NT * SPT));                                  \
    MAG_HALF_VARIANTS(k_su2_half, true, NT, SPT, grid, st, U, G, g, parity, (T)omega, pp,  \
                      ctl, o.slot, o.U0, o.part_rb);                                       \
  } while (0)
#define LS_(NT, SPT)                                                                         \
  do {                                                                                       \
    const dim3 grid((lanes + NT * SPT - 1) / (NT * SPT));                                    \
    MAG_HALF_VARIANTS(k_su2_half_st, true, NT, SPT, grid, st, U, G, g, parity, (T)omega, pp, \
                      ctl, o.slot, o.U0, o.part_rb);                                         \
  } while (0)
  if (stream_links(g, 4 * sizeof(T))) {
    MAG_SHAPE_SWITCH(sh, LS_)
  } else {
    MAG_SHAPE_SWITCH(sh, L_)
  }
#undef LS_
#undef L_
}

// ================================================================== SU(3)
#ifndef MAG_SU3_PTR_REMAT
#define MAG_SU3_PTR_REMAT 1
#endif
#ifndef MAG_SU3_ROW2_REMAT
#define MAG_SU3_ROW2_REMAT 1
#endif
// a value the compiler must treat as changed (no instruction is emitted)
template <typename T>
__device__ __forceinline__ void opaque(M2<T>& m) {
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int c = 0; c < 3; c++) asm volatile("" : "+v"(m.r[i][c].re), "+v"(m.r[i][c].im));
}

// Cabibbo-Marinari subgroup (I,J) update of the 2*NL links a lane holds
// (forward F[k], backward B[k], mu = NL*q + k; NL = 2 for lane pairs, 1 for
// lane quads), rows 0,1 in registers, row 2 rebuilt where the subgroup
// touches it.  Returns 1 - cos(theta).  The field is summed
// as t_mu = f_mu + b_mu, then (t_0 + t_1) + (t_2 + t_3), the same
// operations for either split (a + b == b + a exactly), in the order of the
// oracle.
template <typename T, int I, int J, bool GA, int NL>
__device__ __forceinline__ T cm_subgroup(M2<T>* F, M2<T>* B, M2<T>& Gx, T omega) {
  T h0 = T(0), h1 = T(0), h2 = T(0);
  C<T> Fr2[NL][3], Br2[NL][3];
#pragma unroll
  for (int k = 0; k < NL; k++) {
    if (J == 2) {
      row2_of(F[k], Fr2[k]);
      row2_of(B[k], Br2[k]);
    }
    auto fe = [&](int a, int b) -> C<T> { return a < 2 ? F[k].r[a][b] : Fr2[k][b]; };
    auto be = [&](int a, int b) -> C<T> { return a < 2 ? B[k].r[a][b] : Br2[k][b]; };
    T t0 = T(0), t1 = T(0), t2 = T(0);
    blk_fwd(fe(I, I), fe(I, J), fe(J, I), fe(J, J), t0, t1, t2);
    blk_bwd(be(I, I), be(I, J), be(J, I), be(J, J), t0, t1, t2);
    if (k == 0) {
      h0 = t0;
      h1 = t1;
      h2 = t2;
    } else {
      h0 += t0;
      h1 += t1;
      h2 += t2;
    }
  }
  h0 += pair_xchg(h0);
  h1 += pair_xchg(h1);
  h2 += pair_xchg(h2);
  if (NL == 1) {
    h0 += quad_xchg2(h0);
    h1 += quad_xchg2(h1);
    h2 += quad_xchg2(h2);
  }
  T res;
  const Q<T> gq = omega_gauge(h0, h1, h2, omega, res);
  // r = [[g0 + i g3, g2 + i g1], [-g2 + i g1, g0 - i g3]], g3 = 0
  const C<T> r00{gq.a0, T(0)}, r01{gq.a2, gq.a1}, r10{-gq.a2, gq.a1}, r11{gq.a0, T(0)};
#if MAG_SU3_ROW2_REMAT
  // row 2 of the forward links again rather than kept live across the
  // gauge computation (the empty asm hides that F is unchanged)
  if (J == 2) {
#pragma unroll
    for (int k = 0; k < NL; k++) {
      opaque(F[k]);
      row2_of(F[k], Fr2[k]);
    }
  }
#endif
#pragma unroll
  for (int k = 0; k < NL; k++) {
    // forward: rows I,J <- r (rows I,J); only stored rows are written
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const C<T> a = F[k].r[I][c];
      const C<T> b = (J < 2) ? F[k].r[J][c] : Fr2[k][c];
      F[k].r[I][c] = cadd(cmul(r00, a), cmul(r01, b));
      if (J < 2) F[k].r[J][c] = cadd(cmul(r10, a), cmul(r11, b));
    }
    // backward: cols I,J of rows 0,1 <- (cols I,J) r^+
#pragma unroll
    for (int c = 0; c < 2; c++) {
      const C<T> a = B[k].r[c][I], b = B[k].r[c][J];
      B[k].r[c][I] = cadd(cmulc(a, r00), cmulc(b, r01));
      B[k].r[c][J] = cadd(cmulc(a, r10), cmulc(b, r11));
    }
  }
  if (GA) {
    C<T> g2[3];
    if (J == 2) row2_of(Gx, g2);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const C<T> a = Gx.r[I][c];
      const C<T> b = (J < 2) ? Gx.r[J][c] : g2[c];
      Gx.r[I][c] = cadd(cmul(r00, a), cmul(r01, b));
      if (J < 2) Gx.r[J][c] = cadd(cmul(r10, a), cmul(r11, b));
    }
  }
  return res;
}

// rows 0,1 of a lane's matrix replaced by those of the first lane of its
// site's lane group (LS = 2 or 4 lanes, DPP within the quad)
template <int LS>
__device__ __forceinline__ void group_broadcast(M2<double>& m, int q) {
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      double re = m.r[i][c].re, im = m.r[i][c].im;
      const double pre = pair_xchg(re), pim = pair_xchg(im);
      if (q & 1) { re = pre; im = pim; }
      if (LS == 4) {
        const double qre = quad_xchg2(re), qim = quad_xchg2(im);
        if (q & 2) { re = qre; im = qim; }
      }
      m.r[i][c] = C<double>{re, im};
    }
}

// FP64 links, and FP32 links past the MALL, are stored with the streaming hint
template <typename T, bool ST>
__device__ __forceinline__ void su3_store(LRef<T> r, const M2<T>& m) {
  if constexpr (sizeof(T) == 8 || ST) stm_stream(r, m);
  else stm(r, m);
}

// Plain half-sweeps (every shape), as for SU(2): every tile's links (and
// red gauge) are loaded before the first tile's update.  Against the
// one-tile-at-a-time code, 32^4: FP64 256x1 135 us per half-sweep (137),
// FP32 256x1 65.3 (66.1); 24^3x48 FP64 90.1 (92.8).  Two tiles per thread
// gain nothing here (FP64: 254 VGPRs, 2 waves per SIMD, 160 us).
template <typename T, bool GA, int NT, int SPT, bool ST>
__device__ __forceinline__ void su3_half_tiles(T* __restrict__ U, T* __restrict__ G,
                                               const Geom& g, int p, T omega,
                                               double* __restrict__ part, uint32_t lb) {
  constexpr int LS = kSu3Lanes, NL = 4 / kSu3Lanes;
  M2<T> F[SPT][NL], B[SPT][NL], Gx[SPT];
  uint32_t hbm[SPT][NL];
#pragma unroll
  for (int kk = 0; kk < SPT; kk++) {
    const uint32_t tid = (lb * SPT + kk) * NT + threadIdx.x;
    const uint32_t h = tid / LS;
    const int q = (int)(tid % LS);
    if (h < g.Vh) {
      const SiteIdx s = decode(g, h, p);
      uint32_t hb[4];
      back_all(g, h, s, hb);
#pragma unroll
      for (int k = 0; k < NL; k++) {
        const int mu = NL * q + k;
        hbm[kk][k] = mu == 0 ? hb[0] : mu == 1 ? hb[1] : mu == 2 ? hb[2] : hb[3];
        const LRef<T> pf = link_ptr<12>(U, g, p, mu, h);
        const LRef<T> pb = link_ptr<12>(U, g, 1 - p, mu, hbm[kk][k]);
        if constexpr (ST) {
          F[kk][k] = ldm_nt(LRef<const T>{pf.p, pf.cs});
          B[kk][k] = ldm_nt(LRef<const T>{pb.p, pb.cs});
        } else {
          F[kk][k] = ldm(pf);
          B[kk][k] = ldm(pb);
        }
      }
      if constexpr (GA) Gx[kk] = ldm(site_ptr<12>(G, g, p, h));
    }
  }
#pragma unroll
  for (int kk = 0; kk < SPT; kk++) {
    const uint32_t base = (lb * SPT + kk) * NT;
    const uint32_t tid = base + threadIdx.x;
    const uint32_t h = tid / LS;
    const int q = (int)(tid % LS);
    double fs = 0.0, rs = 0.0, rm = 0.0;
    if (h < g.Vh) {
      T res = cm_subgroup<T, 0, 1, GA, NL>(F[kk], B[kk], Gx[kk], omega);
      res += cm_subgroup<T, 1, 2, GA, NL>(F[kk], B[kk], Gx[kk], omega);
      res += cm_subgroup<T, 0, 2, GA, NL>(F[kk], B[kk], Gx[kk], omega);
#pragma unroll
      for (int k = 0; k < NL; k++) {
        const int mu = NL * q + k;
        su3_store<T, ST>(link_ptr<12>(U, g, p, mu, h), F[kk][k]);
        su3_store<T, ST>(link_ptr<12>(U, g, 1 - p, mu, hbm[kk][k]), B[kk][k]);
      }
      if (GA && q == 0) stm(site_ptr<12>(G, g, p, h), Gx[kk]);
      if (p == 1) {
#pragma unroll
        for (int k = 0; k < NL; k++) fs += su3_link_F(F[kk][k]) + su3_link_F(B[kk][k]);
      }
      if (q == 0) {
        rs = (double)res;
        rm = (double)res;
      }
    }
    wave_partial_sweep(part, p, base + (threadIdx.x & ~63u), LS * g.Vh, fs, rs, rm);
  }
}

// One workgroup lb of an SU(3) half-sweep, LS = kSu3Lanes lanes per site:
// lane q of the group holds the forward and backward links of the NL = 4/LS
// directions mu = NL*q .. NL*q + NL - 1, the group exchanges the partial
// fields with DPP and every lane derives the same g.  GA / MODE as for SU(2):
// the red FP32 half accumulates the site gauge (renormalised first in
// HALF_RENORM); HALF_REBUILD recovers the black gauge from the swept
// W(x,0) on the group's first lane, broadcasts it, and each lane rebuilds
// its links from the master.
// Plain half-sweeps go to su3_half_tiles; this loop runs the other modes.
template <typename T, bool GA, int NT, int SPT, int MODE, bool ST>
__device__ __forceinline__ void su3_half_block_st(T* __restrict__ U, T* __restrict__ G,
                                                  const Geom& g, int p, T omega,
                                                  double* __restrict__ part, uint32_t lb,
                                                  const double* __restrict__ U0,
                                                  double* __restrict__ part_rb) {
  constexpr int LS = kSu3Lanes, NL = 4 / kSu3Lanes;
  if constexpr (MODE == HALF_PLAIN) {
    su3_half_tiles<T, GA, NT, SPT, ST>(U, G, g, p, omega, part, lb);
    return;
  }
#pragma unroll
  for (int kk = 0; kk < SPT; kk++) {
    const uint32_t base = (lb * SPT + kk) * NT;
    const uint32_t tid = base + threadIdx.x;
    const uint32_t h = tid / LS;
    const int q = (int)(tid % LS);
    double fs = 0.0, rs = 0.0, rm = 0.0, fr = 0.0;
    if (h < g.Vh) {
      const SiteIdx s = decode(g, h, p);
      uint32_t hb[4];
      back_all(g, h, s, hb);  // all four, then a select: no divergent switch
      LRef<T> pf[NL];
      LRef<T> pb[NL];
      M2<T> F[NL], B[NL];
#pragma unroll
      for (int k = 0; k < NL; k++) {
        const int mu = NL * q + k;
        const uint32_t hbm = mu == 0 ? hb[0] : mu == 1 ? hb[1] : mu == 2 ? hb[2] : hb[3];
        pf[k] = link_ptr<12>(U, g, p, mu, h);
        pb[k] = link_ptr<12>(U, g, 1 - p, mu, hbm);
        if constexpr (ST) {
          F[k] = ldm_nt(LRef<const T>{pf[k].p, pf[k].cs});
          B[k] = ldm_nt(LRef<const T>{pb[k].p, pb[k].cs});
        } else {
          F[k] = ldm(pf[k]);
          B[k] = ldm(pb[k]);
        }
      }
      M2<T> Gx;
      const LRef<T> gp = site_ptr<12>(G, g, p, h);
      if constexpr (GA) {
        Gx = ldm(gp);
        if constexpr (MODE == HALF_RENORM) Gx = to_f(gs3(to_d(Gx)));
      }
      T res = cm_subgroup<T, 0, 1, GA, NL>(F, B, Gx, omega);
      res += cm_subgroup<T, 1, 2, GA, NL>(F, B, Gx, omega);
      res += cm_subgroup<T, 0, 2, GA, NL>(F, B, Gx, omega);
      if constexpr (MODE == HALF_REBUILD) {
        // F of the swept links for the history, then the links from the master
#pragma unroll
        for (int k = 0; k < NL; k++) fs += su3_link_F(F[k]) + su3_link_F(B[k]);
        const int mu0 = NL * q;
        const uint32_t hf0 = fwd_h(g, h, s, mu0);
        // the master is read once per period: loaded with the nontemporal
        // hint it leaves the MALL to the FP32 links and the red gauges the
        // next half-sweeps read (32^4 per sweep 146.6 -> 144.4 us; the same
        // hint cost SU(2) 4%, which keeps plain loads)
        M2<double> y = black_gauge_rows(to_d(F[0]), gs3(to_d(ldm(site_ptr<12>(G, g, 0, hf0)))),
                                        ldm_nt(link_ptr<12>(U0, g, 1, mu0, h)));
        group_broadcast<LS>(y, q);  // W(x,0) is on the group's first lane
        const M3<double> gx = gs3(y);
#pragma unroll
        for (int k = 0; k < NL; k++) {
          const int mu = NL * q + k;
          const uint32_t hbm = mu == 0 ? hb[0] : mu == 1 ? hb[1] : mu == 2 ? hb[2] : hb[3];
          const M2<double> wf =
              rebuild_link(gx, ldm_nt(link_ptr<12>(U0, g, 1, mu, h)),
                           gs3(to_d(ldm(site_ptr<12>(G, g, 0, fwd_h(g, h, s, mu))))));
          // each rebuilt link is stored at once: the swept links are dead
          // here (holding the rebuilt ones to the common stores below took
          // 182 VGPRs, now 164; the duration is the same, bound by the
          // launch's bytes: master, swept links and neighbour gauges)
          const double Ff = su3_link_F(wf);
          su3_store<T, ST>(link_ptr<12>(U, g, p, mu, h), to_f(wf));
          const M2<double> wb = rebuild_link(gs3(to_d(ldm(site_ptr<12>(G, g, 0, hbm)))),
                                             ldm_nt(link_ptr<12>(U0, g, 0, mu, hbm)), gx);
          fr += Ff + su3_link_F(wb);
          su3_store<T, ST>(link_ptr<12>(U, g, 1 - p, mu, hbm), to_f(wb));
        }
      }
#if MAG_SU3_PTR_REMAT
      // the store addresses again from the 32-bit indices rather than 64-bit
      // pointers kept live through the three subgroups
#pragma unroll
      for (int k = 0; k < NL; k++) {
        const int mu = NL * q + k;
        uint32_t hh = h, hbm = mu == 0 ? hb[0] : mu == 1 ? hb[1] : mu == 2 ? hb[2] : hb[3];
        asm volatile("" : "+v"(hh), "+v"(hbm));
        pf[k] = link_ptr<12>(U, g, p, mu, hh);
        pb[k] = link_ptr<12>(U, g, 1 - p, mu, hbm);
      }
#endif
      if constexpr (MODE != HALF_REBUILD)
#pragma unroll
        for (int k = 0; k < NL; k++) {
          su3_store<T, ST>(pf[k], F[k]);
          su3_store<T, ST>(pb[k], B[k]);
        }
      if (GA && q == 0) stm(gp, Gx);
      if (MODE != HALF_REBUILD && p == 1) {
#pragma unroll
        for (int k = 0; k < NL; k++) fs += su3_link_F(F[k]) + su3_link_F(B[k]);
      }
      if (q == 0) {
        rs = (double)res;
        rm = (double)res;
      }
    }
    wave_partial_sweep(part, p, base + (threadIdx.x & ~63u), LS * g.Vh, fs, rs, rm);
    if constexpr (MODE == HALF_REBUILD)
      wave_partial(part_rb, base + (threadIdx.x & ~63u), LS * g.Vh, fr, 0.0, 0.0);
  }
}

// FP32 links past the MALL (48^3x24) are loaded and stored with the
// nontemporal hint, as for SU(2): 212 -> 191 us per half-sweep.  FP64
// links are always stored with it and always loaded plainly (the hinted
// loads cost 11% at 32^4, 768 MiB of links).
template <typename T>
__host__ __device__ inline bool su3_stream(const Geom& g) {
  return sizeof(T) == 4 && stream_links(g, 12 * sizeof(T));
}
template <typename T, bool GA, int NT, int SPT, int MODE>
__device__ __forceinline__ void su3_half_block(T* __restrict__ U, T* __restrict__ G,
                                               const Geom& g, int p, T omega,
                                               double* __restrict__ part, uint32_t lb,
                                               const double* __restrict__ U0,
                                               double* __restrict__ part_rb) {
  if (su3_stream<T>(g))  // uniform
    su3_half_block_st<T, GA, NT, SPT, MODE, true>(U, G, g, p, omega, part, lb, U0, part_rb);
  else
    su3_half_block_st<T, GA, NT, SPT, MODE, false>(U, G, g, p, omega, part, lb, U0, part_rb);
}

// one job: the launcher picks the plain or the streaming kernel by lattice
template <typename T, bool GA, int NT, int SPT, int MODE>
__global__ __launch_bounds__(NT) void k_su3_half(T* __restrict__ U, T* __restrict__ G, Geom g,
                                                 int p, T omega, double* __restrict__ part,
                                                 const JobCtl* __restrict__ ctl, int slot,
                                                 const double* __restrict__ U0,
                                                 double* __restrict__ part_rb) {
  if (half_skip(ctl, slot)) return;
  su3_half_block_st<T, GA, NT, SPT, MODE, false>(U, G, g, p, omega, part, blockIdx.x, U0,
                                                 part_rb);
}
template <typename T, bool GA, int NT, int SPT, int MODE>
__global__ __launch_bounds__(NT) void k_su3_half_st(T* __restrict__ U, T* __restrict__ G, Geom g,
                                                    int p, T omega, double* __restrict__ part,
                                                    const JobCtl* __restrict__ ctl, int slot,
                                                    const double* __restrict__ U0,
                                                    double* __restrict__ part_rb) {
  if (half_skip(ctl, slot)) return;
  su3_half_block_st<T, GA, NT, SPT, MODE, true>(U, G, g, p, omega, part, blockIdx.x, U0,
                                                part_rb);
}

template <typename T, bool GA, int NT, int SPT, int MODE>
__global__ __launch_bounds__(NT) void k_su3_half_batch(const BatchJob* __restrict__ jobs,
                                                       int njobs, int p, int slot) {
  const BatchJob& j = jobs[batch_job(jobs, njobs, blockIdx.x)];
  if (half_skip(j.ctl, slot)) return;
  su3_half_block<T, GA, NT, SPT, MODE>((T*)j.U, (T*)j.G, j.g, p, (T)j.omega,
                                       j.part + (size_t)slot * j.pstride + (size_t)3 * j.nw * p,
                                       blockIdx.x - j.block0, j.U0, j.part + (size_t)kSlots * j.pstride);
}

template <typename T>
void launch_su3_half(T* U, T* G, const Geom& g, int parity, double omega, double* part,
                     const JobCtl* ctl, Shape sh, hipStream_t st, const HalfOpt& o) {
  const size_t nw = (size_t)sweep_waves(g, 3);
  double* pp = part + (size_t)o.slot * part_stride((int)nw) + 3 * nw * parity;
  const uint32_t lanes = (uint32_t)sweep_lanes(g, 3);
  const int kind = half_kind<T>(G, o.mode);
#define L_(NT, SPT)                                                                        \
  do {                                                                                     \
    const dim3 grid((lanes + NT * SPT - 1) / (NT * SPT));                                  \
    MAG_HALF_VARIANTS(k_su3_half, false, NT, SPT, grid, st, U, G, g, parity, (T)omega, pp, \
                      ctl, o.slot, o.U0, o.part_rb);                                       \
  } while (0)
#define LS_(NT, SPT)                                                                          \
  do {                                                                                        \
    const dim3 grid((lanes + NT * SPT - 1) / (NT * SPT));                                     \
    MAG_HALF_VARIANTS(k_su3_half_st, false, NT, SPT, grid, st, U, G, g, parity, (T)omega, pp, \
                      ctl, o.slot, o.U0, o.part_rb);                                          \
  } while (0)
  if (su3_stream<T>(g)) {
    MAG_SHAPE_SWITCH(sh, LS_)
  } else {
    MAG_SHAPE_SWITCH(sh, L_)
  }
#undef LS_
#undef L_
}

// batched half-sweep over a class table of storage type T (float: the
// mixed FP32 phase, double: FP64): every job of the table has the same
// group; block0 / nblocks_total counted in workgroups of the shape sh; in
// the FP32 class parity 0 accumulates the site gauge and `mode` selects the
// renormalising / rebuilding variants as in HalfOpt
template <typename T>
void launch_half_batch_t(int group, Shape sh, const BatchJob* jobs, int njobs, int nblocks_total,
                         int parity, hipStream_t st, int slot, int mode) {
  if (njobs == 0 || nblocks_total == 0) return;
  const dim3 grid(nblocks_total);
  // red: site gauge (kind 1 / 2), black: none (0) or rebuild (3)
  const int kind = sizeof(T) == 8 ? 0
                   : parity == 0  ? (mode == HALF_RENORM ? 2 : 1)
                                  : (mode == HALF_REBUILD ? 3 : 0);
#define L_(NT, SPT)                                                                          \
  do {                                                                                       \
    if (group == 2)                                                                          \
      MAG_HALF_VARIANTS(k_su2_half_batch, false, NT, SPT, grid, st, jobs, njobs, parity, slot); \
    else                                                                                     \
      MAG_HALF_VARIANTS(k_su3_half_batch, false, NT, SPT, grid, st, jobs, njobs, parity, slot); \
  } while (0)
  MAG_SHAPE_SWITCH(sh, L_)
#undef L_
}

}  // namespace mag
