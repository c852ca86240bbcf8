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
struct Geom {
  int L0, L1, L2, L3;
  uint32_t hx;  // L0/2: half sites per x-row
  uint32_t Vh;  // V/2
  uint32_t s1;  // half-index stride of y  (= hx)
  uint32_t s2;  // half-index stride of z  (= hx*L1)
  uint32_t s3;  // half-index stride of t  (= hx*L1*L2)
};

inline Geom make_geom(const int* d) {
  Geom g;
  g.L0 = d[0]; g.L1 = d[1]; g.L2 = d[2]; g.L3 = d[3];
  g.hx = (uint32_t)(d[0] / 2);
  g.s1 = g.hx;
  g.s2 = g.hx * (uint32_t)d[1];
  g.s3 = g.s2 * (uint32_t)d[2];
  g.Vh = g.s3 * (uint32_t)d[3];
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
  unsigned int arrived;  // black half-sweep workgroups done (fused reduction)
  int pad;
};

// Everything the end-of-sweep reduction needs: partial triples of the red
// half at part[0 .. 3*nb), of the black half at part[3*nb .. 6*nb).
struct SweepRed {
  JobCtl* ctl;
  double* part;
  double* hF;
  double* hR;
  double* hA;
  double invF;  // 1/(4V) (SU(2)) or 1/(12V) (SU(3))
  double invV;
  int nparts;   // 2*nb
};

// reduction kernel modes
//   RED_INIT : F0 = F_prev = F      RED_SWEEP: append a sweep to the history
//   RED_PLAIN: out3 only            RED_RESET: F_prev = F (phase change)
enum { RED_INIT = 0, RED_SWEEP = 1, RED_PLAIN = 2, RED_RESET = 3 };

// ---------------------------------------------------------------- launchers
// Partial buffers hold 3 doubles per workgroup: (F sum, residual sum,
// residual max).  A sweep uses 2*nblocks(g) partial triples (red, black).
inline int nblocks(const Geom& g) { return (int)((g.Vh + kBlock - 1) / kBlock); }

// fused = true: the last black workgroup to finish reduces the sweep's
// partials itself (device-scope fence + arrival counter) and no separate
// reduction kernel is launched for the sweep.
template <typename T>
void launch_su2_half(T* U, T* G, const Geom& g, int parity, double omega,
                     const SweepRed& r, bool fused, hipStream_t st);
template <typename T>
void launch_su3_half(T* U, T* G, const Geom& g, int parity, double omega,
                     const SweepRed& r, bool fused, hipStream_t st);

// per-sweep (or initial) reduction of the partials into the job history
void launch_reduce(const double* part, int nparts, int mode, double invF,
                   double invV, JobCtl* ctl, double* hist_F, double* hist_rmax,
                   double* hist_ravg, double* out3, hipStream_t st);

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

// elementwise FP64 -> FP32 copy (n reals)
void launch_f64_to_f32(const double* a, float* b, size_t n, hipStream_t st);
// site gauges set to the identity (NR reals per site, both parities)
void launch_gauge_identity(float* G, const Geom& g, int group, hipStream_t st);

// Rebuild links from the FP64 master and the accumulated site gauge:
// out = G(x) U0 G(x+mu)^+ with G normalised in FP64 (SU(2): q/|q|, SU(3):
// Gram-Schmidt).  out32 (FP32 working links) or out64 (may alias U0).
void launch_su2_rebuild(const double* U0, const float* G, float* out32,
                        double* out64, const Geom& g, hipStream_t st);
void launch_su3_rebuild(const double* U0, const float* G, float* out32,
                        double* out64, const Geom& g, hipStream_t st);
// G <- float(normalise_fp64(G))
void launch_gauge_normalize(float* G, const Geom& g, int group, hipStream_t st);
// FP64 reunitarisation of links in place (SU(2) q/|q|, SU(3) Gram-Schmidt)
void launch_su2_reunit64(double* U, const Geom& g, hipStream_t st);
void launch_su3_reunit64(double* U, const Geom& g, hipStream_t st);

// seeded synthetic configurations in the reference AoS layout (FP64);
// SU(3) needs V*18 doubles of scratch for the site gauges
void launch_gen_su2(double* aos, const Geom& g, double eps, uint64_t seed, hipStream_t st);
void launch_gen_su3(double* aos, double* gauge_scratch, const Geom& g, double eps,
                    uint64_t seed, hipStream_t st);

// plaquette partials (sum over sites of the 6-plane mean), 2*nblocks triples
void launch_su2_plaquette(const double* U, const Geom& g, double* part, hipStream_t st);
void launch_su3_plaquette(const double* U, const Geom& g, double* part, hipStream_t st);

}  // namespace mag

namespace mag {
// 16-byte streaming copy of bytes (a multiple of 16)
void launch_copy(const void* a, void* b, size_t bytes, hipStream_t st);
}  // namespace mag
