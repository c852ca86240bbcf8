// mag_monopole.hip -- abelian projection of gauge-fixed links and monopole
// currents on the device: what the reference's monopole apps
// (apps/monopoles_su2/monopoles_su2.cpp:110, apps/monopoles_su3/
// monopoles_su3.cpp:103-120) compute from MAG-fixed configurations, so the
// links of a MAG job go from fixing to currents without leaving the GPU.
//
//   angles     SU(2): phi = atan2(a3, a0) per link (MatrixConverter<su2,
//              abelian>, matrix.h:1059-1062); SU(3): the phases of the
//              diagonal less a third of their sum taken into [-pi, pi)
//              (MatrixConverter<su3, su3_angles>, matrix.h:1082-1100)
//   plaquettes phi_mu(x) + phi_nu(x+mu) - phi_mu(x+nu) - phi_nu(x) taken
//              into [-pi, pi] (monopoles.cpp:368-457); SU(3): then the
//              colour sum brought into (-pi, pi) by moving 2 pi off the
//              largest / onto the smallest colour (monopoles.cpp:844-884)
//   currents   J_mu(x) = forward differences at x + mu of the three
//              plaquette planes orthogonal to mu, over 2 pi
//              (get_current, monopoles.cpp:44-122)
//   clusters   connected components of the current links (calculate_clusters,
//              monopoles.cpp:1277-1303) with the sums the apps take over
//              each cluster (cluster_length, length_mu,
//              currents_directions, monopoles.cpp:1343-1449)
// Every kernel runs one thread per lexicographic site (the plaquette and
// current kernels, which read neighbours, in XCD-aware block order).
// Between the kernels
// the angles are planar, th[(colour * 4 + mu) * V + site], so that the
// neighbour loads of a wave are contiguous; the reference's link-major
// layouts ([4V], [4V][3]) are produced or read only at the boundary.  The
// plaquette and current kernels do the reference's additions, subtractions
// and divisions in the reference's order, so from the same angles they give
// the same bits; the angles go through the device atan2.
#include "mag_capi_util.h"
#include "mag_gpu.h"
#include "mag_internal.h"
#include "mag_job_impl.h"
#include "mag_layout.h"
#include "mag_math.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace mag {
namespace {

constexpr double kPi = 3.141592653589793;  // M_PI
constexpr int kMonoBlock = 256;
// input angles accepted by the host entry points: every plaquette sum is
// then within 4000 of 0 and its reduction takes at most 640 steps of 2 pi
constexpr double kMaxAngle = 1000.0;
constexpr int kMaxWrap = 1024;

// lexicographic lattice of any extents (indexing.h:111-123)
struct Lat {
  int L[4];
  uint32_t st[4];  // site strides
  uint32_t V;
};

Lat make_lat(const int* d) {
  Lat l;
  uint32_t s = 1;
  for (int mu = 0; mu < 4; mu++) {
    l.L[mu] = d[mu];
    l.st[mu] = s;
    s *= (uint32_t)d[mu];
  }
  l.V = s;
  return l;
}

__device__ __forceinline__ void coords_of(const Lat& l, uint32_t s, int* c) {
  c[0] = (int)(s % (uint32_t)l.L[0]);
  s /= (uint32_t)l.L[0];
  c[1] = (int)(s % (uint32_t)l.L[1]);
  s /= (uint32_t)l.L[1];
  c[2] = (int)(s % (uint32_t)l.L[2]);
  c[3] = (int)(s / (uint32_t)l.L[2]);
}
// s + mu, periodic (DataPatternLexicographical::move_forward, indexing.h:97-99)
__device__ __forceinline__ uint32_t fwd(const Lat& l, uint32_t s, const int* c, int mu) {
  return c[mu] + 1 == l.L[mu] ? s - (uint32_t)(l.L[mu] - 1) * l.st[mu] : s + l.st[mu];
}
// s + mu + nu
__device__ __forceinline__ uint32_t fwd2(const Lat& l, uint32_t s, const int* c, int mu, int nu) {
  int d[4] = {c[0], c[1], c[2], c[3]};
  const uint32_t a = fwd(l, s, d, mu);
  d[mu] = d[mu] + 1 == l.L[mu] ? 0 : d[mu] + 1;
  return fwd(l, a, d, nu);
}

// into [-pi, pi] as the reference's two while loops (monopoles.cpp:443-448)
__device__ __forceinline__ double wrap_pi(double p) {
  for (int i = 0; i < kMaxWrap && p > kPi; i++) p -= 2 * kPi;
  for (int i = 0; i < kMaxWrap && p < -kPi; i++) p += 2 * kPi;
  return p;
}

// ------------------------------------------------------------------ angles
template <typename T>
__global__ __launch_bounds__(kMonoBlock) void k_su2_angles(const T* __restrict__ U, Geom g,
                                                          double* __restrict__ th) {
  const uint32_t s = blockIdx.x * kMonoBlock + threadIdx.x;
  if (s >= 2 * g.Vh) return;
  int p;
  uint32_t h;
  lex_site(g, s, p, h);
  const size_t V = 2 * (size_t)g.Vh;
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    const Q<T> q = ldq(link_ptr<4>(U, g, p, mu, h));
    th[mu * V + s] = atan2((double)q.a3, (double)q.a0);
  }
}

template <typename T>
__global__ __launch_bounds__(kMonoBlock) void k_su3_angles(const T* __restrict__ U, Geom g,
                                                          double* __restrict__ th3) {
  const uint32_t s = blockIdx.x * kMonoBlock + threadIdx.x;
  if (s >= 2 * g.Vh) return;
  int p;
  uint32_t h;
  lex_site(g, s, p, h);
  const size_t V = 2 * (size_t)g.Vh;
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    const M2<T> o = ldm(link_ptr<12>(U, g, p, mu, h));
    M2<double> m;
#pragma unroll
    for (int k = 0; k < 6; k++)
      m.r[k / 3][k % 3] = {(double)o.r[k / 3][k % 3].re, (double)o.r[k / 3][k % 3].im};
    C<double> r2[3];
    row2_of(m, r2);
    double a[3];
    a[0] = atan2(m.r[0][0].im, m.r[0][0].re);
    a[1] = atan2(m.r[1][1].im, m.r[1][1].re);
    a[2] = atan2(r2[2].im, r2[2].re);
    double sum = a[0] + a[1];
    sum = sum + a[2];
    for (int i = 0; i < kMaxWrap && sum >= kPi; i++) sum -= 2 * kPi;
    for (int i = 0; i < kMaxWrap && sum < -kPi; i++) sum += 2 * kPi;
#pragma unroll
    for (int c = 0; c < 3; c++) th3[(c * 4 + mu) * V + s] = a[c] - sum / 3;
  }
}

// reference link-major angles [4V][ncol] <-> planar [(c * 4 + mu) * V + site]
__global__ __launch_bounds__(kMonoBlock) void k_angles_to_planar(const double* __restrict__ ref,
                                                                int ncol, uint32_t V,
                                                                double* __restrict__ th) {
  const uint32_t s = blockIdx.x * kMonoBlock + threadIdx.x;
  if (s >= V) return;
  for (int mu = 0; mu < 4; mu++)
    for (int c = 0; c < ncol; c++) th[(size_t)(c * 4 + mu) * V + s] = ref[(4 * (size_t)s + mu) * ncol + c];
}
__global__ __launch_bounds__(kMonoBlock) void k_angles_to_ref(const double* __restrict__ th,
                                                             int ncol, uint32_t V,
                                                             double* __restrict__ ref) {
  const uint32_t s = blockIdx.x * kMonoBlock + threadIdx.x;
  if (s >= V) return;
  for (int mu = 0; mu < 4; mu++)
    for (int c = 0; c < ncol; c++) ref[(4 * (size_t)s + mu) * ncol + c] = th[(size_t)(c * 4 + mu) * V + s];
}

// ------------------------------------------------------------------ plaquettes
// plane index of (mu, nu), mu < nu, in the order of the reference's loops
__device__ __forceinline__ int plane_of(int mu, int nu) {
  return mu == 0 ? nu - 1 : (mu == 1 ? nu + 1 : 5);
}

// angle of the plaquette (mu, nu) at s from the planar angles th[mu * V + site]
__device__ __forceinline__ double plaq_angle(const double* __restrict__ th, const Lat& l,
                                             uint32_t s, const int* c, int mu, int nu) {
  const uint32_t smu = fwd(l, s, c, mu), snu = fwd(l, s, c, nu);
  const size_t V = l.V;
  double p = th[mu * V + s];
  p = p + th[nu * V + smu];
  p = p - th[mu * V + snu];
  p = p - th[nu * V + s];
  return wrap_pi(p);
}

__global__ __launch_bounds__(kMonoBlock) void k_u1_plaquettes(const double* __restrict__ th,
                                                             Lat l, double* __restrict__ P) {
  const uint32_t s = xcd_block(blockIdx.x, gridDim.x) * kMonoBlock + threadIdx.x;
  if (s >= l.V) return;
  int c[4];
  coords_of(l, s, c);
#pragma unroll
  for (int mu = 0; mu < 3; mu++)
#pragma unroll
    for (int nu = mu + 1; nu < 4; nu++)
      P[(size_t)plane_of(mu, nu) * l.V + s] = plaq_angle(th, l, s, c, mu, nu);
}

// P3[colour][plane][site]
__global__ __launch_bounds__(kMonoBlock) void k_su3_plaquettes(const double* __restrict__ th3,
                                                              Lat l, double* __restrict__ P3) {
  const uint32_t s = xcd_block(blockIdx.x, gridDim.x) * kMonoBlock + threadIdx.x;
  if (s >= l.V) return;
  int c[4];
  coords_of(l, s, c);
  const size_t n = (size_t)6 * l.V;
#pragma unroll
  for (int mu = 0; mu < 3; mu++)
#pragma unroll
    for (int nu = mu + 1; nu < 4; nu++) {
      double q[3];
#pragma unroll
      for (int k = 0; k < 3; k++) q[k] = plaq_angle(th3 + (size_t)k * 4 * l.V, l, s, c, mu, nu);
      double sum = q[0] + q[1];
      sum = sum + q[2];
      if (sum >= kPi) {
        int at = 0;
        double ext = q[0];
        if (q[1] >= ext) {
          ext = q[1];
          at = 1;
        }
        if (q[2] >= ext) at = 2;
        q[at] -= 2 * kPi;
      }
      if (sum <= -kPi) {
        int at = 0;
        double ext = q[0];
        if (q[1] <= ext) {
          ext = q[1];
          at = 1;
        }
        if (q[2] <= ext) at = 2;
        q[at] += 2 * kPi;
      }
      const size_t o = (size_t)plane_of(mu, nu) * l.V + s;
#pragma unroll
      for (int k = 0; k < 3; k++) P3[k * n + o] = q[k];
    }
}

// ------------------------------------------------------------------ currents
// colour blockIdx.y: planes at P + 6 V colour, currents at J + 4 V colour
__global__ __launch_bounds__(kMonoBlock) void k_monopole_currents(const double* __restrict__ P,
                                                                 Lat l, double* __restrict__ J) {
  const uint32_t s = xcd_block(blockIdx.x, gridDim.x) * kMonoBlock + threadIdx.x;
  if (s >= l.V) return;
  P += (size_t)blockIdx.y * 6 * l.V;
  J += (size_t)blockIdx.y * 4 * l.V;
  int c[4];
  coords_of(l, s, c);
  const size_t V = l.V;
  auto pl = [&](int plane, uint32_t site) { return P[plane * V + site]; };
  // t: planes yz (3), xz (1), xy (0), at x + t
  uint32_t b = fwd(l, s, c, 3);
  const double j3 = pl(3, fwd2(l, s, c, 3, 0)) - pl(3, b) - (pl(1, fwd2(l, s, c, 3, 1)) - pl(1, b)) +
                    pl(0, fwd2(l, s, c, 3, 2)) - pl(0, b);
  // x: planes yz (3), zt (5), yt (4), at x + x
  b = fwd(l, s, c, 0);
  const double j0 = -(pl(3, fwd2(l, s, c, 0, 3)) - pl(3, b)) - (pl(5, fwd2(l, s, c, 0, 1)) - pl(5, b)) +
                    pl(4, fwd2(l, s, c, 0, 2)) - pl(4, b);
  // y: planes xz (1), zt (5), xt (2), at x + y
  b = fwd(l, s, c, 1);
  const double j1 = pl(1, fwd2(l, s, c, 1, 3)) - pl(1, b) + pl(5, fwd2(l, s, c, 1, 0)) - pl(5, b) -
                    (pl(2, fwd2(l, s, c, 1, 2)) - pl(2, b));
  // z: planes xy (0), yt (4), xt (2), at x + z
  b = fwd(l, s, c, 2);
  const double j2 = -(pl(0, fwd2(l, s, c, 2, 3)) - pl(0, b)) - (pl(4, fwd2(l, s, c, 2, 0)) - pl(4, b)) +
                    pl(2, fwd2(l, s, c, 2, 1)) - pl(2, b);
  double* o = J + 4 * (size_t)s;
  o[0] = j0 / 2 / kPi;
  o[1] = j1 / 2 / kPi;
  o[2] = j2 / 2 / kPi;
  o[3] = j3 / 2 / kPi;
}

int grid_of(uint32_t n) { return (int)((n + kMonoBlock - 1) / kMonoBlock); }

// currents from planar angles on the device (SU(2) 4V, SU(3) 12V) -> J
// ([4V] or [3][4V]); P: 6V (SU(2)) or 18V scratch
void launch_currents(int group, const double* th, const Lat& l, double* P, double* J,
                     hipStream_t st) {
  const int nb = grid_of(l.V);
  if (group == MAG_SU2) {
    hipLaunchKernelGGL(k_u1_plaquettes, dim3(nb), dim3(kMonoBlock), 0, st, th, l, P);
    hipLaunchKernelGGL(k_monopole_currents, dim3(nb), dim3(kMonoBlock), 0, st, P, l, J);
  } else {
    hipLaunchKernelGGL(k_su3_plaquettes, dim3(nb), dim3(kMonoBlock), 0, st, th, l, P);
    // the three colours in one launch (blockIdx.y): 24^3x48 angles +
    // plaquettes + currents 133 -> 125 us against three launches
    hipLaunchKernelGGL(k_monopole_currents, dim3(nb, 3), dim3(kMonoBlock), 0, st, P, l, J);
  }
}

// ------------------------------------------------------------------ clusters
// The reference finds a cluster by walking from its first site along the
// links that carry current out of each site reached (|J| > 0.3, charge
// lround(J); find_current / find_paths / find_cluster, monopoles.cpp:
// 1120-1250).  For conserved integer charges (in-flow = out-flow at every
// site, which the currents of k_monopole_currents satisfy) every link of a
// connected set of current links lies on a directed cycle, so that walk
// reaches exactly the connected component: the clusters are the connected
// components of the graph of current links, found here by a lock-free
// union-find over all links at once (roots always the smallest site, so a
// cluster's root is its first site in the reference's scan order).  The
// hook kernel checks the premise and flags any site where it fails.
//
// acc[k V + r] of root r, k = 0..6: length, length_mu[4], time length,
// sites (one array per counter, so that the counters of one root lie on
// different cache lines)
constexpr int kAcc = 7;
constexpr int kSumBlock = 1024;  // threads per workgroup of the sums kernel

__device__ __forceinline__ bool is_current(double j) { return j > 0.3 || j < -0.3; }
// lround: nearest integer, halves away from zero
__device__ __forceinline__ int charge_of(double j) { return (int)round(j); }

__device__ __forceinline__ uint32_t bwd(const Lat& l, uint32_t s, const int* c, int mu) {
  return c[mu] == 0 ? s + (uint32_t)(l.L[mu] - 1) * l.st[mu] : s - l.st[mu];
}
__device__ __forceinline__ uint32_t ld_par(uint32_t* par, uint32_t i) {
  return __hip_atomic_load(par + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// root of i.  parent[i] <= i always and roots point to themselves, so the
// walk ends within V steps (the bound raises the error flag otherwise).
__device__ uint32_t cc_find(uint32_t* par, uint32_t i, uint32_t bound, int* err) {
  for (uint32_t it = 0; it <= bound; it++) {
    const uint32_t p = ld_par(par, i);
    if (p == i) return i;
    const uint32_t gp = ld_par(par, p);
    // path halving: any ancestor is a valid parent, so a plain store does
    // (a store that loses a race leaves a longer but valid path)
    if (gp != p) __hip_atomic_store(par + i, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    i = gp;
  }
  atomicOr(err, 2);
  return i;
}
// join the components of a and b: the larger root is hooked under the
// smaller one; a failed hook means another thread hooked that root first
// (at most V - 1 hooks succeed in all, so the loop ends)
__device__ void cc_unite(uint32_t* par, uint32_t a, uint32_t b, uint32_t bound, int* err) {
  for (uint32_t it = 0; it <= bound; it++) {
    a = cc_find(par, a, bound, err);
    b = cc_find(par, b, bound, err);
    if (a == b) return;
    const uint32_t hi = a > b ? a : b, lo = a > b ? b : a;
    if (atomicCAS(par + hi, hi, lo) == hi) return;
  }
  atomicOr(err, 2);
}

// parent of s: the smallest of s and its neighbours across current links
// (already one union per site, without atomics)
__global__ __launch_bounds__(kMonoBlock) void k_cc_init(const double* __restrict__ J, Lat l,
                                                        uint32_t* __restrict__ par,
                                                        int* __restrict__ acc, int* __restrict__ meta) {
  const uint32_t s = xcd_block(blockIdx.x, gridDim.x) * kMonoBlock + threadIdx.x;
  if (s == 0) {
    meta[0] = 0;  // error flags
    meta[1] = 0;  // clusters
  }
  if (s >= l.V) return;
  int c[4];
  coords_of(l, s, c);
  uint32_t m = s;
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    const uint32_t f = fwd(l, s, c, mu), b = bwd(l, s, c, mu);
    if (is_current(J[4 * (size_t)s + mu]) && f < m) m = f;
    if (is_current(J[4 * (size_t)b + mu]) && b < m) m = b;
  }
  par[s] = m;
#pragma unroll
  for (int k = 0; k < kAcc; k++) acc[(size_t)k * l.V + s] = 0;
}

// err bit 0: a site where the charges do not balance, or a link with
// 0.3 < |J| < 0.5 (a current of charge 0)
__global__ __launch_bounds__(kMonoBlock) void k_cc_hook(const double* __restrict__ J, Lat l,
                                                        uint32_t* __restrict__ par, int* __restrict__ meta) {
  const uint32_t s = xcd_block(blockIdx.x, gridDim.x) * kMonoBlock + threadIdx.x;
  if (s >= l.V) return;
  int c[4];
  coords_of(l, s, c);
  int div = 0;
  bool bad = false;
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    const double jf = J[4 * (size_t)s + mu], jb = J[4 * (size_t)bwd(l, s, c, mu) + mu];
    if (is_current(jf)) {
      bad |= charge_of(jf) == 0;
      div += charge_of(jf);
    }
    if (is_current(jb)) div -= charge_of(jb);
  }
  if (bad || div != 0) atomicOr(meta, 1);
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    if (!is_current(J[4 * (size_t)s + mu])) continue;
    const uint32_t f = fwd(l, s, c, mu);
    // already joined when f hangs under s (usually from k_cc_init: s is the
    // smallest current neighbour of f).  A plain load may return an older
    // parent, but every parent f ever had is an ancestor of f.
    if (par[f] == s) continue;
    cc_unite(par, s, f, l.V, meta);
  }
}

__global__ __launch_bounds__(kMonoBlock) void k_cc_flatten(uint32_t V, uint32_t* __restrict__ par,
                                                           int* __restrict__ meta) {
  const uint32_t s = blockIdx.x * kMonoBlock + threadIdx.x;
  if (s >= V) return;
  par[s] = cc_find(par, s, V, meta);
}

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// per-root sums over the cluster's links (each link counted at its base
// site) and over its sites (sites with a current on one of their 8 links).
// Most current links usually belong to one large cluster, whose counters
// every thread would hit: the lanes of a wave that share the root of its
// first contributing lane add up in the wave, the waves of a workgroup
// that share a root add up in LDS, and that sum takes one atomic per
// counter; other lanes add their own.  Every lane takes part in the wave
// sums (lanes past V contribute nothing).
__global__ __launch_bounds__(kSumBlock) void k_cc_sums(const double* __restrict__ J, Lat l,
                                                       const uint32_t* __restrict__ par,
                                                       int* __restrict__ acc) {
  constexpr int NW = kSumBlock / 64;
  __shared__ uint32_t wr[NW];
  __shared__ int wt[NW][kAcc];
  const uint32_t s = xcd_block(blockIdx.x, gridDim.x) * kSumBlock + threadIdx.x;
  int v[kAcc] = {0, 0, 0, 0, 0, 0, 0};
  uint32_t r = 0;
  if (s < l.V) {
    int c[4];
    coords_of(l, s, c);
    r = par[s];
    bool site = false;
#pragma unroll
    for (int mu = 0; mu < 4; mu++) {
      const double jf = J[4 * (size_t)s + mu];
      site |= is_current(J[4 * (size_t)bwd(l, s, c, mu) + mu]);
      if (!is_current(jf)) continue;
      site = true;
      const int q = charge_of(jf), aq = q < 0 ? -q : q;
      v[0] += aq;
      v[1 + mu] += q;
      if (mu == 3) v[5] += aq;
    }
    v[6] = site ? 1 : 0;
  }
  const bool any = v[6] != 0;  // a site without a current link adds nothing
  const uint64_t mask = __ballot(any);
  const int w = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63u);
  const int lead = mask ? __builtin_ctzll(mask) : 0;
  const uint32_t r0 = __shfl(r, lead);
  const bool mine = any && r == r0;
#pragma unroll
  for (int k = 0; k < kAcc; k++) {
    const int t = wave_sum(mine ? v[k] : 0);
    if (lane == 0) wt[w][k] = t;
    if (any && !mine && v[k] != 0) atomicAdd(acc + (size_t)k * l.V + r, v[k]);
  }
  if (lane == 0) wr[w] = mask ? r0 : 0xffffffffu;
  __syncthreads();
  if (threadIdx.x < NW) {
    // wave w' adds the wave sums of the waves of its root, if it is the
    // first wave with that root
    const uint32_t rw = wr[threadIdx.x];
    bool first = rw != 0xffffffffu;
    for (int u = 0; u < (int)threadIdx.x && first; u++) first = wr[u] != rw;
    if (first) {
      int t[kAcc] = {0, 0, 0, 0, 0, 0, 0};
      for (int u = (int)threadIdx.x; u < NW; u++)
        if (wr[u] == rw)
#pragma unroll
          for (int k = 0; k < kAcc; k++) t[k] += wt[u][k];
#pragma unroll
      for (int k = 0; k < kAcc; k++)
        if (t[k] != 0) atomicAdd(acc + (size_t)k * l.V + rw, t[k]);
    }
  }
}

// one record per root: {first site, length, length_mu[4], space, time, sites}
__global__ __launch_bounds__(kMonoBlock) void k_cc_collect(uint32_t V, const uint32_t* __restrict__ par,
                                                           const int* __restrict__ acc,
                                                           int32_t* __restrict__ rec, int* __restrict__ meta) {
  const uint32_t s = blockIdx.x * kMonoBlock + threadIdx.x;
  if (s >= V || par[s] != s) return;
  int a[kAcc];
#pragma unroll
  for (int k = 0; k < kAcc; k++) a[k] = acc[(size_t)k * V + s];
  if (a[6] == 0) return;
  const int i = atomicAdd(meta + 1, 1);  // distinct roots: i < V
  int32_t* o = rec + (size_t)MAG_CLUSTER_INTS * i;
  o[0] = (int32_t)s;
  o[1] = a[0];
  o[2] = a[1];
  o[3] = a[2];
  o[4] = a[3];
  o[5] = a[4];
  o[6] = a[0] - a[5];
  o[7] = a[5];
  o[8] = a[6];
}

// device memory of one cluster call, freed on every exit path
struct CcBufs {
  uint32_t* par = nullptr;
  int* acc = nullptr;
  int32_t* rec = nullptr;
  int* meta = nullptr;
  ~CcBufs() {
    if (par) (void)hipFree(par);
    if (acc) (void)hipFree(acc);
    if (rec) (void)hipFree(rec);
    if (meta) (void)hipFree(meta);
  }
  int alloc(uint32_t V) {
    HIPCHK(hipMalloc((void**)&par, (size_t)V * sizeof(uint32_t)));
    HIPCHK(hipMalloc((void**)&acc, (size_t)kAcc * V * sizeof(int)));
    HIPCHK(hipMalloc((void**)&rec, (size_t)MAG_CLUSTER_INTS * V * sizeof(int32_t)));
    HIPCHK(hipMalloc((void**)&meta, 2 * sizeof(int)));
    return MAG_OK;
  }
};

// the five cluster kernels for the currents J ([4V]) of one colour
void launch_clusters(const double* J, const Lat& l, CcBufs& b, hipStream_t st) {
  const int nb = grid_of(l.V);
  hipLaunchKernelGGL(k_cc_init, dim3(nb), dim3(kMonoBlock), 0, st, J, l, b.par, b.acc, b.meta);
  hipLaunchKernelGGL(k_cc_hook, dim3(nb), dim3(kMonoBlock), 0, st, J, l, b.par, b.meta);
  hipLaunchKernelGGL(k_cc_flatten, dim3(nb), dim3(kMonoBlock), 0, st, l.V, b.par, b.meta);
  hipLaunchKernelGGL(k_cc_sums, dim3((l.V + kSumBlock - 1) / kSumBlock), dim3(kSumBlock), 0, st, J, l,
                     b.par, b.acc);
  hipLaunchKernelGGL(k_cc_collect, dim3(nb), dim3(kMonoBlock), 0, st, l.V, b.par, b.acc, b.rec, b.meta);
}

// clusters of ncol colours of device currents dJ ([ncol][4V]) -> caller's
// records (colour k at records + k * cap * MAG_CLUSTER_INTS), sorted by
// first site: the reference's order
int device_clusters(const double* dJ, const Lat& l, int ncol, int32_t* records, int64_t cap,
                    int64_t* n, hipStream_t st) {
  CcBufs b;
  CHK(b.alloc(l.V));
  bool short_cap = false;
  for (int k = 0; k < ncol; k++) {
    launch_clusters(dJ + (size_t)k * 4 * l.V, l, b, st);
    HIPCHK(hipGetLastError());
    int meta[2];
    HIPCHK(hipMemcpyAsync(meta, b.meta, sizeof(meta), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (meta[0] & 1)
      return fail(MAG_ERR_ARG, "colour %d: currents not conserved (charges lround(J) of the links "
                  "with |J| > 0.3 must balance at every site, none of them 0)", k);
    if (meta[0] & 2) return fail(MAG_ERR_HIP, "colour %d: cluster union-find did not finish", k);
    const int m = meta[1];
    std::vector<int32_t> h((size_t)m * MAG_CLUSTER_INTS);
    if (m > 0)
      HIPCHK(hipMemcpyAsync(h.data(), b.rec, h.size() * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::vector<int> order(m);
    for (int i = 0; i < m; i++) order[i] = i;
    std::sort(order.begin(), order.end(),
              [&](int x, int y) { return h[(size_t)MAG_CLUSTER_INTS * x] < h[(size_t)MAG_CLUSTER_INTS * y]; });
    int32_t* out = records + (size_t)k * cap * MAG_CLUSTER_INTS;
    for (int i = 0; i < m && i < cap; i++)
      std::copy_n(&h[(size_t)MAG_CLUSTER_INTS * order[i]], MAG_CLUSTER_INTS, out + (size_t)MAG_CLUSTER_INTS * i);
    n[k] = m;
    short_cap |= m > cap;
  }
  if (short_cap) return fail(MAG_ERR_ARG, "more clusters than the capacity %lld (counts returned in n)", (long long)cap);
  return MAG_OK;
}

// device buffers of one call, freed on every exit path
struct Bufs {
  double* p[4] = {nullptr, nullptr, nullptr, nullptr};
  ~Bufs() {
    for (double* q : p)
      if (q) (void)hipFree(q);
  }
};

int check_dims(const int* dims) {
  if (!dims) return fail(MAG_ERR_ARG, "NULL dims");
  uint64_t v = 1;
  for (int mu = 0; mu < 4; mu++) {
    if (dims[mu] < 2) return fail(MAG_ERR_DIMS, "lattice extent %d in direction %d (need >= 2)", dims[mu], mu);
    v *= (uint64_t)dims[mu];
  }
  if (v * 18 >= (1ull << 32)) return fail(MAG_ERR_DIMS, "lattice too large");
  return MAG_OK;
}

// host angles -> device, after checking that every angle is finite and
// within kMaxAngle
int upload_angles(const double* a, size_t n, double** d, hipStream_t st) {
  for (size_t i = 0; i < n; i++) {
    if (!std::isfinite(a[i])) return fail(MAG_ERR_NONFINITE, "angle %zu is %g", i, a[i]);
    if (std::fabs(a[i]) > kMaxAngle)
      return fail(MAG_ERR_ARG, "angle %zu is %g: |angle| must be <= %g", i, a[i], kMaxAngle);
  }
  HIPCHK(hipMalloc((void**)d, n * sizeof(double)));
  HIPCHK(hipMemcpyAsync(*d, a, n * sizeof(double), hipMemcpyHostToDevice, st));
  return MAG_OK;
}

int host_currents(int group, const int* dims, const double* angles, double* J) {
  CHK(check_dims(dims));
  if (!angles || !J) return fail(MAG_ERR_ARG, "NULL argument");
  const Lat l = make_lat(dims);
  const int ncol = group == MAG_SU2 ? 1 : 3;
  Bufs b;
  hipStream_t st = nullptr;  // the null stream: one synchronous call
  CHK(upload_angles(angles, (size_t)4 * l.V * ncol, &b.p[0], st));
  HIPCHK(hipMalloc((void**)&b.p[1], (size_t)6 * l.V * ncol * sizeof(double)));
  HIPCHK(hipMalloc((void**)&b.p[2], (size_t)4 * l.V * ncol * sizeof(double)));
  HIPCHK(hipMalloc((void**)&b.p[3], (size_t)4 * l.V * ncol * sizeof(double)));
  hipLaunchKernelGGL(k_angles_to_planar, dim3(grid_of(l.V)), dim3(kMonoBlock), 0, st, b.p[0], ncol,
                     l.V, b.p[3]);
  launch_currents(group, b.p[3], l, b.p[1], b.p[2], st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(J, b.p[2], (size_t)4 * l.V * ncol * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return MAG_OK;
}

// the job's angles into th (device), FP64 links of copy c taken from the
// job (link form) or formed in tmp (site form, j->nreal doubles)
int job_angles(mag_job* j, int c, double* th, double* tmp) {
  const bool su2 = j->prm.group == MAG_SU2;
  const int nb = grid_of((uint32_t)j->V);
  if (j->spin) {
    job_copy_links64(j, c, tmp);
    if (su2) hipLaunchKernelGGL((k_su2_angles<double>), dim3(nb), dim3(kMonoBlock), 0, j->st, tmp, j->g, th);
    else hipLaunchKernelGGL((k_su3_angles<double>), dim3(nb), dim3(kMonoBlock), 0, j->st, tmp, j->g, th);
  } else if (job_mixed32(j)) {
    if (su2) hipLaunchKernelGGL((k_su2_angles<float>), dim3(nb), dim3(kMonoBlock), 0, j->st, j->W32, j->g, th);
    else hipLaunchKernelGGL((k_su3_angles<float>), dim3(nb), dim3(kMonoBlock), 0, j->st, j->W32, j->g, th);
  } else {
    if (su2) hipLaunchKernelGGL((k_su2_angles<double>), dim3(nb), dim3(kMonoBlock), 0, j->st, j->U64, j->g, th);
    else hipLaunchKernelGGL((k_su3_angles<double>), dim3(nb), dim3(kMonoBlock), 0, j->st, j->U64, j->g, th);
  }
  HIPCHK(hipGetLastError());
  return MAG_OK;
}

int check_job(mag_job* j, int c) {
  if (!j) return fail(MAG_ERR_ARG, "NULL job");
  if (!j->loaded) return fail(MAG_ERR_STATE, "no configuration uploaded");
  if (c < 0 || c >= j->ncopies) return fail(MAG_ERR_ARG, "copy %d out of range 0..%d", c, j->ncopies - 1);
  HIPCHK(hipSetDevice(j->prm.device));
  return MAG_OK;
}

int lat_dims(const mag_job* j, int* d) {
  d[0] = j->g.L0;
  d[1] = j->g.L1;
  d[2] = j->g.L2;
  d[3] = j->g.L3;
  return MAG_OK;
}

}  // namespace
}  // namespace mag

using namespace mag;

extern "C" {

int mag_job_abelian_angles(mag_job* j, int copy, double* angles) {
  CHK(check_job(j, copy));
  if (!angles) return fail(MAG_ERR_ARG, "NULL angles");
  const size_t n = (size_t)4 * j->V * (j->prm.group == MAG_SU2 ? 1 : 3);
  Bufs b;
  HIPCHK(hipMalloc((void**)&b.p[0], n * sizeof(double)));
  HIPCHK(hipMalloc((void**)&b.p[2], n * sizeof(double)));
  if (j->spin) HIPCHK(hipMalloc((void**)&b.p[1], j->nreal * sizeof(double)));
  CHK(job_angles(j, copy, b.p[0], b.p[1]));
  hipLaunchKernelGGL(k_angles_to_ref, dim3(grid_of((uint32_t)j->V)), dim3(kMonoBlock), 0, j->st,
                     b.p[0], j->prm.group == MAG_SU2 ? 1 : 3, (uint32_t)j->V, b.p[2]);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(angles, b.p[2], n * sizeof(double), hipMemcpyDeviceToHost, j->st));
  HIPCHK(hipStreamSynchronize(j->st));
  return MAG_OK;
}

int mag_job_monopole_currents(mag_job* j, int copy, double* J) {
  CHK(check_job(j, copy));
  if (!J) return fail(MAG_ERR_ARG, "NULL J");
  int d[4];
  lat_dims(j, d);
  const Lat l = make_lat(d);
  const int ncol = j->prm.group == MAG_SU2 ? 1 : 3;
  Bufs b;
  HIPCHK(hipMalloc((void**)&b.p[0], (size_t)4 * l.V * ncol * sizeof(double)));
  HIPCHK(hipMalloc((void**)&b.p[1], (size_t)6 * l.V * ncol * sizeof(double)));
  HIPCHK(hipMalloc((void**)&b.p[2], (size_t)4 * l.V * ncol * sizeof(double)));
  if (j->spin) HIPCHK(hipMalloc((void**)&b.p[3], j->nreal * sizeof(double)));
  CHK(job_angles(j, copy, b.p[0], b.p[3]));
  launch_currents(j->prm.group, b.p[0], l, b.p[1], b.p[2], j->st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(J, b.p[2], (size_t)4 * l.V * ncol * sizeof(double), hipMemcpyDeviceToHost, j->st));
  HIPCHK(hipStreamSynchronize(j->st));
  return MAG_OK;
}

int mag_job_monopole_timed(mag_job* j, int reps, double* ms) {
  CHK(check_job(j, 0));
  if (reps < 1 || !ms) return fail(MAG_ERR_ARG, "bad argument");
  int d[4];
  lat_dims(j, d);
  const Lat l = make_lat(d);
  const int ncol = j->prm.group == MAG_SU2 ? 1 : 3;
  Bufs b;
  HIPCHK(hipMalloc((void**)&b.p[0], (size_t)4 * l.V * ncol * sizeof(double)));
  HIPCHK(hipMalloc((void**)&b.p[1], (size_t)6 * l.V * ncol * sizeof(double)));
  HIPCHK(hipMalloc((void**)&b.p[2], (size_t)4 * l.V * ncol * sizeof(double)));
  if (j->spin) HIPCHK(hipMalloc((void**)&b.p[3], j->nreal * sizeof(double)));
  CHK(job_angles(j, 0, b.p[0], b.p[3]));  // warm-up
  launch_currents(j->prm.group, b.p[0], l, b.p[1], b.p[2], j->st);
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  (void)hipEventRecord(e0, j->st);
  for (int i = 0; i < reps; i++) {
    (void)job_angles(j, 0, b.p[0], b.p[3]);
    launch_currents(j->prm.group, b.p[0], l, b.p[1], b.p[2], j->st);
  }
  (void)hipEventRecord(e1, j->st);
  hipError_t e = hipEventSynchronize(e1);
  float t = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return fail(MAG_ERR_HIP, "timing failed: %s", hipGetErrorString(e));
  *ms = (double)t / reps;
  return MAG_OK;
}

int mag_job_monopole_clusters(mag_job* j, int copy, int32_t* records, int64_t cap, int64_t* n) {
  CHK(check_job(j, copy));
  if ((!records && cap > 0) || cap < 0 || !n) return fail(MAG_ERR_ARG, "bad records / capacity / n");
  int d[4];
  lat_dims(j, d);
  const Lat l = make_lat(d);
  const int ncol = j->prm.group == MAG_SU2 ? 1 : 3;
  Bufs b;
  HIPCHK(hipMalloc((void**)&b.p[0], (size_t)4 * l.V * ncol * sizeof(double)));
  HIPCHK(hipMalloc((void**)&b.p[1], (size_t)6 * l.V * ncol * sizeof(double)));
  HIPCHK(hipMalloc((void**)&b.p[2], (size_t)4 * l.V * ncol * sizeof(double)));
  if (j->spin) HIPCHK(hipMalloc((void**)&b.p[3], j->nreal * sizeof(double)));
  CHK(job_angles(j, copy, b.p[0], b.p[3]));
  launch_currents(j->prm.group, b.p[0], l, b.p[1], b.p[2], j->st);
  HIPCHK(hipGetLastError());
  return device_clusters(b.p[2], l, ncol, records, cap, n, j->st);
}

int mag_job_clusters_timed(mag_job* j, int reps, double* ms) {
  CHK(check_job(j, 0));
  if (reps < 1 || !ms) return fail(MAG_ERR_ARG, "bad argument");
  int d[4];
  lat_dims(j, d);
  const Lat l = make_lat(d);
  const int ncol = j->prm.group == MAG_SU2 ? 1 : 3;
  Bufs b;
  HIPCHK(hipMalloc((void**)&b.p[0], (size_t)4 * l.V * ncol * sizeof(double)));
  HIPCHK(hipMalloc((void**)&b.p[1], (size_t)6 * l.V * ncol * sizeof(double)));
  HIPCHK(hipMalloc((void**)&b.p[2], (size_t)4 * l.V * ncol * sizeof(double)));
  if (j->spin) HIPCHK(hipMalloc((void**)&b.p[3], j->nreal * sizeof(double)));
  CHK(job_angles(j, 0, b.p[0], b.p[3]));
  launch_currents(j->prm.group, b.p[0], l, b.p[1], b.p[2], j->st);
  CcBufs cb;
  CHK(cb.alloc(l.V));
  for (int k = 0; k < ncol; k++) launch_clusters(b.p[2] + (size_t)k * 4 * l.V, l, cb, j->st);  // warm-up
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  (void)hipEventRecord(e0, j->st);
  for (int i = 0; i < reps; i++)
    for (int k = 0; k < ncol; k++) launch_clusters(b.p[2] + (size_t)k * 4 * l.V, l, cb, j->st);
  (void)hipEventRecord(e1, j->st);
  hipError_t e = hipEventSynchronize(e1);
  float t = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return fail(MAG_ERR_HIP, "timing failed: %s", hipGetErrorString(e));
  int meta[2];
  HIPCHK(hipMemcpy(meta, cb.meta, sizeof(meta), hipMemcpyDeviceToHost));
  if (meta[0]) return fail(MAG_ERR_HIP, "cluster kernels flagged %d", meta[0]);
  *ms = (double)t / reps;
  return MAG_OK;
}

int mag_current_clusters(const int* dims, int ncol, const double* J, int32_t* records, int64_t cap,
                         int64_t* n) {
  CHK(check_dims(dims));
  if (ncol != 1 && ncol != 3) return fail(MAG_ERR_ARG, "ncol %d (1 or 3)", ncol);
  if (!J || (!records && cap > 0) || cap < 0 || !n) return fail(MAG_ERR_ARG, "bad J / records / capacity / n");
  const Lat l = make_lat(dims);
  const size_t nj = (size_t)4 * l.V * ncol;
  for (size_t i = 0; i < nj; i++)
    if (!std::isfinite(J[i])) return fail(MAG_ERR_NONFINITE, "J[%zu] is %g", i, J[i]);
  Bufs b;
  hipStream_t st = nullptr;  // the null stream: one synchronous call
  HIPCHK(hipMalloc((void**)&b.p[0], nj * sizeof(double)));
  HIPCHK(hipMemcpyAsync(b.p[0], J, nj * sizeof(double), hipMemcpyHostToDevice, st));
  return device_clusters(b.p[0], l, ncol, records, cap, n, st);
}

int mag_u1_monopole_currents(const int* dims, const double* angles, double* J) {
  return host_currents(MAG_SU2, dims, angles, J);
}

int mag_su3_monopole_currents(const int* dims, const double* angles3, double* J3) {
  return host_currents(MAG_SU3, dims, angles3, J3);
}

}  // extern "C"
