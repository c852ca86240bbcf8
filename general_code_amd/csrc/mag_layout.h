// mag_layout.h -- device-side access to the chunk-major parity-split SoA
// link layout.
//
// A link of NR reals of type T is cut into NC = NR / CW chunks of 16 bytes
// (CW = 16 / sizeof(T) reals per chunk).  The links of base parity p and
// direction mu (element a = p*4 + mu) form one field of Vh = V/2 links,
// stored in tiles of 64 consecutive half indices h; a tile holds chunk 0 of
// its 64 links, then chunk 1, ...:
//   chunk k of link (a, h) at  U + (a*NC*hp + (h/64)*NC*64 + k*64 + h%64) * CW
// (hp = Vh rounded up to whole tiles, plus a pad).  One 16-byte vector load
// of 64 lanes that own 64 consecutive h reads 1 KiB of contiguous memory:
// whole cache lines per instruction, whatever NR.  (With the reals of a link
// contiguous instead, an FP64 su3 link of 96 bytes takes 6 loads at 96-byte
// lane stride and the L2 evicts lines between them: measured +43% HBM
// reads.)  Against one array per chunk index (chunk k of all Vh links
// contiguous), the tiles cut the address streams of a half-sweep from
// 8*NC to 8 read-modify-write streams: SU(3) FP64 163 -> 155 us per
// 32^4 half-sweep, the others unchanged.
// Site arrays (the accumulated gauge of mixed precision) use the same rule
// with one field per parity (a = p).
#pragma once
#include <type_traits>

#include "mag_internal.h"
#include "mag_math.h"

namespace mag {

// chunk 0 of one link (or site) and the distance between its chunks
template <typename P>
struct LRef {
  P* p;
  size_t cs;
};

template <typename T>
__host__ __device__ constexpr int chunk_reals() {
  return 16 / (int)sizeof(T);
}

// element a = p*4 + mu (links) or p (sites) of a field of NC-chunk elements
template <int NC, int CW, typename P>
__device__ __forceinline__ LRef<P> chunk_ref(P* base, const Geom& g, uint32_t a, uint32_t h) {
#if MAG_TILE_LAYOUT
  return LRef<P>{base + ((size_t)a * NC * g.hp + (h / kTile) * (NC * kTile) + h % kTile) * CW,
                 (size_t)kTile * CW};
#else
  return LRef<P>{base + ((size_t)a * NC * g.hp + h) * CW, (size_t)g.hp * CW};
#endif
}
template <int NR, typename P>
__device__ __forceinline__ LRef<P> link_ptr(P* U, const Geom& g, int p, int mu, uint32_t h) {
  using T = std::remove_const_t<P>;
  constexpr int CW = chunk_reals<T>(), NC = NR / CW;
  return chunk_ref<NC, CW>(U, g, (uint32_t)(p * 4 + mu), h);
}
// pm = p*4 + mu in [0, 8)
template <int NR, typename P>
__device__ __forceinline__ LRef<P> link_ptr_pm(P* U, const Geom& g, uint32_t pm, uint32_t h) {
  using T = std::remove_const_t<P>;
  constexpr int CW = chunk_reals<T>(), NC = NR / CW;
  return chunk_ref<NC, CW>(U, g, pm, h);
}
template <int NR, typename P>
__device__ __forceinline__ LRef<P> site_ptr(P* G, const Geom& g, int p, uint32_t h) {
  using T = std::remove_const_t<P>;
  constexpr int CW = chunk_reals<T>(), NC = NR / CW;
  return chunk_ref<NC, CW>(G, g, (uint32_t)p, h);
}

// ------------------------------------------------------------------ su2 links
template <typename P>
__device__ __forceinline__ auto ldq(LRef<P> r) {
  using T = std::remove_const_t<P>;
  if constexpr (sizeof(T) == 4) {
    const float4 v = *reinterpret_cast<const float4*>(r.p);
    return Q<float>{v.x, v.y, v.z, v.w};
  } else {
    const double2 a = *reinterpret_cast<const double2*>(r.p);
    const double2 b = *reinterpret_cast<const double2*>(r.p + r.cs);
    return Q<double>{a.x, a.y, b.x, b.y};
  }
}
__device__ __forceinline__ void stq(LRef<float> r, const Q<float>& q) {
  *reinterpret_cast<float4*>(r.p) = make_float4(q.a0, q.a1, q.a2, q.a3);
}
__device__ __forceinline__ void stq(LRef<double> r, const Q<double>& q) {
  *reinterpret_cast<double2*>(r.p) = make_double2(q.a0, q.a1);
  *reinterpret_cast<double2*>(r.p + r.cs) = make_double2(q.a2, q.a3);
}

// ------------------------------------------------------------------ su3 links
// rows 0,1: (re,im) of (0,0) (0,1) (0,2) (1,0) (1,1) (1,2)
template <typename P>
__device__ __forceinline__ auto ldm(LRef<P> r) {
  using T = std::remove_const_t<P>;
  M2<T> m;
  if constexpr (sizeof(T) == 4) {
    const float4 a = *reinterpret_cast<const float4*>(r.p);
    const float4 b = *reinterpret_cast<const float4*>(r.p + r.cs);
    const float4 c = *reinterpret_cast<const float4*>(r.p + 2 * r.cs);
    m.r[0][0] = {a.x, a.y}; m.r[0][1] = {a.z, a.w}; m.r[0][2] = {b.x, b.y};
    m.r[1][0] = {b.z, b.w}; m.r[1][1] = {c.x, c.y}; m.r[1][2] = {c.z, c.w};
  } else {
#pragma unroll
    for (int k = 0; k < 6; k++) {
      const double2 d = *reinterpret_cast<const double2*>(r.p + k * r.cs);
      m.r[k / 3][k % 3] = {d.x, d.y};
    }
  }
  return m;
}
__device__ __forceinline__ void stm(LRef<float> r, const M2<float>& m) {
  *reinterpret_cast<float4*>(r.p) = make_float4(m.r[0][0].re, m.r[0][0].im, m.r[0][1].re, m.r[0][1].im);
  *reinterpret_cast<float4*>(r.p + r.cs) =
      make_float4(m.r[0][2].re, m.r[0][2].im, m.r[1][0].re, m.r[1][0].im);
  *reinterpret_cast<float4*>(r.p + 2 * r.cs) =
      make_float4(m.r[1][1].re, m.r[1][1].im, m.r[1][2].re, m.r[1][2].im);
}
__device__ __forceinline__ void stm(LRef<double> r, const M2<double>& m) {
#pragma unroll
  for (int k = 0; k < 6; k++)
    *reinterpret_cast<double2*>(r.p + k * r.cs) = make_double2(m.r[k / 3][k % 3].re, m.r[k / 3][k % 3].im);
}

// The same store with the nontemporal hint (streamed past the caches).
// Used where it measured faster: the FP64 SU(3) link half-sweep, whose
// 768 MiB per 32^4 half-sweep overflows the 256 MB MALL (153.9 -> 129.4 us
// per 32^4 half-sweep; 16^3x32 21.7 -> 20.4 us; 24^3x48 84.0 -> 85.5 us).
// For the SU(2) and mixed sweeps ordinary stores measured as fast or faster.
typedef double mag_d2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void stm_stream(LRef<double> r, const M2<double>& m) {
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const mag_d2v v = {m.r[k / 3][k % 3].re, m.r[k / 3][k % 3].im};
    __builtin_nontemporal_store(v, reinterpret_cast<mag_d2v*>(r.p + k * r.cs));
  }
}

__device__ __forceinline__ void stm_stream(LRef<float> r, const M2<float>& m) {
  typedef float f4v __attribute__((ext_vector_type(4)));
  const f4v a = {m.r[0][0].re, m.r[0][0].im, m.r[0][1].re, m.r[0][1].im};
  const f4v b = {m.r[0][2].re, m.r[0][2].im, m.r[1][0].re, m.r[1][0].im};
  const f4v c = {m.r[1][1].re, m.r[1][1].im, m.r[1][2].re, m.r[1][2].im};
  __builtin_nontemporal_store(a, reinterpret_cast<f4v*>(r.p));
  __builtin_nontemporal_store(b, reinterpret_cast<f4v*>(r.p + r.cs));
  __builtin_nontemporal_store(c, reinterpret_cast<f4v*>(r.p + 2 * r.cs));
}

// The same load with the nontemporal hint: the links of an SU(3) site-form
// set of several copies, each read once per half-sweep while the copies'
// neighbour gauges are read by up to 8 sites each; marked for early
// eviction, the links leave the caches to the gauges.
typedef float mag_f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ M2<float> ldm_nt(LRef<const float> r) {
  const mag_f4v a = __builtin_nontemporal_load(reinterpret_cast<const mag_f4v*>(r.p));
  const mag_f4v b = __builtin_nontemporal_load(reinterpret_cast<const mag_f4v*>(r.p + r.cs));
  const mag_f4v c = __builtin_nontemporal_load(reinterpret_cast<const mag_f4v*>(r.p + 2 * r.cs));
  M2<float> m;
  m.r[0][0] = {a.x, a.y}; m.r[0][1] = {a.z, a.w}; m.r[0][2] = {b.x, b.y};
  m.r[1][0] = {b.z, b.w}; m.r[1][1] = {c.x, c.y}; m.r[1][2] = {c.z, c.w};
  return m;
}

// su2 links with the nontemporal hint
__device__ __forceinline__ Q<float> ldq_nt(LRef<const float> r) {
  const mag_f4v v = __builtin_nontemporal_load(reinterpret_cast<const mag_f4v*>(r.p));
  return Q<float>{v.x, v.y, v.z, v.w};
}
__device__ __forceinline__ Q<double> ldq_nt(LRef<const double> r) {
  const mag_d2v a = __builtin_nontemporal_load(reinterpret_cast<const mag_d2v*>(r.p));
  const mag_d2v b = __builtin_nontemporal_load(reinterpret_cast<const mag_d2v*>(r.p + r.cs));
  return Q<double>{a.x, a.y, b.x, b.y};
}
__device__ __forceinline__ void stq_nt(LRef<float> r, const Q<float>& q) {
  const mag_f4v v = {q.a0, q.a1, q.a2, q.a3};
  __builtin_nontemporal_store(v, reinterpret_cast<mag_f4v*>(r.p));
}
__device__ __forceinline__ void stq_nt(LRef<double> r, const Q<double>& q) {
  const mag_d2v a = {q.a0, q.a1}, b = {q.a2, q.a3};
  __builtin_nontemporal_store(a, reinterpret_cast<mag_d2v*>(r.p));
  __builtin_nontemporal_store(b, reinterpret_cast<mag_d2v*>(r.p + r.cs));
}
__device__ __forceinline__ M2<double> ldm_nt(LRef<const double> r) {
  M2<double> m;
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const mag_d2v d = __builtin_nontemporal_load(reinterpret_cast<const mag_d2v*>(r.p + k * r.cs));
    m.r[k / 3][k % 3] = {d.x, d.y};
  }
  return m;
}

// The site-form kernels read every link once per half-sweep and the
// neighbours' site variables up to 8 times.  Where the links of both
// parities (8 Vh links of link_bytes) exceed the 256 MB MALL, the next
// half-sweep reads them from HBM whatever the hint, and loading them with
// the nontemporal hint leaves the caches to the site variables: per
// half-sweep SU(2) 48^4 FP64 177 -> 158 us, FP32 87 -> 83 us, SU(3) FP64
// 32^4 114 -> 105 us, SU(3) FP32 48^3x24 152 -> 126 us.  Where they fit, the
// plain load lets the next half-sweep find them in the MALL (the hint cost
// 10% at SU(3) FP32 32^4).
constexpr size_t kMallBytes = (size_t)256 << 20;
__host__ __device__ inline bool stream_links(const Geom& g, size_t link_bytes) {
  return (size_t)g.Vh * 8 * link_bytes > kMallBytes;
}

// ------------------------------------------------------------------ geometry
struct SiteIdx {
  uint32_t x, y, z, t, row;
};

__device__ __forceinline__ SiteIdx decode(const Geom& g, uint32_t h, int p) {
  SiteIdx s;
  const uint32_t xh = h % g.hx;
  s.row = h / g.hx;
  uint32_t r = s.row;
  s.y = r % (uint32_t)g.L1;
  r /= (uint32_t)g.L1;
  s.z = r % (uint32_t)g.L2;
  s.t = r / (uint32_t)g.L2;
  s.x = 2 * xh + ((uint32_t)(p + s.y + s.z + s.t) & 1u);
  return s;
}

// half index of x - mu (opposite parity)
__device__ __forceinline__ uint32_t back_h(const Geom& g, uint32_t h, const SiteIdx& s, int mu) {
  switch (mu) {
    case 0: {
      const uint32_t xb = (s.x == 0) ? (uint32_t)g.L0 - 1 : s.x - 1;
      return s.row * g.hx + (xb >> 1);
    }
    case 1: return (s.y == 0) ? h + ((uint32_t)g.L1 - 1) * g.s1 : h - g.s1;
    case 2: return (s.z == 0) ? h + ((uint32_t)g.L2 - 1) * g.s2 : h - g.s2;
    default: return (s.t == 0) ? h + ((uint32_t)g.L3 - 1) * g.s3 : h - g.s3;
  }
}
// half index of x + mu (opposite parity)
__device__ __forceinline__ uint32_t fwd_h(const Geom& g, uint32_t h, const SiteIdx& s, int mu) {
  switch (mu) {
    case 0: {
      const uint32_t xf = (s.x + 1 == (uint32_t)g.L0) ? 0 : s.x + 1;
      return s.row * g.hx + (xf >> 1);
    }
    case 1: return (s.y + 1 == (uint32_t)g.L1) ? h - ((uint32_t)g.L1 - 1) * g.s1 : h + g.s1;
    case 2: return (s.z + 1 == (uint32_t)g.L2) ? h - ((uint32_t)g.L2 - 1) * g.s2 : h + g.s2;
    default: return (s.t + 1 == (uint32_t)g.L3) ? h - ((uint32_t)g.L3 - 1) * g.s3 : h + g.s3;
  }
}

__device__ __forceinline__ void write_part(double* part, double a, double b, double c) {
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x + 0] = a;
    part[3 * blockIdx.x + 1] = b;
    part[3 * blockIdx.x + 2] = c;
  }
}

// Workgroups are dispatched round-robin over the 8 XCDs (blockIdx % 8),
// each with its own L2.  For kernels that read neighbour data written or
// read by other workgroups, workgroup b processes the logical block that
// gives XCD b % 8 one contiguous range of blocks: neighbouring sites then
// mostly run on the same XCD and share its L2.  A bijection of [0, n); it
// only reorders which workgroup computes which block.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t n) {
  const uint32_t q = n / 8, r = n % 8, x = b % 8, i = b / 8;
  return x * q + (x < r ? x : r) + i;
}

// lexicographic site s (reference order, indexing.h:111-123) -> parity p
// and half index h of the parity-split SoA
__device__ __forceinline__ void lex_site(const Geom& g, uint32_t s, int& p, uint32_t& h) {
  uint32_t x = s % (uint32_t)g.L0, r = s / (uint32_t)g.L0;
  uint32_t y = r % (uint32_t)g.L1;
  r /= (uint32_t)g.L1;
  uint32_t z = r % (uint32_t)g.L2, t = r / (uint32_t)g.L2;
  p = (int)((x + y + z + t) & 1u);
  h = s >> 1;
}

// sum_j |U_jj|^2 of an su3 link (mag.cpp:1837-1848), in FP64;
// U_22 = conj(U00 U11 - U01 U10)
template <typename T>
__device__ __forceinline__ double su3_link_F(const M2<T>& m) {
  C<double> a{(double)m.r[0][0].re, (double)m.r[0][0].im};
  C<double> b{(double)m.r[1][1].re, (double)m.r[1][1].im};
  C<double> c{(double)m.r[0][1].re, (double)m.r[0][1].im};
  C<double> d{(double)m.r[1][0].re, (double)m.r[1][0].im};
  C<double> u22 = csub(cmul(a, b), cmul(c, d));
  return cnorm2(a) + cnorm2(b) + cnorm2(u22);
}

}  // namespace mag
