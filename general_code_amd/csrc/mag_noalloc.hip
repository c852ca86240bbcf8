// mag_noalloc.hip -- the MAG functional of device-resident links with no
// device allocation: mag_functional_noalloc (include/mag_gpu.h).
//
// For callers that cannot allocate (code under HIP graph capture, tight
// serving loops): the same functional as mag_su2_functional /
// mag_su3_functional (MAG_functional_su2, mag.cpp:993-1001;
// mag_functional_su3, mag.cpp:1837-1862), from links already in device
// memory in the reference layout, with its partial sums in a file-scope
// device array instead of buffers the library allocates.
//
// The partial-sum array `scratch` holds kNoAllocParts doubles per device
// (64 KiB, resident until the process ends).  Its device address is taken
// once per device with hipGetSymbolAddress; host code only ever passes that
// address, never the array's name.  Every call on a device shares the
// array: calls that overlap on one device (other streams, other host
// threads) race on it.  Nothing in it has to start at zero: the first
// kernel writes every slot the second one reads.
#include "mag_capi_util.h"
#include "mag_gpu.h"
#include "mag_reduce.h"

#include <mutex>

namespace {

constexpr int kNoAllocThreads = 256;                        // first kernel's workgroup
constexpr int kNoAllocLinksPerThread = 32;
constexpr int kNoAllocLinksPerBlock = kNoAllocThreads * kNoAllocLinksPerThread;
constexpr int kNoAllocParts = 8192;                         // partial sums (one per workgroup)
constexpr int kNoAllocReduce = 1024;                        // second kernel's workgroup
constexpr int64_t kNoAllocMaxLinks = (int64_t)kNoAllocParts * kNoAllocLinksPerBlock;
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) double scratch[kNoAllocParts];

std::mutex g_mu;
double* g_scratch[kNoAllocMaxDevices] = {};

// device address of `scratch` on the current device (taken on first use)
int scratch_ptr(int dev, double** out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return mag::fail(MAG_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", dev,
                       hipGetErrorString(e));
    g_scratch[dev] = static_cast<double*>(p);
  }
  *out = g_scratch[dev];
  return MAG_OK;
}

// per-link functional in the reference's terms:
//   SU(2) a0^2 - a1^2 - a2^2 + a3^2 = tr(sigma3 U sigma3 U^+)/2
//   SU(3) sum_j |U_jj|^2 of the stored 3x3 matrix (row-major re, im)
template <int NR>
__device__ __forceinline__ double link_functional(const double* __restrict__ u) {
  if constexpr (NR == 4) {
    const double a0 = u[0], a1 = u[1], a2 = u[2], a3 = u[3];
    return a0 * a0 - a1 * a1 - a2 * a2 + a3 * a3;
  } else {
    double f = 0.0;
#pragma unroll
    for (int j = 0; j < 3; j++) {
      const double re = u[2 * (4 * j)], im = u[2 * (4 * j) + 1];
      f += re * re + im * im;
    }
    return f;
  }
}

// workgroup b sums links b*kNoAllocLinksPerBlock + k*256 + t (k ascending)
// per thread, then the fixed wave/workgroup tree; one partial per workgroup
template <int NR>
__global__ __launch_bounds__(kNoAllocThreads) void k_noalloc_partials(
    const double* __restrict__ links, int64_t nlinks, double* __restrict__ part) {
  const int64_t base = (int64_t)blockIdx.x * kNoAllocLinksPerBlock + threadIdx.x;
  double f = 0.0;
#pragma unroll 8
  for (int k = 0; k < kNoAllocLinksPerThread; k++) {
    const int64_t i = base + (int64_t)k * kNoAllocThreads;
    if (i < nlinks) f += link_functional<NR>(links + i * NR);
  }
  double b = 0.0, c = 0.0;
  mag::block_reduce3<kNoAllocThreads / 64>(f, b, c);
  if (threadIdx.x == 0) part[blockIdx.x] = f;
}

// one workgroup: thread t sums partials t, t + 1024, ... in order, then the
// fixed tree; out[0] = sum * inv
__global__ __launch_bounds__(kNoAllocReduce) void k_noalloc_reduce(
    const double* __restrict__ part, int nparts, double inv, double* __restrict__ out) {
  double a = 0.0;
  for (int i = threadIdx.x; i < nparts; i += kNoAllocReduce) a += part[i];
  double b = 0.0, c = 0.0;
  mag::block_reduce3<kNoAllocReduce / 64>(a, b, c);
  if (threadIdx.x == 0) out[0] = a * inv;
}

}  // namespace

extern "C" int mag_functional_noalloc(int group, const int* dims, const double* dev_links,
                                      double* dev_F, void* stream) {
  if (group != MAG_SU2 && group != MAG_SU3) return mag::fail(MAG_ERR_ARG, "group must be 2 or 3");
  if (!dims || !dev_links || !dev_F) return mag::fail(MAG_ERR_ARG, "NULL argument");
  int64_t V = 1;
  for (int i = 0; i < 4; i++) {
    if (dims[i] < 2 || (dims[i] & 1))
      return mag::fail(MAG_ERR_DIMS, "extent %d of dimension %d must be even and >= 2", dims[i], i);
    V *= dims[i];
    if (V > kNoAllocMaxLinks) break;
  }
  const int64_t nlinks = 4 * V;
  if (nlinks > kNoAllocMaxLinks)
    return mag::fail(MAG_ERR_DIMS,
                     "lattice too large for mag_functional_noalloc: at most %lld sites "
                     "(%d partial sums of %d links)",
                     (long long)(kNoAllocMaxLinks / 4), kNoAllocParts, kNoAllocLinksPerBlock);
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices)
    return mag::fail(MAG_ERR_ARG, "device %d beyond the %d supported", dev, kNoAllocMaxDevices);
  double* part = nullptr;
  CHK(scratch_ptr(dev, &part));
  const int nparts = (int)((nlinks + kNoAllocLinksPerBlock - 1) / kNoAllocLinksPerBlock);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (group == MAG_SU2)
    hipLaunchKernelGGL(k_noalloc_partials<4>, dim3(nparts), dim3(kNoAllocThreads), 0, st,
                       dev_links, nlinks, part);
  else
    hipLaunchKernelGGL(k_noalloc_partials<18>, dim3(nparts), dim3(kNoAllocThreads), 0, st,
                       dev_links, nlinks, part);
  const double inv = 1.0 / (double)(nlinks * (group == MAG_SU2 ? 1 : 3));
  hipLaunchKernelGGL(k_noalloc_reduce, dim3(1), dim3(kNoAllocReduce), 0, st, part, nparts, inv,
                     dev_F);
  HIPCHK(hipGetLastError());
  return MAG_OK;
}
