// mag_reduce.h -- fixed-order wave64 / workgroup reductions.
//
// Every lane ends with the same value and the combination order is fixed,
// so a per-job reduction gives bit-identical results however the job is
// scheduled (alone, batched, or in the serving loop).
#pragma once
#include <hip/hip_runtime.h>

namespace mag {

// DPP controls (GCN3+/CDNA encoding)
constexpr int DPP_QUAD_XOR1 = 0xB1;     // quad_perm [1,0,3,2]
constexpr int DPP_QUAD_XOR2 = 0x4E;     // quad_perm [2,3,0,1]
constexpr int DPP_ROW_HALF_MIRROR = 0x141;
constexpr int DPP_ROW_MIRROR = 0x140;

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  int2 b = __builtin_bit_cast(int2, v);
  b.x = __builtin_amdgcn_update_dpp(0, b.x, CTRL, 0xF, 0xF, true);
  b.y = __builtin_amdgcn_update_dpp(0, b.y, CTRL, 0xF, 0xF, true);
  return __builtin_bit_cast(double, b);
}

__device__ __forceinline__ double shfl_xor_f64(double v, int m) {
  return __shfl_xor(v, m, 64);
}

// Butterfly sum over the 64 lanes: DPP within rows of 16, then two
// cross-row exchanges.  Requires all 64 lanes active.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_f64<DPP_QUAD_XOR1>(v);
  v += dpp_f64<DPP_QUAD_XOR2>(v);
  v += dpp_f64<DPP_ROW_HALF_MIRROR>(v);
  v += dpp_f64<DPP_ROW_MIRROR>(v);
  v += shfl_xor_f64(v, 16);
  v += shfl_xor_f64(v, 32);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
  v = fmax(v, dpp_f64<DPP_QUAD_XOR1>(v));
  v = fmax(v, dpp_f64<DPP_QUAD_XOR2>(v));
  v = fmax(v, dpp_f64<DPP_ROW_HALF_MIRROR>(v));
  v = fmax(v, dpp_f64<DPP_ROW_MIRROR>(v));
  v = fmax(v, shfl_xor_f64(v, 16));
  v = fmax(v, shfl_xor_f64(v, 32));
  return v;
}

// Workgroup reduction of (sum a, sum b, max c); result valid in thread 0.
// NW = number of waves in the workgroup.
template <int NW>
__device__ __forceinline__ void block_reduce3(double& a, double& b, double& c) {
  __shared__ double sa[NW], sb[NW], sc[NW];
  a = wave_sum(a);
  b = wave_sum(b);
  c = wave_max(c);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    sa[w] = a;
    sb[w] = b;
    sc[w] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    a = sa[0];
    b = sb[0];
    c = sc[0];
    for (int i = 1; i < NW; i++) {
      a += sa[i];
      b += sb[i];
      c = fmax(c, sc[i]);
    }
  }
}

// K independent triples with one barrier; for each triple the operations
// and their order are those of block_reduce3, so the results are identical.
template <int NW, int K>
__device__ __forceinline__ void block_reduce3k(double* a, double* b, double* c) {
  __shared__ double s[3 * K][NW];
#pragma unroll
  for (int k = 0; k < K; k++) {
    a[k] = wave_sum(a[k]);
    b[k] = wave_sum(b[k]);
    c[k] = wave_max(c[k]);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) {
      s[3 * k][w] = a[k];
      s[3 * k + 1][w] = b[k];
      s[3 * k + 2][w] = c[k];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) {
      a[k] = s[3 * k][0];
      b[k] = s[3 * k + 1][0];
      c[k] = s[3 * k + 2][0];
      for (int i = 1; i < NW; i++) {
        a[k] += s[3 * k][i];
        b[k] += s[3 * k + 1][i];
        c[k] = fmax(c[k], s[3 * k + 2][i]);
      }
    }
  }
}

}  // namespace mag
