// Wave64 / workgroup scan helpers shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace gala {

template <class V>
__device__ __forceinline__ V wave_incl_scan(V x) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 1; k < 64; k <<= 1) {
    V y = __shfl_up(x, k, 64);
    if (lane >= k) x += y;
  }
  return x;
}

template <class V>
__device__ __forceinline__ V wave_sum(V x) {
#pragma unroll
  for (int k = 32; k >= 1; k >>= 1) x += __shfl_xor(x, k, 64);
  return x;
}

// Exclusive workgroup scan for blockDim.x <= 1024; returns the prefix and
// *total = the workgroup sum.  sh needs 17 elements.  Contains barriers: every
// thread of the workgroup must call it.
template <class V>
__device__ __forceinline__ V block_excl_scan(V x, V* sh, V* total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
  V inc = wave_incl_scan(x);
  if (lane == 63) sh[wid] = inc;
  __syncthreads();
  if (wid == 0) {
    V s = lane < nw ? sh[lane] : V(0);
    V si = wave_incl_scan(s);
    if (lane < nw) sh[lane] = si - s;
    if (lane == nw - 1) sh[16] = si;
  }
  __syncthreads();
  V r = inc - x + sh[wid];
  *total = sh[16];
  __syncthreads();
  return r;
}

}  // namespace gala
