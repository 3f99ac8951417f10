// Wave64 / workgroup scan helpers shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace gala {

// atomicAdd for signed 64-bit sums (two's complement: the unsigned add is the
// same operation); the runtime's own overloads stay visible beside them.
using ::atomicAdd;
__device__ __forceinline__ long atomicAdd(long* p, long x) {
  return (long)::atomicAdd((unsigned long long*)p, (unsigned long long)x);
}
__device__ __forceinline__ long long atomicAdd(long long* p, long long x) {
  return (long long)::atomicAdd((unsigned long long*)p, (unsigned long long)x);
}

// Lane shuffles of 128-bit gains (two 64-bit halves); the runtime's overloads
// cover the narrower types.
using ::__shfl_xor;
__device__ __forceinline__ __int128 __shfl_xor(__int128 x, int k, int w) {
  const long long lo = ::__shfl_xor((long long)(unsigned long long)x, k, w);
  const long long hi = ::__shfl_xor((long long)(x >> 64), k, w);
  return ((__int128)hi << 64) | (__int128)(unsigned long long)lo;
}
using ::__shfl;
__device__ __forceinline__ __int128 __shfl(__int128 x, int src, int w) {
  const long long lo = ::__shfl((long long)(unsigned long long)x, src, w);
  const long long hi = ::__shfl((long long)(x >> 64), src, w);
  return ((__int128)hi << 64) | (__int128)(unsigned long long)lo;
}
using ::__shfl_up;
using ::__shfl_down;

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

// Equal keys within a wave, combined before an LDS atomic insert: while the
// first pending lane's key is held by at least two pending lanes, their values
// are summed across the wave and that lane inserts the sum once; the lanes
// left (keys not repeated in the wave) insert their own values.  Many lanes
// adding into one LDS slot otherwise serialise on that address (a community
// whose row entries almost all carry a few keys).  Every lane of the wave
// must call it; insert(key, value) is called for valid lanes only.
template <class V, class F>
__device__ __forceinline__ void wave_combine_insert(int key, V w, bool valid, F&& insert) {
  const int lane = threadIdx.x & 63;
  bool pending = valid;
  for (int round = 0; round < 64; round++) {
    const unsigned long long m = __ballot(pending);
    if (!m) return;
    const int src = __ffsll((unsigned long long)m) - 1;
    const int lk = __shfl(key, src, 64);
    const bool mine = pending && key == lk;
    if (__popcll(__ballot(mine)) < 2) break;  // the first pending key is not repeated
    const V s = wave_sum<V>(mine ? w : V(0));
    if (lane == src) insert(lk, s);
    pending = pending && !mine;
  }
  if (pending) insert(key, w);
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

namespace gala {

// Bitonic sort, ascending by key, of NT*E (key, value) pairs held in
// registers: thread t holds positions t*E .. t*E+E-1.  Compare-exchanges
// between positions of one thread stay in registers, partners in another lane
// of the same wave are reached with lane shuffles, and partners in another
// wave (NT > 64) through xk/xv: NT*E elements of LDS each, element-major so a
// wave's accesses are contiguous.  Every thread of the workgroup must call it
// (it has barriers when NT > 64); on return xk/xv are free again.
template <int NT, int E, class V>
__device__ __forceinline__ void reg_bitonic(int (&k)[E], V (&v)[E], int* xk, V* xv) {
  const int t = threadIdx.x;
#pragma unroll
  for (int kk = 2; kk <= NT * E; kk <<= 1) {
#pragma unroll
    for (int j = kk >> 1; j > 0; j >>= 1) {
      if (j < E) {
#pragma unroll
        for (int e = 0; e < E; e++) {
          const int l = e ^ j;
          if (l > e) {
            const bool up = ((t * E + e) & kk) == 0;
            if ((k[e] > k[l]) == up) {
              const int tk = k[e];
              k[e] = k[l];
              k[l] = tk;
              const V tv = v[e];
              v[e] = v[l];
              v[l] = tv;
            }
          }
        }
      } else {
        const int m = j / E;  // partner thread: t ^ m
        const bool keep_min = ((t & m) == 0) == (((t * E) & kk) == 0);
        if (m < 64) {
#pragma unroll
          for (int e = 0; e < E; e++) {
            const int ok = __shfl_xor(k[e], m, 64);
            const V ov = __shfl_xor(v[e], m, 64);
            if (keep_min ? ok < k[e] : ok > k[e]) { k[e] = ok; v[e] = ov; }
          }
        } else {
          __syncthreads();
#pragma unroll
          for (int e = 0; e < E; e++) { xk[e * NT + t] = k[e]; xv[e * NT + t] = v[e]; }
          __syncthreads();
          const int p = t ^ m;
#pragma unroll
          for (int e = 0; e < E; e++) {
            const int ok = xk[e * NT + p];
            const V ov = xv[e * NT + p];
            if (keep_min ? ok < k[e] : ok > k[e]) { k[e] = ok; v[e] = ov; }
          }
        }
      }
    }
  }
  if (NT > 64) __syncthreads();
}

}  // namespace gala
