// Phase-2 (graph contraction) kernels for gfx950.
//
// Follows build_compressed_graph (LinXi-lx/GALA-mg
// src/louvain_gpu/build_compressed_graph.cu:5-173): communities are renumbered
// by the rank of the non-empty ones (:18-24), member rows are merged per new
// community with the neighbour key new_id(comm(u)) and summed weights, and the
// internal weight becomes a self-loop.  Unlike the reference (atomic gather
// and hash insertion order, :1101-1111, :1264) every output row is canonical:
// ascending neighbour id, the order preprocess's std::map produces.
//
//   rows with <= 4096 staged entries: bitonic sort of (key, weight) in LDS,
//     then reduce-by-key;
//   longer rows: one workgroup, a global open-addressing table per row
//     (sized next_pow2(2*min(L, C)), no drops), then an ordered extraction
//     through a C-bit bitmap (LDS when C <= 2^20, else a global slice).
#include "gala_kernels.h"
#include <climits>

namespace gala {

constexpr int SMALL_A = 512;    // one 64-thread workgroup
constexpr int SMALL_B = 4096;   // one 256-thread workgroup
constexpr int BITMAP_LDS_WORDS = 32768;  // 2^20 bits = 128 KB

__device__ __forceinline__ unsigned chash(int key, int lg) {
  return ((unsigned)key * 2654435769u) >> (32 - lg);
}
__device__ __forceinline__ int clog2(long long x) {
  int l = 0;
  while ((1ll << l) < x) l++;
  return l;
}

// community sizes
__global__ void k_hist(const int32_t* __restrict__ comm, int32_t n, int32_t* __restrict__ cnt) {
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    atomicAdd(&cnt[comm[v]], 1);
}

__global__ void k_nonempty(const int32_t* __restrict__ cnt, int32_t n, uint8_t* __restrict__ flag) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
    flag[c] = cnt[c] > 0;
}

template <class T>
__global__ void k_ecount(DevGraph<T> g, const int32_t* __restrict__ comm, const int32_t* __restrict__ nid,
                         int64_t* __restrict__ ecnt) {
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < g.n; v += gridDim.x * blockDim.x) {
    int64_t d = g.off[v + 1] - g.off[v];
    if (d) atomicAdd((unsigned long long*)&ecnt[nid[comm[v]]], (unsigned long long)d);
  }
}

// stage every row entry as (new id of the neighbour's community, weight) in the
// segment of the owner's new community
template <class T>
__global__ void k_stage_small(DevGraph<T> g, const int32_t* __restrict__ comm,
                              const int32_t* __restrict__ nid, const int64_t* __restrict__ eoff,
                              int64_t* __restrict__ ecur, int32_t* __restrict__ skey,
                              typename T::acc_t* __restrict__ sval) {
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < g.n; v += gridDim.x * blockDim.x) {
    int64_t o = g.off[v], e = g.off[v + 1];
    int64_t d = e - o;
    if (d == 0 || d > 64) continue;
    int nc = nid[comm[v]];
    int64_t base = eoff[nc] + (int64_t)atomicAdd((unsigned long long*)&ecur[nc], (unsigned long long)d);
    for (int64_t j = o; j < e; j++) {
      skey[base + (j - o)] = nid[comm[g.nbr[j]]];
      sval[base + (j - o)] = (typename T::acc_t)(g.w ? g.w[j] : (typename T::w_t)1);
    }
  }
}

template <class T>
__global__ void k_stage_wide(DevGraph<T> g, const int32_t* __restrict__ comm,
                             const int32_t* __restrict__ nid, const int64_t* __restrict__ eoff,
                             int64_t* __restrict__ ecur, int32_t* __restrict__ skey,
                             typename T::acc_t* __restrict__ sval) {
  const int lane = threadIdx.x & 63;
  const int nw = gridDim.x * (blockDim.x >> 6);
  for (int v = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); v < g.n; v += nw) {
    int64_t o = g.off[v], e = g.off[v + 1];
    int64_t d = e - o;
    if (d <= 64) continue;
    int nc = nid[comm[v]];
    int64_t base = 0;
    if (lane == 0)
      base = eoff[nc] + (int64_t)atomicAdd((unsigned long long*)&ecur[nc], (unsigned long long)d);
    base = __shfl(base, 0, 64);
    for (int64_t j = o + lane; j < e; j += 64) {
      skey[base + (j - o)] = nid[comm[g.nbr[j]]];
      sval[base + (j - o)] = (typename T::acc_t)(g.w ? g.w[j] : (typename T::w_t)1);
    }
  }
}

// classify rows: long rows are appended to big_list with their table size
__global__ void k_classify(const int64_t* __restrict__ ecnt, int32_t C, int32_t* __restrict__ big_list,
                           int32_t* __restrict__ nbig, int64_t* __restrict__ big_tsz) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < C; c += gridDim.x * blockDim.x) {
    long long L = ecnt[c];
    if (L > SMALL_B) {
      int k = atomicAdd(nbig, 1);
      big_list[k] = c;
      long long lim = L < (long long)C ? L : (long long)C;
      big_tsz[k] = 1ll << clog2(2 * lim);
    }
  }
}

// ---------------------------------------------------------------------------
// short rows: bitonic sort in LDS then reduce-by-key, in place
// ---------------------------------------------------------------------------
template <int NT, int CAP, class T>
__global__ __launch_bounds__(NT) void k_merge_small(const int64_t* __restrict__ ecnt,
                                                     const int64_t* __restrict__ eoff, int32_t C,
                                                     int lo_excl, int32_t* __restrict__ skey,
                                                     typename T::acc_t* __restrict__ sval,
                                                     int32_t* __restrict__ dcnt) {
  using acc = typename T::acc_t;
  __shared__ int k_s[CAP];
  __shared__ acc v_s[CAP];
  __shared__ int h_s[CAP];
  __shared__ int sh[17];
  for (int c = blockIdx.x; c < C; c += gridDim.x) {
    const long long L = ecnt[c];
    if (L <= lo_excl || L > CAP) continue;  // uniform per block
    const int64_t base = eoff[c];
    int P = 1;
    while (P < L) P <<= 1;
    for (int t = threadIdx.x; t < P; t += NT) {
      if (t < L) { k_s[t] = skey[base + t]; v_s[t] = sval[base + t]; }
      else { k_s[t] = INT_MAX; v_s[t] = 0; }
    }
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int t = threadIdx.x; t < P; t += NT) {
          int u = t ^ j;
          if (u > t) {
            bool up = (t & k) == 0;
            int a = k_s[t], b = k_s[u];
            if ((a > b) == up) {
              k_s[t] = b; k_s[u] = a;
              acc x = v_s[t]; v_s[t] = v_s[u]; v_s[u] = x;
            }
          }
        }
        __syncthreads();
      }
    }
    // heads of runs
    for (int t = threadIdx.x; t < P; t += NT)
      h_s[t] = (t < L) && (t == 0 || k_s[t] != k_s[t - 1]);
    __syncthreads();
    // exclusive scan of head flags over P items (each thread a contiguous chunk)
    const int per = (P + NT - 1) / NT;
    const int a0 = threadIdx.x * per, a1 = min(P, a0 + per);
    int s = 0;
    for (int t = a0; t < a1; t++) s += h_s[t];
    // block exclusive scan of s
    {
      const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
      int x = s;
#pragma unroll
      for (int q = 1; q < 64; q <<= 1) {
        int y = __shfl_up(x, q, 64);
        if (lane >= q) x += y;
      }
      if (lane == 63) sh[wid] = x;
      __syncthreads();
      if (threadIdx.x == 0) {
        int run = 0;
        for (int w = 0; w < NT / 64; w++) { int tmp = sh[w]; sh[w] = run; run += tmp; }
        sh[16] = run;
      }
      __syncthreads();
      s = x - s + sh[wid];
    }
    const int D = sh[16];
    __syncthreads();
    int pos = s;
    for (int t = a0; t < a1; t++) {
      if (h_s[t]) {
        acc sum = 0;
        int key = k_s[t];
        for (int u = t; u < L && k_s[u] == key; u++) sum += v_s[u];
        skey[base + pos] = key;
        sval[base + pos] = sum;
        pos++;
      }
    }
    if (threadIdx.x == 0) dcnt[c] = D;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// long rows: global hash + ordered extraction through a bitmap of C bits
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(1024) void k_merge_big(const int64_t* __restrict__ ecnt,
                                                     const int64_t* __restrict__ eoff, int32_t C,
                                                     const int32_t* __restrict__ big_list, int nbig,
                                                     const int64_t* __restrict__ big_off,
                                                     int32_t* __restrict__ arena_keys,
                                                     typename T::acc_t* __restrict__ arena_vals,
                                                     uint32_t* __restrict__ arena_bits,
                                                     int32_t* __restrict__ skey,
                                                     typename T::acc_t* __restrict__ sval,
                                                     int32_t* __restrict__ dcnt, int* overflow) {
  using acc = typename T::acc_t;
  constexpr int NT = 1024;
  __shared__ uint32_t bits_lds[BITMAP_LDS_WORDS];
  __shared__ int sh[17];
  const int nwords = (C + 31) >> 5;
  for (int b = blockIdx.x; b < nbig; b += gridDim.x) {
    const int c = big_list[b];
    const long long L = ecnt[c];
    const int64_t base = eoff[c];
    const int64_t toff = big_off[b];
    const long long tsz = big_off[b + 1] - toff;
    const int lg = clog2(tsz);
    int* keys = arena_keys + toff;
    acc* vals = arena_vals + toff;
    uint32_t* bits = nwords <= BITMAP_LDS_WORDS ? bits_lds : arena_bits + (int64_t)blockIdx.x * nwords;
    for (long long t = threadIdx.x; t < tsz; t += NT) { keys[t] = 0; vals[t] = 0; }
    for (int t = threadIdx.x; t < nwords; t += NT) bits[t] = 0;
    __syncthreads();
    for (long long j = threadIdx.x; j < L; j += NT) {
      int key = skey[base + j];
      acc w = sval[base + j];
      unsigned p = chash(key, lg);
      long long probe = 0;
      for (; probe < tsz; probe++) {
        int k = atomicCAS(&keys[p], 0, key + 1);
        if (k == 0 || k == key + 1) { atomicAdd(&vals[p], w); break; }
        p = (p + 1) & (unsigned)(tsz - 1);
      }
      if (probe == tsz) *overflow = 1;
    }
    __syncthreads();
    for (long long t = threadIdx.x; t < tsz; t += NT) {
      int k = keys[t];
      if (k) atomicOr(&bits[(k - 1) >> 5], 1u << ((k - 1) & 31));
    }
    __syncthreads();
    // ordered extraction: contiguous word chunk per thread
    const int per = (nwords + NT - 1) / NT;
    const int a0 = threadIdx.x * per, a1 = min(nwords, a0 + per);
    int s = 0;
    for (int t = a0; t < a1; t++) s += __popc(bits[t]);
    {
      const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
      int x = s;
#pragma unroll
      for (int q = 1; q < 64; q <<= 1) {
        int y = __shfl_up(x, q, 64);
        if (lane >= q) x += y;
      }
      if (lane == 63) sh[wid] = x;
      __syncthreads();
      if (threadIdx.x == 0) {
        int run = 0;
        for (int w = 0; w < NT / 64; w++) { int tmp = sh[w]; sh[w] = run; run += tmp; }
        sh[16] = run;
      }
      __syncthreads();
      s = x - s + sh[wid];
    }
    const int D = sh[16];
    int pos = s;
    for (int t = a0; t < a1; t++) {
      uint32_t wbits = bits[t];
      while (wbits) {
        int bit = __ffs(wbits) - 1;
        wbits &= wbits - 1;
        int key = (t << 5) + bit;
        unsigned p = chash(key, lg);
        while (keys[p] != key + 1) p = (p + 1) & (unsigned)(tsz - 1);
        skey[base + pos] = key;
        sval[base + pos] = vals[p];
        pos++;
      }
    }
    if (threadIdx.x == 0) dcnt[c] = D;
    __syncthreads();
  }
}

// copy each merged row to its final place
template <class T>
__global__ void k_compact(const int64_t* __restrict__ eoff, const int64_t* __restrict__ noff,
                          const int32_t* __restrict__ dcnt, int32_t C, const int32_t* __restrict__ skey,
                          const typename T::acc_t* __restrict__ sval, int32_t* __restrict__ nbr,
                          typename T::w_t* __restrict__ w) {
  const int lane = threadIdx.x & 63;
  const int nw = gridDim.x * (blockDim.x >> 6);
  for (int c = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); c < C; c += nw) {
    const int64_t src = eoff[c], dst = noff[c];
    const int D = dcnt[c];
    for (int t = lane; t < D; t += 64) {
      nbr[dst + t] = skey[src + t];
      w[dst + t] = (typename T::w_t)sval[src + t];
    }
  }
}

__global__ void k_compose(int32_t* __restrict__ fin, int32_t n0, const int32_t* __restrict__ comm,
                          const int32_t* __restrict__ nid) {
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n0; v += gridDim.x * blockDim.x)
    fin[v] = nid[comm[fin[v]]];
}

void compose_map(int32_t* fin, int32_t n0, const int32_t* comm, const int32_t* nid, hipStream_t s) {
  if (n0 <= 0) return;
  k_compose<<<std::min((n0 + 255) / 256, 8192), 256, 0, s>>>(fin, n0, comm, nid);
}

static inline int nblk(int64_t n, int per = 256, int cap = 8192) {
  int64_t b = (n + per - 1) / per;
  if (b < 1) b = 1;
  return (int)std::min<int64_t>(b, cap);
}

template <class T>
int contract_prepare(const DevGraph<T>& g, const int32_t* comm, ContractBufs<T>& cb, int32_t* C_out,
                     int64_t* staged_out, hipStream_t s) {
  const int32_t n = g.n;
  fill<int32_t>(cb.cnt, n, 0, s);
  k_hist<<<nblk(n), 256, 0, s>>>(comm, n, cb.cnt);
  uint8_t* flag = (uint8_t*)cb.i32_tmp;  // n bytes fit in the n-int scratch
  k_nonempty<<<nblk(n), 256, 0, s>>>(cb.cnt, n, flag);
  scan_exclusive<uint8_t, int32_t>(flag, cb.nid, n, cb.i32_tmp + (n + 3) / 4 + 1, cb.i32_scalar, s);
  int32_t C = 0;
  if (hipMemcpyAsync(&C, cb.i32_scalar, sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
  if (hipStreamSynchronize(s) != hipSuccess) return -1;
  fill<int64_t>(cb.ecnt, C + 1, 0, s);
  k_ecount<T><<<nblk(n), 256, 0, s>>>(g, comm, cb.nid, cb.ecnt);
  scan_exclusive<int64_t, int64_t>(cb.ecnt, cb.eoff, C, cb.i64_tmp, cb.eoff + C, s);
  *C_out = C;
  *staged_out = g.m;
  return 0;
}

template <class T>
int contract_build(const DevGraph<T>& g, const int32_t* comm, ContractBufs<T>& cb, int32_t C,
                   int64_t staged, ContractOut<T>& out, int64_t* m_out, hipStream_t s) {
  const int32_t n = g.n;
  (void)staged;
  fill<int64_t>(cb.ecur, C, 0, s);
  k_stage_small<T><<<nblk(n), 256, 0, s>>>(g, comm, cb.nid, cb.eoff, cb.ecur, cb.skey, cb.sval);
  k_stage_wide<T><<<nblk(n, 4, 16384), 256, 0, s>>>(g, comm, cb.nid, cb.eoff, cb.ecur, cb.skey, cb.sval);
  fill<int32_t>(cb.dcnt, C + 1, 0, s);
  fill<int32_t>(cb.nbig, 1, 0, s);
  k_classify<<<nblk(C), 256, 0, s>>>(cb.ecnt, C, cb.big_list, cb.nbig, cb.i64_tmp);
  k_merge_small<64, SMALL_A, T><<<nblk(C, 1, 16384), 64, 0, s>>>(cb.ecnt, cb.eoff, C, 0, cb.skey, cb.sval, cb.dcnt);
  k_merge_small<256, SMALL_B, T><<<nblk(C, 1, 4096), 256, 0, s>>>(cb.ecnt, cb.eoff, C, SMALL_A, cb.skey, cb.sval, cb.dcnt);
  int32_t nbig = 0;
  if (hipMemcpyAsync(&nbig, cb.nbig, sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
  if (hipStreamSynchronize(s) != hipSuccess) return -1;
  if (nbig > 0) {
    // table offsets: exclusive scan of sizes (written by k_classify into i64_tmp)
    int64_t* big_off = cb.big_tab;
    scan_exclusive<int64_t, int64_t>(cb.i64_tmp, big_off, nbig, cb.i64_tmp + nbig + 1, big_off + nbig, s);
    int64_t slots = 0;
    if (hipMemcpyAsync(&slots, big_off + nbig, sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
    if (hipStreamSynchronize(s) != hipSuccess) return -1;
    const int nwords = (C + 31) / 32;
    int grid = std::min(nbig, 1024);
    if (nwords > BITMAP_LDS_WORDS) {
      if (cb.arena_words < nwords) return -3;  // caller grows the bitmap arena
      grid = (int)std::min<int64_t>(grid, std::max<int64_t>(1, cb.arena_words / nwords));
    }
    if (slots > cb.arena_slots) return -2;  // caller grows the arena and retries
    int* ovf = cb.i32_scalar + 3;
    fill<int32_t>(ovf, 1, 0, s);
    k_merge_big<T><<<grid, 1024, 0, s>>>(cb.ecnt, cb.eoff, C, cb.big_list, nbig, big_off, cb.arena_keys,
                                         cb.arena_vals, cb.arena_bits, cb.skey, cb.sval, cb.dcnt, ovf);
  }
  scan_exclusive<int32_t, int64_t>(cb.dcnt, out.off, C, cb.i64_tmp, out.off + C, s);
  k_compact<T><<<nblk(C, 4, 16384), 256, 0, s>>>(cb.eoff, out.off, cb.dcnt, C, cb.skey, cb.sval, out.nbr, out.w);
  int64_t m = 0;
  if (hipMemcpyAsync(&m, out.off + C, sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
  if (hipStreamSynchronize(s) != hipSuccess) return -1;
  *m_out = m;
  return 0;
}

// host-side helper: arena need of the current contraction (after classify)
#define GALA_CINST(T)                                                                             \
  template int contract_prepare<T>(const DevGraph<T>&, const int32_t*, ContractBufs<T>&, int32_t*, \
                                   int64_t*, hipStream_t);                                        \
  template int contract_build<T>(const DevGraph<T>&, const int32_t*, ContractBufs<T>&, int32_t,   \
                                 int64_t, ContractOut<T>&, int64_t*, hipStream_t);
GALA_CINST(IntW)
GALA_CINST(FltW)

}  // namespace gala
