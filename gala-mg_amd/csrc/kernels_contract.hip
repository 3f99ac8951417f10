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
// Pipeline: label[v] = new_id(comm[v]) -> per-community entry counts -> stage
// every entry (label of the neighbour, weight) in its community's segment ->
// merge each segment into a sorted, duplicate-free row, by segment length L:
//   L <= 16/32/64 one 16-, 32- or 64-lane group per row: bitonic sort across
//                 the group's lanes + segmented scan
//   L <= 512      64-thread workgroup: LDS hash aggregate, then a register
//                 bitonic sort of the distinct keys (8 per lane, no barriers)
//   L <= 2048     128-thread workgroup, same with 16 keys per thread
//   L <= 4096     256-thread workgroup
//   L <= LT       1024-thread workgroup: LT-slot LDS table (LT = 16384, 8192 on
//                 the FP path), overflow into a global slice, register bitonic
//                 sort of <= LT keys (LDS only for exchanges between waves)
//   L >  LT       dense rows: the segment is cut into chunks, one workgroup per
//                 chunk pre-aggregates in LDS and adds into a C-wide accumulator
//                 for the row (HBM is plentiful); a tiled scan of the accumulator
//                 then emits the row already in key order
// -> exclusive scan of the distinct counts -> compaction into the new CSR.
#include "gala_kernels.h"
#include "gala_device.h"
#include <algorithm>
#include <climits>
#include <vector>

namespace gala {

constexpr int DENSE_TILE = 4096;  // keys per workgroup in the dense-row scans

__device__ __forceinline__ unsigned chash(int key, int lg) {
  return ((unsigned)key * 2654435769u) >> (32 - lg);
}
__host__ __device__ __forceinline__ int clog2(long long x) {
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

__global__ void k_label(const int32_t* __restrict__ comm, const int32_t* __restrict__ nid, int32_t n,
                        int32_t* __restrict__ label) {
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    label[v] = nid[comm[v]];
}

__global__ void k_ecount(const int64_t* __restrict__ off, int32_t n, const int32_t* __restrict__ label,
                         int64_t* __restrict__ ecnt) {
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    int64_t d = off[v + 1] - off[v];
    if (d) atomicAdd((unsigned long long*)&ecnt[label[v]], (unsigned long long)d);
  }
}

// staged length of each listed dense row
__global__ void k_dense_lengths(const int32_t* __restrict__ list, int32_t nlist, const int64_t* __restrict__ ecnt,
                                int64_t* __restrict__ len) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nlist; i += gridDim.x * blockDim.x) len[i] = ecnt[list[i]];
}

// classify rows: 4096 < L <= lt -> big_list (with the size of its global
// overflow slice); L > lt -> dense_list
// Row length tiers of the merge: 0: 1..16, 1: 17..32, 2: 33..64, 3: 65..512,
// 4: 513..2048, 5: 2049..4096; longer rows go to the big and dense lists,
// empty rows nowhere.
constexpr int TIER_T16 = 0, TIER_T32 = 1, TIER_T64 = 2, TIER_MID1 = 3, TIER_MID2 = 4, TIER_MID3 = 5;

__global__ void k_classify(const int64_t* __restrict__ ecnt, int32_t C, long long lt,
                           int32_t* __restrict__ big_list, int32_t* __restrict__ counts,
                           int64_t* __restrict__ big_tsz, int32_t* __restrict__ dense_list,
                           int8_t* __restrict__ tier) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < C; c += gridDim.x * blockDim.x) {
    const long long L = ecnt[c];
    tier[c] = L == 0 ? -1 : L <= 16 ? TIER_T16 : L <= 32 ? TIER_T32 : L <= 64 ? TIER_T64
            : L <= 512 ? TIER_MID1 : L <= 2048 ? TIER_MID2 : L <= 4096 ? TIER_MID3 : -1;
    if (L > lt) {
      dense_list[atomicAdd(&counts[1], 1)] = c;
    } else if (L > 4096) {
      const int k = atomicAdd(&counts[0], 1);
      big_list[k] = c;
      const long long lim = L < (long long)C ? L : (long long)C;
      big_tsz[k] = 1ll << clog2(2 * lim);
    }
  }
}

// ---------------------------------------------------------------------------
// staging: (label of the neighbour, weight) into the owner's segment.
// Rows of fewer than 512 entries: a block owns 256 consecutive vertices and
// reserves space with one global atomic per distinct community in the block
// (LDS aggregation first, so a community with many members is not a hot spot),
// then all threads copy the block's rows.  Longer rows: k_stage_rows.
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_stage(DevGraph<T> g, const int32_t* __restrict__ label,
                                               const int64_t* __restrict__ eoff,
                                               unsigned long long* __restrict__ ecur,
                                               int32_t* __restrict__ skey,
                                               typename T::acc_t* __restrict__ sval) {
  using acc = typename T::acc_t;
  constexpr int HS = 512;
  __shared__ int64_t s_ofs[257];
  __shared__ int64_t s_dst[256];
  __shared__ int64_t s_row[256];
  __shared__ int s_hk[HS];
  __shared__ unsigned long long s_hv[HS];
  __shared__ unsigned long long s_hb[HS];
  __shared__ int64_t s_sh[17];
  for (int vb = blockIdx.x * 256; vb < g.n; vb += gridDim.x * 256) {
    const int nloc = min(256, g.n - vb);
    const int v = vb + threadIdx.x;
    const bool valid = threadIdx.x < nloc;
    const int64_t o = valid ? g.off[v] : 0;
    const int64_t d0 = valid ? g.off[v + 1] - o : 0;
    const int64_t d = d0 < 512 ? d0 : 0;  // longer rows (degree bins 6-8): k_stage_rows
    const int nc = (valid && d) ? label[v] : -1;
    for (int t = threadIdx.x; t < HS; t += 256) { s_hk[t] = 0; s_hv[t] = 0; }
    int64_t tot;
    const int64_t pre = block_excl_scan<int64_t>(d, s_sh, &tot);
    s_ofs[threadIdx.x] = pre;
    if (threadIdx.x == 0) s_ofs[nloc] = tot;
    __syncthreads();
    unsigned p = 0;
    unsigned long long intra = 0;
    if (nc >= 0) {
      p = chash(nc, 9);
      for (;;) {
        const int k = atomicCAS(&s_hk[p], 0, nc + 1);
        if (k == 0 || k == nc + 1) break;
        p = (p + 1) & (HS - 1);
      }
      intra = atomicAdd(&s_hv[p], (unsigned long long)d);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < HS; t += 256)
      if (s_hk[t]) s_hb[t] = atomicAdd(&ecur[s_hk[t] - 1], s_hv[t]);
    __syncthreads();
    if (nc >= 0) s_dst[threadIdx.x] = eoff[nc] + (int64_t)(s_hb[p] + intra);
    __syncthreads();
    if (valid) s_row[threadIdx.x] = o;
    __syncthreads();
    for (int64_t q = threadIdx.x; q < tot; q += 256) {
      int a = 0, z = nloc;  // largest i with s_ofs[i] <= q
      while (z - a > 1) {
        const int mid = (a + z) >> 1;
        if (s_ofs[mid] <= q) a = mid; else z = mid;
      }
      const int64_t j = s_row[a] + (q - s_ofs[a]);
      const int64_t dst = s_dst[a] + (q - s_ofs[a]);
      skey[dst] = label[g.nbr[j]];
      sval[dst] = g.w ? (acc)g.w[j] : acc(1);
    }
    __syncthreads();
  }
}

// rows of the vertices order[bstart[blo] .. bstart[bhi]) (the round's degree
// bins): NT threads per vertex, one reservation each
template <int NT, class T>
__global__ __launch_bounds__(256) void k_stage_rows(DevGraph<T> g, const int32_t* __restrict__ order,
                                                    const int32_t* __restrict__ bstart, int blo, int bhi,
                                                    const int32_t* __restrict__ label,
                                                    const int64_t* __restrict__ eoff,
                                                    unsigned long long* __restrict__ ecur,
                                                    int32_t* __restrict__ skey,
                                                    typename T::acc_t* __restrict__ sval) {
  using acc = typename T::acc_t;
  constexpr int G = 256 / NT;  // vertices per block step
  const int lo = bstart[blo], hi = bstart[bhi];
  const int lane = threadIdx.x % NT;
  for (int i = lo + (blockIdx.x * blockDim.x + threadIdx.x) / NT; i < hi; i += gridDim.x * G) {
    const int v = order[i];
    const int64_t o = g.off[v];
    const int64_t d = g.off[v + 1] - o;
    const int nc = label[v];
    int64_t base = 0;
    if (lane == 0) base = eoff[nc] + (int64_t)atomicAdd(&ecur[nc], (unsigned long long)d);
    base = __shfl(base, 0, NT > 64 ? 64 : NT);
    if (NT > 64) {  // broadcast across the waves of the group through LDS
      __shared__ int64_t s_b;
      if (threadIdx.x == 0) s_b = base;
      __syncthreads();
      base = s_b;
      __syncthreads();
    }
    for (int64_t j = lane; j < d; j += NT) {
      skey[base + j] = label[g.nbr[o + j]];
      sval[base + j] = g.w ? (acc)g.w[o + j] : acc(1);
    }
  }
}

// ---------------------------------------------------------------------------
// L <= W (16, 32 or 64): a W-lane group per community, everything in registers
// ---------------------------------------------------------------------------
template <int W, int TR, class T>
__global__ __launch_bounds__(256) void k_merge_tiny(const int64_t* __restrict__ ecnt,
                                                     const int64_t* __restrict__ eoff,
                                                     const int32_t* __restrict__ tlist,
                                                     const int32_t* __restrict__ tstart,
                                                     int32_t* __restrict__ skey,
                                                     typename T::acc_t* __restrict__ sval,
                                                     int32_t* __restrict__ dcnt) {
  using acc = typename T::acc_t;
  const int lane = threadIdx.x & (W - 1);
  const int gbase = (threadIdx.x & 63) & ~(W - 1);  // first lane of the group in its wave
  const unsigned long long gmask = W == 64 ? ~0ull : ((1ull << W) - 1);
  const int ng = gridDim.x * (blockDim.x / W);
  const int lo = tstart[TR], hi = tstart[TR + 1];
  for (int i = lo + (blockIdx.x * blockDim.x + threadIdx.x) / W; i < hi; i += ng) {  // uniform per group
    const int c = tlist[i];
    const int L = (int)ecnt[c];
    const int64_t base = eoff[c];
    int key = INT_MAX;
    acc val = 0;
    if (lane < L) { key = skey[base + lane]; val = sval[base + lane]; }
    // bitonic sort over the group's W lanes (ascending)
#pragma unroll
    for (int k = 2; k <= W; k <<= 1) {
#pragma unroll
      for (int j = k >> 1; j > 0; j >>= 1) {
        int pk = __shfl_xor(key, j, W);
        acc pv = __shfl_xor(val, j, W);
        const bool up = (lane & k) == 0, lower = (lane & j) == 0;
        const bool take = (lower == up) ? (pk < key) : (pk > key);
        if (take) { key = pk; val = pv; }
      }
    }
    const int prevk = __shfl_up(key, 1, W);
    const bool head = lane < L && (lane == 0 || prevk != key);
    // segmented inclusive scan of val (runs start at heads)
    bool f = head;
#pragma unroll
    for (int k = 1; k < W; k <<= 1) {
      acc y = __shfl_up(val, k, W);
      int fy = __shfl_up((int)f, k, W);
      if (lane >= k && !f) { val += y; f = fy; }
    }
    const unsigned long long heads = (__ballot(head) >> gbase) & gmask;
    const int nextk = __shfl_down(key, 1, W);
    const bool tail = lane < L && (lane == L - 1 || nextk != key);
    if (tail) {
      const unsigned long long le = (lane == 63) ? ~0ull : ((2ull << lane) - 1);
      const int rank = __popcll(heads & le) - 1;
      skey[base + rank] = key;
      sval[base + rank] = val;
    }
    if (lane == 0) dcnt[c] = __popcll(heads);
  }
}

// ---------------------------------------------------------------------------
// Sort the D distinct (key, value) pairs s_k/s_v[0, D) of a row (D <= NT*E)
// in registers and write them to skey/sval[base, base + D).  xk/xv are the
// exchange buffers of the sort (NT*E elements each, may alias s_k/s_v).  Run
// by threads 0..NT-1; for NT > 64 that must be the whole workgroup.
// ---------------------------------------------------------------------------
template <int NT, int E, class V>
__device__ __forceinline__ void sort_row_out(const int* s_k, const V* s_v, int D, int* xk, V* xv,
                                             int32_t* __restrict__ skey, V* __restrict__ sval, int64_t base) {
  int k[E];
  V v[E];
  const int t = threadIdx.x;
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int i = t * E + e;
    k[e] = i < D ? s_k[i] : INT_MAX;
    v[e] = i < D ? s_v[i] : V(0);
  }
  reg_bitonic<NT, E, V>(k, v, xk, xv);
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int i = t * E + e;
    if (i < D) {
      skey[base + i] = k[e];
      sval[base + i] = v[e];
    }
  }
}

// The smallest register sort that holds the D distinct pairs of a row: one
// wave with 1, 2, 4 or 8 keys per lane for D <= 512, else the whole NT-thread
// workgroup with 1..EMAX keys per thread (xk/xv hold NT*EMAX elements).
template <int NT, int EMAX, class V>
__device__ __forceinline__ void sort_row(const int* s_k, const V* s_v, int D, int* xk, V* xv,
                                         int32_t* __restrict__ skey, V* __restrict__ sval, int64_t base) {
  if (D <= 512 || NT == 64) {
    if (threadIdx.x < 64) {
      if (D <= 64) sort_row_out<64, 1, V>(s_k, s_v, D, xk, xv, skey, sval, base);
      else if (D <= 128) sort_row_out<64, 2, V>(s_k, s_v, D, xk, xv, skey, sval, base);
      else if (D <= 256) sort_row_out<64, 4, V>(s_k, s_v, D, xk, xv, skey, sval, base);
      else sort_row_out<64, 8, V>(s_k, s_v, D, xk, xv, skey, sval, base);
    }
    return;
  }
  if constexpr (NT > 64) {
    if (D <= NT) { sort_row_out<NT, 1, V>(s_k, s_v, D, xk, xv, skey, sval, base); return; }
    if constexpr (EMAX >= 2)
      if (D <= NT * 2) { sort_row_out<NT, 2, V>(s_k, s_v, D, xk, xv, skey, sval, base); return; }
    if constexpr (EMAX >= 4)
      if (D <= NT * 4) { sort_row_out<NT, 4, V>(s_k, s_v, D, xk, xv, skey, sval, base); return; }
    if constexpr (EMAX >= 8)
      if (D <= NT * 8) { sort_row_out<NT, 8, V>(s_k, s_v, D, xk, xv, skey, sval, base); return; }
    if constexpr (EMAX >= 16) sort_row_out<NT, 16, V>(s_k, s_v, D, xk, xv, skey, sval, base);
  }
}

// ---------------------------------------------------------------------------
// LO < L <= LMAX: one workgroup per community: LDS hash aggregate, then sort
// the distinct keys (unique, so no reduction after the sort)
// ---------------------------------------------------------------------------
template <int NT, int LO, int LMAX, class T>
__global__ __launch_bounds__(NT) void k_merge_mid(const int64_t* __restrict__ ecnt,
                                                   const int64_t* __restrict__ eoff, int32_t C,
                                                   const int32_t* __restrict__ tlist,
                                                   const int32_t* __restrict__ tstart, int tr,
                                                   int32_t* __restrict__ skey,
                                                   typename T::acc_t* __restrict__ sval,
                                                   int32_t* __restrict__ dcnt) {
  using acc = typename T::acc_t;
  constexpr int TMAX = 2 * LMAX;
  __shared__ int h_k[TMAX];
  __shared__ acc h_v[TMAX];
  __shared__ int s_k[LMAX];
  __shared__ acc s_v[LMAX];
  __shared__ int s_sh[17];
  const int lo = tstart[tr], hi = tstart[tr + 1];
  for (int i = lo + blockIdx.x; i < hi; i += gridDim.x) {
    const int c = tlist[i];
    const long long L = ecnt[c];
    const int64_t base = eoff[c];
    const long long lim = L < (long long)C ? L : (long long)C;
    const int lg = clog2(2 * lim) < 1 ? 1 : clog2(2 * lim);
    const int tsz = 1 << lg;
    for (int t = threadIdx.x; t < tsz; t += NT) { h_k[t] = 0; h_v[t] = 0; }
    __syncthreads();
    for (int j = threadIdx.x; j < L; j += NT) {
      const int key = skey[base + j];
      const acc w = sval[base + j];
      unsigned p = chash(key, lg);
      for (;;) {
        const int k = atomicCAS(&h_k[p], 0, key + 1);
        if (k == 0 || k == key + 1) { atomicAdd(&h_v[p], w); break; }
        p = (p + 1) & (tsz - 1);
      }
    }
    __syncthreads();
    int cnt = 0;
    for (int t = threadIdx.x; t < tsz; t += NT) cnt += h_k[t] != 0;
    int D;
    int q = block_excl_scan<int>(cnt, s_sh, &D);
    for (int t = threadIdx.x; t < tsz; t += NT) {
      const int k = h_k[t];
      if (k) { s_k[q] = k - 1; s_v[q] = h_v[t]; q++; }
    }
    __syncthreads();
    sort_row<NT, LMAX / NT, acc>(s_k, s_v, D, h_k, h_v, skey, sval, base);
    if (threadIdx.x == 0) dcnt[c] = D;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// 4096 < L <= LT: one 1024-thread workgroup per community.  An LT-slot LDS
// table is tried first (bounded probing); entries that find no LDS slot go to
// the community's slice of the global arena (next_pow2(2 min(L, C)) slots, so
// nothing is dropped).  A key lives in exactly one of the two tables, and
// D <= L <= LT keys are then sorted in LDS.  Global slices are kept all-zero
// between uses (the reader clears what it read).
// ---------------------------------------------------------------------------
template <int LT, class T>
__global__ __launch_bounds__(1024) void k_merge_big(const int64_t* __restrict__ ecnt,
                                                     const int64_t* __restrict__ eoff, int32_t C,
                                                     const int32_t* __restrict__ big_list, int nbig,
                                                     const int64_t* __restrict__ big_off,
                                                     int32_t* __restrict__ arena_keys,
                                                     typename T::acc_t* __restrict__ arena_vals,
                                                     int32_t* __restrict__ skey,
                                                     typename T::acc_t* __restrict__ sval,
                                                     int32_t* __restrict__ dcnt) {
  using acc = typename T::acc_t;
  constexpr int NT = 1024, PMAX = 32, PER = LT / NT;
  constexpr int LLG = __builtin_ctz(LT);
  __shared__ int s_k[LT];
  __shared__ acc s_v[LT];
  __shared__ int s_sh[17];
  __shared__ int s_ovf;
  for (int b = blockIdx.x; b < nbig; b += gridDim.x) {
    const int c = big_list[b];
    const long long L = ecnt[c];
    const int64_t base = eoff[c];
    const int64_t toff = big_off[b];
    const long long gsz = big_off[b + 1] - toff;
    const int glg = clog2(gsz);
    int* keys = arena_keys + toff;
    acc* vals = arena_vals + toff;
    for (int t = threadIdx.x; t < LT; t += NT) { s_k[t] = 0; s_v[t] = 0; }
    if (threadIdx.x == 0) s_ovf = 0;
    __syncthreads();
    for (long long j = threadIdx.x; j < L; j += NT) {
      const int key = skey[base + j];
      const acc w = sval[base + j];
      unsigned p = chash(key, LLG);
      bool placed = false;
      for (int probe = 0; probe < PMAX; probe++) {
        const int k = atomicCAS(&s_k[p], 0, key + 1);
        if (k == 0 || k == key + 1) { atomicAdd(&s_v[p], w); placed = true; break; }
        p = (p + 1) & (LT - 1);
      }
      if (!placed) {
        s_ovf = 1;
        unsigned q = chash(key, glg);
        for (;;) {
          const int k = atomicCAS(&keys[q], 0, key + 1);
          if (k == 0 || k == key + 1) { atomicAdd(&vals[q], w); break; }
          q = (q + 1) & (unsigned)(gsz - 1);
        }
      }
    }
    __syncthreads();
    const bool ovf = s_ovf != 0;
    int cl = 0, cg = 0;
    for (int t = threadIdx.x; t < LT; t += NT) cl += s_k[t] != 0;
    if (ovf)
      for (long long t = threadIdx.x; t < gsz; t += NT) cg += keys[t] != 0;
    int Dl, Dg;
    const int ql = block_excl_scan<int>(cl, s_sh, &Dl);
    const int qg = block_excl_scan<int>(cg, s_sh, &Dg);
    const int D = Dl + Dg;
    // compact: LDS pairs to [0, Dl), overflow pairs to [Dl, D)
    int rk[PER];
    acc rv[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
      rk[i] = s_k[threadIdx.x + i * NT];
      rv[i] = s_v[threadIdx.x + i * NT];
    }
    __syncthreads();
    int q = ql;
#pragma unroll
    for (int i = 0; i < PER; i++)
      if (rk[i]) { s_k[q] = rk[i] - 1; s_v[q] = rv[i]; q++; }
    if (ovf) {
      q = Dl + qg;
      for (long long t = threadIdx.x; t < gsz; t += NT) {
        const int k = keys[t];
        if (k) {
          s_k[q] = k - 1;
          s_v[q] = vals[t];
          q++;
          keys[t] = 0;
          vals[t] = 0;
        }
      }
    }
    __syncthreads();
    sort_row<NT, LT / NT, acc>(s_k, s_v, D, s_k, s_v, skey, sval, base);
    if (threadIdx.x == 0) dcnt[c] = D;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// dense rows (L > LT).  chunk table: (dense slot, community, start, length)
// ---------------------------------------------------------------------------
template <int LT, class T>
__global__ __launch_bounds__(1024) void k_dense_accum(const DenseChunk* __restrict__ chunks, int nchunks,
                                                      const int64_t* __restrict__ eoff, int32_t C,
                                                      const int32_t* __restrict__ skey,
                                                      const typename T::acc_t* __restrict__ sval,
                                                      typename T::acc_t* __restrict__ dense,
                                                      uint32_t* __restrict__ dbits, int64_t words) {
  using acc = typename T::acc_t;
  constexpr int NT = 1024, LLG = __builtin_ctz(LT);
  __shared__ int s_k[LT];
  __shared__ acc s_v[LT];
  for (int ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    const DenseChunk cc = chunks[ch];
    const int64_t base = eoff[cc.c] + cc.start;
    acc* row = dense + (int64_t)cc.slot * C;
    uint32_t* bits = dbits + (int64_t)cc.slot * words;
    for (int t = threadIdx.x; t < LT; t += NT) { s_k[t] = 0; s_v[t] = 0; }
    __syncthreads();
    // a chunk holds at most LT/2 entries, so the table never fills
    for (int j = threadIdx.x; j < cc.len; j += NT) {
      const int key = skey[base + j];
      const acc w = sval[base + j];
      unsigned p = chash(key, LLG);
      for (;;) {
        const int k = atomicCAS(&s_k[p], 0, key + 1);
        if (k == 0 || k == key + 1) { atomicAdd(&s_v[p], w); break; }
        p = (p + 1) & (LT - 1);
      }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < LT; t += NT) {
      const int k = s_k[t];
      if (k) {
        atomicAdd(&row[k - 1], s_v[t]);
        atomicOr(&bits[((k - 1) >> 5)], 1u << ((k - 1) & 31));
      }
    }
    __syncthreads();
  }
}

// present-key count of every DENSE_TILE-key tile of every dense row (bitmap popcounts)
__global__ __launch_bounds__(128) void k_dense_count(const uint32_t* __restrict__ dbits, int nslots, int64_t words,
                                                     int ntiles, int32_t* __restrict__ tcnt) {
  __shared__ int32_t sh[17];
  constexpr int TW = DENSE_TILE / 32;  // words per tile (one per thread)
  const int total = nslots * ntiles;
  for (int t = blockIdx.x; t < total; t += gridDim.x) {
    const int slot = t / ntiles, tile = t % ntiles;
    const int64_t wi = (int64_t)tile * TW + threadIdx.x;
    const int cnt = wi < words ? __popc(dbits[(int64_t)slot * words + wi]) : 0;
    int tot;
    block_excl_scan<int32_t>(cnt, sh, &tot);
    if (threadIdx.x == 0) tcnt[t] = tot;
  }
}

// ordered extraction: each tile writes its present keys (ascending) at the
// row's running position and clears the accumulator and bitmap words it read
template <class T>
__global__ __launch_bounds__(128) void k_dense_scatter(typename T::acc_t* __restrict__ dense,
                                                       uint32_t* __restrict__ dbits, int nslots, int32_t C,
                                                       int64_t words, int ntiles,
                                                       const int32_t* __restrict__ slot_c,
                                                       const int64_t* __restrict__ toff,
                                                       const int64_t* __restrict__ eoff,
                                                       int32_t* __restrict__ skey,
                                                       typename T::acc_t* __restrict__ sval,
                                                       int32_t* __restrict__ dcnt) {
  using acc = typename T::acc_t;
  constexpr int TW = DENSE_TILE / 32;
  __shared__ int32_t sh[17];
  const int total = nslots * ntiles;
  for (int t = blockIdx.x; t < total; t += gridDim.x) {
    const int slot = t / ntiles, tile = t % ntiles;
    acc* row = dense + (int64_t)slot * C;
    uint32_t* bits = dbits + (int64_t)slot * words;
    const int c = slot_c[slot];
    const int64_t rbase = toff[slot * ntiles];
    const int64_t wi = (int64_t)tile * TW + threadIdx.x;
    uint32_t wbits = wi < words ? bits[wi] : 0u;
    int tot;
    const int pos = block_excl_scan<int32_t>(__popc(wbits), sh, &tot);
    int64_t out = eoff[c] + (toff[t] - rbase) + pos;
    if (wbits) bits[wi] = 0;
    while (wbits) {
      const int bit = __ffs(wbits) - 1;
      wbits &= wbits - 1;
      const int k = (int)(wi * 32 + bit);
      skey[out] = k;
      sval[out] = row[k];
      row[k] = 0;
      out++;
    }
    if (tile == 0 && threadIdx.x == 0) dcnt[c] = (int32_t)(toff[(slot + 1) * ntiles] - rbase);
  }
}

// copy each merged row to its final place: 16-lane groups for rows of at
// most 4096 staged entries, one workgroup per row for the listed long rows
template <class T>
__global__ __launch_bounds__(256) void k_compact(const int32_t* __restrict__ tlist, const int32_t* __restrict__ tstart,
                                                  const int64_t* __restrict__ eoff,
                                                  const int64_t* __restrict__ noff, const int32_t* __restrict__ dcnt,
                                                  const int32_t* __restrict__ skey,
                                                  const typename T::acc_t* __restrict__ sval,
                                                  int32_t* __restrict__ nbr, typename T::w_t* __restrict__ w) {
  const int lane = threadIdx.x & 15;
  const int ng = gridDim.x * (blockDim.x >> 4);
  // rows of the tiers up to 4096 entries; longer ones: k_compact_list
  const int lo = tstart[TIER_T16], hi = tstart[TIER_MID3 + 1];
  for (int i = lo + blockIdx.x * (blockDim.x >> 4) + (threadIdx.x >> 4); i < hi; i += ng) {
    const int c = tlist[i];
    const int64_t src = eoff[c], dst = noff[c];
    const int D = dcnt[c];
    for (int t = lane; t < D; t += 16) {
      nbr[dst + t] = skey[src + t];
      w[dst + t] = (typename T::w_t)sval[src + t];
    }
  }
}

template <class T>
__global__ __launch_bounds__(256) void k_compact_list(const int32_t* __restrict__ list, int nlist,
                                                       const int64_t* __restrict__ eoff,
                                                       const int64_t* __restrict__ noff,
                                                       const int32_t* __restrict__ dcnt,
                                                       const int32_t* __restrict__ skey,
                                                       const typename T::acc_t* __restrict__ sval,
                                                       int32_t* __restrict__ nbr, typename T::w_t* __restrict__ w) {
  for (int i = blockIdx.x; i < nlist; i += gridDim.x) {
    const int c = list[i];
    const int64_t src = eoff[c], dst = noff[c];
    const int D = dcnt[c];
    for (int t = threadIdx.x; t < D; t += 256) {
      nbr[dst + t] = skey[src + t];
      w[dst + t] = (typename T::w_t)sval[src + t];
    }
  }
}

__global__ void k_compose(int32_t* __restrict__ fin, int32_t n0, const int32_t* __restrict__ label) {
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n0; v += gridDim.x * blockDim.x)
    fin[v] = label[fin[v]];
}

void compose_map(int32_t* fin, int32_t n0, const int32_t* label, hipStream_t s) {
  if (n0 <= 0) return;
  k_compose<<<std::min((n0 + 255) / 256, 8192), 256, 0, s>>>(fin, n0, label);
}

static inline int nblk(int64_t n, int per = 256, int cap = 8192) {
  int64_t b = (n + per - 1) / per;
  if (b < 1) b = 1;
  return (int)std::min<int64_t>(b, cap);
}

template <class T>
int contract_prepare(const DevGraph<T>& g, const int32_t* comm, ContractBufs<T>& cb, ContractPlan* plan,
                     hipStream_t s) {
  constexpr long long LT = T::FP ? 8192 : 16384;
  const int32_t n = g.n;
  fill<int32_t>(cb.cnt, n, 0, s);
  k_hist<<<nblk(n), 256, 0, s>>>(comm, n, cb.cnt);
  uint8_t* flag = (uint8_t*)cb.i32_tmp;  // n bytes fit in the n-int scratch
  k_nonempty<<<nblk(n), 256, 0, s>>>(cb.cnt, n, flag);
  scan_exclusive<uint8_t, int32_t>(flag, cb.nid, n, cb.i32_tmp + (n + 3) / 4 + 1, cb.i32_scalar, s);
  k_label<<<nblk(n), 256, 0, s>>>(comm, cb.nid, n, cb.label);
  int32_t* hp32 = (int32_t*)cb.hpin;
  int64_t* hp64 = (int64_t*)cb.hpin;
  if (hipMemcpyAsync(hp32, cb.i32_scalar, sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
  if (hipStreamSynchronize(s) != hipSuccess) return -1;
  const int32_t C = hp32[0];
  fill<int64_t>(cb.ecnt, C + 1, 0, s);
  k_ecount<<<nblk(n), 256, 0, s>>>(g.off, n, cb.label, cb.ecnt);
  scan_exclusive<int64_t, int64_t>(cb.ecnt, cb.eoff, C, cb.i64_tmp, cb.eoff + C, s);
  fill<int32_t>(cb.nbig, 2, 0, s);  // big rows, dense rows
  int32_t* dense_list = cb.big_list + n;
  k_classify<<<nblk(C), 256, 0, s>>>(cb.ecnt, C, LT, cb.big_list, cb.nbig, cb.i64_tmp, dense_list, cb.tier);
  partition_bins(0, nullptr, cb.tier, nullptr, C, cb.tlist, cb.tstart, cb.part_tmp, s);
  // hp32[0..1] = big and dense row counts, hp32[2..] = the tier starts
  if (hipMemcpyAsync(hp32, cb.nbig, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
  if (hipMemcpyAsync(hp32 + 2, cb.tstart, sizeof plan->tstart, hipMemcpyDeviceToHost, s) != hipSuccess)
    return -1;
  if (hipStreamSynchronize(s) != hipSuccess) return -1;
  const int32_t nbig = hp32[0], ndense = hp32[1];
  std::copy(hp32 + 2, hp32 + 2 + NBINS + 1, plan->tstart);
  int64_t slots = 0;
  if (nbig > 0) {
    scan_exclusive<int64_t, int64_t>(cb.i64_tmp, cb.big_tab, nbig, cb.i64_tmp + nbig + 1, cb.big_tab + nbig, s);
    if (hipMemcpyAsync(hp64, cb.big_tab + nbig, sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess)
      return -1;
    if (hipStreamSynchronize(s) != hipSuccess) return -1;
    slots = hp64[0];
  }
  plan->C = C;
  plan->nbig = nbig;
  plan->slots = slots;
  plan->ndense = ndense;
  plan->dense_c.clear();
  plan->chunks.clear();
  if (ndense > 0) {
    // (community, staged length) of every dense row, lengths gathered on the device
    int64_t* lens = cb.i64_tmp;  // free again after the big-row scan
    k_dense_lengths<<<nblk(ndense), 256, 0, s>>>(dense_list, ndense, cb.ecnt, lens);
    std::vector<char> pv;
    const int64_t bytes = (int64_t)(sizeof(int64_t) + sizeof(int32_t)) * ndense;
    char* hv = cb.hpin;
    if (bytes > cb.hpin_bytes) {
      pv.resize(bytes);
      hv = pv.data();
    }
    int64_t* hl = (int64_t*)hv;
    int32_t* hc = (int32_t*)(hv + sizeof(int64_t) * ndense);
    if (hipMemcpyAsync(hl, lens, sizeof(int64_t) * ndense, hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
    if (hipMemcpyAsync(hc, dense_list, sizeof(int32_t) * ndense, hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
    if (hipStreamSynchronize(s) != hipSuccess) return -1;
    // dense rows in ascending community order (deterministic slot assignment)
    std::vector<int32_t> idx(ndense);
    for (int32_t i = 0; i < ndense; i++) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) { return hc[a] < hc[b]; });
    plan->dense_c.resize(ndense);
    const int64_t CH = LT / 2;  // entries per chunk: the LDS table stays at most half full
    for (int32_t i = 0; i < ndense; i++) {
      const int32_t c = hc[idx[i]];
      const int64_t L = hl[idx[i]];
      plan->dense_c[i] = c;
      for (int64_t st = 0; st < L; st += CH)
        plan->chunks.push_back({i, c, st, (int32_t)std::min<int64_t>(CH, L - st)});
    }
  }
  return 0;
}

template <class T>
int contract_build(const DevGraph<T>& g, ContractBufs<T>& cb, const ContractPlan& plan,
                   ContractOut<T>& out, int64_t* m_out, hipStream_t s) {
  constexpr int LT = T::FP ? 8192 : 16384;
  const int32_t n = g.n;
  const int32_t C = plan.C;
  fill<int64_t>(cb.ecur, C, 0, s);
  unsigned long long* ecur = (unsigned long long*)cb.ecur;
  k_stage<T><<<nblk(n, 256, 16384), 256, 0, s>>>(g, cb.label, cb.eoff, ecur, cb.skey, cb.sval);
  k_stage_rows<64, T><<<nblk(plan.caps[6], 4, 8192), 256, 0, s>>>(g, cb.order, cb.bstart, 6, 7, cb.label, cb.eoff,
                                                                    ecur, cb.skey, cb.sval);
  k_stage_rows<256, T><<<nblk(plan.caps[7] + plan.caps[8], 1, 4096), 256, 0, s>>>(
      g, cb.order, cb.bstart, 7, NBINS, cb.label, cb.eoff, ecur, cb.skey, cb.sval);
  fill<int32_t>(cb.dcnt, C + 1, 0, s);
  const int32_t* ts = plan.tstart;
  auto tcount = [&](int t) { return ts[t + 1] - ts[t]; };
  if (tcount(TIER_T16) > 0)
    k_merge_tiny<16, TIER_T16, T><<<nblk(tcount(TIER_T16), 16, 16384), 256, 0, s>>>(
        cb.ecnt, cb.eoff, cb.tlist, cb.tstart, cb.skey, cb.sval, cb.dcnt);
  if (tcount(TIER_T32) > 0)
    k_merge_tiny<32, TIER_T32, T><<<nblk(tcount(TIER_T32), 8, 16384), 256, 0, s>>>(
        cb.ecnt, cb.eoff, cb.tlist, cb.tstart, cb.skey, cb.sval, cb.dcnt);
  if (tcount(TIER_T64) > 0)
    k_merge_tiny<64, TIER_T64, T><<<nblk(tcount(TIER_T64), 4, 16384), 256, 0, s>>>(
        cb.ecnt, cb.eoff, cb.tlist, cb.tstart, cb.skey, cb.sval, cb.dcnt);
  if (tcount(TIER_MID1) > 0)
    k_merge_mid<64, 64, 512, T><<<nblk(tcount(TIER_MID1), 1, 16384), 64, 0, s>>>(
        cb.ecnt, cb.eoff, C, cb.tlist, cb.tstart, TIER_MID1, cb.skey, cb.sval, cb.dcnt);
  if (tcount(TIER_MID2) > 0)
    k_merge_mid<128, 512, 2048, T><<<nblk(tcount(TIER_MID2), 1, 8192), 128, 0, s>>>(
        cb.ecnt, cb.eoff, C, cb.tlist, cb.tstart, TIER_MID2, cb.skey, cb.sval, cb.dcnt);
  if (tcount(TIER_MID3) > 0)
    k_merge_mid<256, 2048, 4096, T><<<nblk(tcount(TIER_MID3), 1, 4096), 256, 0, s>>>(
        cb.ecnt, cb.eoff, C, cb.tlist, cb.tstart, TIER_MID3, cb.skey, cb.sval, cb.dcnt);
  if (plan.nbig > 0)
    k_merge_big<LT, T><<<std::min(plan.nbig, 2048), 1024, 0, s>>>(cb.ecnt, cb.eoff, C, cb.big_list, plan.nbig,
                                                                   cb.big_tab, cb.arena_keys, cb.arena_vals,
                                                                   cb.skey, cb.sval, cb.dcnt);
  if (plan.ndense > 0) {
    const int ntiles = (C + DENSE_TILE - 1) / DENSE_TILE;
    const int per_batch = (int)std::max<int64_t>(1, std::min<int64_t>(plan.ndense, cb.dense_slots));
    size_t ci = 0;
    for (int s0 = 0; s0 < plan.ndense; s0 += per_batch) {
      const int nb = std::min(per_batch, plan.ndense - s0);
      // chunks of this batch, slots renumbered from 0
      std::vector<DenseChunk> ch;
      while (ci < plan.chunks.size() && plan.chunks[ci].slot < s0 + nb) {
        DenseChunk d = plan.chunks[ci++];
        d.slot -= s0;
        ch.push_back(d);
      }
      if (hipMemcpyAsync(cb.chunks, ch.data(), sizeof(DenseChunk) * ch.size(), hipMemcpyHostToDevice, s) != hipSuccess)
        return -1;
      if (hipMemcpyAsync(cb.dense_c, plan.dense_c.data() + s0, sizeof(int32_t) * nb, hipMemcpyHostToDevice, s) !=
          hipSuccess)
        return -1;
      const int64_t words = (C + 31) / 32;
      k_dense_accum<LT, T><<<(int)std::min<size_t>(ch.size(), 4096), 1024, 0, s>>>(
          cb.chunks, (int)ch.size(), cb.eoff, C, cb.skey, cb.sval, cb.dense, cb.dense_bits, words);
      k_dense_count<<<std::min(nb * ntiles, 16384), 128, 0, s>>>(cb.dense_bits, nb, words, ntiles, cb.dense_tcnt);
      scan_exclusive<int32_t, int64_t>(cb.dense_tcnt, cb.dense_toff, (int64_t)nb * ntiles, cb.i64_tmp,
                                       cb.dense_toff + (int64_t)nb * ntiles, s);
      k_dense_scatter<T><<<std::min(nb * ntiles, 16384), 128, 0, s>>>(cb.dense, cb.dense_bits, nb, C, words,
                                                                       ntiles, cb.dense_c, cb.dense_toff, cb.eoff,
                                                                       cb.skey, cb.sval, cb.dcnt);
      if (hipStreamSynchronize(s) != hipSuccess) return -1;  // host chunk vector reused
    }
  }
  scan_exclusive<int32_t, int64_t>(cb.dcnt, out.off, C, cb.i64_tmp, out.off + C, s);
  const int32_t nsmall = ts[TIER_MID3 + 1] - ts[TIER_T16];
  if (nsmall > 0)
    k_compact<T><<<nblk(nsmall, 16, 16384), 256, 0, s>>>(cb.tlist, cb.tstart, cb.eoff, out.off, cb.dcnt, cb.skey,
                                                         cb.sval, out.nbr, out.w);
  if (plan.nbig > 0)
    k_compact_list<T><<<std::min(plan.nbig, 8192), 256, 0, s>>>(cb.big_list, plan.nbig, cb.eoff, out.off, cb.dcnt,
                                                                 cb.skey, cb.sval, out.nbr, out.w);
  if (plan.ndense > 0)
    k_compact_list<T><<<std::min(plan.ndense, 8192), 256, 0, s>>>(cb.big_list + n, plan.ndense, cb.eoff, out.off,
                                                                   cb.dcnt, cb.skey, cb.sval, out.nbr, out.w);
  int64_t* hp64 = (int64_t*)cb.hpin;
  if (hipMemcpyAsync(hp64, out.off + C, sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
  if (hipStreamSynchronize(s) != hipSuccess) return -1;
  *m_out = hp64[0];
  return 0;
}

#define GALA_CINST(T)                                                                               \
  template int contract_prepare<T>(const DevGraph<T>&, const int32_t*, ContractBufs<T>&, ContractPlan*, \
                                   hipStream_t);                                                    \
  template int contract_build<T>(const DevGraph<T>&, ContractBufs<T>&, const ContractPlan&,         \
                                 ContractOut<T>&, int64_t*, hipStream_t);
GALA_CINST(IntW)
GALA_CINST(FltW)

}  // namespace gala
