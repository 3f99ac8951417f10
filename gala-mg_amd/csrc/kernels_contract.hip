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
//                 the wide path), overflow into a global slice, register bitonic
//                 sort of <= LT keys (LDS only for exchanges between waves)
//   L >  LT       dense rows: the segment is cut into chunks, one workgroup per
//                 chunk pre-aggregates in LDS; the chunks' pairs are packed row
//                 after row, radix sorted by key within each row and their runs
//                 merged, so the row comes out in key order
// -> exclusive scan of the distinct counts -> compaction into the new CSR.
#include "gala_kernels.h"
#include "gala_device.h"
#include <algorithm>
#include <climits>
#include <vector>

namespace gala {

__device__ __forceinline__ unsigned chash(int key, int lg) {
  return ((unsigned)key * 2654435769u) >> (32 - lg);
}
__host__ __device__ __forceinline__ int clog2(long long x) {
  int l = 0;
  while ((1ll << l) < x) l++;
  return l;
}

// member count and row-length sum of every community, and the offset of every
// vertex's row inside its community's staging segment.  A workgroup takes a
// tile of CT vertices at a time and pre-aggregates them in an LDS hash keyed by
// community: each vertex's row length is added to its community's tile sum
// (the returned partial sum is the row's offset inside the tile's share), and
// each community's tile sum is added to the global sum with one atomic whose
// returned value is the tile's share's offset inside the segment.  So a large
// community costs one global atomic per tile instead of one per member
// (per-member atomics on a few hot addresses serialise), and staging needs no
// atomics at all.
constexpr int CT = 1024, CTS = 2 * CT, CT_PT = CT / 256;
// packed (every row-length sum < 2^32): dsum[c] holds count * 2^32 + row-length
// sum, so a (community, tile) pair costs one global atomic (k_unpack_sums
// separates them); otherwise cnt and dsum take one atomic each.
__global__ __launch_bounds__(256) void k_comm_sums(const int32_t* __restrict__ comm, const int64_t* __restrict__ off,
                                                   const int64_t* __restrict__ end, int32_t n, int32_t* __restrict__ cnt,
                                                   unsigned long long* __restrict__ dsum, int64_t* __restrict__ voff,
                                                   int packed) {
  __shared__ int s_key[CTS];
  __shared__ int s_cnt[CTS];
  __shared__ unsigned long long s_sum[CTS];
  __shared__ unsigned long long s_base[CTS];
  for (int t = threadIdx.x; t < CTS; t += blockDim.x) { s_key[t] = 0; s_cnt[t] = 0; s_sum[t] = 0; }
  __syncthreads();
  constexpr int LG = 11;  // CTS = 2^LG
  for (int64_t base = (int64_t)blockIdx.x * CT; base < n; base += (int64_t)gridDim.x * CT) {
    int slot[CT_PT];
    unsigned long long loc[CT_PT];
#pragma unroll
    for (int r = 0; r < CT_PT; r++) {
      slot[r] = -1;
      const int64_t v = base + threadIdx.x + r * 256;
      if (v >= n) continue;
      const int c = comm[v];
      const unsigned long long d = (unsigned long long)(end[v] - off[v]);
      unsigned p = ((unsigned)c * 2654435769u) >> (32 - LG);
      for (int probe = 0; probe < CTS; probe++) {
        const int k = atomicCAS(&s_key[p], 0, c + 1);
        if (k == 0 || k == c + 1) {
          atomicAdd(&s_cnt[p], 1);
          loc[r] = atomicAdd(&s_sum[p], d);
          slot[r] = (int)p;
          break;
        }
        p = (p + 1) & (CTS - 1);
      }
    }
    __syncthreads();
    // the tile's communities: every global atomic of a thread is issued before
    // any of their returned offsets is used (one L2 round trip, not one per slot)
    constexpr int FL = CTS / 256;
    int fk[FL];
    unsigned long long fb[FL];
#pragma unroll
    for (int r = 0; r < FL; r++) fk[r] = s_key[threadIdx.x + r * 256];
#pragma unroll
    for (int r = 0; r < FL; r++) {
      const int t = threadIdx.x + r * 256;
      fb[r] = 0;
      if (fk[r]) {
        if (packed) {
          fb[r] = atomicAdd(&dsum[fk[r] - 1], ((unsigned long long)s_cnt[t] << 32) + s_sum[t]) & 0xffffffffull;
        } else {
          atomicAdd(&cnt[fk[r] - 1], s_cnt[t]);
          fb[r] = atomicAdd(&dsum[fk[r] - 1], s_sum[t]);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < FL; r++)
      if (fk[r]) s_base[threadIdx.x + r * 256] = fb[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < CT_PT; r++)
      if (slot[r] >= 0) voff[base + threadIdx.x + r * 256] = (int64_t)(s_base[slot[r]] + loc[r]);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < FL; r++) {
      const int t = threadIdx.x + r * 256;
      if (fk[r]) {
        s_key[t] = 0;
        s_cnt[t] = 0;
        s_sum[t] = 0;
      }
    }
    __syncthreads();
  }
}

// row ends of a gapped contracted graph: end[c] = off[c] + distinct count.
// The number of entries of the new graph is the sum of the MTOTAL_SLOTS
// slots of total (zero on entry; slot k at total[16 k], one L2 line each),
// workgroup b adding its sum to slot b % MTOTAL_SLOTS.
__global__ __launch_bounds__(256) void k_row_ends(const int64_t* __restrict__ off, const int32_t* __restrict__ dcnt,
                                                   int32_t C, int64_t* __restrict__ end,
                                                   unsigned long long* __restrict__ total) {
  __shared__ unsigned long long s_t[4];
  unsigned long long t = 0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < C; c += gridDim.x * blockDim.x) {
    end[c] = off[c] + dcnt[c];
    t += (unsigned long long)dcnt[c];
  }
  t = wave_sum(t);
  if ((threadIdx.x & 63) == 0) s_t[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    t = s_t[0] + s_t[1] + s_t[2] + s_t[3];
    if (t) atomicAdd(&total[16 * (blockIdx.x % MTOTAL_SLOTS)], t);
  }
}

// entries of every new community: ecnt[nid[c]] = row-length sum of c
__global__ void k_ecount_comm(const int32_t* __restrict__ cnt, const int32_t* __restrict__ nid,
                              const unsigned long long* __restrict__ dsum, int32_t n, int64_t* __restrict__ ecnt) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
    if (cnt[c] > 0) ecnt[nid[c]] = (int64_t)dsum[c];
}

// packed community sums -> member counts, non-empty flags, row-length sums
__global__ void k_unpack_sums(unsigned long long* __restrict__ dsum, int32_t n, int32_t* __restrict__ cnt,
                              uint8_t* __restrict__ flag) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    const unsigned long long x = dsum[c];
    const int32_t k = (int32_t)(x >> 32);
    cnt[c] = k;
    flag[c] = k > 0;
    dsum[c] = x & 0xffffffffull;
  }
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

// staged length of each listed dense row
__global__ void k_dense_lengths(const int32_t* __restrict__ list, int32_t nlist, const int64_t* __restrict__ ecnt,
                                int64_t* __restrict__ len) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nlist; i += gridDim.x * blockDim.x) len[i] = ecnt[list[i]];
}

// classify rows: 4096 < L <= lt -> big_list (with the size of its global
// overflow slice); L > lt -> dense_list
// Row length tiers of the merge: 0: 1..16, 1: 17..32, 2: 33..64, 3: 65..128,
// 4: 129..512, 5: 513..1024, 6: 1025..2048, 7: 2049..4096; longer rows go to
// the big and dense lists, empty rows nowhere.
constexpr int TIER_T16 = 0, TIER_T32 = 1, TIER_T64 = 2, TIER_MID0 = 3, TIER_MID1 = 4, TIER_MID1K = 5,
              TIER_MID2 = 6, TIER_MID3 = 7;
static_assert(TIER_MID3 < NBINS, "the tiers are partitioned like degree bins");

// C is read on the device (rows c >= C get tier -1), so the host needs no
// readback before the launch; n is an upper bound of C.
__global__ void k_classify(const int64_t* __restrict__ ecnt, const int32_t* __restrict__ Cdev, int32_t n,
                           long long lt, int32_t* __restrict__ big_list, int32_t* __restrict__ counts,
                           int64_t* __restrict__ big_tsz, int32_t* __restrict__ dense_list,
                           int8_t* __restrict__ tier) {
  const int32_t C = *Cdev;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    if (c >= C) {
      tier[c] = -1;
      continue;
    }
    const long long L = ecnt[c];
    tier[c] = L == 0 ? -1 : L <= 16 ? TIER_T16 : L <= 32 ? TIER_T32 : L <= 64 ? TIER_T64
            : L <= 128 ? TIER_MID0 : L <= 512 ? TIER_MID1 : L <= 1024 ? TIER_MID1K : L <= 2048 ? TIER_MID2 : L <= 4096 ? TIER_MID3 : -1;
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
// ---------------------------------------------------------------------------
// entries [o, o + d) of the CSR to skey/sval[base ..] as (label[nbr], w), NT
// lanes from `lane`: the row and label loads of STG_U entries per lane are
// issued before their stores (the stores may alias the CSR for the compiler)
constexpr int STG_U = 4;
template <int NT, class T>
__device__ __forceinline__ void stage_span(const DevGraph<T>& g, int64_t o, int64_t d, int lane,
                                           const int32_t* __restrict__ label, int32_t* __restrict__ skey,
                                           typename T::acc_t* __restrict__ sval, int64_t base) {
  using acc = typename T::acc_t;
  for (int64_t j0 = lane; j0 < d; j0 += (int64_t)NT * STG_U) {
    int u[STG_U];
    acc w[STG_U];
#pragma unroll
    for (int r = 0; r < STG_U; r++) {
      const int64_t j = j0 + (int64_t)r * NT;
      u[r] = j < d ? g.nbr[o + j] : 0;
      w[r] = j < d ? (acc)load_w<T>(g, o + j) : acc(1);
    }
#pragma unroll
    for (int r = 0; r < STG_U; r++) u[r] = j0 + (int64_t)r * NT < d ? label[u[r]] : 0;
    // unit weights (an unweighted input round): the merges take every staged
    // value as 1, so none is written
    const bool unit = !g.weighted();
#pragma unroll
    for (int r = 0; r < STG_U; r++) {
      const int64_t j = j0 + (int64_t)r * NT;
      if (j < d) {
        skey[base + j] = u[r];
        if (!unit) sval[base + j] = w[r];
      }
    }
  }
}

// rows of the vertices order[bstart[blo] .. bstart[bhi]) (the round's degree
// bins): NT lanes per vertex, each row at its offset from k_comm_sums
template <int NT, class T>
__global__ __launch_bounds__(256) void k_stage_rows(DevGraph<T> g, const int32_t* __restrict__ order,
                                                    const int32_t* __restrict__ bstart, int blo, int bhi,
                                                    const int32_t* __restrict__ label,
                                                    const int64_t* __restrict__ eoff,
                                                    const int64_t* __restrict__ voff,
                                                    int32_t* __restrict__ skey,
                                                    typename T::acc_t* __restrict__ sval) {
  constexpr int G = 256 / NT;  // vertices per block step
  const int lo = bstart[blo], hi = bstart[bhi];
  const int lane = threadIdx.x % NT;
  for (int i = lo + (blockIdx.x * blockDim.x + threadIdx.x) / NT; i < hi; i += gridDim.x * G) {
    const int v = order[i];
    const int64_t o = g.off[v];
    const int64_t d = g.end[v] - o;
    const int64_t base = eoff[label[v]] + voff[v];  // the row's place in its community's segment
    stage_span<NT>(g, o, d, lane, label, skey, sval, base);
  }
}

// rows of order[bstart[blo] .. bstart[bhi]) cut into STAGE_CH-entry chunks:
// workgroup (x, y) stages chunks y, y + gridDim.y, ... of row x, so a hub row
// of 10^5 entries is spread over gridDim.y workgroups instead of one
constexpr int STAGE_CH = 4096;
template <class T>
__global__ __launch_bounds__(256) void k_stage_chunks(DevGraph<T> g, const int32_t* __restrict__ order,
                                                      const int32_t* __restrict__ bstart, int blo, int bhi,
                                                      const int32_t* __restrict__ label,
                                                      const int64_t* __restrict__ eoff,
                                                      const int64_t* __restrict__ voff,
                                                      int32_t* __restrict__ skey,
                                                      typename T::acc_t* __restrict__ sval) {
  const int i = bstart[blo] + blockIdx.x;
  if (i >= bstart[bhi]) return;  // uniform per workgroup
  const int v = order[i];
  const int64_t o = g.off[v];
  const int64_t d = g.end[v] - o;
  const int64_t row = eoff[label[v]] + voff[v];
  for (int64_t c0 = (int64_t)blockIdx.y * STAGE_CH; c0 < d; c0 += (int64_t)gridDim.y * STAGE_CH) {
    const int64_t len = d - c0 < STAGE_CH ? d - c0 : STAGE_CH;
    stage_span<256>(g, o + c0, len, (int)threadIdx.x, label, skey, sval, row + c0);
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
                                                     int32_t* __restrict__ dcnt, int unit) {
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
    if (lane < L) { key = skey[base + lane]; val = unit ? acc(1) : sval[base + lane]; }
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
                                                   int32_t* __restrict__ dcnt, int canonical, int unit) {
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
    // L <= LMAX = NT * U: every entry of a thread is loaded before the
    // inserts; equal keys of a wave are combined first
    constexpr int U = LMAX / NT;
    int key[U];
    acc w[U];
#pragma unroll
    for (int r = 0; r < U; r++) {
      const long long j = threadIdx.x + r * NT;
      key[r] = j < L ? skey[base + j] : -1;
      w[r] = j < L ? (unit ? acc(1) : sval[base + j]) : acc(0);
    }
#pragma unroll
    for (int r = 0; r < U; r++) {
      wave_combine_insert<acc>(key[r], w[r], key[r] >= 0, [&](int kk, acc ww) {
        unsigned p = chash(kk, lg);
        for (;;) {
          const int k = atomicCAS(&h_k[p], 0, kk + 1);
          if (k == 0 || k == kk + 1) { atomicAdd(&h_v[p], ww); break; }
          p = (p + 1) & (tsz - 1);
        }
      });
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
    if (canonical) {
      sort_row<NT, LMAX / NT, acc>(s_k, s_v, D, h_k, h_v, skey, sval, base);
    } else {
      for (int t = threadIdx.x; t < D; t += NT) { skey[base + t] = s_k[t]; sval[base + t] = s_v[t]; }
    }
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
                                                     int32_t* __restrict__ dcnt, int canonical, int unit) {
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
    // L <= LT = NT * PER: every entry of a thread is loaded before the
    // inserts; equal keys of a wave are combined first
    int key[PER];
    acc w[PER];
#pragma unroll
    for (int r = 0; r < PER; r++) {
      const long long j = threadIdx.x + r * NT;
      key[r] = j < L ? skey[base + j] : -1;
      w[r] = j < L ? (unit ? acc(1) : sval[base + j]) : acc(0);
    }
#pragma unroll
    for (int r = 0; r < PER; r++) {
      wave_combine_insert<acc>(key[r], w[r], key[r] >= 0, [&](int kk, acc ww) {
        unsigned p = chash(kk, LLG);
        bool placed = false;
        for (int probe = 0; probe < PMAX; probe++) {
          const int k = atomicCAS(&s_k[p], 0, kk + 1);
          if (k == 0 || k == kk + 1) { atomicAdd(&s_v[p], ww); placed = true; break; }
          p = (p + 1) & (LT - 1);
        }
        if (!placed) {
          s_ovf = 1;
          unsigned q = chash(kk, glg);
          for (;;) {
            const int k = atomicCAS(&keys[q], 0, kk + 1);
            if (k == 0 || k == kk + 1) { atomicAdd(&vals[q], ww); break; }
            q = (q + 1) & (unsigned)(gsz - 1);
          }
        }
      });
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
    if (canonical) {
      sort_row<NT, LT / NT, acc>(s_k, s_v, D, s_k, s_v, skey, sval, base);
    } else {
      for (int t = threadIdx.x; t < D; t += NT) { skey[base + t] = s_k[t]; sval[base + t] = s_v[t]; }
    }
    if (threadIdx.x == 0) dcnt[c] = D;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// dense rows (L > LT): sort-based merge.
//  1. k_dense_agg: every chunk (<= LT/2 staged entries) is aggregated in an
//     LDS hash table and its distinct (key, value) pairs are written back to
//     the start of the chunk's span (ccnt[ch] of them); k_dense_pack then
//     packs the chunks' pairs row after row into one buffer.
//  2. the packed rows are cut into tiles of <= LT/2 pairs (a tile table
//     built on the host from the row lengths) and sorted by a stable LSD radix
//     sort over RB-bit digits of the key (RB = 10 when the keys fit in 20
//     bits, two passes; else 8), segmented by row: per pass k_rs_hist counts
//     the digits of every tile into hist[2^RB*first + digit*nt + tile-in-row],
//     one exclusive scan of that array gives every (row, digit, tile) its
//     absolute output position, and k_rs_scatter ranks each tile's pairs
//     stably (per wave with RB ballots, then across waves) and writes them to
//     the other packed buffer.
//  3. k_rs_heads / scan over tiles / k_rs_emit: equal keys are adjacent;
//     each run head writes its key and the run's sum at the run's rank from
//     the start of the row's staged segment, so the row's distinct pairs end
//     up in ascending key order, dcnt[c] = their count.
// A tile table entry (DenseChunk) holds: slot = row, c = community, start =
// packed offset of the tile, len = pairs, first/nch = the row's first tile
// and tile count.
// ---------------------------------------------------------------------------
// A chunk holds at most LT/2 = NT * U entries, so the table never fills.
// Every entry of a thread is in registers before the inserts, and the next
// chunk's entries are loaded before this chunk's table is read out, so a
// workgroup (one per CU: the table takes 128 KB of LDS) overlaps its HBM
// reads with its LDS work.  Equal keys of a wave are combined before the LDS
// atomics (a dense row's entries often carry only a few distinct keys); the
// table is cleared as it is read out.
template <int LT, class T>
__global__ __launch_bounds__(1024) void k_dense_agg(const DenseChunk* __restrict__ chunks, int nch,
                                                    const int64_t* __restrict__ eoff, int32_t* __restrict__ skey,
                                                    typename T::acc_t* __restrict__ sval, int32_t* __restrict__ ccnt,
                                                    int unit) {
  using acc = typename T::acc_t;
  constexpr int NT = 1024, LLG = __builtin_ctz(LT), U = LT / 2 / NT;
  __shared__ int s_k[LT];
  __shared__ acc s_v[LT];
  __shared__ int s_sh[17];
  for (int t = threadIdx.x; t < LT; t += NT) { s_k[t] = 0; s_v[t] = 0; }
  int key[U];
  acc w[U];
  int ch = blockIdx.x;
  int64_t base = 0;
  auto load = [&](int c2, int64_t& b2) {
    const DenseChunk cc = chunks[c2];
    b2 = eoff[cc.c] + cc.start;
#pragma unroll
    for (int r = 0; r < U; r++) {
      const int j = threadIdx.x + r * NT;
      key[r] = j < cc.len ? skey[b2 + j] : -1;
      w[r] = j < cc.len ? (unit ? acc(1) : sval[b2 + j]) : acc(0);
    }
  };
  if (ch < nch) load(ch, base);
  __syncthreads();  // table cleared
  for (; ch < nch; ch += gridDim.x) {
#pragma unroll
    for (int r = 0; r < U; r++) {
      wave_combine_insert<acc>(key[r], w[r], key[r] >= 0, [&](int kk, acc ww) {
        unsigned p = chash(kk, LLG);
        for (;;) {
          const int k = atomicCAS(&s_k[p], 0, kk + 1);
          if (k == 0 || k == kk + 1) { atomicAdd(&s_v[p], ww); break; }
          p = (p + 1) & (LT - 1);
        }
      });
    }
    const int64_t cbase = base;
    if (ch + (int)gridDim.x < nch) load(ch + gridDim.x, base);  // in flight during the read-out
    __syncthreads();
    int cnt = 0;
    for (int t = threadIdx.x; t < LT; t += NT) cnt += s_k[t] != 0;
    int D;
    int q = block_excl_scan<int>(cnt, s_sh, &D);
    for (int t = threadIdx.x; t < LT; t += NT) {
      const int k = s_k[t];
      if (k) {
        skey[cbase + q] = k - 1;
        sval[cbase + q] = s_v[t];
        q++;
        s_k[t] = 0;
        s_v[t] = acc(0);
      }
    }
    if (threadIdx.x == 0) ccnt[ch] = D;
    __syncthreads();
  }
}

// chunk ch's pairs to [pstart[ch], pstart[ch] + ccnt[ch]) of the packed buffer;
// bound[row] = packed start of every row, bound[nrows] = packed total
template <class T>
__global__ __launch_bounds__(256) void k_dense_pack(const DenseChunk* __restrict__ chunks, int nch,
                                                    const int32_t* __restrict__ ccnt, const int64_t* __restrict__ pstart,
                                                    const int64_t* __restrict__ eoff, const int32_t* __restrict__ skey,
                                                    const typename T::acc_t* __restrict__ sval,
                                                    int32_t* __restrict__ kout, typename T::acc_t* __restrict__ vout,
                                                    int nrows, int64_t* __restrict__ bound) {
  for (int ch = blockIdx.x; ch < nch; ch += gridDim.x) {
    const DenseChunk cc = chunks[ch];
    const int64_t src = eoff[cc.c] + cc.start, dst = pstart[ch];
    const int n = ccnt[ch];
    for (int i = threadIdx.x; i < n; i += 256) {
      kout[dst + i] = skey[src + i];
      vout[dst + i] = sval[src + i];
    }
    if (threadIdx.x == 0) {
      if (ch == cc.first) bound[cc.slot] = dst;
      if (ch == nch - 1) bound[nrows] = pstart[nch];
    }
  }
}

// digit counts of every tile (layout: see above)
// RB-bit digits (NB = 2^RB buckets): 8 bits, or 10 when the keys have at most
// 20 bits, so that two passes cover them
template <int RB>
__global__ __launch_bounds__(256) void k_rs_hist(const DenseChunk* __restrict__ tiles, int ntiles,
                                                 const int32_t* __restrict__ kin, int shift,
                                                 int32_t* __restrict__ hist) {
  constexpr int NB = 1 << RB;
  __shared__ int cnt[NB];
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const DenseChunk tt = tiles[t];
    const int32_t* src = kin + tt.start;
    for (int d = threadIdx.x; d < NB; d += 256) cnt[d] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < tt.len; i += 256) atomicAdd(&cnt[(src[i] >> shift) & (NB - 1)], 1);
    __syncthreads();
    for (int d = threadIdx.x; d < NB; d += 256)
      hist[(int64_t)NB * tt.first + (int64_t)d * tt.nch + (t - tt.first)] = cnt[d];
    __syncthreads();
  }
}

// stable scatter of every tile's pairs to their digit positions.  The tile is
// first ordered by digit in LDS, so that the global writes of a digit's run
// come from consecutive threads to consecutive addresses (written straight
// from the ranking, every lane of a wave would hit a different line).
template <int LT, int RB, class T>
__global__ __launch_bounds__(1024) void k_rs_scatter(const DenseChunk* __restrict__ tiles, int ntiles,
                                                     const int32_t* __restrict__ kin,
                                                     const typename T::acc_t* __restrict__ vin, int shift,
                                                     const int64_t* __restrict__ hscan, int32_t* __restrict__ kout,
                                                     typename T::acc_t* __restrict__ vout) {
  using acc = typename T::acc_t;
  constexpr int NW = 16, R = (LT / 2) / (NW * 64);  // rounds of 64 pairs per wave
  constexpr int NB = 1 << RB;
  static_assert(NB <= 1024, "one digit per thread in the digit scan");
  __shared__ int wcnt[NW][NB];
  __shared__ int64_t gofs[NB];
  __shared__ int lstart[NB];  // the tile's first position of each digit
  __shared__ int s_key[LT / 2];
  __shared__ acc s_val[LT / 2];
  __shared__ int sh[17];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const unsigned long long below_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const DenseChunk tt = tiles[t];
    const int n = tt.len;
    const int64_t b0 = tt.start;
    for (int q = threadIdx.x; q < NW * NB; q += 1024) (&wcnt[0][0])[q] = 0;
    for (int d = threadIdx.x; d < NB; d += 1024)
      gofs[d] = hscan[(int64_t)NB * tt.first + (int64_t)d * tt.nch + (t - tt.first)];
    __syncthreads();
    int key[R], rk[R];
    acc val[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int i = (w * R + r) * 64 + lane;
      const bool ok = i < n;
      key[r] = ok ? kin[b0 + i] : 0;
      val[r] = ok ? vin[b0 + i] : acc(0);
      const int d = (key[r] >> shift) & (NB - 1);
      unsigned long long peers = __ballot(ok);
#pragma unroll
      for (int bit = 0; bit < RB; bit++) {
        const bool on = (d >> bit) & 1;
        const unsigned long long bb = __ballot(on);
        peers &= on ? bb : ~bb;
      }
      const int below = __popcll(peers & below_mask);
      const int base = wcnt[w][d];
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
      __builtin_amdgcn_wave_barrier();
      if (ok && below == 0) wcnt[w][d] = base + __popcll(peers);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
      __builtin_amdgcn_wave_barrier();
      rk[r] = base + below;
    }
    __syncthreads();
    int dtot = 0;  // the tile's pairs of digit threadIdx.x
    if (threadIdx.x < NB) {  // exclusive prefix over the waves, per digit
      const int d = threadIdx.x;
      for (int q = 0; q < NW; q++) {
        const int x = wcnt[q][d];
        wcnt[q][d] = dtot;
        dtot += x;
      }
    }
    int32_t tot;
    const int dstart = block_excl_scan<int32_t>(dtot, sh, &tot);
    if (threadIdx.x < NB) lstart[threadIdx.x] = dstart;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int i = (w * R + r) * 64 + lane;
      if (i < n) {
        const int d = (key[r] >> shift) & (NB - 1);
        const int lp = lstart[d] + wcnt[w][d] + rk[r];
        s_key[lp] = key[r];
        s_val[lp] = val[r];
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += 1024) {
      const int k = s_key[i];
      const int d = (k >> shift) & (NB - 1);
      const int64_t pos = gofs[d] + (i - lstart[d]);
      kout[pos] = k;
      vout[pos] = s_val[i];
    }
    __syncthreads();
  }
}

// Run merge of the sorted packed pairs: equal keys are adjacent, and the
// first pair of a row or of a new key is a run head.  k_rs_heads counts each
// tile's heads; after an exclusive scan over the tiles, k_rs_emit ranks the
// heads (per wave segment of the tile, with ballots) and writes each head's
// key at its rank from the start of the row's staged segment with the sum of
// its run inside the head's 64-pair window (a segmented reduction over the
// window).  A window's leading pairs whose head lies before the window (a run
// longer than the rest of its window: hot keys of long rows meet one pair per
// chunk) leave their partial sum and the run's slot in cont; k_rs_cont adds
// those once every head has been written.  The row's first tile writes its
// distinct count.
constexpr int RS_T = 1024;  // k_rs_emit threads
constexpr int RS_WIN = 128;  // 64-pair windows per tile (tiles <= LT/2 = 8192 pairs)
__global__ __launch_bounds__(256) void k_rs_heads(const DenseChunk* __restrict__ tiles, int ntiles,
                                                  const int32_t* __restrict__ keys, int64_t* __restrict__ tcnt) {
  __shared__ int64_t sh[17];
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const DenseChunk tt = tiles[t];
    const int64_t seg0 = tiles[tt.first].start;
    int64_t c = 0;
    for (int i = threadIdx.x; i < tt.len; i += 256) {
      const int64_t g = tt.start + i;
      c += (g == seg0 || keys[g] != keys[g - 1]) ? 1 : 0;
    }
    int64_t tot;
    (void)block_excl_scan<int64_t>(c, sh, &tot);
    if (threadIdx.x == 0) tcnt[t] = tot;
  }
}

template <class T>
__global__ __launch_bounds__(RS_T) void k_rs_emit(const DenseChunk* __restrict__ tiles, int ntiles,
                                                  const int32_t* __restrict__ keys,
                                                  const typename T::acc_t* __restrict__ vals,
                                                  const int64_t* __restrict__ tpos, const int64_t* __restrict__ eoff,
                                                  int32_t* __restrict__ skey, typename T::acc_t* __restrict__ sval,
                                                  int32_t* __restrict__ dcnt, int64_t* __restrict__ cont_slot,
                                                  typename T::acc_t* __restrict__ cont_val) {
  using acc = typename T::acc_t;
  constexpr int NW = RS_T / 64;
  __shared__ int32_t wh[NW];  // heads in each wave's segment of the tile
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const DenseChunk tt = tiles[t];
    const int64_t seg0 = tiles[tt.first].start;
    const int64_t row0 = tpos[tt.first];
    if (t == tt.first && threadIdx.x == 0) dcnt[tt.c] = (int32_t)(tpos[tt.first + tt.nch] - row0);
    // wave w owns the tile's pairs [s0, s1), in windows of 64 consecutive pairs
    const int seg = (tt.len + NW * 64 - 1) / (NW * 64) * 64;
    const int s0 = w * seg, s1 = min(tt.len, s0 + seg);
    int cnt = 0;
    for (int b = s0; b < s1; b += 64) {  // uniform per wave
      const int i = b + lane;
      const int64_t g = tt.start + i;
      const bool hd = i < s1 && (g == seg0 || keys[g] != keys[g - 1]);
      cnt += __popcll(__ballot(hd));
    }
    if (lane == 0) wh[w] = cnt;
    __syncthreads();
    int64_t r = eoff[tt.c] + (tpos[t] - row0);  // output slot of the next head
    for (int q = 0; q < w; q++) r += wh[q];
    for (int b = s0; b < s1; b += 64) {
      const int i = b + lane;
      const int64_t g = tt.start + i;
      const bool ok = i < s1;
      const bool hd = ok && (g == seg0 || keys[g] != keys[g - 1]);
      const unsigned long long bal = __ballot(hd);
      // segment of the lane inside the window: heads at or before it (0: the
      // pairs before the window's first head)
      const int sid = __popcll(bal & (below | (1ull << lane)));
      acc x = ok ? vals[g] : acc(0);
#pragma unroll
      for (int k = 1; k < 64; k <<= 1) {  // suffix sums within the segment
        const acc xo = __shfl_down(x, k, 64);
        const int so = __shfl_down(sid, k, 64);
        if (lane + k < 64 && so == sid) x += xo;
      }
      if (hd) {
        const int64_t o = r + __popcll(bal & below);
        skey[o] = keys[g];
        sval[o] = x;
      }
      if (lane == 0) {
        const int win = t * RS_WIN + b / 64;
        cont_slot[win] = hd ? -1 : r - 1;
        cont_val[win] = x;
      }
      r += __popcll(bal);
    }
    __syncthreads();
  }
}

// the continuation sums of the windows, once every head's slot holds its key
template <class T>
__global__ __launch_bounds__(256) void k_rs_cont(const DenseChunk* __restrict__ tiles, int ntiles,
                                                 const int64_t* __restrict__ cont_slot,
                                                 const typename T::acc_t* __restrict__ cont_val,
                                                 typename T::acc_t* __restrict__ sval) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < (int64_t)ntiles * RS_WIN;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int t = (int)(q / RS_WIN), j = (int)(q % RS_WIN);
    if (j * 64 >= tiles[t].len) continue;
    const int64_t slot = cont_slot[q];
    if (slot >= 0) atomicAdd(&sval[slot], cont_val[q]);
  }
}

// Scratch layout of the dense-row sort (offsets in bytes, 256-byte aligned).
constexpr int64_t RS_NB_MAX = 1024;  // buckets of the widest radix digit (10 bits)
struct DenseLayout {
  int64_t ccnt, pstart, bound, tiles, hist, hscan, k0, k1, v0, v1, tcnt, tpos, cslot, cval, tmp, total;
};
// nch chunks, nrows rows, P staged entries; tiles of the packed rows are at
// most nch + nrows
static DenseLayout dense_layout(int64_t nch, int64_t nrows, int64_t P, int64_t vsz) {
  auto al = [](int64_t x) { return (x + 255) & ~(int64_t)255; };
  const int64_t nt = nch + nrows;
  DenseLayout L{};
  int64_t o = 0;
  L.ccnt = o; o = al(o + 4 * nch);
  L.pstart = o; o = al(o + 8 * (nch + 1));
  L.bound = o; o = al(o + 8 * (nrows + 1));
  L.tiles = o; o = al(o + (int64_t)sizeof(DenseChunk) * nt);
  L.hist = o; o = al(o + 4 * RS_NB_MAX * nt);
  L.hscan = o; o = al(o + 8 * (RS_NB_MAX * nt + 1));
  L.k0 = o; o = al(o + 4 * P);
  L.k1 = o; o = al(o + 4 * P);
  L.v0 = o; o = al(o + vsz * P);
  L.v1 = o; o = al(o + vsz * P);
  L.tcnt = o; o = al(o + 8 * nt);
  L.tpos = o; o = al(o + 8 * (nt + 1));
  L.cslot = o; o = al(o + 8 * RS_WIN * nt);
  L.cval = o; o = al(o + vsz * RS_WIN * nt);
  L.tmp = o; o = al(o + 8 * (scan_tmp_elems(std::max<int64_t>(P, RS_NB_MAX * nt)) + 1));
  L.total = o;
  return L;
}

int64_t dense_scratch_bytes(int64_t nchunks, int64_t nrows, int64_t entries, bool fp) {
  return dense_layout(nchunks, nrows, entries, fp ? 8 : 4).total;
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

// gathers the contraction plan readback into one block (layout: contract_prepare)
template <int SLOT64>
__global__ void k_plan_stage(const int32_t* __restrict__ C, const int32_t* __restrict__ nbig,
                             const int32_t* __restrict__ tstart, const int64_t* __restrict__ slots,
                             int64_t* __restrict__ stage) {
  int32_t* s32 = (int32_t*)stage;
  const int t = threadIdx.x;
  if (t == 0) s32[0] = *C;
  if (t < 2) s32[1 + t] = nbig[t];
  if (t <= NBINS) s32[3 + t] = tstart[t];
  if (t == 0) stage[SLOT64] = *slots;
}

template <class T>
int contract_prepare(const DevGraph<T>& g, const int32_t* comm, ContractBufs<T>& cb, ContractPlan* plan,
                     hipStream_t s) {
  constexpr long long LT = T::WIDE ? 8192 : 16384;
  const int32_t n = g.n;
  // count and row-length sum of a community in one word while every sum is
  // below 2^32 (g.m bounds the round's entries)
  const int packed = g.m < (1ll << 32) ? 1 : 0;
  if (!packed) fill<int32_t>(cb.cnt, n, 0, s);
  // the staging cursors are free until contract_build: row-length sums by community
  unsigned long long* dsum = (unsigned long long*)cb.ecur;
  fill<int64_t>(cb.ecur, n, 0, s);
  k_comm_sums<<<std::max(1, std::min((n + CT - 1) / CT, 2048)), 256, 0, s>>>(comm, g.off, g.end, n, cb.cnt, dsum,
                                                                             cb.voff, packed);
  uint8_t* flag = (uint8_t*)cb.i32_tmp;  // n bytes fit in the n-int scratch
  if (packed) k_unpack_sums<<<nblk(n), 256, 0, s>>>(dsum, n, cb.cnt, flag);
  else k_nonempty<<<nblk(n), 256, 0, s>>>(cb.cnt, n, flag);
  scan_exclusive<uint8_t, int32_t>(flag, cb.nid, n, cb.i32_tmp + (n + 3) / 4 + 1, cb.i32_scalar, s);
  k_label<<<nblk(n), 256, 0, s>>>(comm, cb.nid, n, cb.label);
  // Everything up to the one readback below is sized by n >= C, so C, the
  // tier starts, the big/dense counts and the big rows' overflow slots come
  // back with a single synchronisation.  Rows c >= C are empty: ecnt[c] = 0,
  // eoff[c] = the total, tier -1, big_tsz 0.
  fill<int64_t>(cb.ecnt, (int64_t)n + 1, 0, s);
  k_ecount_comm<<<nblk(n), 256, 0, s>>>(cb.cnt, cb.nid, dsum, n, cb.ecnt);
  scan_exclusive<int64_t, int64_t>(cb.ecnt, cb.eoff, n, cb.i64_tmp, cb.eoff + n, s);
  fill<int32_t>(cb.nbig, 2, 0, s);  // big rows, dense rows
  int32_t* dense_list = cb.big_list + n;
  int64_t* big_tsz = cb.i64_tmp;  // free again after the ecnt scan
  fill<int64_t>(big_tsz, n, 0, s);
  k_classify<<<nblk(n), 256, 0, s>>>(cb.ecnt, cb.i32_scalar, n, LT, cb.big_list, cb.nbig, big_tsz, dense_list,
                                     cb.tier);
  partition_bins(0, nullptr, nullptr, cb.tier, nullptr, n, cb.tlist, cb.tstart, cb.part_tmp, s);
  scan_exclusive<int64_t, int64_t>(big_tsz, cb.big_tab, n, cb.i64_tmp + n + 1, cb.big_tab + n, s);
  // pinned readback: [0] C, [1..2] big and dense row counts, [3..] tier starts,
  // then (8-byte aligned) the overflow slots of the big rows
  int32_t* hp32 = (int32_t*)cb.hpin;
  int64_t* hp64 = (int64_t*)cb.hpin;
  constexpr int SLOT64 = (3 + NBINS + 1 + 1) / 2;
  k_plan_stage<SLOT64><<<1, 64, 0, s>>>(cb.i32_scalar, cb.nbig, cb.tstart, cb.big_tab + n, cb.i64_scalar);
  if (hipMemcpyAsync(hp64, cb.i64_scalar, sizeof(int64_t) * (SLOT64 + 1), hipMemcpyDeviceToHost, s) != hipSuccess)
    return -1;
  if (hipStreamSynchronize(s) != hipSuccess) return -1;
  const int32_t C = hp32[0];
  const int32_t nbig = hp32[1], ndense = hp32[2];
  std::copy(hp32 + 3, hp32 + 3 + NBINS + 1, plan->tstart);
  const int64_t slots = nbig > 0 ? hp64[SLOT64] : 0;
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
    plan->dense_entries = 0;
    const int64_t CH = LT / 2;  // entries per chunk: the LDS table stays at most half full
    for (int32_t i = 0; i < ndense; i++) {
      const int32_t c = hc[idx[i]];
      const int64_t L = hl[idx[i]];
      plan->dense_c[i] = c;
      plan->dense_entries += L;
      const int32_t first = (int32_t)plan->chunks.size(), nc = (int32_t)((L + CH - 1) / CH);
      for (int64_t st = 0; st < L; st += CH)
        plan->chunks.push_back({i, c, st, (int32_t)std::min<int64_t>(CH, L - st), first, nc, 0});
    }
  }
  return 0;
}

template <class T>
int contract_build(const DevGraph<T>& g, ContractBufs<T>& cb, const ContractPlan& plan,
                   ContractOut<T>& out, int64_t* m_out, hipStream_t s) {
  constexpr int LT = T::WIDE ? 8192 : 16384;
  const int32_t n = g.n;
  const int32_t C = plan.C;
  // W lanes per row, W matched to the round's degree bins; every row goes to
  // its offset from k_comm_sums
  const int32_t* cp = plan.caps;
  const int64_t* voff = cb.voff;
  k_stage_rows<4, T><<<nblk(cp[0], 64, 8192), 256, 0, s>>>(g, cb.order, cb.bstart, 0, 1, cb.label, cb.eoff, voff,
                                                           cb.skey, cb.sval);
  k_stage_rows<8, T><<<nblk(cp[1], 32, 8192), 256, 0, s>>>(g, cb.order, cb.bstart, 1, 2, cb.label, cb.eoff, voff,
                                                           cb.skey, cb.sval);
  k_stage_rows<16, T><<<nblk(cp[2], 16, 8192), 256, 0, s>>>(g, cb.order, cb.bstart, 2, 3, cb.label, cb.eoff, voff,
                                                            cb.skey, cb.sval);
  k_stage_rows<32, T><<<nblk(cp[3] + cp[4], 8, 8192), 256, 0, s>>>(g, cb.order, cb.bstart, 3, 5, cb.label, cb.eoff,
                                                                   voff, cb.skey, cb.sval);
  k_stage_rows<64, T><<<nblk(cp[5] + cp[6], 4, 8192), 256, 0, s>>>(g, cb.order, cb.bstart, 5, 7, cb.label, cb.eoff,
                                                                   voff, cb.skey, cb.sval);
  if (cp[7] + cp[8] > 0)
    k_stage_chunks<T><<<dim3(cp[7] + cp[8], 16), 256, 0, s>>>(g, cb.order, cb.bstart, 7, NBINS, cb.label, cb.eoff,
                                                              voff, cb.skey, cb.sval);
  fill<int32_t>(cb.dcnt, C + 1, 0, s);
  const int canon = plan.canonical ? 1 : 0;
  const int unit = g.weighted() ? 0 : 1;  // staging wrote keys only: every staged value is 1
  const int32_t* ts = plan.tstart;
  auto tcount = [&](int t) { return ts[t + 1] - ts[t]; };
  if (tcount(TIER_T16) > 0)
    k_merge_tiny<16, TIER_T16, T><<<nblk(tcount(TIER_T16), 16, 16384), 256, 0, s>>>(
        cb.ecnt, cb.eoff, cb.tlist, cb.tstart, cb.skey, cb.sval, cb.dcnt, unit);
  if (tcount(TIER_T32) > 0)
    k_merge_tiny<32, TIER_T32, T><<<nblk(tcount(TIER_T32), 8, 16384), 256, 0, s>>>(
        cb.ecnt, cb.eoff, cb.tlist, cb.tstart, cb.skey, cb.sval, cb.dcnt, unit);
  if (tcount(TIER_T64) > 0)
    k_merge_tiny<64, TIER_T64, T><<<nblk(tcount(TIER_T64), 4, 16384), 256, 0, s>>>(
        cb.ecnt, cb.eoff, cb.tlist, cb.tstart, cb.skey, cb.sval, cb.dcnt, unit);
  if (tcount(TIER_MID0) > 0)
    k_merge_mid<64, 64, 128, T><<<nblk(tcount(TIER_MID0), 1, 16384), 64, 0, s>>>(
        cb.ecnt, cb.eoff, C, cb.tlist, cb.tstart, TIER_MID0, cb.skey, cb.sval, cb.dcnt, canon, unit);
  if (tcount(TIER_MID1) > 0)
    k_merge_mid<64, 128, 512, T><<<nblk(tcount(TIER_MID1), 1, 16384), 64, 0, s>>>(
        cb.ecnt, cb.eoff, C, cb.tlist, cb.tstart, TIER_MID1, cb.skey, cb.sval, cb.dcnt, canon, unit);
  if (tcount(TIER_MID1K) > 0)
    k_merge_mid<128, 512, 1024, T><<<nblk(tcount(TIER_MID1K), 1, 8192), 128, 0, s>>>(
        cb.ecnt, cb.eoff, C, cb.tlist, cb.tstart, TIER_MID1K, cb.skey, cb.sval, cb.dcnt, canon, unit);
  if (tcount(TIER_MID2) > 0)
    k_merge_mid<128, 1024, 2048, T><<<nblk(tcount(TIER_MID2), 1, 8192), 128, 0, s>>>(
        cb.ecnt, cb.eoff, C, cb.tlist, cb.tstart, TIER_MID2, cb.skey, cb.sval, cb.dcnt, canon, unit);
  if (tcount(TIER_MID3) > 0)
    k_merge_mid<256, 2048, 4096, T><<<nblk(tcount(TIER_MID3), 1, 4096), 256, 0, s>>>(
        cb.ecnt, cb.eoff, C, cb.tlist, cb.tstart, TIER_MID3, cb.skey, cb.sval, cb.dcnt, canon, unit);
  if (plan.nbig > 0)
    k_merge_big<LT, T><<<std::min(plan.nbig, 2048), 1024, 0, s>>>(cb.ecnt, cb.eoff, C, cb.big_list, plan.nbig,
                                                                   cb.big_tab, cb.arena_keys, cb.arena_vals,
                                                                   cb.skey, cb.sval, cb.dcnt, canon, unit);
  if (plan.ndense > 0) {
    using acc = typename T::acc_t;
    const int nch = (int)plan.chunks.size();
    const int nrows = plan.ndense;
    const int64_t P = plan.dense_entries;
    const DenseLayout L = dense_layout(nch, nrows, P, sizeof(acc));
    char* d = cb.dscratch;
    int32_t* ccnt = (int32_t*)(d + L.ccnt);
    int64_t* pstart = (int64_t*)(d + L.pstart);
    int64_t* bound = (int64_t*)(d + L.bound);
    DenseChunk* tiles = (DenseChunk*)(d + L.tiles);
    int32_t* hist = (int32_t*)(d + L.hist);
    int64_t* hscan = (int64_t*)(d + L.hscan);
    int32_t* kb[2] = {(int32_t*)(d + L.k0), (int32_t*)(d + L.k1)};
    acc* vb[2] = {(acc*)(d + L.v0), (acc*)(d + L.v1)};
    int64_t* tcnt = (int64_t*)(d + L.tcnt);
    int64_t* tpos = (int64_t*)(d + L.tpos);
    int64_t* cslot = (int64_t*)(d + L.cslot);
    acc* cval = (acc*)(d + L.cval);
    int64_t* tmp = (int64_t*)(d + L.tmp);
    if (hipMemcpyAsync(cb.chunks, plan.chunks.data(), sizeof(DenseChunk) * nch, hipMemcpyHostToDevice, s) !=
        hipSuccess)
      return -1;
    k_dense_agg<LT, T><<<std::min(nch, 4096), 1024, 0, s>>>(cb.chunks, nch, cb.eoff, cb.skey, cb.sval, ccnt,
                                                              unit);
    scan_exclusive<int32_t, int64_t>(ccnt, pstart, nch, tmp, pstart + nch, s);
    // pass 0 reads kb[1]; pass p writes kb[p & 1]
    k_dense_pack<T><<<std::min(nch, 8192), 256, 0, s>>>(cb.chunks, nch, ccnt, pstart, cb.eoff, cb.skey, cb.sval,
                                                         kb[1], vb[1], nrows, bound);
    // tile table of the packed rows
    std::vector<int64_t> hbv;
    int64_t* hb = (int64_t*)cb.hpin;
    if ((int64_t)sizeof(int64_t) * (nrows + 1) > cb.hpin_bytes) {
      hbv.resize(nrows + 1);
      hb = hbv.data();
    }
    if (hipMemcpyAsync(hb, bound, sizeof(int64_t) * (nrows + 1), hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
    if (hipStreamSynchronize(s) != hipSuccess) return -1;
    constexpr int64_t TL = LT / 2;
    std::vector<DenseChunk> tv;
    tv.reserve(nch + nrows);
    for (int32_t r = 0; r < nrows; r++) {
      const int64_t len = hb[r + 1] - hb[r];
      const int32_t first = (int32_t)tv.size(), nt = (int32_t)((len + TL - 1) / TL);
      for (int64_t o = 0; o < len; o += TL)
        tv.push_back({r, plan.dense_c[r], hb[r] + o, (int32_t)std::min<int64_t>(TL, len - o), first, nt, 0});
    }
    const int ntl = (int)tv.size();
    if (hipMemcpyAsync(tiles, tv.data(), sizeof(DenseChunk) * ntl, hipMemcpyHostToDevice, s) != hipSuccess)
      return -1;
    const int bits = C > 1 ? 32 - __builtin_clz((unsigned)(C - 1)) : 1;
    const int rb = bits <= 20 ? 10 : 8;  // two passes of 10 bits cover C <= 2^20
    const int npass = (bits + rb - 1) / rb;
    const int64_t nb = 1ll << rb;
    for (int p = 0; p < npass; p++) {
      const int32_t* kin = kb[(p + 1) & 1];
      const acc* vin = vb[(p + 1) & 1];
      if (rb == 10)
        k_rs_hist<10><<<std::min(ntl, 8192), 256, 0, s>>>(tiles, ntl, kin, rb * p, hist);
      else
        k_rs_hist<8><<<std::min(ntl, 8192), 256, 0, s>>>(tiles, ntl, kin, rb * p, hist);
      scan_exclusive<int32_t, int64_t>(hist, hscan, nb * ntl, tmp, hscan + nb * ntl, s);
      if (rb == 10)
        k_rs_scatter<LT, 10, T><<<std::min(ntl, 4096), 1024, 0, s>>>(tiles, ntl, kin, vin, rb * p, hscan, kb[p & 1],
                                                                    vb[p & 1]);
      else
        k_rs_scatter<LT, 8, T><<<std::min(ntl, 4096), 1024, 0, s>>>(tiles, ntl, kin, vin, rb * p, hscan, kb[p & 1],
                                                                   vb[p & 1]);
    }
    const int32_t* kf = kb[(npass - 1) & 1];
    const acc* vf = vb[(npass - 1) & 1];
    k_rs_heads<<<std::min(ntl, 8192), 256, 0, s>>>(tiles, ntl, kf, tcnt);
    scan_exclusive<int64_t, int64_t>(tcnt, tpos, ntl, tmp, tpos + ntl, s);
    static_assert(LT / 2 <= RS_WIN * 64, "a tile's windows fit its continuation slots");
    k_rs_emit<T><<<std::min(ntl, 4096), RS_T, 0, s>>>(tiles, ntl, kf, vf, tpos, cb.eoff, cb.skey, cb.sval, cb.dcnt,
                                                      cslot, cval);
    k_rs_cont<T><<<(int)std::min<int64_t>(((int64_t)ntl * RS_WIN + 255) / 256, 8192), 256, 0, s>>>(tiles, ntl, cslot,
                                                                                                   cval, cb.sval);
    if (hipStreamSynchronize(s) != hipSuccess) return -1;  // the host tile vector is in flight until here
  }
  if (out.end) {  // gapped: every row stays at the start of its staging segment
    fill<int64_t>(cb.mtotal, 16 * MTOTAL_SLOTS, 0, s);
    k_row_ends<<<nblk(C), 256, 0, s>>>(cb.eoff, cb.dcnt, C, out.end, (unsigned long long*)cb.mtotal);
    return 0;
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
GALA_CINST(WideW)

}  // namespace gala
