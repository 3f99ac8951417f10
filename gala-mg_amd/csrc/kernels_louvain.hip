// Phase-1 kernels for gfx950: community init, decide-and-move (four kernel
// families over nine degree bins), Tot/size apply with the modularity sums,
// and the weight update (In + sound prune).
//
// Semantics follow LinXi-lx/GALA-mg src/louvain_gpu/kernel_functions.cu:
//   init_communities :34-73, build_and_select_in_warp :289-441,
//   build_and_select :473-593, build_and_select_blk :640-813,
//   compute_In / compute_In_blk :889-1034,
// with gains compared exactly:  H(d) = w_d*m2 - K*(Tot_d - Tot_c + K)  (= dq*m2/2),
// best = argmax H over d != c with H > 0, ties to the lowest id, move iff
// H(best) - m2*(w_c - Self) > 0 and not (size[best]==size[c]==1 && best > c).
#include "gala_kernels.h"
#include "gala_device.h"
#include <climits>

namespace gala {

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned hash_slot(int key, int lg) {
  // multiplicative hashing onto a power-of-two table
  return ((unsigned)key * 2654435769u) >> (32 - lg);
}

__device__ __forceinline__ int ilog2_ceil(int x) {  // x >= 1
  return x <= 1 ? 0 : 32 - __clz(x - 1);
}

template <class G>
__device__ __forceinline__ bool better(G h, int d, G bh, int bd) {
  return h > bh || (h == bh && d < bd);
}

// All lanes of the wave (or sub-group of width W) reach the same (H, d).
template <class G, int W>
__device__ __forceinline__ void argmax_xor(G& h, int& d) {
#pragma unroll
  for (int k = W / 2; k >= 1; k >>= 1) {
    G ho = __shfl_xor(h, k, W);
    int dd = __shfl_xor(d, k, W);
    if (better(ho, dd, h, d)) { h = ho; d = dd; }
  }
}

template <class V, int W>
__device__ __forceinline__ V sum_xor(V x) {
#pragma unroll
  for (int k = W / 2; k >= 1; k >>= 1) x += __shfl_xor(x, k, W);
  return x;
}

__device__ __forceinline__ void lds_fence_wave() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

// Scaled gain H(d) = w_d*m2 - K*(Tot_d - Tot_c + K), exact: int64 on the
// integer path (m2 <= 1e9), int128 on the wide path (m2 < 2^62), where the
// differences of sums stay in int64 and only the two products are widened.
template <class T>
__device__ __forceinline__ typename T::gain_t gain_h(typename T::gain_t a, typename T::gain_t m2,
                                                     typename T::gain_t K, typename T::gain_t Td,
                                                     typename T::gain_t Tc) {
  using G = typename T::gain_t;
  if constexpr (T::WIDE) {
    const int64_t t = (int64_t)Td - (int64_t)Tc + (int64_t)K;
    return (G)(int64_t)a * (G)(int64_t)m2 - (G)(int64_t)K * (G)t;
  } else {
    return a * m2 - K * (Td - Tc + K);
  }
}

// Move test H(best) - m2*(w_c - S) > 0.
template <class T>
__device__ __forceinline__ typename T::gain_t gain_net(typename T::gain_t h, typename T::gain_t m2,
                                                       typename T::gain_t wc, typename T::gain_t S) {
  using G = typename T::gain_t;
  if constexpr (T::WIDE) return h - (G)(int64_t)m2 * (G)((int64_t)wc - (int64_t)S);
  else return h - m2 * (wc - S);
}

// Decision and state write-back shared by every decide kernel (one thread).
template <class T>
__device__ __forceinline__ bool commit_move(int v, int c, typename T::gain_t bestH, int bestD,
                                            typename T::acc_t wc, typename T::tot_t K,
                                            typename T::tot_t S, typename T::tot_t m2,
                                            const DevState<T>& st) {
  using G = typename T::gain_t;
  const int32_t* __restrict__ size = st.size;
  int nc = c;
  G gain = -(G)m2 * ((G)wc - (G)S);  // no positive candidate: the cost of leaving
  if (bestD != INT_MAX) {
    gain = gain_net<T>(bestH, (G)m2, (G)wc, (G)S);
    if (gain > 0 && !(size[bestD] == 1 && size[c] == 1 && bestD > c)) nc = bestD;
  }
  st.cur[v] = nc;
  if (st.cp) st.cp[v] = make_int2(nc, c);
  if (st.cm) st.cm[v] = nc != c ? (int)((unsigned)nc | 0x80000000u) : nc;
  if (st.track) {
    st.last_gain[v] = (float)((double)gain / ((double)m2 * (double)K));
    if (nc != c) st.moved_it[v] = st.iter;
  }
  return nc != c;  // Tot and size take the movers' deltas after the sweep (launch_apply)
}

// commit_move with the sizes of the best and the own community already at hand
// (from the community view).
template <class T>
__device__ __forceinline__ bool commit_move_sz(int v, int c, typename T::gain_t bestH, int bestD, int szD,
                                               int szC, typename T::acc_t wc, typename T::tot_t K,
                                               typename T::tot_t S, typename T::tot_t m2, const DevState<T>& st) {
  using G = typename T::gain_t;
  int nc = c;
  G gain = -(G)m2 * ((G)wc - (G)S);
  if (bestD != INT_MAX) {
    gain = gain_net<T>(bestH, (G)m2, (G)wc, (G)S);
    if (gain > 0 && !(szD == 1 && szC == 1 && bestD > c)) nc = bestD;
  }
  st.cur[v] = nc;
  if (st.cp) st.cp[v] = make_int2(nc, c);
  if (st.cm) st.cm[v] = nc != c ? (int)((unsigned)nc | 0x80000000u) : nc;
  if (st.track) {
    st.last_gain[v] = (float)((double)gain / ((double)m2 * (double)K));
    if (nc != c) st.moved_it[v] = st.iter;
  }
  return nc != c;
}

// A mover invalidates In of itself and of every neighbour (plain byte stores:
// every writer stores 1).
__device__ __forceinline__ void mark_row(uint8_t* dirty, const int32_t* __restrict__ nbr, int64_t o, int d,
                                         int v, int lane, int width) {
  for (int j = lane; j < d; j += width) dirty[nbr[o + j]] = 1;
  if (lane == 0) dirty[v] = 1;
}

template <class V>
__device__ __forceinline__ void atomic_add_ull(unsigned long long* p, V x) {
  if (x) atomicAdd(p, (unsigned long long)x);
}

// Decide-and-move counts of one workgroup, added to slot blockIdx.x % CTR_SLOTS
// of the counters: thousands of workgroups adding to one address serialise in
// L2 at the end of every launch; the host sums the slots when it reads them.
template <class A, class B, class C>
__device__ __forceinline__ void add_decide_counts(Counters* c, int bin, A moved, B swept, C edges) {
  DecideSlot& sl = c->slot[blockIdx.x % CTR_SLOTS];
  if (moved) atomic_add_ull(&sl.moved, moved);
  if (swept) {
    atomic_add_ull(&sl.swept, swept);
    atomic_add_ull(&sl.bin_swept[bin], swept);
  }
  if (edges) {
    atomic_add_ull(&sl.swept_edges, edges);
    atomic_add_ull(&sl.bin_edges[bin], edges);
  }
}

// m2 of the vertex's graph: the batch's per-graph m2 when graphs are batched.
template <class T>
__device__ __forceinline__ typename T::tot_t vertex_m2(const DevState<T>& st, int v, typename T::tot_t m2_all) {
  return st.gid ? st.m2g[st.gid[v]] : m2_all;
}

// Community view of u.  In the first sweep of a round every vertex is a
// singleton with Tot = K, so the view is built from one 4-byte Tot gather
// (8 bytes on the wide path) instead of reading the full view, which is then
// not built for that sweep.
template <class T, bool FIRST>
__device__ __forceinline__ typename T::cv_t view_of(const DevState<T>& st, int u) {
  if constexpr (FIRST) {
    typename T::cv_t e;
    cv_set(e, u, 1, st.Tot[u]);
    return e;
  } else {
    return st.cv[u];
  }
}

// ---------------------------------------------------------------------------
// init: K, Self, Tot=K, size=1, prev=cur=v, bin id, active=1
// (init_communities, kernel_functions.cu:34-73)
// ---------------------------------------------------------------------------
// degree_only: unit weights and no self-loop (the input graph, checked on
// upload): K = Tot = degree and Self = 0 without reading the rows
template <class T>
__global__ __launch_bounds__(256) void k_init_small(DevGraph<T> g, DevState<T> st, int degree_only) {
  using tt = typename T::tot_t;
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < g.n; v += gridDim.x * blockDim.x) {
    int64_t o = g.off[v], e = g.end[v];
    int64_t d = e - o;
    st.binid[v] = (int8_t)bin_of_degree(d, st.cuts);
    st.size[v] = 1;
    st.prev[v] = v;
    st.cur[v] = v;
    if (st.cp) st.cp[v] = make_int2(v, v);
    if (st.cm) st.cm[v] = v;
    st.active[v] = 1;
    st.dirty[v] = 0;
    if (st.track) {
      st.last_gain[v] = 0.f;
      st.moved_it[v] = -1;
      st.nbr_moved[v] = 0.f;
    }
    if (degree_only) {
      st.K[v] = (tt)d;
      st.Tot[v] = (tt)d;
      st.Self[v] = tt(0);
      continue;
    }
    if (d >= 8) continue;  // bins 2 and up: launch_init_rows, once the bins are partitioned
    tt k = 0, self = 0;
    for (int64_t j = o; j < e; j++) {
      tt w = (tt)load_w<T>(g, j);
      k += w;
      if (g.nbr[j] == v) self = w;
    }
    st.K[v] = k;
    st.Tot[v] = k;
    st.Self[v] = self;
  }
}

// K and Self of the rows of degree bins BLO .. BHI-1 (order[bstart[BLO] ..
// bstart[BHI])), W lanes per row.
template <int W, int BLO, int BHI, class T>
__global__ __launch_bounds__(256) void k_init_wide(DevGraph<T> g, DevState<T> st, const int32_t* __restrict__ order,
                                                    const int32_t* __restrict__ bstart) {
  using tt = typename T::tot_t;
  const int lane = threadIdx.x & (W - 1);
  const int ng = gridDim.x * (blockDim.x / W);
  const int lo = bstart[BLO], hi = bstart[BHI];
  for (int i = lo + (blockIdx.x * blockDim.x + threadIdx.x) / W; i < hi; i += ng) {  // uniform per group
    const int v = order[i];
    const int64_t o = g.off[v], e = g.end[v];
    tt k = 0, self = 0;
    for (int64_t j = o + lane; j < e; j += W) {
      tt w = (tt)load_w<T>(g, j);
      k += w;
      if (g.nbr[j] == v) self = w;
    }
    k = sum_xor<tt, W>(k);
    self = sum_xor<tt, W>(self);  // at most one self entry per row
    if (lane == 0) {
      st.K[v] = k;
      st.Tot[v] = k;
      st.Self[v] = self;
    }
  }
}

// K and Self of rows of 4096 entries and more (degree bins 7 and 8), one
// 1024-thread workgroup per row, four independent loads per thread in flight;
// fixed reduction order, so FP sums are reproducible.
template <class T>
__global__ __launch_bounds__(1024) void k_init_huge(DevGraph<T> g, DevState<T> st, const int32_t* __restrict__ order,
                                                     const int32_t* __restrict__ bstart) {
  using tt = typename T::tot_t;
  constexpr int NT = 1024, U = 4;
  __shared__ tt r_k[NT / 64], r_s[NT / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int lo = bstart[7], hi = bstart[NBINS];
  for (int i = lo + blockIdx.x; i < hi; i += gridDim.x) {
    const int v = order[i];
    const int64_t o = g.off[v], e = g.end[v];
    tt k = 0, self = 0;
    for (int64_t j0 = o + threadIdx.x; j0 < e; j0 += (int64_t)NT * U) {
      int u[U];
      tt w[U];
#pragma unroll
      for (int q = 0; q < U; q++) {
        const int64_t j = j0 + (int64_t)q * NT;
        u[q] = j < e ? g.nbr[j] : -1;
        w[q] = j < e ? (tt)load_w<T>(g, j) : tt(0);
      }
#pragma unroll
      for (int q = 0; q < U; q++) {
        k += w[q];
        if (u[q] == v) self = w[q];
      }
    }
    k = sum_xor<tt, 64>(k);
    self = sum_xor<tt, 64>(self);  // at most one self entry per row
    if (lane == 0) { r_k[wid] = k; r_s[wid] = self; }
    __syncthreads();
    if (threadIdx.x == 0) {
      k = 0;
      self = 0;
      for (int q = 0; q < NT / 64; q++) { k += r_k[q]; self += r_s[q]; }
      st.K[v] = k;
      st.Tot[v] = k;
      st.Self[v] = self;
    }
    __syncthreads();
  }
}

// rows of degree bins 2-3 (8..31 entries): 16 lanes each; bins 4-6 (32..4095):
// one wave each; bins 7-8: one workgroup each
template <class T>
void launch_init_rows(const DevGraph<T>& g, DevState<T>& st, const int32_t* order, const int32_t* bstart,
                      const int32_t* caps, hipStream_t s) {
  const int c23 = caps[2] + caps[3], c46 = caps[4] + caps[5] + caps[6], c78 = caps[7] + caps[8];
  if (c23 > 0) k_init_wide<16, 2, 4, T><<<std::min((c23 + 15) / 16, 16384), 256, 0, s>>>(g, st, order, bstart);
  if (c46 > 0) k_init_wide<64, 4, 7, T><<<std::min((c46 + 3) / 4, 8192), 256, 0, s>>>(g, st, order, bstart);
  if (c78 > 0) k_init_huge<T><<<std::min(c78, 2048), 1024, 0, s>>>(g, st, order, bstart);
}

template <class T>
void launch_init(const DevGraph<T>& g, DevState<T>& st, hipStream_t s, bool degree_only) {
  if (g.n <= 0) return;
  int nb = std::min((g.n + 255) / 256, 16384);
  k_init_small<T><<<nb, 256, 0, s>>>(g, st, degree_only ? 1 : 0);
}

// ---------------------------------------------------------------------------
// community view for the sweep: cv[u] = {prev[u], size[prev[u]], Tot[prev[u]]}
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_build_cv(DevState<T> st, int32_t n, const int32_t* __restrict__ vlist,
                                                  const int32_t* __restrict__ vcnt) {
  if (vlist) n = *vcnt;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int u = vlist ? vlist[i] : i;
    const int c = st.prev[u];
    typename T::cv_t e;
    cv_set(e, c, st.size[c], st.Tot[c]);
    st.cv[u] = e;
  }
}

template <class T>
void launch_build_cv(DevState<T>& st, int32_t n, hipStream_t s, const int32_t* vlist, const int32_t* vcnt) {
  if (n <= 0) return;
  k_build_cv<T><<<std::min((n + 255) / 256, 8192), 256, 0, s>>>(st, n, vlist, vcnt);
}

// ---------------------------------------------------------------------------
// decide-and-move, sub-wave ring kernels (degree < 32)
// Each group of W lanes owns one vertex, lane l holds entry l.  A xor
// exchange over every other lane of the group sums the weight of equal
// communities; the lowest lane of each community ("leader") evaluates its
// gain once, and an xor argmax selects the best (ties -> lowest id).
// (build_and_select_in_warp, kernel_functions.cu:289-441)
// ---------------------------------------------------------------------------
template <int W, class T>
__global__ __launch_bounds__(256) void k_decide_ring(DevGraph<T> g, DevState<T> st,
                                                      const int32_t* __restrict__ front,
                                                      const int32_t* __restrict__ bstart, int bin,
                                                      typename T::tot_t m2_all) {
  using acc = typename T::acc_t;
  using tt = typename T::tot_t;
  using G = typename T::gain_t;
  __shared__ unsigned long long s_moved, s_swept, s_edges;
  if (threadIdx.x == 0) { s_moved = 0; s_swept = 0; s_edges = 0; }
  __syncthreads();
  const int lo = bstart[bin], hi = bstart[bin + 1];
  const int lane = threadIdx.x & (W - 1);
  const int groups = gridDim.x * (blockDim.x / W);
  unsigned moved = 0, swept = 0, edges = 0;
  for (int i = lo + (blockIdx.x * blockDim.x + threadIdx.x) / W; i < hi; i += groups) {
    const int v = front[i];
    const typename T::tot_t m2 = vertex_m2<T>(st, v, m2_all);
    const int64_t o = g.off[v];
    const int d = (int)(g.end[v] - o);
    const typename T::cv_t ev = st.cv[v];
    const int c = ev.c;
    int cl = -1, ul = 0;
    acc wl = 0;
    tt tl = 0;
    if (lane < d) {
      ul = g.nbr[o + lane];
      const typename T::cv_t eu = st.cv[ul];
      cl = eu.c;
      tl = cv_tot(eu);
      wl = (acc)load_w<T>(g, o + lane);
    }
    acc s = wl, wc = (cl == c) ? wl : acc(0);
    bool leader = cl >= 0;
#pragma unroll
    for (int k = 1; k < W; k++) {
      int co = __shfl_xor(cl, k, W);
      acc wo = __shfl_xor(wl, k, W);
      if (co == cl) {
        s += wo;
        if ((lane ^ k) < lane) leader = false;
      }
      if (co == c) wc += wo;
    }
    const tt K = st.K[v];
    const tt Tc = cv_tot(ev);
    G h = 0;
    int bd = INT_MAX;
    if (leader && cl != c) {
      G hh = gain_h<T>((G)s, (G)m2, (G)K, (G)tl, (G)Tc);
      if (hh > 0) { h = hh; bd = cl; }
    }
    argmax_xor<G, W>(h, bd);
    int mv = 0;
    if (lane == 0) {
      swept++;
      edges += d;
      mv = commit_move<T>(v, c, h, bd, wc, K, st.Self[v], m2, st);
      moved += mv;
    }
    if (st.dirty && __shfl(mv, 0, W)) {  // In of the mover and its neighbours must be redone
      if (lane < d) st.dirty[ul] = 1;
      if (lane == 0) st.dirty[v] = 1;
    }
  }
  if (lane == 0) {
    if (moved) atomicAdd(&s_moved, (unsigned long long)moved);
    if (swept) atomicAdd(&s_swept, (unsigned long long)swept);
    if (edges) atomicAdd(&s_edges, (unsigned long long)edges);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    add_decide_counts(st.ctr, bin, s_moved, s_swept, s_edges);
  }
}

// ---------------------------------------------------------------------------
// decide-and-move, pipelined sub-wave ring kernels (degree < 32)
// Same decision as k_decide_ring.  The kernels are bound by the latency of a
// chain of dependent loads per vertex (front -> offsets -> row -> community
// view), so each W-lane group keeps four vertices in flight: in one iteration
// it decides the oldest while the loads of the three behind it, one chain link
// each, are outstanding.  The community view carries the sizes the singleton
// guard needs, so the commit waits on nothing.
// ---------------------------------------------------------------------------
template <int W, bool FIRST, class T>
__global__ __launch_bounds__(256) void k_decide_pipe(DevGraph<T> g, DevState<T> st,
                                                      const int32_t* __restrict__ front,
                                                      const int32_t* __restrict__ bstart, int bin,
                                                      typename T::tot_t m2_all) {
  using acc = typename T::acc_t;
  using tt = typename T::tot_t;
  using G = typename T::gain_t;
  using CV = typename T::cv_t;
  __shared__ unsigned long long s_moved, s_swept, s_edges;
  if (threadIdx.x == 0) { s_moved = 0; s_swept = 0; s_edges = 0; }
  __syncthreads();
  const int lo = bstart[bin], hi = bstart[bin + 1];
  const int lane = threadIdx.x & (W - 1);
  const int stride = gridDim.x * (blockDim.x / W);
  unsigned moved = 0, swept = 0, edges = 0;
  // stage A: vertex id loaded
  int iA = lo + (blockIdx.x * blockDim.x + threadIdx.x) / W;
  bool okA = iA < hi;
  int vA = okA ? front[iA] : 0;
  // stage B: row header and own view loaded
  bool okB = false;
  int vB = 0, dB = 0;
  int64_t oB = 0;
  CV evB{};
  tt KB = 0, SB = 0;
  // stage C: row entry loaded
  bool okC = false;
  int vC = 0, dC = 0, uC = 0;
  CV evC{};
  tt KC = 0, SC = 0;
  acc wC = 0;
  // stage D: neighbour's community view loaded
  bool okD = false;
  int vD = 0, dD = 0, uD = 0;
  CV evD{}, euD{};
  tt KD = 0, SD = 0;
  acc wD = 0;
  while (okA || okB || okC || okD) {  // uniform per group
    if (okD) {
      const typename T::tot_t m2 = vertex_m2<T>(st, vD, m2_all);
      const int c = evD.c;
      const bool has = lane < dD;
      const int cl = has ? euD.c : -1;
      const acc wl = has ? wD : acc(0);
      acc s = wl, wc = (cl == c) ? wl : acc(0);
      bool leader = has;
      if constexpr (FIRST) {
        // singleton partition with duplicate-free rows: every entry is its own
        // community, and the own community's weight is the self-loop
        wc = (acc)SD;
      } else {
#pragma unroll 4
        for (int k = 1; k < W; k++) {
          const int co = __shfl_xor(cl, k, W);
          const acc wo = __shfl_xor(wl, k, W);
          if (co == cl) {
            s += wo;
            if ((lane ^ k) < lane) leader = false;
          }
          if (co == c) wc += wo;
        }
      }
      G h = 0;
      int bd = INT_MAX, bsz = 0;
      if (leader && cl != c) {
        const G hh = gain_h<T>((G)s, (G)m2, (G)KD, (G)cv_tot(euD), (G)cv_tot(evD));
        if (hh > 0) { h = hh; bd = cl; bsz = cv_sz(euD); }
      }
#pragma unroll
      for (int k = W / 2; k >= 1; k >>= 1) {
        const G ho = __shfl_xor(h, k, W);
        const int dd = __shfl_xor(bd, k, W);
        const int so = __shfl_xor(bsz, k, W);
        if (better(ho, dd, h, bd)) { h = ho; bd = dd; bsz = so; }
      }
      int mv = 0;
      if (lane == 0) {
        swept++;
        edges += dD;
        mv = commit_move_sz<T>(vD, c, h, bd, bsz, cv_sz(evD), wc, KD, SD, m2, st);
        moved += mv;
      }
      if (st.dirty && __shfl(mv, 0, W)) {
        if (has) st.dirty[uD] = 1;
        if (lane == 0) st.dirty[vD] = 1;
      }
    }
    // advance, issuing the load each stage needs for the next iteration
    okD = okC;
    if (okC) {
      vD = vC; dD = dC; uD = uC; evD = evC; KD = KC; SD = SC; wD = wC;
      if (lane < dC) euD = view_of<T, FIRST>(st, uC);
    }
    okC = okB;
    if (okB) {
      vC = vB; dC = dB; evC = evB; KC = KB; SC = SB;
      if (lane < dB) {
        uC = g.nbr[oB + lane];
        wC = (acc)load_w<T>(g, oB + lane);
      }
    }
    okB = okA;
    if (okA) {
      vB = vA;
      oB = g.off[vA];
      dB = (int)(g.end[vA] - oB);
      evB = view_of<T, FIRST>(st, vA);
      KB = st.K[vA];
      SB = st.Self[vA];
      iA += stride;
      okA = iA < hi;
      if (okA) vA = front[iA];
    }
  }
  if (lane == 0) {
    if (moved) atomicAdd(&s_moved, (unsigned long long)moved);
    if (swept) atomicAdd(&s_swept, (unsigned long long)swept);
    if (edges) atomicAdd(&s_edges, (unsigned long long)edges);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    add_decide_counts(st.ctr, bin, s_moved, s_swept, s_edges);
  }
}

// ---------------------------------------------------------------------------
// decide-and-move, one thread per vertex (degree <= DMAX, integer path)
// Same decision as k_decide_ring.  Each lane loads its whole row, then the
// community views of all DMAX entries, so a lane has up to DMAX independent
// gathers in flight and a wave 64 * DMAX; the aggregation is a predicated
// O(DMAX^2) pass in registers: an entry whose community occurs at no lower
// index sums the weights of the equal communities after it and is evaluated.
// The view's sign bit carries the singleton flag the guard needs.
// ---------------------------------------------------------------------------
// A row of up to DMAX int32 entries read by one lane with 16-byte loads from
// the aligned address at or below its start (the CSR arrays carry 16 entries
// of padding, so the last vector may run past the final row).  Four entries
// per load instead of one: a wave's row loads touch the same lines 4x fewer
// times, which keeps them from being evicted and fetched again while the
// lane's later entries are still outstanding.
template <int DMAX>
__device__ __forceinline__ void load_row_vec(const int32_t* __restrict__ p, int64_t o, int d, int (&out)[DMAX]) {
  constexpr int NV = (DMAX + 3 + 3) / 4;
  const int mis = (int)(o & 3);
  const int4* __restrict__ q = reinterpret_cast<const int4*>(p + (o - mis));
  const int nv = (mis + d + 3) >> 2;
  int e[NV * 4];
#pragma unroll
  for (int j = 0; j < NV; j++) {
    const int4 x = j < nv ? q[j] : make_int4(0, 0, 0, 0);
    e[4 * j] = x.x;
    e[4 * j + 1] = x.y;
    e[4 * j + 2] = x.z;
    e[4 * j + 3] = x.w;
  }
#pragma unroll
  for (int k = 0; k < DMAX; k++) {
    const int val = mis == 0 ? e[k] : mis == 1 ? e[k + 1] : mis == 2 ? e[k + 2] : e[k + 3];
    out[k] = k < d ? val : 0;
  }
}

template <int DMAX, bool FIRST, class T>
__global__ __launch_bounds__(256) void k_decide_tpv(DevGraph<T> g, DevState<T> st,
                                                     const int32_t* __restrict__ front,
                                                     const int32_t* __restrict__ bstart, int bin,
                                                     typename T::tot_t m2_all) {
  static_assert(!T::WIDE, "integer path only");
  using acc = typename T::acc_t;
  using tt = typename T::tot_t;
  using G = typename T::gain_t;
  using CV = typename T::cv_t;
  __shared__ unsigned long long s_moved, s_swept, s_edges;
  if (threadIdx.x == 0) { s_moved = 0; s_swept = 0; s_edges = 0; }
  __syncthreads();
  const int lo = bstart[bin], hi = bstart[bin + 1];
  unsigned moved = 0, swept = 0, edges = 0;
  for (int i = lo + blockIdx.x * blockDim.x + threadIdx.x; i < hi; i += gridDim.x * blockDim.x) {
    const int v = front[i];
    const tt m2 = vertex_m2<T>(st, v, m2_all);
    const int64_t o = g.off[v];
    const int d = (int)(g.end[v] - o);
    const CV ev = view_of<T, FIRST>(st, v);
    const int c = ev.c;
    const tt K = st.K[v], S = st.Self[v], Tc = cv_tot(ev);
    int cs[DMAX];
    int32_t ts[DMAX];  // Tot with the singleton flag in the sign bit
    acc ws[DMAX];
    load_row_vec<DMAX>(g.nbr, o, d, cs);
    if (g.w) {
      load_row_vec<DMAX>((const int32_t*)g.w, o, d, ws);
    } else {
#pragma unroll
      for (int k = 0; k < DMAX; k++) ws[k] = k < d ? acc(1) : acc(0);
    }
#pragma unroll
    for (int k = 0; k < DMAX; k++) {
      CV e{};
      if (k < d) e = view_of<T, FIRST>(st, cs[k]);
      cs[k] = k < d ? e.c : -1;
      ts[k] = e.ts;
    }
    G h = 0;
    int bd = INT_MAX, bsz = 0;
    acc wc = 0;
    if constexpr (FIRST) {
      // singleton partition with duplicate-free rows: every entry is its own
      // community, and the own community's weight is the self-loop
      wc = (acc)S;
#pragma unroll
      for (int k = 0; k < DMAX; k++) {
        if (cs[k] >= 0 && cs[k] != c) {
          const G hh = gain_h<T>((G)ws[k], (G)m2, (G)K, (G)(ts[k] & 0x7fffffff), (G)Tc);
          if (hh > 0 && better(hh, cs[k], h, bd)) { h = hh; bd = cs[k]; bsz = ts[k] < 0 ? 1 : 2; }
        }
      }
    } else {
#pragma unroll
      for (int k = 0; k < DMAX; k++) {
        bool dup = cs[k] < 0;
#pragma unroll
        for (int j = 0; j < k; j++) dup |= cs[j] == cs[k];
        acc sum = ws[k];
#pragma unroll
        for (int j = k + 1; j < DMAX; j++) sum += cs[j] == cs[k] ? ws[j] : acc(0);
        if (!dup) {
          if (cs[k] == c) {
            wc = sum;
          } else {
            const G hh = gain_h<T>((G)sum, (G)m2, (G)K, (G)(ts[k] & 0x7fffffff), (G)Tc);
            if (hh > 0 && better(hh, cs[k], h, bd)) { h = hh; bd = cs[k]; bsz = ts[k] < 0 ? 1 : 2; }
          }
        }
      }
    }
    swept++;
    edges += d;
    const int mv = commit_move_sz<T>(v, c, h, bd, bsz, cv_sz(ev), wc, K, S, m2, st);
    moved += mv;
    if (st.dirty && mv) mark_row(st.dirty, g.nbr, o, d, v, 0, 1);
  }
  if (moved) atomicAdd(&s_moved, (unsigned long long)moved);
  if (swept) atomicAdd(&s_swept, (unsigned long long)swept);
  if (edges) atomicAdd(&s_edges, (unsigned long long)edges);
  __syncthreads();
  if (threadIdx.x == 0) {
    add_decide_counts(st.ctr, bin, s_moved, s_swept, s_edges);
  }
}

// ---------------------------------------------------------------------------
// decide-and-move, pipelined per-group LDS hash (32..511)
// W lanes per vertex, EPL row entries per lane (degree <= W*EPL), four
// vertices in flight per group as in k_decide_pipe.  The table records each
// community's Tot and size from the community view, so the scan and the
// commit need no further gathers.
// ---------------------------------------------------------------------------
template <int TMAX, int W, int EPL, class T>
__global__ __launch_bounds__(256) void k_decide_hpipe(DevGraph<T> g, DevState<T> st,
                                                       const int32_t* __restrict__ front,
                                                       const int32_t* __restrict__ bstart, int bin,
                                                       typename T::tot_t m2_all) {
  using acc = typename T::acc_t;
  using tt = typename T::tot_t;
  using G = typename T::gain_t;
  using CV = typename T::cv_t;
  constexpr int GPB = 256 / W;
  __shared__ int s_keys[GPB][TMAX];
  __shared__ acc s_vals[GPB][TMAX];
  __shared__ tt s_tots[GPB][TMAX];
  __shared__ int s_sizes[GPB][TMAX];
  __shared__ unsigned long long s_moved, s_swept, s_edges;
  if (threadIdx.x == 0) { s_moved = 0; s_swept = 0; s_edges = 0; }
  __syncthreads();
  const int lo = bstart[bin], hi = bstart[bin + 1];
  const int lane = threadIdx.x & (W - 1), grp = threadIdx.x / W;
  int* keys = s_keys[grp];
  acc* vals = s_vals[grp];
  tt* tots = s_tots[grp];
  int* sizes = s_sizes[grp];
  const int stride = gridDim.x * GPB;
  unsigned moved = 0, swept = 0, edges = 0;
  int iA = lo + blockIdx.x * GPB + grp;
  bool okA = iA < hi;
  int vA = okA ? front[iA] : 0;
  bool okB = false;
  int vB = 0, dB = 0;
  int64_t oB = 0;
  CV evB{};
  tt KB = 0, SB = 0;
  bool okC = false;
  int vC = 0, dC = 0;
  CV evC{};
  tt KC = 0, SC = 0;
  int uC[EPL];
  acc wC[EPL];
  bool okD = false;
  int vD = 0, dD = 0;
  CV evD{};
  tt KD = 0, SD = 0;
  int uD[EPL];
  acc wD[EPL];
  CV euD[EPL];
#pragma unroll
  for (int k = 0; k < EPL; k++) { uC[k] = 0; wC[k] = 0; uD[k] = 0; wD[k] = 0; euD[k] = CV{}; }
  while (okA || okB || okC || okD) {  // uniform per group
    if (okD) {
      const typename T::tot_t m2 = vertex_m2<T>(st, vD, m2_all);
      const int c = evD.c;
      int lg = ilog2_ceil(2 * dD);
      if (lg < 1) lg = 1;
      const int tsz = 1 << lg;  // <= TMAX by bin construction
      for (int t = lane; t < tsz; t += W) { keys[t] = 0; vals[t] = 0; }
      lds_fence_wave();
#pragma unroll
      for (int k = 0; k < EPL; k++) {
        if (lane + k * W < dD) {
          const int cl = euD[k].c;
          unsigned p = hash_slot(cl, lg);
          for (int probe = 0; probe < tsz; probe++) {
            const int kk = atomicCAS(&keys[p], 0, cl + 1);
            if (kk == 0) { tots[p] = cv_tot(euD[k]); sizes[p] = cv_sz(euD[k]); }
            if (kk == 0 || kk == cl + 1) { atomicAdd(&vals[p], wD[k]); break; }
            p = (p + 1) & (tsz - 1);
          }
        }
      }
      lds_fence_wave();
      G h = 0;
      int bd = INT_MAX, bsz = 0;
      acc wc = 0;
      for (int t = lane; t < tsz; t += W) {
        const int kk = keys[t];
        if (kk) {
          const int dd = kk - 1;
          const acc a = vals[t];
          if (dd == c) {
            wc = a;
          } else {
            const G hh = gain_h<T>((G)a, (G)m2, (G)KD, (G)tots[t], (G)cv_tot(evD));
            if (hh > 0 && better(hh, dd, h, bd)) { h = hh; bd = dd; bsz = sizes[t]; }
          }
        }
      }
#pragma unroll
      for (int k = W / 2; k >= 1; k >>= 1) {
        const G ho = __shfl_xor(h, k, W);
        const int dd = __shfl_xor(bd, k, W);
        const int so = __shfl_xor(bsz, k, W);
        if (better(ho, dd, h, bd)) { h = ho; bd = dd; bsz = so; }
      }
      wc = sum_xor<acc, W>(wc);
      int mv = 0;
      if (lane == 0) {
        swept++;
        edges += dD;
        mv = commit_move_sz<T>(vD, c, h, bd, bsz, cv_sz(evD), wc, KD, SD, m2, st);
        moved += mv;
      }
      if (st.dirty && __shfl(mv, 0, W)) {
#pragma unroll
        for (int k = 0; k < EPL; k++)
          if (lane + k * W < dD) st.dirty[uD[k]] = 1;
        if (lane == 0) st.dirty[vD] = 1;
      }
      lds_fence_wave();
    }
    okD = okC;
    if (okC) {
      vD = vC; dD = dC; evD = evC; KD = KC; SD = SC;
#pragma unroll
      for (int k = 0; k < EPL; k++) {
        uD[k] = uC[k];
        wD[k] = wC[k];
        if (lane + k * W < dC) euD[k] = st.cv[uC[k]];
      }
    }
    okC = okB;
    if (okB) {
      vC = vB; dC = dB; evC = evB; KC = KB; SC = SB;
#pragma unroll
      for (int k = 0; k < EPL; k++) {
        if (lane + k * W < dB) {
          uC[k] = g.nbr[oB + lane + k * W];
          wC[k] = (acc)load_w<T>(g, oB + lane + k * W);
        }
      }
    }
    okB = okA;
    if (okA) {
      vB = vA;
      oB = g.off[vA];
      dB = (int)(g.end[vA] - oB);
      evB = st.cv[vA];
      KB = st.K[vA];
      SB = st.Self[vA];
      iA += stride;
      okA = iA < hi;
      if (okA) vA = front[iA];
    }
  }
  if (lane == 0) {
    if (moved) atomicAdd(&s_moved, (unsigned long long)moved);
    if (swept) atomicAdd(&s_swept, (unsigned long long)swept);
    if (edges) atomicAdd(&s_edges, (unsigned long long)edges);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    add_decide_counts(st.ctr, bin, s_moved, s_swept, s_edges);
  }
}

// ---------------------------------------------------------------------------
// decide-and-move, pipelined per-group LDS hash with leader evaluation
// (8..511).  Same pipeline as k_decide_hpipe (W lanes per vertex, EPL row
// entries per lane, four vertices in flight per group), but no table scan:
// the lane whose insert claimed a slot ("leader") keeps the slot index, reads
// the completed sum after the inserts, evaluates that community's gain with
// the Tot and size of its own community-view entry, and clears the slot.  The
// tables are zeroed once at kernel start and stay zero between vertices, so
// the per-vertex cost is O(d / W) per lane instead of O(next_pow2(2d) / W),
// and a slot is 8 bytes of LDS (12 on the wide path).
// ---------------------------------------------------------------------------
// FIRST: the first sweep of a round (singleton partition, duplicate-free
// rows): every entry other than the self-loop is its own candidate with the
// entry's weight, the own-community weight is Self, and no table is used.
template <int TMAX, int W, int EPL, bool FIRST, class T>
__global__ __launch_bounds__(256) void k_decide_hlead(DevGraph<T> g, DevState<T> st,
                                                       const int32_t* __restrict__ front,
                                                       const int32_t* __restrict__ bstart, int bin,
                                                       typename T::tot_t m2_all) {
  using acc = typename T::acc_t;
  using tt = typename T::tot_t;
  using G = typename T::gain_t;
  using CV = typename T::cv_t;
  constexpr int GPB = 256 / W;
  constexpr int LEAD = 1 << 30;
  __shared__ int s_keys[GPB][TMAX];
  __shared__ acc s_vals[GPB][TMAX];
  __shared__ unsigned long long s_moved, s_swept, s_edges;
  for (int t = threadIdx.x; t < GPB * TMAX; t += blockDim.x) {
    (&s_keys[0][0])[t] = 0;
    (&s_vals[0][0])[t] = acc(0);
  }
  if (threadIdx.x == 0) { s_moved = 0; s_swept = 0; s_edges = 0; }
  __syncthreads();
  const int lo = bstart[bin], hi = bstart[bin + 1];
  const int lane = threadIdx.x & (W - 1), grp = threadIdx.x / W;
  int* keys = s_keys[grp];
  acc* vals = s_vals[grp];
  const int stride = gridDim.x * GPB;
  unsigned moved = 0, swept = 0, edges = 0;
  int iA = lo + blockIdx.x * GPB + grp;
  bool okA = iA < hi;
  int vA = okA ? front[iA] : 0;
  bool okB = false;
  int vB = 0, dB = 0;
  int64_t oB = 0;
  CV evB{};
  tt KB = 0, SB = 0;
  bool okC = false;
  int vC = 0, dC = 0;
  CV evC{};
  tt KC = 0, SC = 0;
  int uC[EPL];
  acc wC[EPL];
  bool okD = false;
  int vD = 0, dD = 0;
  CV evD{};
  tt KD = 0, SD = 0;
  int uD[EPL];
  acc wD[EPL];
  CV euD[EPL];
#pragma unroll
  for (int k = 0; k < EPL; k++) { uC[k] = 0; wC[k] = 0; uD[k] = 0; wD[k] = 0; euD[k] = CV{}; }
  while (okA || okB || okC || okD) {  // uniform per group
    if (okD) {
      const typename T::tot_t m2 = vertex_m2<T>(st, vD, m2_all);
      const int c = evD.c;
      int lg = ilog2_ceil(2 * dD);
      if (lg < 1) lg = 1;
      const int tmask = (1 << lg) - 1;  // next_pow2(2d) <= TMAX by bin construction
      G h = 0;
      int bd = INT_MAX, bsz = 0;
      acc wc = 0;
      if constexpr (FIRST) {
        (void)tmask;
#pragma unroll
        for (int k = 0; k < EPL; k++) {
          const int cl = euD[k].c;
          if (lane + k * W < dD && cl != c) {
            const G hh = gain_h<T>((G)wD[k], (G)m2, (G)KD, (G)cv_tot(euD[k]), (G)cv_tot(evD));
            if (hh > 0 && better(hh, cl, h, bd)) { h = hh; bd = cl; bsz = cv_sz(euD[k]); }
          }
        }
      } else {
        int sl[EPL];
#pragma unroll
        for (int k = 0; k < EPL; k++) {
          sl[k] = -1;
          if (lane + k * W < dD) {
            const int cl = euD[k].c;
            unsigned p = hash_slot(cl, lg);
            for (int probe = 0; probe <= tmask; probe++) {
              const int kk = atomicCAS(&keys[p], 0, cl + 1);
              if (kk == 0 || kk == cl + 1) {
                atomicAdd(&vals[p], wD[k]);
                sl[k] = (int)p | (kk == 0 ? LEAD : 0);
                break;
              }
              p = (p + 1) & tmask;
            }
          }
        }
        lds_fence_wave();
#pragma unroll
        for (int k = 0; k < EPL; k++) {
          if (sl[k] >= LEAD) {
            const int p = sl[k] & (LEAD - 1);
            const acc a = vals[p];
            keys[p] = 0;
            vals[p] = acc(0);
            const int cl = euD[k].c;
            if (cl == c) {
              wc = a;
            } else {
              const G hh = gain_h<T>((G)a, (G)m2, (G)KD, (G)cv_tot(euD[k]), (G)cv_tot(evD));
              if (hh > 0 && better(hh, cl, h, bd)) { h = hh; bd = cl; bsz = cv_sz(euD[k]); }
            }
          }
        }
      }
#pragma unroll
      for (int k = W / 2; k >= 1; k >>= 1) {
        const G ho = __shfl_xor(h, k, W);
        const int dd = __shfl_xor(bd, k, W);
        const int so = __shfl_xor(bsz, k, W);
        if (better(ho, dd, h, bd)) { h = ho; bd = dd; bsz = so; }
      }
      if constexpr (FIRST) wc = (acc)SD;
      else wc = sum_xor<acc, W>(wc);
      int mv = 0;
      if (lane == 0) {
        swept++;
        edges += dD;
        mv = commit_move_sz<T>(vD, c, h, bd, bsz, cv_sz(evD), wc, KD, SD, m2, st);
        moved += mv;
      }
      if (st.dirty && __shfl(mv, 0, W)) {
#pragma unroll
        for (int k = 0; k < EPL; k++)
          if (lane + k * W < dD) st.dirty[uD[k]] = 1;
        if (lane == 0) st.dirty[vD] = 1;
      }
      lds_fence_wave();  // every slot cleared before the next vertex inserts
    }
    okD = okC;
    if (okC) {
      vD = vC; dD = dC; evD = evC; KD = KC; SD = SC;
#pragma unroll
      for (int k = 0; k < EPL; k++) {
        uD[k] = uC[k];
        wD[k] = wC[k];
        if (lane + k * W < dC) euD[k] = view_of<T, FIRST>(st, uC[k]);
      }
    }
    okC = okB;
    if (okB) {
      vC = vB; dC = dB; evC = evB; KC = KB; SC = SB;
#pragma unroll
      for (int k = 0; k < EPL; k++) {
        if (lane + k * W < dB) {
          uC[k] = g.nbr[oB + lane + k * W];
          wC[k] = (acc)load_w<T>(g, oB + lane + k * W);
        }
      }
    }
    okB = okA;
    if (okA) {
      vB = vA;
      oB = g.off[vA];
      dB = (int)(g.end[vA] - oB);
      evB = view_of<T, FIRST>(st, vA);
      KB = st.K[vA];
      SB = st.Self[vA];
      iA += stride;
      okA = iA < hi;
      if (okA) vA = front[iA];
    }
  }
  if (lane == 0) {
    if (moved) atomicAdd(&s_moved, (unsigned long long)moved);
    if (swept) atomicAdd(&s_swept, (unsigned long long)swept);
    if (edges) atomicAdd(&s_edges, (unsigned long long)edges);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    add_decide_counts(st.ctr, bin, s_moved, s_swept, s_edges);
  }
}

// ---------------------------------------------------------------------------
// decide-and-move, leader-hash with one vertex at a time per group (32..511).
// W lanes per vertex, EPL row entries per lane: the lanes load their entries
// and the entries' community views together, insert (community -> summed
// weight) into the group's LDS table, and the lane whose insert claimed a slot
// evaluates that community once and clears the slot (as k_decide_hlead), but
// without the four-stage software pipeline: a workgroup stays resident (the
// grid is sized to the CUs) and its groups loop over the frontier, so many
// groups' independent gathers are in flight per CU while each group's
// instruction stream stays short.  FIRST: the first sweep of a round
// (singleton partition, duplicate-free rows): every entry but the self-loop is
// its own candidate and no table is used (TMAX = 1).
// ---------------------------------------------------------------------------
template <int TMAX, int W, int EPL, bool FIRST, class T>
__global__ __launch_bounds__(256) void k_decide_glead(DevGraph<T> g, DevState<T> st,
                                                       const int32_t* __restrict__ front,
                                                       const int32_t* __restrict__ bstart, int bin,
                                                       typename T::tot_t m2_all) {
  using acc = typename T::acc_t;
  using tt = typename T::tot_t;
  using G = typename T::gain_t;
  using CV = typename T::cv_t;
  constexpr int GPB = 256 / W;
  constexpr int LEAD = 1 << 30;
  __shared__ int s_keys[GPB][TMAX];
  __shared__ acc s_vals[GPB][TMAX];
  __shared__ unsigned long long s_moved, s_swept, s_edges;
  for (int t = threadIdx.x; t < GPB * TMAX; t += blockDim.x) {
    (&s_keys[0][0])[t] = 0;
    (&s_vals[0][0])[t] = acc(0);
  }
  if (threadIdx.x == 0) { s_moved = 0; s_swept = 0; s_edges = 0; }
  __syncthreads();
  const int lo = bstart[bin], hi = bstart[bin + 1];
  const int lane = threadIdx.x & (W - 1), grp = threadIdx.x / W;
  int* keys = s_keys[grp];
  acc* vals = s_vals[grp];
  unsigned moved = 0, swept = 0, edges = 0;
  for (int i = lo + blockIdx.x * GPB + grp; i < hi; i += gridDim.x * GPB) {  // uniform per group
    const int v = front[i];
    const tt m2 = vertex_m2<T>(st, v, m2_all);
    const int64_t o = g.off[v];
    const int d = (int)(g.end[v] - o);  // < W * EPL by bin construction
    const CV ev = view_of<T, FIRST>(st, v);
    const tt K = st.K[v];
    const tt S = st.Self[v];
    int u[EPL];
    acc wv[EPL];
#pragma unroll
    for (int k = 0; k < EPL; k++) {
      const int j = lane + k * W;
      u[k] = j < d ? g.nbr[o + j] : -1;
      wv[k] = j < d ? (acc)load_w<T>(g, o + j) : acc(0);
    }
    CV eu[EPL];
#pragma unroll
    for (int k = 0; k < EPL; k++) eu[k] = u[k] >= 0 ? view_of<T, FIRST>(st, u[k]) : CV{};
    const int c = ev.c;
    G h = 0;
    int bd = INT_MAX, bsz = 0;
    acc wc = 0;
    if constexpr (FIRST) {
#pragma unroll
      for (int k = 0; k < EPL; k++) {
        const int cl = eu[k].c;
        if (u[k] >= 0 && cl != c) {
          const G hh = gain_h<T>((G)wv[k], (G)m2, (G)K, (G)cv_tot(eu[k]), (G)cv_tot(ev));
          if (hh > 0 && better(hh, cl, h, bd)) { h = hh; bd = cl; bsz = cv_sz(eu[k]); }
        }
      }
    } else {
      int lg = ilog2_ceil(2 * d);
      if (lg < 1) lg = 1;
      const int tmask = (1 << lg) - 1;  // next_pow2(2d) <= TMAX by bin construction
      int sl[EPL];
#pragma unroll
      for (int k = 0; k < EPL; k++) {
        sl[k] = -1;
        if (u[k] >= 0) {
          const int cl = eu[k].c;
          unsigned p = hash_slot(cl, lg);
          for (int probe = 0; probe <= tmask; probe++) {
            const int kk = atomicCAS(&keys[p], 0, cl + 1);
            if (kk == 0 || kk == cl + 1) {
              atomicAdd(&vals[p], wv[k]);
              sl[k] = (int)p | (kk == 0 ? LEAD : 0);
              break;
            }
            p = (p + 1) & tmask;
          }
        }
      }
      lds_fence_wave();
#pragma unroll
      for (int k = 0; k < EPL; k++) {
        if (sl[k] >= LEAD) {
          const int p = sl[k] & (LEAD - 1);
          const acc a = vals[p];
          keys[p] = 0;
          vals[p] = acc(0);
          const int cl = eu[k].c;
          if (cl == c) {
            wc = a;
          } else {
            const G hh = gain_h<T>((G)a, (G)m2, (G)K, (G)cv_tot(eu[k]), (G)cv_tot(ev));
            if (hh > 0 && better(hh, cl, h, bd)) { h = hh; bd = cl; bsz = cv_sz(eu[k]); }
          }
        }
      }
    }
#pragma unroll
    for (int k = W / 2; k >= 1; k >>= 1) {
      const G ho = __shfl_xor(h, k, W);
      const int dd = __shfl_xor(bd, k, W);
      const int so = __shfl_xor(bsz, k, W);
      if (better(ho, dd, h, bd)) { h = ho; bd = dd; bsz = so; }
    }
    if constexpr (FIRST) wc = (acc)S;
    else wc = sum_xor<acc, W>(wc);
    int mv = 0;
    if (lane == 0) {
      swept++;
      edges += d;
      mv = commit_move_sz<T>(v, c, h, bd, bsz, cv_sz(ev), wc, K, S, m2, st);
      moved += mv;
    }
    if (st.dirty && __shfl(mv, 0, W)) {
#pragma unroll
      for (int k = 0; k < EPL; k++)
        if (u[k] >= 0) st.dirty[u[k]] = 1;
      if (lane == 0) st.dirty[v] = 1;
    }
    if constexpr (!FIRST) lds_fence_wave();  // every slot cleared before the group's next vertex inserts
  }
  if (lane == 0) {
    if (moved) atomicAdd(&s_moved, (unsigned long long)moved);
    if (swept) atomicAdd(&s_swept, (unsigned long long)swept);
    if (edges) atomicAdd(&s_edges, (unsigned long long)edges);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    add_decide_counts(st.ctr, bin, s_moved, s_swept, s_edges);
  }
}

// ---------------------------------------------------------------------------
// decide-and-move, W lanes per vertex with a per-group LDS hash (8..511)
// Phase 1 inserts (community -> summed weight) with LDS CAS/add; phase 2 scans
// the table so each distinct community is evaluated once (one Tot gather per
// community rather than per entry).  W < 64 runs 64/W vertices per wave, so a
// wave keeps more independent row and community gathers in flight.
// (build_and_select, kernel_functions.cu:473-593)
// ---------------------------------------------------------------------------
template <int TMAX, int W, class T>
__global__ __launch_bounds__(256) void k_decide_wave(DevGraph<T> g, DevState<T> st,
                                                      const int32_t* __restrict__ front,
                                                      const int32_t* __restrict__ bstart, int bin,
                                                      typename T::tot_t m2_all) {
  using acc = typename T::acc_t;
  using tt = typename T::tot_t;
  using G = typename T::gain_t;
  constexpr int GPB = 256 / W;  // groups per workgroup
  constexpr int LGMIN = W >= 64 ? 6 : 1;  // tables hold next_pow2(2d) <= TMAX slots
  __shared__ int s_keys[GPB][TMAX];
  __shared__ acc s_vals[GPB][TMAX];
  __shared__ tt s_tots[GPB][TMAX];  // Tot of each slot's community, from the community view
  __shared__ unsigned long long s_moved, s_swept, s_edges;
  if (threadIdx.x == 0) { s_moved = 0; s_swept = 0; s_edges = 0; }
  __syncthreads();
  const int lo = bstart[bin], hi = bstart[bin + 1];
  const int lane = threadIdx.x & (W - 1), grp = threadIdx.x / W;
  int* keys = s_keys[grp];
  acc* vals = s_vals[grp];
  tt* tots = s_tots[grp];
  const int ngroups = gridDim.x * GPB;
  unsigned moved = 0, swept = 0, edges = 0;
  for (int i = lo + blockIdx.x * GPB + grp; i < hi; i += ngroups) {
    const int v = front[i];
    const typename T::tot_t m2 = vertex_m2<T>(st, v, m2_all);
    const int64_t o = g.off[v];
    const int d = (int)(g.end[v] - o);
    const typename T::cv_t ev = st.cv[v];
    const int c = ev.c;
    int lg = ilog2_ceil(2 * d);
    if (lg < LGMIN) lg = LGMIN;
    const int tsz = 1 << lg;  // <= TMAX by bin construction
    for (int t = lane; t < tsz; t += W) { keys[t] = 0; vals[t] = 0; }
    lds_fence_wave();
    // U entries per lane are loaded before any is inserted (U independent
    // row loads, then U independent community-view gathers in flight)
    constexpr int U = 4;
    for (int j0 = lane; j0 < d; j0 += W * U) {
      int un[U];
      acc wv[U];
      typename T::cv_t ev4[U];
#pragma unroll
      for (int r = 0; r < U; r++) {
        const int j = j0 + r * W;
        un[r] = j < d ? g.nbr[o + j] : -1;
        wv[r] = j < d ? (acc)load_w<T>(g, o + j) : acc(0);
      }
#pragma unroll
      for (int r = 0; r < U; r++)
        if (un[r] >= 0) ev4[r] = st.cv[un[r]];
#pragma unroll
      for (int r = 0; r < U; r++) {
        if (un[r] < 0) continue;
        const int cl = ev4[r].c;
        unsigned p = hash_slot(cl, lg);
        for (int probe = 0; probe < tsz; probe++) {
          int k = atomicCAS(&keys[p], 0, cl + 1);
          if (k == 0) tots[p] = cv_tot(ev4[r]);  // the claiming lane records the community's Tot
          if (k == 0 || k == cl + 1) { atomicAdd(&vals[p], wv[r]); break; }
          p = (p + 1) & (tsz - 1);
        }
      }
    }
    lds_fence_wave();
    const tt K = st.K[v];
    const tt Tc = cv_tot(ev);
    G h = 0;
    int bd = INT_MAX;
    acc wc = 0;
    for (int t = lane; t < tsz; t += W) {
      const int k = keys[t];
      if (k) {
        const int dd = k - 1;
        const acc a = vals[t];
        if (dd == c) {
          wc = a;
        } else {
          G hh = gain_h<T>((G)a, (G)m2, (G)K, (G)tots[t], (G)Tc);
          if (hh > 0 && better(hh, dd, h, bd)) { h = hh; bd = dd; }
        }
      }
    }
    argmax_xor<G, W>(h, bd);
    wc = sum_xor<acc, W>(wc);
    int mv = 0;
    if (lane == 0) {
      swept++;
      edges += d;
      mv = commit_move<T>(v, c, h, bd, wc, K, st.Self[v], m2, st);
      moved += mv;
    }
    if (st.dirty && __shfl(mv, 0, W)) mark_row(st.dirty, g.nbr, o, d, v, lane, W);
    lds_fence_wave();
  }
  if (lane == 0) {
    if (moved) atomicAdd(&s_moved, (unsigned long long)moved);
    if (swept) atomicAdd(&s_swept, (unsigned long long)swept);
    if (edges) atomicAdd(&s_edges, (unsigned long long)edges);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    add_decide_counts(st.ctr, bin, s_moved, s_swept, s_edges);
  }
}

// ---------------------------------------------------------------------------
// decide-and-move, one workgroup per vertex (512 and up).  TMAX > 0: LDS hash
// of up to TMAX slots; TMAX == 0: the vertex's slice of the global hash arena
// (hub_off, sized next_pow2(2*deg) per hub per round -- no overlap, no drop).
// (build_and_select_blk, kernel_functions.cu:640-813 and :815-887)
// ---------------------------------------------------------------------------
template <int NT, int TMAX, class T>
__global__ __launch_bounds__(NT) void k_decide_block(DevGraph<T> g, DevState<T> st,
                                                      const int32_t* __restrict__ front,
                                                      const int32_t* __restrict__ bstart, int bin,
                                                      typename T::tot_t m2_all) {
  using acc = typename T::acc_t;
  using tt = typename T::tot_t;
  using G = typename T::gain_t;
  constexpr int TL = TMAX > 0 ? TMAX : 1;
  __shared__ int s_keys[TL];
  __shared__ acc s_vals[TL];
  __shared__ G r_h[NT / 64];
  __shared__ int r_d[NT / 64];
  __shared__ acc r_wc[NT / 64];
  __shared__ int s_mv;
  const int lo = bstart[bin], hi = bstart[bin + 1];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  unsigned moved = 0, swept = 0;
  unsigned long long edges = 0;
  for (int i = lo + blockIdx.x; i < hi; i += gridDim.x) {
    const int v = front[i];
    const typename T::tot_t m2 = vertex_m2<T>(st, v, m2_all);
    const int64_t o = g.off[v];
    const int d = (int)(g.end[v] - o);
    const int c = st.prev[v];
    int lg = ilog2_ceil(2 * d);
    const int tsz = 1 << lg;
    int* keys;
    acc* vals;
    if (TMAX > 0) {
      keys = s_keys;
      vals = s_vals;
    } else {
      keys = st.hub_keys + st.hub_off[v];
      vals = st.hub_vals + st.hub_off[v];
    }
    for (int t = threadIdx.x; t < tsz; t += NT) { keys[t] = 0; vals[t] = 0; }
    __syncthreads();
    // U entries per thread are loaded before any is inserted: U independent
    // row loads, then U independent community gathers in flight
    constexpr int U = 4;
    for (int j0 = threadIdx.x; j0 < d; j0 += NT * U) {
      int cl[U];
      acc wv[U];
#pragma unroll
      for (int q = 0; q < U; q++) {
        const int j = j0 + q * NT;
        cl[q] = j < d ? g.nbr[o + j] : -1;
        wv[q] = j < d ? (acc)load_w<T>(g, o + j) : acc(0);
      }
#pragma unroll
      for (int q = 0; q < U; q++) cl[q] = cl[q] >= 0 ? st.prev[cl[q]] : -1;
#pragma unroll
      for (int q = 0; q < U; q++) {
        if (cl[q] < 0) continue;
        unsigned p = hash_slot(cl[q], lg);
        int probe = 0;
        for (; probe < tsz; probe++) {
          int k = atomicCAS(&keys[p], 0, cl[q] + 1);
          if (k == 0 || k == cl[q] + 1) { atomicAdd(&vals[p], wv[q]); break; }
          p = (p + 1) & (tsz - 1);
        }
        if (probe == tsz) st.ctr->overflow = 1;
      }
    }
    __syncthreads();
    const tt K = st.K[v];
    const tt Tc = st.Tot[c];
    G h = 0;
    int bd = INT_MAX;
    acc wc = 0;
    for (int t = threadIdx.x; t < tsz; t += NT) {
      const int k = keys[t];
      if (k) {
        const int dd = k - 1;
        const acc a = vals[t];
        if (dd == c) {
          wc = a;
        } else {
          G hh = gain_h<T>((G)a, (G)m2, (G)K, (G)st.Tot[dd], (G)Tc);
          if (hh > 0 && better(hh, dd, h, bd)) { h = hh; bd = dd; }
        }
      }
    }
    argmax_xor<G, 64>(h, bd);
    wc = sum_xor<acc, 64>(wc);
    if (lane == 0) { r_h[wid] = h; r_d[wid] = bd; r_wc[wid] = wc; }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int q = 1; q < NT / 64; q++) {
        if (better(r_h[q], r_d[q], h, bd)) { h = r_h[q]; bd = r_d[q]; }
        wc += r_wc[q];
      }
      swept++;
      edges += d;
      const int mv = commit_move<T>(v, c, h, bd, wc, K, st.Self[v], m2, st);
      moved += mv;
      s_mv = mv;
    }
    __syncthreads();
    if (st.dirty && s_mv) mark_row(st.dirty, g.nbr, o, d, v, threadIdx.x, NT);
  }
  if (threadIdx.x == 0) {
    add_decide_counts(st.ctr, bin, moved, swept, edges);
  }
}

// ---------------------------------------------------------------------------
// decide-and-move, one workgroup per vertex with leader evaluation (512 and
// up, d <= NT * EPL).  Every thread loads its EPL row entries and their
// community views up front, inserts them into the LDS table, and after a
// barrier each slot's claiming thread evaluates that community once (Tot and
// size from its own view entry) and clears the slot.  The table is zeroed
// once per kernel, so a vertex costs O(d / NT) per thread, not
// O(next_pow2(2d) / NT), and needs no Tot gathers.
// ---------------------------------------------------------------------------
// NPASS > 1 (rows of up to TMAX - 1 entries): the entries are split into
// NPASS groups by a second hash of their community, and the table is filled,
// evaluated and cleared once per group, so it stays at most about half full
// while every entry stays in registers.  split_above > 0: longer rows are
// left to the split hub kernels.
template <int NT, int TMAX, int EPL, int NPASS, class T>
__global__ __launch_bounds__(NT) void k_decide_blead(DevGraph<T> g, DevState<T> st,
                                                      const int32_t* __restrict__ front,
                                                      const int32_t* __restrict__ bstart, int bin,
                                                      typename T::tot_t m2_all, int split_above) {
  using acc = typename T::acc_t;
  using tt = typename T::tot_t;
  using G = typename T::gain_t;
  using CV = typename T::cv_t;
  constexpr int LEAD = 1 << 30;
  __shared__ int s_keys[TMAX];
  __shared__ acc s_vals[TMAX];
  __shared__ G r_h[NT / 64];
  __shared__ int r_d[NT / 64];
  __shared__ int r_sz[NT / 64];
  __shared__ acc r_wc[NT / 64];
  __shared__ int s_mv;
  for (int t = threadIdx.x; t < TMAX; t += NT) { s_keys[t] = 0; s_vals[t] = acc(0); }
  __syncthreads();
  const int lo = bstart[bin], hi = bstart[bin + 1];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  unsigned moved = 0, swept = 0;
  unsigned long long edges = 0;
  for (int i = lo + blockIdx.x; i < hi; i += gridDim.x) {
    const int v = front[i];
    const tt m2 = vertex_m2<T>(st, v, m2_all);
    const int64_t o = g.off[v];
    const int d = (int)(g.end[v] - o);  // <= NT * EPL by bin construction
    if (split_above > 0 && d > split_above) continue;  // uniform per workgroup
    const CV ev = st.cv[v];
    const int c = ev.c;
    int lg = ilog2_ceil(2 * d);
    constexpr int LGMAX = __builtin_ctz(TMAX);
    if (lg > LGMAX) lg = LGMAX;  // NPASS > 1 only: d < TMAX keeps a free slot
    const int tmask = (1 << lg) - 1;
    int u[EPL];
    acc wv[EPL];
    CV eu[EPL];
#pragma unroll
    for (int k = 0; k < EPL; k++) {
      const int j = threadIdx.x + k * NT;
      u[k] = j < d ? g.nbr[o + j] : -1;
      wv[k] = j < d ? (acc)load_w<T>(g, o + j) : acc(0);
    }
#pragma unroll
    for (int k = 0; k < EPL; k++) eu[k] = u[k] >= 0 ? st.cv[u[k]] : CV{};
    G h = 0;
    int bd = INT_MAX, bsz = 0;
    acc wc = 0;
    const tt K = st.K[v];
#pragma unroll
    for (int pass = 0; pass < NPASS; pass++) {
      int sl[EPL];
#pragma unroll
      for (int k = 0; k < EPL; k++) {
        sl[k] = -1;
        const int cl = eu[k].c;
        const bool mine = NPASS == 1 || (int)((((unsigned)cl * 0x85EBCA6Bu) >> 16) % NPASS) == pass;
        if (u[k] >= 0 && mine) {
          unsigned p = hash_slot(cl, lg);
          for (int probe = 0; probe <= tmask; probe++) {
            const int kk = atomicCAS(&s_keys[p], 0, cl + 1);
            if (kk == 0 || kk == cl + 1) {
              atomicAdd(&s_vals[p], wv[k]);
              sl[k] = (int)p | (kk == 0 ? LEAD : 0);
              break;
            }
            p = (p + 1) & tmask;
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < EPL; k++) {
        if (sl[k] >= LEAD) {
          const int p = sl[k] & (LEAD - 1);
          const acc a = s_vals[p];
          s_keys[p] = 0;
          s_vals[p] = acc(0);
          const int cl = eu[k].c;
          if (cl == c) {
            wc = a;
          } else {
            const G hh = gain_h<T>((G)a, (G)m2, (G)K, (G)cv_tot(eu[k]), (G)cv_tot(ev));
            if (hh > 0 && better(hh, cl, h, bd)) { h = hh; bd = cl; bsz = cv_sz(eu[k]); }
          }
        }
      }
      if (NPASS > 1) __syncthreads();  // slots cleared before the next group inserts
    }
#pragma unroll
    for (int k = 32; k >= 1; k >>= 1) {
      const G ho = __shfl_xor(h, k, 64);
      const int dd = __shfl_xor(bd, k, 64);
      const int so = __shfl_xor(bsz, k, 64);
      if (better(ho, dd, h, bd)) { h = ho; bd = dd; bsz = so; }
    }
    wc = sum_xor<acc, 64>(wc);
    if (lane == 0) { r_h[wid] = h; r_d[wid] = bd; r_sz[wid] = bsz; r_wc[wid] = wc; }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int q = 1; q < NT / 64; q++) {
        if (better(r_h[q], r_d[q], h, bd)) { h = r_h[q]; bd = r_d[q]; bsz = r_sz[q]; }
        wc += r_wc[q];
      }
      swept++;
      edges += d;
      const int mv = commit_move_sz<T>(v, c, h, bd, bsz, cv_sz(ev), wc, K, st.Self[v], m2, st);
      moved += mv;
      s_mv = mv;
    }
    __syncthreads();
    if (st.dirty && s_mv) {
#pragma unroll
      for (int k = 0; k < EPL; k++)
        if (u[k] >= 0) st.dirty[u[k]] = 1;
      if (threadIdx.x == 0) st.dirty[v] = 1;
    }
  }
  if (threadIdx.x == 0) {
    add_decide_counts(st.ctr, bin, moved, swept, edges);
  }
}

// ---------------------------------------------------------------------------
// decide-and-move for hubs: one 1024-thread workgroup per vertex, a LT-slot
// LDS table tried first (bounded probing) and the vertex's slice of the global
// arena for the overflow.  Global slices are kept all-zero between uses: the
// scan that reads an overflow slot clears it.
// ---------------------------------------------------------------------------
// split_above > 0: rows longer than that are left to the split hub kernels.
template <int LT, class T>
__global__ __launch_bounds__(1024) void k_decide_hub(DevGraph<T> g, DevState<T> st,
                                                      const int32_t* __restrict__ front,
                                                      const int32_t* __restrict__ bstart, int bin,
                                                      typename T::tot_t m2_all, int split_above) {
  using acc = typename T::acc_t;
  using tt = typename T::tot_t;
  using G = typename T::gain_t;
  constexpr int NT = 1024, PMAX = 32;
  __shared__ int s_keys[LT];
  __shared__ acc s_vals[LT];
  __shared__ G r_h[NT / 64];
  __shared__ int r_d[NT / 64];
  __shared__ acc r_wc[NT / 64];
  __shared__ int s_ovf, s_mv;
  const int lo = bstart[bin], hi = bstart[bin + 1];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  unsigned moved = 0, swept = 0;
  unsigned long long edges = 0;
  for (int i = lo + blockIdx.x; i < hi; i += gridDim.x) {
    const int v = front[i];
    const typename T::tot_t m2 = vertex_m2<T>(st, v, m2_all);
    const int64_t o = g.off[v];
    const int d = (int)(g.end[v] - o);
    if (split_above > 0 && d > split_above) continue;  // uniform per workgroup
    const int c = st.prev[v];
    const int glg = ilog2_ceil(2 * d);
    const int gsz = 1 << glg;
    const int llg = glg < __builtin_ctz(LT) ? glg : __builtin_ctz(LT);
    const int lsz = 1 << llg;
    int* gkeys = st.hub_keys + st.hub_off[v];
    acc* gvals = st.hub_vals + st.hub_off[v];
    for (int t = threadIdx.x; t < lsz; t += NT) { s_keys[t] = 0; s_vals[t] = 0; }
    if (threadIdx.x == 0) s_ovf = 0;
    __syncthreads();
    // U entries per thread are loaded before any is inserted: U independent
    // row loads, then U independent community gathers in flight
    constexpr int U = 8;
    for (int j0 = threadIdx.x; j0 < d; j0 += NT * U) {
      int cu[U];
      acc wv[U];
#pragma unroll
      for (int r = 0; r < U; r++) {
        const int j = j0 + r * NT;
        cu[r] = j < d ? g.nbr[o + j] : -1;
        wv[r] = j < d ? (acc)load_w<T>(g, o + j) : acc(0);
      }
#pragma unroll
      for (int r = 0; r < U; r++) cu[r] = cu[r] >= 0 ? st.prev[cu[r]] : -1;
#pragma unroll
      for (int r = 0; r < U; r++) {
        const int cl = cu[r];
        if (cl < 0) continue;
        const acc w = wv[r];
        unsigned p = hash_slot(cl, llg);
        bool placed = false;
        for (int probe = 0; probe < PMAX; probe++) {
          const int k = atomicCAS(&s_keys[p], 0, cl + 1);
          if (k == 0 || k == cl + 1) { atomicAdd(&s_vals[p], w); placed = true; break; }
          p = (p + 1) & (lsz - 1);
        }
        if (!placed) {
          s_ovf = 1;
          unsigned q = hash_slot(cl, glg);
          int probe = 0;
          for (; probe < gsz; probe++) {
            const int k = atomicCAS(&gkeys[q], 0, cl + 1);
            if (k == 0 || k == cl + 1) { atomicAdd(&gvals[q], w); break; }
            q = (q + 1) & (gsz - 1);
          }
          if (probe == gsz) st.ctr->overflow = 1;
        }
      }
    }
    __syncthreads();
    const tt K = st.K[v];
    const tt Tc = st.Tot[c];
    G h = 0;
    int bd = INT_MAX;
    acc wc = 0;
    for (int t = threadIdx.x; t < lsz; t += NT) {
      const int k = s_keys[t];
      if (k) {
        const int dd = k - 1;
        const acc a = s_vals[t];
        if (dd == c) {
          wc += a;
        } else {
          G hh = gain_h<T>((G)a, (G)m2, (G)K, (G)st.Tot[dd], (G)Tc);
          if (hh > 0 && better(hh, dd, h, bd)) { h = hh; bd = dd; }
        }
      }
    }
    if (s_ovf) {
      for (int t = threadIdx.x; t < gsz; t += NT) {
        const int k = gkeys[t];
        if (k) {
          const int dd = k - 1;
          const acc a = gvals[t];
          gkeys[t] = 0;
          gvals[t] = 0;
          if (dd == c) {
            wc += a;
          } else {
            G hh = gain_h<T>((G)a, (G)m2, (G)K, (G)st.Tot[dd], (G)Tc);
            if (hh > 0 && better(hh, dd, h, bd)) { h = hh; bd = dd; }
          }
        }
      }
    }
    argmax_xor<G, 64>(h, bd);
    wc = sum_xor<acc, 64>(wc);
    if (lane == 0) { r_h[wid] = h; r_d[wid] = bd; r_wc[wid] = wc; }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int q = 1; q < NT / 64; q++) {
        if (better(r_h[q], r_d[q], h, bd)) { h = r_h[q]; bd = r_d[q]; }
        wc += r_wc[q];
      }
      swept++;
      edges += d;
      const int mv = commit_move<T>(v, c, h, bd, wc, K, st.Self[v], m2, st);
      moved += mv;
      s_mv = mv;
    }
    __syncthreads();
    if (st.dirty && s_mv) mark_row(st.dirty, g.nbr, o, d, v, threadIdx.x, NT);
  }
  if (threadIdx.x == 0) {
    add_decide_counts(st.ctr, bin, moved, swept, edges);
  }
}

// ---------------------------------------------------------------------------
// decide-and-move for hubs, split over many workgroups.  One workgroup per
// hub leaves most of the chip idle when a few contracted communities hold
// rows of 10^5 entries or more.  Four launches instead:
//   plan:   prefix sums over the frontier's hubs of their row chunks (HUB_CH
//           entries) and of their table chunks (HUB_TCH slots of the hub's
//           next_pow2(2d)-slot slice of the global arena)
//   agg:    one workgroup per row chunk: its entries' communities are summed
//           in an LDS hash (<= HUB_CH distinct keys in 2*HUB_CH slots, never
//           full), then each distinct (community, sum) is added into the
//           hub's global slice with one CAS/add
//   scan:   one workgroup per table chunk: gains of the chunk's communities
//           (clearing every slot it reads, so the slices are all-zero again),
//           best (H, id) and the own-community weight -> one partial
//   commit: one workgroup per hub: combines its partials in chunk order,
//           commits the move and marks the row dirty.
// The decision is the same as k_decide_hub's: the global table holds the
// exact per-community sums and the argmax order is total (H, then lowest id).
// ---------------------------------------------------------------------------
// locate the hub whose chunk range [start[i], start[i+1]) holds q
__device__ __forceinline__ int hub_find(const int32_t* __restrict__ start, int nh, int q) {
  int a = 0, b = nh;  // start[a] <= q < start[b]
  while (b - a > 1) {
    const int mid = (a + b) >> 1;
    if (start[mid] <= q) a = mid; else b = mid;
  }
  return a;
}

// part_keys > 0: instead of table chunks, plan ceil(d / part_keys) community
// partitions per hub (k_hub_part) for rows of up to part_max entries; longer
// rows (whose P workgroups would each stream the whole row) are marked in
// fail_out for the global-table split, every other hub's flag is cleared.
// only != null: plan only the hubs marked there.
template <class T>
__global__ __launch_bounds__(1024) void k_hub_plan(DevGraph<T> g, const int32_t* __restrict__ front,
                                                   const int32_t* __restrict__ bstart, int bin,
                                                   int32_t* __restrict__ plan, int split_above, int part_keys,
                                                   const uint8_t* __restrict__ only, int64_t part_max,
                                                   uint8_t* __restrict__ fail_out) {
  __shared__ int sh[17];
  const int lo = bstart[bin], hi = bstart[bin + 1], nh = hi - lo;
  int32_t* cstart = plan;           // nh + 1 row-chunk starts
  int32_t* tstart = plan + nh + 1;  // nh + 1 table-chunk starts
  int carry_c = 0, carry_t = 0;
  for (int base = 0; base < nh; base += blockDim.x) {  // uniform trip count
    const int i = base + threadIdx.x;
    int nc = 0, nt = 0;
    if (i < nh) {
      const int v = front[lo + i];
      const int64_t d = g.end[v] - g.off[v];
      uint8_t f = 0;
      if (d > split_above && (!only || only[i])) {
        if (part_keys > 0) {
          if (d <= part_max) nt = (int)((d + part_keys - 1) / part_keys);
          else f = 1;
        } else {
          int64_t tsz = 1;
          while (tsz < 2 * d) tsz <<= 1;
          nc = (int)((d + HUB_CH - 1) / HUB_CH);
          nt = (int)((tsz + HUB_TCH - 1) / HUB_TCH);
        }
      }
      if (fail_out) fail_out[i] = f;
    }
    int tot_c, tot_t;
    const int xc = block_excl_scan<int>(nc, sh, &tot_c);
    const int xt = block_excl_scan<int>(nt, sh, &tot_t);
    if (i < nh) { cstart[i] = carry_c + xc; tstart[i] = carry_t + xt; }
    carry_c += tot_c;
    carry_t += tot_t;
  }
  if (threadIdx.x == 0) { cstart[nh] = carry_c; tstart[nh] = carry_t; }
}

template <class T>
__global__ __launch_bounds__(256) void k_hub_agg(DevGraph<T> g, DevState<T> st, const int32_t* __restrict__ front,
                                                 const int32_t* __restrict__ bstart, int bin,
                                                 const int32_t* __restrict__ plan) {
  using acc = typename T::acc_t;
  constexpr int NT = 256, TS = 2 * HUB_CH, LG = 13, PER = HUB_CH / NT;
  static_assert((1 << LG) == TS, "table size");
  __shared__ int s_keys[TS];
  __shared__ acc s_vals[TS];
  for (int t = threadIdx.x; t < TS; t += NT) { s_keys[t] = 0; s_vals[t] = acc(0); }
  __syncthreads();
  const int lo = bstart[bin], nh = bstart[bin + 1] - lo;
  const int32_t* cstart = plan;
  const int nq = cstart[nh];
  for (int q = blockIdx.x; q < nq; q += gridDim.x) {
    const int i = hub_find(cstart, nh, q);
    const int v = front[lo + i];
    const int64_t o = g.off[v];
    const int d = (int)(g.end[v] - o);
    const int j0 = (q - cstart[i]) * HUB_CH;
    const int j1 = min(d, j0 + HUB_CH);
    // PER entries per thread: all row loads, then all community gathers, in flight
    int cu[PER];
    acc wv[PER];
#pragma unroll
    for (int r = 0; r < PER; r++) {
      const int j = j0 + r * NT + threadIdx.x;
      cu[r] = j < j1 ? g.nbr[o + j] : -1;
      wv[r] = j < j1 ? (acc)load_w<T>(g, o + j) : acc(0);
    }
#pragma unroll
    for (int r = 0; r < PER; r++) cu[r] = cu[r] >= 0 ? st.prev[cu[r]] : -1;
#pragma unroll
    for (int r = 0; r < PER; r++) {
      const int cl = cu[r];
      if (cl < 0) continue;
      unsigned p = hash_slot(cl, LG);
      for (int probe = 0; probe < TS; probe++) {
        const int k = atomicCAS(&s_keys[p], 0, cl + 1);
        if (k == 0 || k == cl + 1) { atomicAdd(&s_vals[p], wv[r]); break; }
        p = (p + 1) & (TS - 1);
      }
    }
    __syncthreads();
    // flush the distinct communities into the hub's global slice
    const int glg = ilog2_ceil(2 * d);
    const int gsz = 1 << glg;
    int* gkeys = st.hub_keys + st.hub_off[v];
    acc* gvals = st.hub_vals + st.hub_off[v];
    for (int t = threadIdx.x; t < TS; t += NT) {
      const int key = s_keys[t];
      if (key) {
        const acc a = s_vals[t];
        s_keys[t] = 0;
        s_vals[t] = acc(0);
        unsigned p = hash_slot(key - 1, glg);
        int probe = 0;
        for (; probe < gsz; probe++) {
          const int k = atomicCAS(&gkeys[p], 0, key);
          if (k == 0 || k == key) { atomicAdd(&gvals[p], a); break; }
          p = (p + 1) & (gsz - 1);
        }
        if (probe == gsz) st.ctr->overflow = 1;
      }
    }
    __syncthreads();
  }
}

template <class T>
__global__ __launch_bounds__(256) void k_hub_scan(DevGraph<T> g, DevState<T> st, const int32_t* __restrict__ front,
                                                  const int32_t* __restrict__ bstart, int bin,
                                                  const int32_t* __restrict__ plan, HubPart<T>* __restrict__ part,
                                                  typename T::tot_t m2_all) {
  using acc = typename T::acc_t;
  using tt = typename T::tot_t;
  using G = typename T::gain_t;
  constexpr int NT = 256;
  __shared__ G r_h[NT / 64];
  __shared__ int r_d[NT / 64];
  __shared__ acc r_wc[NT / 64];
  const int lo = bstart[bin], nh = bstart[bin + 1] - lo;
  const int32_t* tstart = plan + nh + 1;
  const int nt = tstart[nh];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int q = blockIdx.x; q < nt; q += gridDim.x) {
    const int i = hub_find(tstart, nh, q);
    const int v = front[lo + i];
    const tt m2 = vertex_m2<T>(st, v, m2_all);
    const int d = (int)(g.end[v] - g.off[v]);
    const int gsz = 1 << ilog2_ceil(2 * d);
    const int s0 = (q - tstart[i]) * HUB_TCH;
    const int s1 = min(gsz, s0 + HUB_TCH);
    int* gkeys = st.hub_keys + st.hub_off[v];
    acc* gvals = st.hub_vals + st.hub_off[v];
    const int c = st.prev[v];
    const tt K = st.K[v];
    const tt Tc = st.Tot[c];
    G h = 0;
    int bd = INT_MAX;
    acc wc = 0;
    for (int t = s0 + threadIdx.x; t < s1; t += NT) {
      const int k = gkeys[t];
      if (k) {
        const int dd = k - 1;
        const acc a = gvals[t];
        gkeys[t] = 0;
        gvals[t] = acc(0);
        if (dd == c) {
          wc = a;
        } else {
          const G hh = gain_h<T>((G)a, (G)m2, (G)K, (G)st.Tot[dd], (G)Tc);
          if (hh > 0 && better(hh, dd, h, bd)) { h = hh; bd = dd; }
        }
      }
    }
    argmax_xor<G, 64>(h, bd);
    wc = sum_xor<acc, 64>(wc);
    if (lane == 0) { r_h[wid] = h; r_d[wid] = bd; r_wc[wid] = wc; }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < NT / 64; w++) {
        if (better(r_h[w], r_d[w], h, bd)) { h = r_h[w]; bd = r_d[w]; }
        wc += r_wc[w];
      }
      part[q].h = h;
      part[q].d = bd;
      part[q].wc = wc;
    }
    __syncthreads();
  }
}

// Hubs split by community: workgroup (hub, p) of P streams the whole row and
// aggregates only the entries whose community c has part(c, P) == p, in an
// LT-slot LDS table (P is chosen so that a part expects at most 5/8 LT
// distinct communities: fewer parts, fewer re-reads of the row, against
// longer probes), then evaluates them: no global table, no global
// atomics; the row and the prev[] lines it gathers are re-read from L2 by the
// P workgroups.  A part whose table still fills up (an adversarial skew of the
// hash) flags its hub in fail[]; those hubs are redone by the global-table
// split kernels.
__device__ __forceinline__ int hub_part_of(int c, int P) {
  const unsigned x = ((unsigned)c ^ ((unsigned)c >> 15)) * 0x85EBCA6Bu;
  return (int)((x >> 8) % (unsigned)P);
}

// The rows of the hubs with at least HUB_RESOLVE_P parts, resolved once for
// their P part workgroups: hub_rc/hub_rw[hub_off[v] + j] = (prev[nbr[j]],
// w[j]); hubs with fewer parts gather in k_hub_part directly.  Workgroup
// (x, y) covers hubs x, x + gridDim.x, ... and every gridDim.y-th run of 1024
// entries of their rows.
constexpr int HUB_RESOLVE_P = 3;
template <class T>
__global__ __launch_bounds__(1024) void k_hub_resolve(DevGraph<T> g, DevState<T> st,
                                                       const int32_t* __restrict__ front,
                                                       const int32_t* __restrict__ bstart, int bin,
                                                       const int32_t* __restrict__ plan) {
  using acc = typename T::acc_t;
  constexpr int NT = 1024, U = 4;
  const int lo = bstart[bin], nh = bstart[bin + 1] - lo;
  const int32_t* tstart = plan + nh + 1;
  for (int i = blockIdx.x; i < nh; i += gridDim.x) {
    if (tstart[i + 1] - tstart[i] < HUB_RESOLVE_P) continue;  // uniform per workgroup
    const int v = front[lo + i];
    const int64_t o = g.off[v];
    const int d = (int)(g.end[v] - o);
    const int64_t rb = st.hub_off[v];
    for (int j0 = blockIdx.y * NT + threadIdx.x; j0 < d; j0 += gridDim.y * NT * U) {
      int u[U];
      acc w[U];
#pragma unroll
      for (int r = 0; r < U; r++) {
        const int j = j0 + r * gridDim.y * NT;
        u[r] = j < d ? g.nbr[o + j] : -1;
        w[r] = j < d ? (acc)load_w<T>(g, o + j) : acc(0);
      }
#pragma unroll
      for (int r = 0; r < U; r++) u[r] = u[r] >= 0 ? st.prev[u[r]] : -1;
#pragma unroll
      for (int r = 0; r < U; r++) {
        const int j = j0 + r * gridDim.y * NT;
        if (j < d) {
          st.hub_rc[rb + j] = u[r];
          st.hub_rw[rb + j] = w[r];
        }
      }
    }
  }
}

// The table is evaluated by a scan, the Tot gathers of eight slots per lane
// issued together.  (Holding a row of up to 16,384 entries in registers with
// leader evaluation from 8-byte community views measured slower on the
// power-law workload: the views' gathers hit L2 less often than prev[].)
template <int LT, class T>
__global__ __launch_bounds__(1024) void k_hub_part(DevGraph<T> g, DevState<T> st, const int32_t* __restrict__ front,
                                                    const int32_t* __restrict__ bstart, int bin,
                                                    const int32_t* __restrict__ plan, HubPart<T>* __restrict__ part,
                                                    uint8_t* __restrict__ fail, typename T::tot_t m2_all) {
  using acc = typename T::acc_t;
  using tt = typename T::tot_t;
  using G = typename T::gain_t;
  // PMAX: probes before a part gives up (its hub then takes the fallback);
  // at most 5/8 of the table is expected to fill, so only a skewed partition
  // reaches it
  constexpr int NT = 1024, PMAX = 256, U = 8, LG = __builtin_ctz(LT);
  constexpr int SB = 8;  // table scan: slots per lane per batch of Tot gathers
  static_assert(LT % (NT * SB) == 0, "table scan batches");
  __shared__ int s_keys[LT];
  __shared__ acc s_vals[LT];
  __shared__ G r_h[NT / 64];
  __shared__ int r_d[NT / 64];
  __shared__ acc r_wc[NT / 64];
  __shared__ int s_fail;
  for (int t = threadIdx.x; t < LT; t += NT) { s_keys[t] = 0; s_vals[t] = acc(0); }
  const int lo = bstart[bin], nh = bstart[bin + 1] - lo;
  const int32_t* tstart = plan + nh + 1;
  const int nt = tstart[nh];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int q = blockIdx.x; q < nt; q += gridDim.x) {
    const int i = hub_find(tstart, nh, q);
    const int pp = q - tstart[i], P = tstart[i + 1] - tstart[i];
    const int v = front[lo + i];
    const tt m2 = vertex_m2<T>(st, v, m2_all);
    const int64_t o = g.off[v];
    const int d = (int)(g.end[v] - o);
    const tt K = st.K[v];
    G h = 0;
    int bd = INT_MAX;
    acc wc = 0;
    if (threadIdx.x == 0) s_fail = 0;
    __syncthreads();  // also orders the table clears before the inserts
    const bool res = P >= HUB_RESOLVE_P;  // uniform per workgroup
    const int64_t rb = res ? st.hub_off[v] : 0;  // the row resolved by k_hub_resolve
    for (int j0 = threadIdx.x; j0 < d; j0 += NT * U) {
      int cu[U];
      acc wv[U];
      if (res) {
#pragma unroll
        for (int r = 0; r < U; r++) {
          const int j = j0 + r * NT;
          cu[r] = j < d ? st.hub_rc[rb + j] : -1;
          wv[r] = j < d ? st.hub_rw[rb + j] : acc(0);
        }
      } else {
#pragma unroll
        for (int r = 0; r < U; r++) {
          const int j = j0 + r * NT;
          cu[r] = j < d ? g.nbr[o + j] : -1;
          wv[r] = j < d ? (acc)load_w<T>(g, o + j) : acc(0);
        }
#pragma unroll
        for (int r = 0; r < U; r++) cu[r] = cu[r] >= 0 ? st.prev[cu[r]] : -1;
      }
#pragma unroll
      for (int r = 0; r < U; r++) {
        const int cl = cu[r];
        if (cl < 0 || hub_part_of(cl, P) != pp) continue;
        unsigned p = hash_slot(cl, LG);
        bool placed = false;
        for (int probe = 0; probe < PMAX; probe++) {
          const int k = atomicCAS(&s_keys[p], 0, cl + 1);
          if (k == 0 || k == cl + 1) { atomicAdd(&s_vals[p], wv[r]); placed = true; break; }
          p = (p + 1) & (LT - 1);
        }
        if (!placed) s_fail = 1;
      }
    }
    __syncthreads();
    const bool failed = s_fail != 0;
    const int c = st.prev[v];
    const tt Tc = st.Tot[c];
    for (int t0 = threadIdx.x; t0 < LT; t0 += NT * SB) {
      int kk[SB];
      acc av[SB];
      tt tv[SB];
#pragma unroll
      for (int r = 0; r < SB; r++) {
        const int t = t0 + r * NT;
        kk[r] = s_keys[t];
        av[r] = acc(0);
        if (kk[r]) {
          av[r] = s_vals[t];
          s_keys[t] = 0;
          s_vals[t] = acc(0);
        }
      }
#pragma unroll
      for (int r = 0; r < SB; r++) tv[r] = (kk[r] && !failed && kk[r] - 1 != c) ? st.Tot[kk[r] - 1] : tt(0);
#pragma unroll
      for (int r = 0; r < SB; r++) {
        if (kk[r] && !failed) {
          const int dd = kk[r] - 1;
          if (dd == c) {
            wc = av[r];
          } else {
            const G hh = gain_h<T>((G)av[r], (G)m2, (G)K, (G)tv[r], (G)Tc);
            if (hh > 0 && better(hh, dd, h, bd)) { h = hh; bd = dd; }
          }
        }
      }
    }
    argmax_xor<G, 64>(h, bd);
    wc = sum_xor<acc, 64>(wc);
    if (lane == 0) { r_h[wid] = h; r_d[wid] = bd; r_wc[wid] = wc; }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < NT / 64; w++) {
        if (better(r_h[w], r_d[w], h, bd)) { h = r_h[w]; bd = r_d[w]; }
        wc += r_wc[w];
      }
      part[q].h = h;
      part[q].d = bd;
      part[q].wc = wc;
      if (s_fail) fail[i] = 1;
    }
    __syncthreads();
  }
}

// Reads a slot of a global table and leaves it zero (one atomic at L2).
template <class V>
__device__ __forceinline__ V atomic_take(V* p) {
  if constexpr (sizeof(V) == 8) return (V)atomicExch((unsigned long long*)p, 0ull);
  else return (V)atomicExch((int*)p, 0);
}

// One workgroup aggregates the whole row of hub v in its global slice
// (next_pow2(2d) slots, so every distinct community fits), evaluates every
// community and leaves the slice zero; thread 0 receives the best candidate
// and w_c.  The fallback of a hub whose community part overflowed its LDS
// table: rare (a skewed hash), so one workgroup suffices.
template <int NT, class T>
__device__ void hub_global_decide(const DevGraph<T>& g, const DevState<T>& st, int v, int64_t o, int d,
                                  typename T::tot_t m2, typename T::gain_t* r_h, int* r_d,
                                  typename T::acc_t* r_wc, typename T::gain_t& h, int& bd,
                                  typename T::acc_t& wc) {
  using acc = typename T::acc_t;
  using tt = typename T::tot_t;
  using G = typename T::gain_t;
  const int glg = ilog2_ceil(2 * d);
  const int gsz = 1 << glg;
  int* gkeys = st.hub_keys + st.hub_off[v];
  acc* gvals = st.hub_vals + st.hub_off[v];
  for (int j = threadIdx.x; j < d; j += NT) {
    const int cl = st.prev[g.nbr[o + j]];
    const acc w = (acc)load_w<T>(g, o + j);
    unsigned p = hash_slot(cl, glg);
    int probe = 0;
    for (; probe < gsz; probe++) {
      const int k = atomicCAS(&gkeys[p], 0, cl + 1);
      if (k == 0 || k == cl + 1) { atomicAdd(&gvals[p], w); break; }
      p = (p + 1) & (gsz - 1);
    }
    if (probe == gsz) st.ctr->overflow = 1;
  }
  __threadfence();  // every insert performed at L2 before the table is read
  __syncthreads();
  const int c = st.prev[v];
  const tt K = st.K[v];
  const tt Tc = st.Tot[c];
  h = 0;
  bd = INT_MAX;
  wc = 0;
  for (int t = threadIdx.x; t < gsz; t += NT) {
    const int k = atomic_take(&gkeys[t]);
    if (k) {
      const acc a = atomic_take(&gvals[t]);
      const int dd = k - 1;
      if (dd == c) {
        wc = a;
      } else {
        const G hh = gain_h<T>((G)a, (G)m2, (G)K, (G)st.Tot[dd], (G)Tc);
        if (hh > 0 && better(hh, dd, h, bd)) { h = hh; bd = dd; }
      }
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  argmax_xor<G, 64>(h, bd);
  wc = sum_xor<acc, 64>(wc);
  if (lane == 0) { r_h[wid] = h; r_d[wid] = bd; r_wc[wid] = wc; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < NT / 64; w++) {
      if (better(r_h[w], r_d[w], h, bd)) { h = r_h[w]; bd = r_d[w]; }
      wc += r_wc[w];
    }
  }
}

// Combines each split hub's part results in part order (the same sums every
// run) and commits the move.  skip[i] set: a part of hub i overflowed, or the
// row was too long to split; with inline_fb (no too-long rows this sweep) the
// hub is decided here from its global slice, otherwise the global-table split
// kernels redo it.
template <class T>
__global__ __launch_bounds__(256) void k_hub_commit(DevGraph<T> g, DevState<T> st, const int32_t* __restrict__ front,
                                                    const int32_t* __restrict__ bstart, int bin,
                                                    const int32_t* __restrict__ plan,
                                                    const HubPart<T>* __restrict__ part, typename T::tot_t m2_all,
                                                    const uint8_t* __restrict__ skip, int inline_fb) {
  using acc = typename T::acc_t;
  using G = typename T::gain_t;
  constexpr int NT = 256;
  __shared__ int s_mv;
  __shared__ G r_h[NT / 64];
  __shared__ int r_d[NT / 64];
  __shared__ acc r_wc[NT / 64];
  const int lo = bstart[bin], nh = bstart[bin + 1] - lo;
  const int32_t* tstart = plan + nh + 1;
  unsigned moved = 0, swept = 0;
  unsigned long long edges = 0;
  for (int i = blockIdx.x; i < nh; i += gridDim.x) {
    // no chunks: left to k_decide_hub; skipped: redone by the fallback
    const bool failed = skip && skip[i];
    if (tstart[i] == tstart[i + 1] || (failed && !inline_fb)) continue;  // uniform per workgroup
    const int v = front[lo + i];
    const int64_t o = g.off[v];
    const int d = (int)(g.end[v] - o);
    G h = 0;
    int bd = INT_MAX;
    acc wc = 0;
    if (failed)
      hub_global_decide<NT, T>(g, st, v, o, d, vertex_m2<T>(st, v, m2_all), r_h, r_d, r_wc, h, bd, wc);
    if (threadIdx.x == 0) {
      if (!failed) {
        for (int q = tstart[i]; q < tstart[i + 1]; q++) {  // chunk order: the same sums every run
          if (better(part[q].h, part[q].d, h, bd)) { h = part[q].h; bd = part[q].d; }
          wc += part[q].wc;
        }
      }
      const int c = st.prev[v];
      swept++;
      edges += d;
      const int mv = commit_move<T>(v, c, h, bd, wc, st.K[v], st.Self[v], vertex_m2<T>(st, v, m2_all), st);
      moved += mv;
      s_mv = mv;
    }
    __syncthreads();
    if (st.dirty && s_mv) mark_row(st.dirty, g.nbr, o, d, v, threadIdx.x, NT);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    add_decide_counts(st.ctr, bin, moved, swept, edges);
  }
}

// ---------------------------------------------------------------------------
// First sweep of a round (every vertex its own community, rows duplicate-free):
// each entry is its own candidate, so no aggregation and no prev[] gathers:
// H(u) = w(v,u) m2 - K (Tot_u - Tot_v + K), w_c = Self[v].  Streaming argmax.
// ---------------------------------------------------------------------------
template <int W, class T>
__global__ __launch_bounds__(256) void k_decide_first(DevGraph<T> g, DevState<T> st,
                                                       const int32_t* __restrict__ front,
                                                       const int32_t* __restrict__ bstart, int bin,
                                                       typename T::tot_t m2_all) {
  using tt = typename T::tot_t;
  using G = typename T::gain_t;
  __shared__ unsigned long long s_moved, s_swept, s_edges;
  if (threadIdx.x == 0) { s_moved = 0; s_swept = 0; s_edges = 0; }
  __syncthreads();
  const int lo = bstart[bin], hi = bstart[bin + 1];
  const int lane = threadIdx.x & (W - 1);
  const int groups = gridDim.x * (blockDim.x / W);
  unsigned moved = 0, swept = 0, edges = 0;
  for (int i = lo + (blockIdx.x * blockDim.x + threadIdx.x) / W; i < hi; i += groups) {
    const int v = front[i];
    const typename T::tot_t m2 = vertex_m2<T>(st, v, m2_all);
    const int64_t o = g.off[v];
    const int d = (int)(g.end[v] - o);
    const tt K = st.K[v];
    const tt Tc = st.Tot[v];
    G h = 0;
    int bd = INT_MAX;
    for (int j = lane; j < d; j += W) {
      const int u = g.nbr[o + j];
      if (u != v) {
        const G hh = gain_h<T>((G)load_w<T>(g, o + j), (G)m2, (G)K, (G)st.Tot[u], (G)Tc);
        if (hh > 0 && better(hh, u, h, bd)) { h = hh; bd = u; }
      }
    }
    argmax_xor<G, W>(h, bd);
    int mv = 0;
    if (lane == 0) {
      swept++;
      edges += d;
      const tt S = st.Self[v];
      mv = commit_move<T>(v, v, h, bd, (typename T::acc_t)S, K, S, m2, st);
      moved += mv;
    }
    if (st.dirty && __shfl(mv, 0, W)) mark_row(st.dirty, g.nbr, o, d, v, lane, W);
  }
  if (lane == 0) {
    if (moved) atomicAdd(&s_moved, (unsigned long long)moved);
    if (swept) atomicAdd(&s_swept, (unsigned long long)swept);
    if (edges) atomicAdd(&s_edges, (unsigned long long)edges);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    add_decide_counts(st.ctr, bin, s_moved, s_swept, s_edges);
  }
}

template <int NT, class T>
__global__ __launch_bounds__(NT) void k_decide_first_block(DevGraph<T> g, DevState<T> st,
                                                            const int32_t* __restrict__ front,
                                                            const int32_t* __restrict__ bstart, int bin,
                                                            typename T::tot_t m2_all) {
  using tt = typename T::tot_t;
  using G = typename T::gain_t;
  __shared__ G r_h[NT / 64];
  __shared__ int r_d[NT / 64];
  __shared__ int s_mv;
  const int lo = bstart[bin], hi = bstart[bin + 1];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  unsigned moved = 0, swept = 0;
  unsigned long long edges = 0;
  for (int i = lo + blockIdx.x; i < hi; i += gridDim.x) {
    const int v = front[i];
    const typename T::tot_t m2 = vertex_m2<T>(st, v, m2_all);
    const int64_t o = g.off[v];
    const int d = (int)(g.end[v] - o);
    const tt K = st.K[v];
    const tt Tc = st.Tot[v];
    G h = 0;
    int bd = INT_MAX;
    for (int j = threadIdx.x; j < d; j += NT) {
      const int u = g.nbr[o + j];
      if (u != v) {
        const G hh = gain_h<T>((G)load_w<T>(g, o + j), (G)m2, (G)K, (G)st.Tot[u], (G)Tc);
        if (hh > 0 && better(hh, u, h, bd)) { h = hh; bd = u; }
      }
    }
    argmax_xor<G, 64>(h, bd);
    if (lane == 0) { r_h[wid] = h; r_d[wid] = bd; }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int q = 1; q < NT / 64; q++)
        if (better(r_h[q], r_d[q], h, bd)) { h = r_h[q]; bd = r_d[q]; }
      swept++;
      edges += d;
      const tt S = st.Self[v];
      const int mv = commit_move<T>(v, v, h, bd, (typename T::acc_t)S, K, S, m2, st);
      moved += mv;
      s_mv = mv;
    }
    __syncthreads();
    if (st.dirty && s_mv) mark_row(st.dirty, g.nbr, o, d, v, threadIdx.x, NT);
  }
  if (threadIdx.x == 0) {
    add_decide_counts(st.ctr, bin, moved, swept, edges);
  }
}

static int grid_for(int cap, int per_block, int max_blocks) {
  int nb = (cap + per_block - 1) / per_block;
  if (nb < 1) nb = 1;
  return std::min(nb, max_blocks);
}

// Decide kernel variants per bin (every variant takes the same exact decision;
// they differ only in how lanes are mapped to vertices):
//   bins 0-3: 0 = xor ring over W lanes, 1 = LDS hash over W lanes,
//             2 = xor ring over W lanes, four vertices in flight per group,
//             3 = one thread per vertex (integer path; FP falls back to 2),
//             4 (bins 2-3) = pipelined leader hash, W/2 lanes x 2 entries
//   bin 4:    0 = one wave per vertex, 1 = 32 lanes, 2 = 16 lanes,
//             3 = pipelined, 32 lanes x 4 entries, 4 = pipelined, 16 lanes x 8 entries,
//             5 = pipelined leader hash, 32 lanes x 4 entries, 6 = the same, 16 x 8,
//             7 = the same, 8 x 16, 8 = leader hash without the pipeline, 16 x 8
//             (resident workgroups loop over the frontier)
//   bin 5:    0 = one wave per vertex, 1 = 32 lanes, 2 = pipelined, 64 lanes x 8 entries,
//             3 = pipelined leader hash, 64 lanes x 8 entries, 4 = the same, 32 x 16,
//             5 = leader hash without the pipeline, 64 x 8
//   bins 6-7: 0 = one workgroup per vertex, table scan, 1 = one workgroup per
//             vertex, leader evaluation (bin 7 FP: the hub kernel)
//   bin 8:    0 = one workgroup per hub, 1 = hubs split over row and table
//             chunks (plan / agg / scan / commit), 2 = rows longer than the
//             LDS table split, the others one workgroup per hub, 3 = rows
//             longer than the LDS table split, the others one workgroup per
//             hub with two-pass leader evaluation, 4 = rows longer than the
//             LDS table split by community over workgroups (k_hub_part),
//             5 = the same with one partition per hub (overflow fallback),
//             6 = every hub split by community
// First sweeps: bins 0-3 have 0 = streaming argmax, 1 = pipelined (four
// vertices in flight per group), 2 = one thread per vertex (integer path; FP
// falls back to 1), 3 (bin 3) = pipelined leader kernel, 16 lanes x 2 entries;
// bin 4 has 0 = one wave per vertex, 1 = 32 lanes, 2 = pipelined, 16 lanes x
// 8 entries, 3 = without the pipeline, 16 x 8; bin 5 has 0 = one wave per
// vertex, 1 = pipelined, 64 lanes x 8, 2 = without the pipeline, 64 x 8.
int decide_variants(int bin) {
  static const int nv[NBINS] = {4, 4, 5, 5, 9, 6, 2, 2, 7};
  return bin >= 0 && bin < NBINS ? nv[bin] : 1;
}
int decide_first_variants(int bin) {
  static const int nv[NBINS] = {3, 3, 3, 4, 4, 3, 1, 1, 1};
  return bin >= 0 && bin < NBINS ? nv[bin] : 1;
}

// one thread per vertex on the integer path, the pipelined W-lane kernel on
// the wide path (its per-entry 8-byte totals would not fit the registers)
#define TPV_OR(DMAX, FIRST, W, PER)                                                                       \
  do {                                                                                                     \
    if constexpr (T::WIDE)                                                                                   \
      k_decide_pipe<W, FIRST, T><<<grid_for(cap, PER, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2);   \
    else                                                                                                   \
      k_decide_tpv<DMAX, FIRST, T><<<grid_for(cap, 256, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); \
  } while (0)

template <class T>
void launch_decide(int bin, const DevGraph<T>& g, DevState<T>& st, const int32_t* front,
                   const int32_t* bstart, int cap, typename T::tot_t m2, hipStream_t s, int variant) {
  if (cap <= 0) return;
  constexpr int HUB_LT = T::WIDE ? 8192 : 16384;
  if (variant < 0 || variant >= decide_variants(bin)) variant = 0;
  const int b = bin;
  switch (bin * 16 + variant) {
    // bin 0: 1..3
    case 0: k_decide_ring<4, T><<<grid_for(cap, 64, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 1: k_decide_wave<8, 4, T><<<grid_for(cap, 64, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 2: k_decide_pipe<4, false, T><<<grid_for(cap, 256, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 3: TPV_OR(3, false, 4, 256); break;
    // bin 1: 4..7
    case 16: k_decide_ring<8, T><<<grid_for(cap, 32, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 17: k_decide_wave<16, 8, T><<<grid_for(cap, 32, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 18: k_decide_pipe<8, false, T><<<grid_for(cap, 128, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 19: TPV_OR(7, false, 8, 128); break;
    // bin 2: 8..15
    case 32: k_decide_ring<16, T><<<grid_for(cap, 16, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 33: k_decide_wave<32, 16, T><<<grid_for(cap, 16, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 34: k_decide_pipe<16, false, T><<<grid_for(cap, 64, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 35: TPV_OR(15, false, 16, 64); break;
    case 36: k_decide_hlead<32, 8, 2, false, T><<<grid_for(cap, 128, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    // bin 3: 16..31
    case 48: k_decide_ring<32, T><<<grid_for(cap, 8, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 49: k_decide_wave<64, 32, T><<<grid_for(cap, 8, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 50: k_decide_pipe<32, false, T><<<grid_for(cap, 32, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 51: TPV_OR(31, false, 32, 32); break;
    case 52: k_decide_hlead<64, 16, 2, false, T><<<grid_for(cap, 64, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    // bin 4: 32..127
    case 64: k_decide_wave<256, 64, T><<<grid_for(cap, 4, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 65: k_decide_wave<256, 32, T><<<grid_for(cap, 8, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 66: k_decide_wave<256, 16, T><<<grid_for(cap, 16, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 67: k_decide_hpipe<256, 32, 4, T><<<grid_for(cap, 32, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 68: k_decide_hpipe<256, 16, 8, T><<<grid_for(cap, 64, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 69: k_decide_hlead<256, 32, 4, false, T><<<grid_for(cap, 32, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 70: k_decide_hlead<256, 16, 8, false, T><<<grid_for(cap, 64, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 71: k_decide_hlead<256, 8, 16, false, T><<<grid_for(cap, 128, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 72: k_decide_glead<256, 16, 8, false, T><<<grid_for(cap, 16, 1024), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    // bin 5: 128..511
    case 80: k_decide_wave<1024, 64, T><<<grid_for(cap, 4, 4096), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 81:
      if constexpr (T::WIDE)  // 8 groups x 1024 slots of 8-byte sums and totals exceed the LDS
        k_decide_wave<1024, 64, T><<<grid_for(cap, 4, 4096), 256, 0, s>>>(g, st, front, bstart, b, m2);
      else
        k_decide_wave<1024, 32, T><<<grid_for(cap, 8, 4096), 256, 0, s>>>(g, st, front, bstart, b, m2);
      break;
    case 82:
      if constexpr (T::WIDE)
        k_decide_wave<1024, 64, T><<<grid_for(cap, 4, 4096), 256, 0, s>>>(g, st, front, bstart, b, m2);
      else
        k_decide_hpipe<1024, 64, 8, T><<<grid_for(cap, 16, 4096), 256, 0, s>>>(g, st, front, bstart, b, m2);
      break;
    case 83: k_decide_hlead<1024, 64, 8, false, T><<<grid_for(cap, 16, 4096), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 84:
      if constexpr (T::WIDE)  // 8 groups x 1024 8-byte slots exceed the LDS with the keys
        k_decide_hlead<1024, 64, 8, false, T><<<grid_for(cap, 16, 4096), 256, 0, s>>>(g, st, front, bstart, b, m2);
      else
        k_decide_hlead<1024, 32, 16, false, T><<<grid_for(cap, 32, 4096), 256, 0, s>>>(g, st, front, bstart, b, m2);
      break;
    case 85: k_decide_glead<1024, 64, 8, false, T><<<grid_for(cap, 4, 1024), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    // bin 6: 512..4095
    case 96: k_decide_block<256, 8192, T><<<grid_for(cap, 1, 2048), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 97: k_decide_blead<256, 8192, 16, 1, T><<<grid_for(cap, 1, 2048), 256, 0, s>>>(g, st, front, bstart, b, m2, 0); break;
    // bin 7: 4096..8191
    case 112:
      if constexpr (T::WIDE)
        k_decide_hub<HUB_LT, T><<<grid_for(cap, 1, 512), 1024, 0, s>>>(g, st, front, bstart, b, m2, 0);
      else
        k_decide_block<1024, 16384, T><<<grid_for(cap, 1, 512), 1024, 0, s>>>(g, st, front, bstart, b, m2);
      break;
    case 113:
      if constexpr (T::WIDE)  // 16384 8-byte slots exceed the LDS: the hub kernel
        k_decide_hub<HUB_LT, T><<<grid_for(cap, 1, 512), 1024, 0, s>>>(g, st, front, bstart, b, m2, 0);
      else
        k_decide_blead<1024, 16384, 8, 1, T><<<grid_for(cap, 1, 512), 1024, 0, s>>>(g, st, front, bstart, b, m2, 0);
      break;
    // bin 8: >= 8192
    case 128: k_decide_hub<HUB_LT, T><<<grid_for(cap, 1, 1024), 1024, 0, s>>>(g, st, front, bstart, b, m2, 0); break;
    case 129:
    case 130:
    case 131: {
      // 1: every hub split; 2: only rows longer than the LDS table, the rest
      // one workgroup per hub; 3: the same with the two-pass leader kernel
      // for rows shorter than the LDS table (integer path; FP: as 2)
      const int above = variant == 1 ? 0 : variant == 2 ? HUB_LT : HUB_LT - 1;
      if (variant == 2)
        k_decide_hub<HUB_LT, T><<<grid_for(cap, 1, 1024), 1024, 0, s>>>(g, st, front, bstart, b, m2, HUB_LT);
      if (variant == 3) {
        if constexpr (T::WIDE)
          k_decide_hub<HUB_LT, T><<<grid_for(cap, 1, 1024), 1024, 0, s>>>(g, st, front, bstart, b, m2, above);
        else
          k_decide_blead<1024, 16384, 16, 2, T><<<grid_for(cap, 1, 1024), 1024, 0, s>>>(g, st, front, bstart, b,
                                                                                        m2, above);
      }
      k_hub_plan<T><<<1, 1024, 0, s>>>(g, front, bstart, b, st.hub_plan, above, 0, nullptr, 0, nullptr);
      k_hub_agg<T><<<2048, 256, 0, s>>>(g, st, front, bstart, b, st.hub_plan);
      k_hub_scan<T><<<2048, 256, 0, s>>>(g, st, front, bstart, b, st.hub_plan, (HubPart<T>*)st.hub_part, m2);
      k_hub_commit<T><<<grid_for(cap, 1, 1024), 256, 0, s>>>(g, st, front, bstart, b, st.hub_plan,
                                                              (const HubPart<T>*)st.hub_part, m2, nullptr, 0);
      break;
    }
    case 132:
    case 133:
    case 134: {
      // 4: rows longer than the LDS table split by community over workgroups
      // (k_hub_part), the others one workgroup per hub; hubs whose partition
      // overflowed are redone by the global-table split.  5: the same with
      // one partition per hub, so every hub with more distinct communities
      // than the table holds takes the fallback (kept for its tests).
      // 6: every hub split by community.
      const int PART_KEYS = variant == 5 ? INT_MAX : HUB_LT / 8 * 5;
      const int above = variant == 6 ? 0 : HUB_LT;
      HubPart<T>* hp = (HubPart<T>*)st.hub_part;
      if (above > 0)
        k_decide_hub<HUB_LT, T><<<grid_for(cap, 1, 1024), 1024, 0, s>>>(g, st, front, bstart, b, m2, above);
      // rows beyond 16 partitions go to the global-table split straight away
      const int64_t part_max = variant == 5 ? INT64_MAX : 16ll * PART_KEYS;
      // from the round's longest hub row: whether any row can be too long to
      // split, or have enough parts to be resolved first (unknown: both)
      const bool long_rows = st.hub_dmax < 0 || st.hub_dmax > part_max;
      const bool resolve = st.hub_dmax < 0 || st.hub_dmax > (int64_t)(HUB_RESOLVE_P - 1) * PART_KEYS;
      k_hub_plan<T><<<1, 1024, 0, s>>>(g, front, bstart, b, st.hub_plan, above, PART_KEYS, nullptr, part_max,
                                        st.hub_fail);  // also clears the flags
      if (resolve)
        k_hub_resolve<T><<<dim3(std::min(cap, 512), 8), 1024, 0, s>>>(g, st, front, bstart, b, st.hub_plan);
      // at most cap hubs of at most ceil(dmax / PART_KEYS) parts each: no
      // workgroup (128 KB of LDS, one per CU) is launched only to find no part
      int64_t pmax = 16;
      if (st.hub_dmax >= 0) pmax = std::min<int64_t>(16, std::max<int64_t>(1, (st.hub_dmax + PART_KEYS - 1) / PART_KEYS));
      const int part_grid = (int)std::min<int64_t>(2048, std::max<int64_t>(1, (int64_t)cap * pmax));
      if (st.kev) (void)hipEventRecord(st.kev[0], s);
      k_hub_part<HUB_LT, T><<<part_grid, 1024, 0, s>>>(g, st, front, bstart, b, st.hub_plan, hp, st.hub_fail, m2);
      if (st.kev) (void)hipEventRecord(st.kev[1], s);
      // without too-long rows an overflowed hub is decided by the commit itself
      k_hub_commit<T><<<grid_for(cap, 1, 1024), 256, 0, s>>>(g, st, front, bstart, b, st.hub_plan, hp, m2,
                                                              st.hub_fail, long_rows ? 0 : 1);
      if (long_rows) {
        k_hub_plan<T><<<1, 1024, 0, s>>>(g, front, bstart, b, st.hub_plan, above, 0, st.hub_fail, 0, nullptr);
        k_hub_agg<T><<<2048, 256, 0, s>>>(g, st, front, bstart, b, st.hub_plan);
        k_hub_scan<T><<<2048, 256, 0, s>>>(g, st, front, bstart, b, st.hub_plan, hp, m2);
        k_hub_commit<T><<<grid_for(cap, 1, 1024), 256, 0, s>>>(g, st, front, bstart, b, st.hub_plan, hp, m2, nullptr,
                                                                0);
      }
      break;
    }
  }
}

template <class T>
void launch_decide_first(int bin, const DevGraph<T>& g, DevState<T>& st, const int32_t* front,
                         const int32_t* bstart, int cap, typename T::tot_t m2, hipStream_t s, int variant) {
  if (cap <= 0) return;
  if (variant < 0 || variant >= decide_first_variants(bin)) variant = 0;
  const int b = bin;
  switch (bin * 8 + variant) {
    case 0: k_decide_first<4, T><<<grid_for(cap, 64, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 1: k_decide_pipe<4, true, T><<<grid_for(cap, 256, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 2: TPV_OR(3, true, 4, 256); break;
    case 8: k_decide_first<8, T><<<grid_for(cap, 32, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 9: k_decide_pipe<8, true, T><<<grid_for(cap, 128, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 10: TPV_OR(7, true, 8, 128); break;
    case 16: k_decide_first<16, T><<<grid_for(cap, 16, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 17: k_decide_pipe<16, true, T><<<grid_for(cap, 64, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 18: TPV_OR(15, true, 16, 64); break;
    case 24: k_decide_first<32, T><<<grid_for(cap, 8, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 25: k_decide_pipe<32, true, T><<<grid_for(cap, 32, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 26: TPV_OR(31, true, 32, 32); break;
    case 27: k_decide_hlead<1, 16, 2, true, T><<<grid_for(cap, 64, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 32: k_decide_first<64, T><<<grid_for(cap, 4, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 33: k_decide_first<32, T><<<grid_for(cap, 8, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 34: k_decide_hlead<1, 16, 8, true, T><<<grid_for(cap, 64, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 35: k_decide_glead<1, 16, 8, true, T><<<grid_for(cap, 16, 1024), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 40: k_decide_first<64, T><<<grid_for(cap, 4, 8192), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 41: k_decide_hlead<1, 64, 8, true, T><<<grid_for(cap, 16, 4096), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 42: k_decide_glead<1, 64, 8, true, T><<<grid_for(cap, 4, 1024), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 48: k_decide_first_block<256, T><<<grid_for(cap, 1, 4096), 256, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 56: k_decide_first_block<1024, T><<<grid_for(cap, 1, 1024), 1024, 0, s>>>(g, st, front, bstart, b, m2); break;
    case 64: k_decide_first_block<1024, T><<<grid_for(cap, 1, 1024), 1024, 0, s>>>(g, st, front, bstart, b, m2); break;
  }
}

// ---------------------------------------------------------------------------
// Tot and size of every community after a sweep (louvain_main_process.cu:
// 949-953 adds the movers' deltas, which the reference's decide kernels gather
// with per-mover atomics, kernel_functions.cu:397-400).  Here one pass reads
// the {cur, prev} pair of every vertex; a workgroup takes tiles of TD_T
// vertices, sums the movers' +K/+1 (new community) and -K/-1 (old community)
// per community in an LDS hash, and adds each community's tile sum with one
// global atomic.  When many movers join one community (late rounds of
// power-law graphs: tens of thousands into one community in one sweep),
// per-mover atomics on that address serialise; the tile sums keep it to one
// atomic per tile.  Exact integer sums.
// ---------------------------------------------------------------------------
constexpr int TD_T = 512, TD_S = 4 * TD_T, TD_LG = 11;  // <= 2 keys per vertex: half load at most
template <class T>
__global__ __launch_bounds__(256) void k_tot_delta(DevState<T> st, int32_t n, const int32_t* __restrict__ vlist,
                                                   const int32_t* __restrict__ vcnt) {
  using tt = typename T::tot_t;
  __shared__ int s_key[TD_S];
  __shared__ int s_cnt[TD_S];
  __shared__ tt s_tot[TD_S];
  __shared__ int s_list[TD_S];
  __shared__ int s_n;
  for (int t = threadIdx.x; t < TD_S; t += blockDim.x) { s_key[t] = 0; s_cnt[t] = 0; s_tot[t] = tt(0); }
  if (threadIdx.x == 0) s_n = 0;
  __syncthreads();
  auto add = [&](int c, tt k, int one) {
    unsigned p = hash_slot(c, TD_LG);
    for (int probe = 0; probe < TD_S; probe++) {  // at most 2 * TD_T keys: a free slot exists
      const int kk = atomicCAS(&s_key[p], 0, c + 1);
      if (kk == 0 || kk == c + 1) {
        if (kk == 0) s_list[atomicAdd(&s_n, 1)] = (int)p;
        atomicAdd(&s_cnt[p], one);
        if (k) atomicAdd(&s_tot[p], one > 0 ? k : -k);
        return;
      }
      p = (p + 1) & (TD_S - 1);
    }
  };
  if (vlist) n = *vcnt;
  for (int64_t base = (int64_t)blockIdx.x * TD_T; base < n; base += (int64_t)gridDim.x * TD_T) {
    for (int j = threadIdx.x; j < TD_T; j += blockDim.x) {
      const int64_t i = base + j;
      if (i >= n) break;
      const int v = vlist ? vlist[i] : (int)i;
      const int2 cp = st.cp[v];  // {cur, prev}
      if (cp.x != cp.y) {
        const tt k = st.K[v];
        add(cp.x, k, 1);
        add(cp.y, k, -1);
      }
    }
    __syncthreads();
    const int ns = s_n;
    for (int i = threadIdx.x; i < ns; i += blockDim.x) {
      const int p = s_list[i];
      const int c = s_key[p] - 1;
      bool packed = false;
      if constexpr (!T::WIDE) {
        if (st.tsd) {  // one atomic for both deltas (|deltas| < 2^31: the product is exact)
          const long long pk = (long long)s_tot[p] * 4294967296ll + (long long)s_cnt[p];
          if (pk) atomicAdd(&st.tsd[c], (unsigned long long)pk);
          packed = true;
        }
      }
      if (!packed) {
        if (s_cnt[p]) atomicAdd(&st.size[c], s_cnt[p]);
        if (s_tot[p]) atomicAdd(&st.Tot[c], s_tot[p]);
      }
      s_key[p] = 0;
      s_cnt[p] = 0;
      s_tot[p] = tt(0);
    }
    __syncthreads();
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// min non-zero Tot (Tot_op, kernel_functions.cuh:112-121), sum Tot^2 and the
// non-empty count over the rebuilt Tot and size.
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_apply(DevState<T> st, int32_t n, typename T::tot_t m2,
                                               const int32_t* __restrict__ vlist, const int32_t* __restrict__ vcnt) {
  using tt = typename T::tot_t;
  __shared__ __int128 sh_t2[4];
  __shared__ unsigned long long sh_min[4];
  __shared__ int sh_ne[4];
  __int128 t2 = 0;  // sum Tot^2: <= m2^2 < 2^124 on the wide path
  unsigned long long mn = (unsigned long long)(long long)m2;
  int ne = 0;
  if (vlist) n = *vcnt;  // communities are vertex ids: those of the listed vertices
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int c = vlist ? vlist[i] : i;
    tt t = st.Tot[c];
    int sz = st.size[c];
    if constexpr (!T::WIDE) {
      if (st.tsd) {  // the sweep's packed deltas (k_tot_delta)
        const unsigned long long D = st.tsd[c];
        if (D) {
          st.tsd[c] = 0;
          // D = dTot * 2^32 + dsize modulo 2^64 with |dTot|, |dsize| < 2^31: exact
          const int32_t ds = (int32_t)(uint32_t)D;
          const int32_t dt = (int32_t)(((long long)D - (long long)ds) / 4294967296ll);
          t += dt;
          sz += ds;
          st.Tot[c] = t;
          st.size[c] = sz;
        }
      }
    }
    ne += sz > 0;
    const long long ti = (long long)t;
    t2 += (__int128)ti * ti;
    if (ti != 0 && (unsigned long long)ti < mn) mn = (unsigned long long)ti;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  t2 = sum_xor<__int128, 64>(t2);
  ne = sum_xor<int, 64>(ne);
#pragma unroll
  for (int k = 32; k >= 1; k >>= 1) {
    unsigned long long o = __shfl_xor(mn, k, 64);
    mn = o < mn ? o : mn;
  }
  if (lane == 0) { sh_t2[wid] = t2; sh_min[wid] = mn; sh_ne[wid] = ne; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < 4; q++) {
      t2 += sh_t2[q]; ne += sh_ne[q];
      mn = sh_min[q] < mn ? sh_min[q] : mn;
    }
    BlockPart bp{};
    bp.s0 = (unsigned long long)t2;
    bp.s0hi = (unsigned long long)(t2 >> 64);
    bp.s1 = mn;
    bp.i0 = ne;
    st.part[blockIdx.x] = bp;
  }
}

// Combines the nb workgroup partials of k_apply into the counters (the
// counters were reset before the sweep; this one workgroup is the only writer
// of sum Tot^2, so its two halves are stored, not added atomically).
template <class T>
__global__ __launch_bounds__(256) void k_apply_fin(DevState<T> st, int nb) {
  __shared__ __int128 sh_t2[4];
  __shared__ unsigned long long sh_min[4];
  __shared__ long long sh_ne[4];
  __int128 t2 = 0;
  unsigned long long mn = ~0ull;
  long long ne = 0;
  for (int b = threadIdx.x; b < nb; b += 256) {
    const BlockPart p = st.part[b];
    t2 += (__int128)(((unsigned __int128)p.s0hi << 64) | p.s0);
    ne += p.i0;
    mn = p.s1 < mn ? p.s1 : mn;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  t2 = sum_xor<__int128, 64>(t2);
  ne = sum_xor<long long, 64>(ne);
#pragma unroll
  for (int k = 32; k >= 1; k >>= 1) {
    unsigned long long o = __shfl_xor(mn, k, 64);
    mn = o < mn ? o : mn;
  }
  if (lane == 0) { sh_t2[wid] = t2; sh_min[wid] = mn; sh_ne[wid] = ne; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < 4; q++) {
      t2 += sh_t2[q]; ne += sh_ne[q];
      mn = sh_min[q] < mn ? sh_min[q] : mn;
    }
    st.ctr->sum_tot2 = (unsigned long long)t2;
    st.ctr->sum_tot2_hi = (unsigned long long)(t2 >> 64);
    atomicMin((unsigned long long*)&st.ctr->min_tot, mn);
    if (ne) atomicAdd(&st.ctr->nonempty, (int)ne);
  }
}

template <class T>
void launch_apply(DevState<T>& st, int32_t n, typename T::tot_t m2, hipStream_t s, const int32_t* vlist,
                  const int32_t* vcnt, const int32_t* dlist, const int32_t* dcnt, bool moves) {
  if (n <= 0) return;
  const int nb = std::min((n + 255) / 256, MAX_PARTS);
  if (!dlist) { dlist = vlist; dcnt = vcnt; }
  if (moves) k_tot_delta<T><<<std::min((n + TD_T - 1) / TD_T, 2048), 256, 0, s>>>(st, n, dlist, dcnt);
  k_apply<T><<<nb, 256, 0, s>>>(st, n, m2, vlist, vcnt);
  k_apply_fin<T><<<1, 256, 0, s>>>(st, nb);
}

// ---------------------------------------------------------------------------
// weight update: In[v] = sum w over cur[u]==cur[v], prune flag
// (compute_In / compute_In_blk, kernel_functions.cu:889-1034).  The prune
// m2*(2In - S - K) + K*(minTot - (Tot_c - K)) > 0 proves every gain < 0.
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ uint8_t prune_keep(typename T::tot_t in, typename T::tot_t K,
                                              typename T::tot_t S, typename T::tot_t Tc,
                                              typename T::tot_t minTot, typename T::tot_t m2) {
  using G = typename T::gain_t;
  // the bracketed differences are sums of at most 2*m2 in magnitude: int64
  const int64_t a = 2 * (int64_t)in - (int64_t)S - (int64_t)K;
  const int64_t b = (int64_t)minTot - ((int64_t)Tc - (int64_t)K);
  const G P = (G)(int64_t)m2 * (G)a + (G)(int64_t)K * (G)b;
  return P > 0 ? 0 : 1;
}

template <class T>
__device__ __forceinline__ typename T::tot_t read_min_tot(const Counters* c) {
  return (typename T::tot_t)c->min_tot;
}

// One row's part of In (entries whose neighbour is in community c) and of the
// moved-neighbour count, NT lanes from `lane`: the row loads and community
// gathers of WR_U entries per lane are issued together; each lane adds its
// entries in row order, so FP sums are unchanged.
constexpr int WR_U = 4;
template <int NT, class T>
__device__ __forceinline__ void row_in_span(const DevGraph<T>& g, const DevState<T>& st, int64_t o, int d,
                                            int lane, int c, typename T::tot_t& x, int& mvn) {
  using tt = typename T::tot_t;
  for (int j0 = lane; j0 < d; j0 += NT * WR_U) {
    int u[WR_U], cu[WR_U];
    tt w[WR_U];
#pragma unroll
    for (int r = 0; r < WR_U; r++) {
      const int j = j0 + r * NT;
      u[r] = j < d ? g.nbr[o + j] : -1;
      w[r] = j < d ? (tt)load_w<T>(g, o + j) : tt(0);
    }
#pragma unroll
    for (int r = 0; r < WR_U; r++) cu[r] = u[r] >= 0 ? st.cur[u[r]] : -1;
#pragma unroll
    for (int r = 0; r < WR_U; r++) {
      if (u[r] < 0) continue;
      if (cu[r] == c) x += w[r];
      if (st.track) mvn += st.moved_it[u[r]] == st.iter;
    }
  }
}

template <int W, class T>
__global__ __launch_bounds__(256) void k_weight_group(DevGraph<T> g, DevState<T> st,
                                                       const int32_t* __restrict__ order,
                                                       const int32_t* __restrict__ bstart, int bin,
                                                       int only_dirty) {
  using tt = typename T::tot_t;
  const int lo = bstart[bin], hi = bstart[bin + 1];
  const int lane = threadIdx.x & (W - 1);
  const int groups = gridDim.x * (blockDim.x / W);
  for (int i = lo + (blockIdx.x * blockDim.x + threadIdx.x) / W; i < hi; i += groups) {
    const int v = order[i];
    if (only_dirty && !st.dirty[v]) continue;  // uniform per group
    const int64_t o = g.off[v];
    const int d = (int)(g.end[v] - o);
    const int c = st.cur[v];
    tt x = 0;
    int mvn = 0;
    row_in_span<W, T>(g, st, o, d, lane, c, x, mvn);
    x = sum_xor<tt, W>(x);
    if (st.track) mvn = sum_xor<int, W>(mvn);
    if (lane == 0) {
      st.In[v] = x;
      if (st.track) st.nbr_moved[v] = d ? (float)mvn / (float)d : 0.f;
    }
  }
}

// Weight update, one thread per vertex (degree <= DMAX): the lane loads its
// row, then the communities of all entries, DMAX independent gathers in flight.
template <int DMAX, class T>
__global__ __launch_bounds__(256) void k_weight_tpv(DevGraph<T> g, DevState<T> st,
                                                     const int32_t* __restrict__ order,
                                                     const int32_t* __restrict__ bstart, int bin, int only_dirty) {
  using tt = typename T::tot_t;
  const int lo = bstart[bin], hi = bstart[bin + 1];
  for (int i = lo + blockIdx.x * blockDim.x + threadIdx.x; i < hi; i += gridDim.x * blockDim.x) {
    const int v = order[i];
    if (only_dirty && !st.dirty[v]) continue;
    const int64_t o = g.off[v];
    const int d = (int)(g.end[v] - o);
    const int c = st.cur[v];
    int us[DMAX];
    tt ws[DMAX];
#pragma unroll
    for (int k = 0; k < DMAX; k++) {
      us[k] = k < d ? g.nbr[o + k] : v;
      ws[k] = k < d ? (tt)load_w<T>(g, o + k) : tt(0);
    }
    tt x = 0;
    int mvn = 0;
#pragma unroll
    for (int k = 0; k < DMAX; k++) {
      if (k < d) {
        if (st.cur[us[k]] == c) x += ws[k];
        if (st.track) mvn += st.moved_it[us[k]] == st.iter;
      }
    }
    st.In[v] = x;
    if (st.track) st.nbr_moved[v] = d ? (float)mvn / (float)d : 0.f;
  }
}

// Pipelined weight update: W lanes x EPL entries per vertex (degree <= W*EPL),
// five items in flight per group, one link of the load chain each
// (order -> dirty flag -> offsets and own community -> row -> neighbour
// communities).  Clean vertices pass through without further loads.
template <int W, int EPL, class T>
__global__ __launch_bounds__(256) void k_weight_pipe(DevGraph<T> g, DevState<T> st,
                                                      const int32_t* __restrict__ order,
                                                      const int32_t* __restrict__ bstart, int bin,
                                                      int only_dirty) {
  using tt = typename T::tot_t;
  using acc = typename T::acc_t;
  const int lo = bstart[bin], hi = bstart[bin + 1];
  const int lane = threadIdx.x & (W - 1);
  const int stride = gridDim.x * (blockDim.x / W);
  const int track = st.track;
  // A: vertex id
  int iA = lo + (blockIdx.x * blockDim.x + threadIdx.x) / W;
  bool okA = iA < hi;
  int vA = okA ? order[iA] : 0;
  // B: dirty flag
  bool okB = false, dyB = false;
  int vB = 0;
  // C: row header and own community
  bool okC = false;
  int vC = 0, dC = 0, cC = 0;
  int64_t oC = 0;
  // D: row entries
  bool okD = false;
  int vD = 0, dD = 0, cD = 0;
  int uD[EPL];
  acc wD[EPL];
  // E: neighbour communities
  bool okE = false;
  int vE = 0, dE = 0, cE = 0;
  int cuE[EPL], mtE[EPL];
  acc wE[EPL];
#pragma unroll
  for (int k = 0; k < EPL; k++) { uD[k] = 0; wD[k] = 0; cuE[k] = -1; mtE[k] = 0; wE[k] = 0; }
  while (okA || okB || okC || okD || okE) {  // uniform per group
    if (okE) {
      tt x = 0;
      int mvn = 0;
#pragma unroll
      for (int k = 0; k < EPL; k++) {
        if (lane + k * W < dE) {
          if (cuE[k] == cE) x += (tt)wE[k];
          mvn += mtE[k] == st.iter;
        }
      }
      x = sum_xor<tt, W>(x);
      if (track) mvn = sum_xor<int, W>(mvn);
      if (lane == 0) {
        st.In[vE] = x;
        if (track) st.nbr_moved[vE] = dE ? (float)mvn / (float)dE : 0.f;
      }
    }
    okE = okD;
    if (okD) {
      vE = vD; dE = dD; cE = cD;
#pragma unroll
      for (int k = 0; k < EPL; k++) {
        wE[k] = wD[k];
        if (lane + k * W < dD) {
          cuE[k] = st.cur[uD[k]];
          if (track) mtE[k] = st.moved_it[uD[k]];
        }
      }
    }
    okD = okC;
    if (okC) {
      vD = vC; dD = dC; cD = cC;
#pragma unroll
      for (int k = 0; k < EPL; k++) {
        if (lane + k * W < dC) {
          uD[k] = g.nbr[oC + lane + k * W];
          wD[k] = (acc)load_w<T>(g, oC + lane + k * W);
        }
      }
    }
    okC = okB && dyB;  // clean vertices leave the pipeline here
    if (okC) {
      vC = vB;
      oC = g.off[vB];
      dC = (int)(g.end[vB] - oC);
      cC = st.cur[vB];
    }
    okB = okA;
    if (okA) {
      vB = vA;
      dyB = !only_dirty || st.dirty[vA];
      iA += stride;
      okA = iA < hi;
      if (okA) vA = order[iA];
    }
  }
}

template <int NT, class T>
__global__ __launch_bounds__(NT) void k_weight_block(DevGraph<T> g, DevState<T> st,
                                                      const int32_t* __restrict__ order,
                                                      const int32_t* __restrict__ bstart, int bin,
                                                      int only_dirty) {
  using tt = typename T::tot_t;
  __shared__ tt r[NT / 64];
  __shared__ int rm[NT / 64];
  const int lo = bstart[bin], hi = bstart[bin + 1];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int i = lo + blockIdx.x; i < hi; i += gridDim.x) {
    const int v = order[i];
    if (only_dirty && !st.dirty[v]) continue;  // uniform per block
    const int64_t o = g.off[v];
    const int d = (int)(g.end[v] - o);
    const int c = st.cur[v];
    tt x = 0;
    int mvn = 0;
    row_in_span<NT, T>(g, st, o, d, (int)threadIdx.x, c, x, mvn);
    x = sum_xor<tt, 64>(x);
    mvn = sum_xor<int, 64>(mvn);
    if (lane == 0) { r[wid] = x; rm[wid] = mvn; }
    __syncthreads();
    if (threadIdx.x == 0) {
      x = 0;
      mvn = 0;
      for (int q = 0; q < NT / 64; q++) { x += r[q]; mvn += rm[q]; }
      st.In[v] = x;
      if (st.track) st.nbr_moved[v] = d ? (float)mvn / (float)d : 0.f;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Incremental weight update: only the rows of the vertices that
// moved in the sweep (cur != prev, read as one {cur, prev} pair) are read.  A mover v recomputes In[v] over
// its row; for every neighbour u that did not move, the entry's contribution
// to In[u] changes by w * ([cur[v] == c_u] - [prev[v] == c_u]), which v adds
// to In[u] (exact integer atomics).  A vertex none of whose neighbours moved
// keeps its In.  Work per sweep: the movers' rows, against the rows of the
// movers and all their neighbours for the dirty-row update.  (Summing the
// deltas for hubs per workgroup in LDS first, or rescanning staying hubs
// whole, both measured slower on the power-law workloads.)
// ---------------------------------------------------------------------------
// One mover's row, NT lanes from `lane`: returns the lane's part of the
// mover's new In (entries whose {cur, prev} pair has cur == c) and adds each
// staying neighbour's change.  The row loads and pair gathers of MV_U entries
// per lane are issued before their atomics (which may alias them for the
// compiler).
constexpr int MV_U = 4;
template <int NT, class T>
__device__ __forceinline__ typename T::tot_t moved_row_span(const DevGraph<T>& g, DevState<T>& st, int64_t o,
                                                            int d, int lane, int c, int p) {
  using tt = typename T::tot_t;
  tt x = 0;
  for (int j0 = lane; j0 < d; j0 += NT * MV_U) {
    int u[MV_U];
    tt w[MV_U];
    int cmu[MV_U];
#pragma unroll
    for (int r = 0; r < MV_U; r++) {
      const int j = j0 + r * NT;
      u[r] = j < d ? g.nbr[o + j] : -1;
      w[r] = j < d ? (tt)load_w<T>(g, o + j) : tt(0);
    }
    // cur[u] with the moved flag in bit 31; -1 past the row (community
    // 0x7fffffff, moved: no contribution)
#pragma unroll
    for (int r = 0; r < MV_U; r++) cmu[r] = u[r] >= 0 ? st.cm[u[r]] : -1;
#pragma unroll
    for (int r = 0; r < MV_U; r++) {
      const int cu = cmu[r] & 0x7fffffff;
      if (cu == c) x += w[r];
      if (cmu[r] >= 0) {  // u stayed: its In gains or loses this entry
        const tt dl = (cu == c ? w[r] : tt(0)) - (cu == p ? w[r] : tt(0));
        if (dl) atomicAdd(&st.In[u[r]], dl);
      }
    }
  }
  return x;
}

template <int W, class T>
__global__ __launch_bounds__(256) void k_weight_moved(DevGraph<T> g, DevState<T> st,
                                                       const int32_t* __restrict__ order,
                                                       const int32_t* __restrict__ bstart, int bin) {
  using tt = typename T::tot_t;
  const int lo = bstart[bin], hi = bstart[bin + 1];
  const int lane = threadIdx.x & (W - 1);
  const int groups = gridDim.x * (blockDim.x / W);
  for (int i = lo + (blockIdx.x * blockDim.x + threadIdx.x) / W; i < hi; i += groups) {
    const int v = order[i];
    const int2 cpv = st.cp[v];
    const int c = cpv.x, p = cpv.y;
    if (c == p) continue;  // uniform per group
    const int64_t o = g.off[v];
    const int d = (int)(g.end[v] - o);
    const tt x = sum_xor<tt, W>(moved_row_span<W, T>(g, st, o, d, lane, c, p));
    if (lane == 0) st.In[v] = x;
  }
}

template <int NT, class T>
__global__ __launch_bounds__(NT) void k_weight_moved_block(DevGraph<T> g, DevState<T> st,
                                                            const int32_t* __restrict__ order,
                                                            const int32_t* __restrict__ bstart, int bin) {
  using tt = typename T::tot_t;
  __shared__ tt r[NT / 64];
  const int lo = bstart[bin], hi = bstart[bin + 1];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int i = lo + blockIdx.x; i < hi; i += gridDim.x) {
    const int v = order[i];
    const int2 cpv = st.cp[v];
    const int c = cpv.x, p = cpv.y;
    if (c == p) continue;  // uniform per block
    const int64_t o = g.off[v];
    const int d = (int)(g.end[v] - o);
    tt x = moved_row_span<NT, T>(g, st, o, d, (int)threadIdx.x, c, p);
    x = sum_xor<tt, 64>(x);
    if (lane == 0) r[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
      x = 0;
      for (int q = 0; q < NT / 64; q++) x += r[q];
      st.In[v] = x;
    }
    __syncthreads();
  }
}

// variant: lanes per row for bins 0-5 -- 0: 4, 8, 16, 32, 64, 64; 1: 4, 4, 8,
// 16, 32, 32; 2: 4, 4, 8, 8, 16, 32 (narrower groups: more rows in flight per
// wave, fewer idle lanes on short rows)
template <class T>
void launch_weight_moved(int bin, const DevGraph<T>& g, DevState<T>& st, const int32_t* order,
                         const int32_t* bstart, int cap, hipStream_t s, int variant) {
  if (cap <= 0) return;
  static const int lanes[3][6] = {{4, 8, 16, 32, 64, 64}, {4, 4, 8, 16, 32, 32}, {4, 4, 8, 8, 16, 32}};
  if (variant < 0 || variant > 2) variant = 0;
  if (bin <= 5) {
    switch (lanes[variant][bin]) {
      case 4: k_weight_moved<4, T><<<grid_for(cap, 64, 8192), 256, 0, s>>>(g, st, order, bstart, bin); break;
      case 8: k_weight_moved<8, T><<<grid_for(cap, 32, 8192), 256, 0, s>>>(g, st, order, bstart, bin); break;
      case 16: k_weight_moved<16, T><<<grid_for(cap, 16, 8192), 256, 0, s>>>(g, st, order, bstart, bin); break;
      case 32: k_weight_moved<32, T><<<grid_for(cap, 8, 8192), 256, 0, s>>>(g, st, order, bstart, bin); break;
      default: k_weight_moved<64, T><<<grid_for(cap, 4, 8192), 256, 0, s>>>(g, st, order, bstart, bin); break;
    }
    return;
  }
  switch (bin) {
    case 6: k_weight_moved_block<256, T><<<grid_for(cap, 1, 4096), 256, 0, s>>>(g, st, order, bstart, 6); break;
    case 7: k_weight_moved_block<1024, T><<<grid_for(cap, 1, 1024), 1024, 0, s>>>(g, st, order, bstart, 7); break;
    case 8: k_weight_moved_block<1024, T><<<grid_for(cap, 1, 1024), 1024, 0, s>>>(g, st, order, bstart, 8); break;
  }
}

template <class T>
void launch_weight(int bin, const DevGraph<T>& g, DevState<T>& st, const int32_t* order,
                   const int32_t* bstart, int cap, bool only_dirty, hipStream_t s, int variant) {
  if (cap <= 0) return;
  const int od = only_dirty ? 1 : 0;
  if (variant == 2 && bin <= 3) {  // one thread per vertex
    switch (bin) {
      case 0: k_weight_tpv<3, T><<<grid_for(cap, 256, 8192), 256, 0, s>>>(g, st, order, bstart, 0, od); break;
      case 1: k_weight_tpv<7, T><<<grid_for(cap, 256, 8192), 256, 0, s>>>(g, st, order, bstart, 1, od); break;
      case 2: k_weight_tpv<15, T><<<grid_for(cap, 256, 8192), 256, 0, s>>>(g, st, order, bstart, 2, od); break;
      case 3: k_weight_tpv<31, T><<<grid_for(cap, 256, 8192), 256, 0, s>>>(g, st, order, bstart, 3, od); break;
    }
    return;
  }
  if (variant == 1 && bin <= 4) {  // pipelined
    switch (bin) {
      case 0: k_weight_pipe<4, 1, T><<<grid_for(cap, 256, 8192), 256, 0, s>>>(g, st, order, bstart, 0, od); break;
      case 1: k_weight_pipe<8, 1, T><<<grid_for(cap, 128, 8192), 256, 0, s>>>(g, st, order, bstart, 1, od); break;
      case 2: k_weight_pipe<16, 1, T><<<grid_for(cap, 64, 8192), 256, 0, s>>>(g, st, order, bstart, 2, od); break;
      case 3: k_weight_pipe<32, 1, T><<<grid_for(cap, 32, 8192), 256, 0, s>>>(g, st, order, bstart, 3, od); break;
      case 4: k_weight_pipe<32, 4, T><<<grid_for(cap, 32, 8192), 256, 0, s>>>(g, st, order, bstart, 4, od); break;
    }
    return;
  }
  switch (bin) {
    case 0: k_weight_group<4, T><<<grid_for(cap, 64, 8192), 256, 0, s>>>(g, st, order, bstart, 0, od); break;
    case 1: k_weight_group<8, T><<<grid_for(cap, 32, 8192), 256, 0, s>>>(g, st, order, bstart, 1, od); break;
    case 2: k_weight_group<16, T><<<grid_for(cap, 16, 8192), 256, 0, s>>>(g, st, order, bstart, 2, od); break;
    case 3: k_weight_group<32, T><<<grid_for(cap, 8, 8192), 256, 0, s>>>(g, st, order, bstart, 3, od); break;
    case 4: k_weight_group<64, T><<<grid_for(cap, 4, 8192), 256, 0, s>>>(g, st, order, bstart, 4, od); break;
    case 5: k_weight_group<64, T><<<grid_for(cap, 4, 8192), 256, 0, s>>>(g, st, order, bstart, 5, od); break;
    case 6: k_weight_block<256, T><<<grid_for(cap, 1, 4096), 256, 0, s>>>(g, st, order, bstart, 6, od); break;
    case 7: k_weight_block<1024, T><<<grid_for(cap, 1, 1024), 1024, 0, s>>>(g, st, order, bstart, 7, od); break;
    case 8: k_weight_block<1024, T><<<grid_for(cap, 1, 1024), 1024, 0, s>>>(g, st, order, bstart, 8, od); break;
  }
}

// ---------------------------------------------------------------------------
// prune + sum In over every vertex of positive degree, clears the dirty marks.
// (the prune test of compute_In, kernel_functions.cu:944,1028)
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_prune(DevState<T> st, int32_t n, typename T::tot_t m2,
                                               const int32_t* __restrict__ vlist, const int32_t* __restrict__ vcnt) {
  using tt = typename T::tot_t;
  __shared__ unsigned long long s_in[4], s_act[4];
  const tt minTot = read_min_tot<T>(st.ctr);
  unsigned long long insum = 0, act = 0;  // sum In <= m2 < 2^62
  if (vlist) n = *vcnt;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int v = vlist ? vlist[i] : i;
    if (st.binid[v] < 0) continue;
    const tt x = st.In[v];
    const uint8_t keep = prune_keep<T>(x, st.K[v], st.Self[v], st.Tot[st.cur[v]], minTot, vertex_m2<T>(st, v, m2));
    st.active[v] = keep;
    if (st.track && st.dirty && !st.dirty[v]) st.nbr_moved[v] = 0.f;  // no neighbour moved
    st.dirty[v] = 0;
    act += keep;
    insum += (unsigned long long)(long long)x;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  insum = sum_xor<unsigned long long, 64>(insum);
  act = sum_xor<unsigned long long, 64>(act);
  if (lane == 0) { s_in[wid] = insum; s_act[wid] = act; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < 4; q++) { insum += s_in[q]; act += s_act[q]; }
    BlockPart bp{};
    bp.s0 = insum;
    bp.s1 = act;
    st.part[blockIdx.x] = bp;
  }
}

// Combines the nb workgroup partials of k_prune into the counters.
template <class T>
__global__ __launch_bounds__(256) void k_prune_fin(DevState<T> st, int nb) {
  __shared__ unsigned long long s_in[4], s_act[4];
  unsigned long long insum = 0, act = 0;
  for (int b = threadIdx.x; b < nb; b += 256) {
    const BlockPart p = st.part[b];
    insum += p.s0;
    act += p.s1;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  insum = sum_xor<unsigned long long, 64>(insum);
  act = sum_xor<unsigned long long, 64>(act);
  if (lane == 0) { s_in[wid] = insum; s_act[wid] = act; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < 4; q++) { insum += s_in[q]; act += s_act[q]; }
    atomic_add_ull(&st.ctr->sum_in, insum);
    atomic_add_ull(&st.ctr->active, act);
  }
}

template <class T>
void launch_prune(DevState<T>& st, int32_t n, typename T::tot_t m2, hipStream_t s, const int32_t* vlist,
                  const int32_t* vcnt) {
  if (n <= 0) return;
  const int nb = std::min((n + 255) / 256, MAX_PARTS);
  k_prune<T><<<nb, 256, 0, s>>>(st, n, m2, vlist, vcnt);
  k_prune_fin<T><<<1, 256, 0, s>>>(st, nb);
}

// explicit instantiations
#define GALA_INST(T)                                                                              \
  template void launch_init<T>(const DevGraph<T>&, DevState<T>&, hipStream_t, bool);                    \
  template void launch_build_cv<T>(DevState<T>&, int32_t, hipStream_t, const int32_t*, const int32_t*);                           \
  template void launch_init_rows<T>(const DevGraph<T>&, DevState<T>&, const int32_t*, const int32_t*,         \
                                    const int32_t*, hipStream_t);                                  \
  template void launch_decide<T>(int, const DevGraph<T>&, DevState<T>&, const int32_t*,           \
                                 const int32_t*, int, typename T::tot_t, hipStream_t, int);       \
  template void launch_decide_first<T>(int, const DevGraph<T>&, DevState<T>&, const int32_t*,     \
                                       const int32_t*, int, typename T::tot_t, hipStream_t, int); \
  template void launch_apply<T>(DevState<T>&, int32_t, typename T::tot_t, hipStream_t, const int32_t*,        \
                                const int32_t*, const int32_t*, const int32_t*, bool);            \
  template void launch_weight<T>(int, const DevGraph<T>&, DevState<T>&, const int32_t*,           \
                                 const int32_t*, int, bool, hipStream_t, int);                    \
  template void launch_prune<T>(DevState<T>&, int32_t, typename T::tot_t, hipStream_t, const int32_t*,        \
                                const int32_t*);
GALA_INST(IntW)
GALA_INST(WideW)
template void launch_weight_moved<IntW>(int, const DevGraph<IntW>&, DevState<IntW>&, const int32_t*, const int32_t*,
                                        int, hipStream_t, int);
template void launch_weight_moved<WideW>(int, const DevGraph<WideW>&, DevState<WideW>&, const int32_t*,
                                         const int32_t*, int, hipStream_t, int);

}  // namespace gala
