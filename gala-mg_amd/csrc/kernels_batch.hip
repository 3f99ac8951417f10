// Batched Louvain: many independent graphs laid out as one block-diagonal
// union (graph g owns a contiguous vertex range, and community ids are vertex
// ids, so no community ever spans two graphs).  Every graph keeps its own m2,
// modularity, convergence and round count; the rules applied per graph are the
// plain ones (louvain_main_process.cu:1004-1007,1026-1031 for the iteration
// break with revert, louvain.cu:193-226 for the round loop), decided on the
// device so the iteration loop needs no host round trip per graph.
#include "gala_kernels.h"
#include "gala_device.h"

namespace gala {

// Q = sum_in/m2 - sum_tot2/m2^2 with every operation rounded once, the same
// sequence the host evaluates for a single graph.
__device__ __forceinline__ double q_exact_dev(unsigned long long sin, unsigned long long st2, double m2) {
  return __dsub_rn(__ddiv_rn((double)sin, m2), __ddiv_rn((double)st2, __dmul_rn(m2, m2)));
}

// Adds x into tab[key] for the 64 lanes of a wave.  Vertex ranges of graphs
// are contiguous, so a wave mostly carries one key: then one atomic per wave.
// Lanes with key < 0 contribute nothing.  Called by all lanes.
__device__ __forceinline__ void wave_seg_add(unsigned long long* tab, int key, unsigned long long x) {
  const int k0 = __shfl(key, 0, 64);
  if (__all(key == k0)) {
    const unsigned long long s = wave_sum<unsigned long long>(x);
    if ((threadIdx.x & 63) == 0 && k0 >= 0 && s) atomicAdd(&tab[k0], s);
  } else if (key >= 0 && x) {
    atomicAdd(&tab[key], x);
  }
}

// Per-graph sum of In over vertices of positive degree and of Tot^2 over
// communities (community c belongs to the graph of vertex c).
__global__ __launch_bounds__(256) void k_graph_sums(const int32_t* __restrict__ gid, const int8_t* __restrict__ binid,
                                                     const int32_t* __restrict__ In, const int32_t* __restrict__ Tot,
                                                     int32_t n, unsigned long long* g_in, unsigned long long* g_tot2,
                                                     const int32_t* __restrict__ vlist,
                                                     const int32_t* __restrict__ vcnt) {
  const int lane = threadIdx.x & 63;
  if (vlist) n = *vcnt;  // listed vertices (grouped by bin, ascending ids within a bin)
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x - lane); base < n;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = base + lane;
    int key = -1;
    unsigned long long xi = 0, xt = 0;
    if (i < n) {
      const int v = vlist ? vlist[i] : (int)i;
      key = gid[v];
      const long long t = Tot[v];
      xt = (unsigned long long)(t * t);
      if (binid[v] >= 0) xi = (unsigned long long)(long long)In[v];
    }
    wave_seg_add(g_in, key, xi);
    wave_seg_add(g_tot2, key, xt);
  }
}

// One thread per graph: Q of the sweep just done, then the plain rule.
// it == 0 is the singleton partition of the round (Q becomes the reference).
// *p += the number of lanes of the wave with flag set; every lane calls it
// (one atomic per wave: thousands of per-graph atomics on one counter queue
// behind each other in L2)
__device__ __forceinline__ void wave_count_add(int32_t* p, bool flag) {
  const unsigned long long m = __ballot(flag);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(p, (int32_t)__popcll(m));
}

__global__ void k_batch_step(BatchDev b, const int32_t* __restrict__ it_dev, double min_mod, int max_it) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  // returns whether graph g iterates on (counted in *b.nlive)
  auto step = [&]() -> bool {
    if (g >= b.G) return false;
    const int it = *it_dev;
    const unsigned long long si = b.g_in[g], st2 = b.g_tot2[g];
    b.g_in[g] = 0;
    b.g_tot2[g] = 0;
    b.accept[g] = 0;
    if (!b.live[g]) return false;
    const double q = q_exact_dev(si, st2, (double)b.m2g[g]);
    if (it == 0) {
      b.qprev[g] = q;
      return true;
    }
    b.iters[g] = it;
    if (__dsub_rn(q, b.qprev[g]) < min_mod || it >= max_it) {
      b.live[g] = 0;  // revert: prev keeps the previous assignment
      return false;
    }
    b.qprev[g] = q;
    b.accept[g] = 1;
    return true;
  };
  wave_count_add(b.nlive, step());
}

// prev = cur for the graphs whose sweep was accepted; vertices of graphs that
// stopped leave the frontier.  first: the round's first frontier is every
// vertex of a live graph.  Advances the iteration counter.  vlive (if given)
// receives whether the vertex's graph iterates on: the next sweep's list.
__global__ __launch_bounds__(256) void k_batch_mask(BatchDev b, const int32_t* __restrict__ gid,
                                                     const int32_t* __restrict__ cur, int32_t* __restrict__ prev,
                                                     uint8_t* __restrict__ active, int32_t n, int first,
                                                     int32_t* __restrict__ it_dev, int2* __restrict__ cp,
                                                     int32_t* __restrict__ cm, const int32_t* __restrict__ vlist,
                                                     const int32_t* __restrict__ vcnt, uint8_t* __restrict__ vlive) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *it_dev += 1;
  if (vlist) n = *vcnt;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int v = vlist ? vlist[i] : i;
    const int g = gid[v];
    if (vlive) vlive[v] = b.live[g];
    if (first) {
      active[v] = b.live[g];
      continue;
    }
    int p = prev[v];
    if (b.accept[g]) prev[v] = p = cur[v];
    if (cp) cp[v] = make_int2(p, p);  // no vertex has moved in the next sweep yet
    if (cm) cm[v] = p;
    if (!b.live[g]) active[v] = 0;
  }
}

// Round start: a graph iterates unless it is done or has no edge.
__global__ void k_batch_live(BatchDev b, int32_t* __restrict__ it_dev) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g == 0) *it_dev = 0;
  if (g == 0) *b.nlive = 0;
  if (g >= b.G) return;
  b.live[g] = (!b.done[g] && b.m2g[g] > 0) ? 1 : 0;
  b.iters[g] = 0;
  b.g_in[g] = 0;
  b.g_tot2[g] = 0;
}

// Round end per graph (louvain.cu:193-226 restated per graph): the accepted Q
// of the round, the round-level stop test, totals.
__global__ void k_batch_round(BatchDev b, int round, double min_mod, int max_rounds, int32_t* __restrict__ nopen) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  // returns whether graph g goes on to another round (counted in *nopen)
  auto end_round = [&]() -> bool {
    if (g >= b.G || b.done[g]) return false;
    if (!(b.m2g[g] > 0)) {  // no edge: every vertex stays alone, Q = 0
      b.done[g] = 1;
      b.cur_mod[g] = 0.0;
      return false;
    }
    const double cm = b.qprev[g];
    const double pm = round > 0 ? b.cur_mod[g] : -1.0;
    b.cur_mod[g] = cm;
    b.rounds[g] = round + 1;
    b.tot_iters[g] += b.iters[g];
    if (!(__dsub_rn(cm, pm) >= min_mod) || round + 1 >= max_rounds) {
      b.done[g] = 1;
      return false;
    }
    return true;
  };
  wave_count_add(nopen, end_round());
}

// Graph of every vertex of the contracted union: new id nid[c] of each
// non-empty community c inherits the graph of c.
__global__ __launch_bounds__(256) void k_batch_gid_next(const int32_t* __restrict__ gid, const int32_t* __restrict__ cnt,
                                                         const int32_t* __restrict__ nid, int32_t n,
                                                         int32_t* __restrict__ gid_next) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
    if (cnt[c] > 0) gid_next[nid[c]] = gid[c];
}

// Communities of the final union graph per graph (gid_final is non-decreasing).
__global__ __launch_bounds__(256) void k_batch_count(const int32_t* __restrict__ gid, int32_t n,
                                                      int32_t* __restrict__ per_graph) {
  // graphs own contiguous vertex ranges: a wave mostly adds to one graph
  for (int v0 = blockIdx.x * blockDim.x; v0 < n; v0 += gridDim.x * blockDim.x) {
    const int v = v0 + threadIdx.x;
    const int key = v < n ? gid[v] : -1;
    const int k0 = __shfl(key, 0, 64);
    if (__all(key == k0)) {
      if ((threadIdx.x & 63) == 0 && k0 >= 0) atomicAdd(&per_graph[k0], 64);
    } else if (key >= 0) {
      atomicAdd(&per_graph[key], 1);
    }
  }
}

// Per-iteration counters of the union (the decide counters keep accumulating).
__global__ void k_batch_reset(Counters* c, long long min_init, int32_t* nlive) {
  if (threadIdx.x == 0) {
    c->sum_in = 0;
    c->sum_tot2 = 0;
    c->min_tot = min_init;
    c->active = 0;
    c->nonempty = 0;
    *nlive = 0;
  }
}

static int blocks_for(int64_t n, int per, int cap) {
  int64_t nb = (n + per - 1) / per;
  if (nb < 1) nb = 1;
  return (int)std::min<int64_t>(nb, cap);
}

void launch_graph_sums(const int32_t* gid, const int8_t* binid, const int32_t* In, const int32_t* Tot, int32_t n,
                       unsigned long long* g_in, unsigned long long* g_tot2, hipStream_t s, const int32_t* vlist,
                       const int32_t* vcnt) {
  if (n <= 0) return;
  k_graph_sums<<<blocks_for(n, 256, 8192), 256, 0, s>>>(gid, binid, In, Tot, n, g_in, g_tot2, vlist, vcnt);
}
void launch_batch_step(const BatchDev& b, const int32_t* it_dev, double min_mod, int max_it, hipStream_t s) {
  k_batch_step<<<blocks_for(b.G, 256, 1 << 20), 256, 0, s>>>(b, it_dev, min_mod, max_it);
}
void launch_batch_mask(const BatchDev& b, const int32_t* gid, const int32_t* cur, int32_t* prev, uint8_t* active,
                       int32_t n, bool first, int32_t* it_dev, int2* cp, int32_t* cm, hipStream_t s,
                       const int32_t* vlist, const int32_t* vcnt, uint8_t* vlive) {
  k_batch_mask<<<blocks_for(n, 256, 8192), 256, 0, s>>>(b, gid, cur, prev, active, n, first ? 1 : 0, it_dev, cp, cm,
                                                        vlist, vcnt, vlive);
}
void launch_batch_live(const BatchDev& b, int32_t* it_dev, hipStream_t s) {
  k_batch_live<<<blocks_for(b.G, 256, 1 << 20), 256, 0, s>>>(b, it_dev);
}
void launch_batch_round(const BatchDev& b, int round, double min_mod, int max_rounds, int32_t* nopen, hipStream_t s) {
  k_batch_round<<<blocks_for(b.G, 256, 1 << 20), 256, 0, s>>>(b, round, min_mod, max_rounds, nopen);
}
void launch_batch_gid_next(const int32_t* gid, const int32_t* cnt, const int32_t* nid, int32_t n, int32_t* gid_next,
                           hipStream_t s) {
  if (n <= 0) return;
  k_batch_gid_next<<<blocks_for(n, 256, 8192), 256, 0, s>>>(gid, cnt, nid, n, gid_next);
}
void launch_batch_count(const int32_t* gid, int32_t n, int32_t* per_graph, hipStream_t s) {
  if (n <= 0) return;
  k_batch_count<<<blocks_for(n, 256, 8192), 256, 0, s>>>(gid, n, per_graph);
}
void launch_batch_reset(Counters* c, long long min_init, int32_t* nlive, hipStream_t s) {
  k_batch_reset<<<1, 64, 0, s>>>(c, min_init, nlive);
}

}  // namespace gala
