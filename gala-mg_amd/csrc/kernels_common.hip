// Scans, stable partition, fills and small reductions used by every phase.
// All kernels are written for gfx950 wave64; block sizes are multiples of 64.
#include "gala_kernels.h"
#include "gala_device.h"

namespace gala {

// ---------------------------------------------------------------------------
// device-wide exclusive scan (reduce-then-scan, three launches)
// ---------------------------------------------------------------------------
constexpr int SCAN_T = 256, SCAN_IPT = 16, SCAN_TILE = SCAN_T * SCAN_IPT;

template <class In, class V>
__global__ __launch_bounds__(SCAN_T) void k_scan_reduce(const In* __restrict__ in, int64_t n,
                                                        V* __restrict__ part) {
  __shared__ V sh[17];
  int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  // all loads first, at indices clamped into [0, n) (n >= 1 for a launched
  // scan), so that none waits behind a branch for the one before it
  In x[SCAN_IPT];
#pragma unroll
  for (int k = 0; k < SCAN_IPT; k++) {
    const int64_t i = base + (int64_t)k * SCAN_T + threadIdx.x;
    x[k] = in[i < n ? i : n - 1];
  }
  V s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_IPT; k++) {
    const int64_t i = base + (int64_t)k * SCAN_T + threadIdx.x;
    s += i < n ? (V)x[k] : V(0);
  }
  V tot;
  block_excl_scan(s, sh, &tot);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

template <class V>
__global__ __launch_bounds__(1024) void k_scan_top(V* __restrict__ part, int nb, V* __restrict__ total) {
  __shared__ V sh[17];
  // each thread owns a contiguous chunk
  int per = (nb + 1023) / 1024;
  int lo = threadIdx.x * per, hi = min(nb, lo + per);
  V s = 0;
  for (int i = lo; i < hi; i++) s += part[i];
  V tot;
  V pre = block_excl_scan(s, sh, &tot);
  for (int i = lo; i < hi; i++) { V t = part[i]; part[i] = pre; pre += t; }
  if (threadIdx.x == 0 && total) *total = tot;
}

template <class In, class V>
__global__ __launch_bounds__(SCAN_T) void k_scan_down(const In* __restrict__ in, int64_t n,
                                                      const V* __restrict__ part, V* __restrict__ out) {
  __shared__ V sh[17];
  int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_IPT;
  In x[SCAN_IPT];
#pragma unroll
  for (int k = 0; k < SCAN_IPT; k++) {
    const int64_t i = base + k;
    x[k] = in[i < n ? i : n - 1];  // clamped, unconditional (as in k_scan_reduce)
  }
  V loc[SCAN_IPT];
  V s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_IPT; k++) {
    const int64_t i = base + k;
    loc[k] = i < n ? (V)x[k] : V(0);
    s += loc[k];
  }
  V tot;
  V pre = block_excl_scan(s, sh, &tot) + part[blockIdx.x];
#pragma unroll
  for (int k = 0; k < SCAN_IPT; k++) {
    int64_t i = base + k;
    if (i < n) out[i] = pre;
    pre += loc[k];
  }
}

int64_t scan_tmp_elems(int64_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE + 1; }

template <class In, class V>
void scan_exclusive(const In* in, V* out, int64_t n, V* tmp, V* total, hipStream_t s) {
  int nb = (int)((n + SCAN_TILE - 1) / SCAN_TILE);
  if (nb == 0) {
    (void)hipMemsetAsync(total, 0, sizeof(V), s);
    return;
  }
  k_scan_reduce<In, V><<<nb, SCAN_T, 0, s>>>(in, n, tmp);
  k_scan_top<V><<<1, 1024, 0, s>>>(tmp, nb, total);
  k_scan_down<In, V><<<nb, SCAN_T, 0, s>>>(in, n, tmp, out);
}

template void scan_exclusive<int32_t, int32_t>(const int32_t*, int32_t*, int64_t, int32_t*, int32_t*, hipStream_t);
template void scan_exclusive<int32_t, int64_t>(const int32_t*, int64_t*, int64_t, int64_t*, int64_t*, hipStream_t);
template void scan_exclusive<int64_t, int64_t>(const int64_t*, int64_t*, int64_t, int64_t*, int64_t*, hipStream_t);
template void scan_exclusive<uint8_t, int32_t>(const uint8_t*, int32_t*, int64_t, int32_t*, int32_t*, hipStream_t);
template void scan_exclusive<uint8_t, int64_t>(const uint8_t*, int64_t*, int64_t, int64_t*, int64_t*, hipStream_t);

// ---------------------------------------------------------------------------
// Stable multi-way partition into NBINS bins (key < 0 drops the item).
//   mode 0: items are 0..n-1, key = binid[i]
//   mode 1: items are order[i], key = active[v] ? binid[v] : -1
// out receives the kept items grouped by bin in input order; bstart[NBINS+1].
// ---------------------------------------------------------------------------
constexpr int PART_T = 256, PART_IPT = 8, PART_TILE = PART_T * PART_IPT;

// The items base + k * stride (k < IPT): item id and key (-1: past n or
// dropped).  Every load is issued without a branch, at indices clamped into
// [0, n), so a thread's loads are in flight together instead of one behind
// another.  n >= 1.
template <int IPT>
__device__ __forceinline__ void part_items(int mode, const int32_t* __restrict__ order,
                                           const int8_t* __restrict__ binid, const uint8_t* __restrict__ active,
                                           int64_t base, int64_t stride, int64_t n, int (&v)[IPT], int (&key)[IPT]) {
  int64_t idx[IPT];
#pragma unroll
  for (int k = 0; k < IPT; k++) {
    const int64_t i = base + k * stride;
    idx[k] = i < n ? i : n - 1;
  }
  int8_t b[IPT];
  if (mode == 0) {
#pragma unroll
    for (int k = 0; k < IPT; k++) b[k] = binid[idx[k]];
#pragma unroll
    for (int k = 0; k < IPT; k++) {
      v[k] = (int)idx[k];
      key[k] = base + k * stride < n ? (int)b[k] : -1;
    }
  } else {
    // mode 1 keeps the flagged vertices of `order`, mode 2 the unflagged ones
#pragma unroll
    for (int k = 0; k < IPT; k++) v[k] = order[idx[k]];
    uint8_t a[IPT];
#pragma unroll
    for (int k = 0; k < IPT; k++) {
      a[k] = active[v[k]];
      b[k] = binid[v[k]];
    }
#pragma unroll
    for (int k = 0; k < IPT; k++)
      key[k] = base + k * stride < n && (a[k] != 0) == (mode == 1) ? (int)b[k] : -1;
  }
}

__global__ __launch_bounds__(PART_T) void k_part_count(int mode, const int32_t* __restrict__ order,
                                                       const int8_t* __restrict__ binid,
                                                       const uint8_t* __restrict__ active, int64_t n,
                                                       int32_t* __restrict__ cnt, int nblk) {
  __shared__ int sc[NBINS];
  if (threadIdx.x < NBINS) sc[threadIdx.x] = 0;
  __syncthreads();
  int64_t base = (int64_t)blockIdx.x * PART_TILE + (int64_t)threadIdx.x * PART_IPT;
  int c[NBINS];
#pragma unroll
  for (int b = 0; b < NBINS; b++) c[b] = 0;
  if (n > 0) {  // uniform
    int vs[PART_IPT], ks[PART_IPT];
    part_items<PART_IPT>(mode, order, binid, active, base, 1, n, vs, ks);
#pragma unroll
    for (int k = 0; k < PART_IPT; k++)
#pragma unroll
      for (int b = 0; b < NBINS; b++) c[b] += (ks[k] == b);
  }
#pragma unroll
  for (int b = 0; b < NBINS; b++) {
    int s = wave_sum(c[b]);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(&sc[b], s);
  }
  __syncthreads();
  if (threadIdx.x < NBINS) cnt[threadIdx.x * nblk + blockIdx.x] = sc[threadIdx.x];
}

// exclusive scan of the bin-major counts, staged through LDS in chunks of
// PART_TOP_CH so that the global loads and stores stay coalesced
constexpr int PART_TOP_CH = 16384;
__global__ __launch_bounds__(1024) void k_part_top(int32_t* __restrict__ cnt, int nblk,
                                                   int32_t* __restrict__ bstart) {
  constexpr int PER = PART_TOP_CH / 1024;
  __shared__ int32_t buf[PART_TOP_CH];
  __shared__ int32_t sh[17];
  const int total = NBINS * nblk;
  int32_t carry = 0;
  for (int c0 = 0; c0 < total; c0 += PART_TOP_CH) {
    const int len = min(PART_TOP_CH, total - c0);
    for (int i = threadIdx.x; i < len; i += 1024) buf[i] = cnt[c0 + i];
    __syncthreads();
    const int lo = threadIdx.x * PER;
    int32_t s = 0;
#pragma unroll
    for (int k = 0; k < PER; k++) s += lo + k < len ? buf[lo + k] : 0;
    int32_t tot;
    int32_t pre = block_excl_scan(s, sh, &tot) + carry;
#pragma unroll
    for (int k = 0; k < PER; k++) {
      if (lo + k < len) {
        const int32_t t = buf[lo + k];
        buf[lo + k] = pre;
        pre += t;
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < len; i += 1024) {
      const int g = c0 + i;
      cnt[g] = buf[i];
      if (g % nblk == 0) bstart[g / nblk] = buf[i];
    }
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) bstart[NBINS] = carry;
}

__global__ __launch_bounds__(PART_T) void k_part_scatter(int mode, const int32_t* __restrict__ order,
                                                         const int8_t* __restrict__ binid,
                                                         const uint8_t* __restrict__ active, int64_t n,
                                                         const int32_t* __restrict__ cnt, int nblk,
                                                         int32_t* __restrict__ out) {
  __shared__ int32_t sh[17];
  int64_t base = (int64_t)blockIdx.x * PART_TILE + (int64_t)threadIdx.x * PART_IPT;
  int vs[PART_IPT], ks[PART_IPT];
  int c[NBINS];
#pragma unroll
  for (int b = 0; b < NBINS; b++) c[b] = 0;
  if (n > 0) {  // uniform
    part_items<PART_IPT>(mode, order, binid, active, base, 1, n, vs, ks);
  } else {
#pragma unroll
    for (int k = 0; k < PART_IPT; k++) { vs[k] = 0; ks[k] = -1; }
  }
#pragma unroll
  for (int k = 0; k < PART_IPT; k++)
#pragma unroll
    for (int b = 0; b < NBINS; b++) c[b] += (ks[k] == b);
  int pos[NBINS];
#pragma unroll
  for (int b = 0; b < NBINS; b++) {
    int32_t tot;
    pos[b] = block_excl_scan<int32_t>(c[b], sh, &tot) + cnt[b * nblk + blockIdx.x];
  }
#pragma unroll
  for (int k = 0; k < PART_IPT; k++) {
    int key = ks[k];
    if (key >= 0) {
      int p = 0;
#pragma unroll
      for (int b = 0; b < NBINS; b++)
        if (key == b) { p = pos[b]; pos[b] = p + 1; }
      out[p] = vs[k];
    }
  }
}

int64_t part_tmp_elems(int64_t n) { return (int64_t)NBINS * ((n + PART_TILE - 1) / PART_TILE + 1); }

// ---------------------------------------------------------------------------
// Stable compaction of a bin-grouped list (modes 1 and 2): kept items keep
// their order, so the bins stay grouped; the new bin starts are the kept
// counts before the old ones.  Each wave owns 64 * CMP_IPT consecutive items
// and ranks them with ballots.
// ---------------------------------------------------------------------------
constexpr int CMP_T = 256, CMP_IPT = 16, CMP_TILE = CMP_T * CMP_IPT;

__global__ __launch_bounds__(CMP_T) void k_cmp_count(int mode, const int32_t* __restrict__ order,
                                                     const uint8_t* __restrict__ flag, int64_t n,
                                                     int32_t* __restrict__ cnt, const int32_t* __restrict__ n_dev) {
  __shared__ int32_t sh[17];
  if (n_dev && *n_dev < n) n = *n_dev;  // the list's length on the device (n is its bound)
  const int64_t base = (int64_t)blockIdx.x * CMP_TILE;
  int c = 0;
  if (n > 0) {  // uniform; loads at clamped indices, none behind a branch
    int vs[CMP_IPT];
#pragma unroll
    for (int k = 0; k < CMP_IPT; k++) {
      const int64_t i = base + k * CMP_T + threadIdx.x;
      vs[k] = order[i < n ? i : n - 1];
    }
    uint8_t f[CMP_IPT];
#pragma unroll
    for (int k = 0; k < CMP_IPT; k++) f[k] = flag[vs[k]];
#pragma unroll
    for (int k = 0; k < CMP_IPT; k++)
      c += (base + k * CMP_T + threadIdx.x < n && (f[k] != 0) == (mode == 1)) ? 1 : 0;
  }
  int32_t tot;
  (void)block_excl_scan<int32_t>(c, sh, &tot);
  if (threadIdx.x == 0) cnt[blockIdx.x] = tot;
}

// exclusive scan of the per-tile counts; bins that start at or past n start
// at the total
__global__ __launch_bounds__(1024) void k_cmp_top(int32_t* __restrict__ cnt, int nblk,
                                                  const int32_t* __restrict__ bstart_in, int64_t n,
                                                  int32_t* __restrict__ bstart, const int32_t* __restrict__ n_dev) {
  __shared__ int32_t sh[17];
  if (n_dev && *n_dev < n) n = *n_dev;
  int32_t carry = 0;
  for (int c0 = 0; c0 < nblk; c0 += 1024) {
    const int i = c0 + threadIdx.x;
    const int32_t x = i < nblk ? cnt[i] : 0;
    int32_t tot;
    const int32_t pre = block_excl_scan(x, sh, &tot);
    if (i < nblk) cnt[i] = pre + carry;
    carry += tot;
  }
  if (threadIdx.x <= NBINS && bstart_in[threadIdx.x] >= n) bstart[threadIdx.x] = carry;
}

__global__ __launch_bounds__(CMP_T) void k_cmp_scatter(int mode, const int32_t* __restrict__ order,
                                                       const uint8_t* __restrict__ flag, int64_t n,
                                                       const int32_t* __restrict__ cnt,
                                                       const int32_t* __restrict__ bstart_in,
                                                       int32_t* __restrict__ out, int32_t* __restrict__ bstart,
                                                       const int32_t* __restrict__ n_dev) {
  __shared__ int32_t wsum[CMP_T / 64];
  if (n_dev && *n_dev < n) n = *n_dev;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const int64_t wbase = (int64_t)blockIdx.x * CMP_TILE + (int64_t)w * 64 * CMP_IPT;
  int vs[CMP_IPT];
  uint8_t f[CMP_IPT];
  if (n > 0) {  // uniform; loads at clamped indices, none behind a branch
#pragma unroll
    for (int k = 0; k < CMP_IPT; k++) {
      const int64_t i = wbase + k * 64 + lane;
      vs[k] = order[i < n ? i : n - 1];
    }
#pragma unroll
    for (int k = 0; k < CMP_IPT; k++) f[k] = flag[vs[k]];
  } else {
#pragma unroll
    for (int k = 0; k < CMP_IPT; k++) { vs[k] = 0; f[k] = 0; }
  }
  unsigned keep = 0;
  int wc = 0;
#pragma unroll
  for (int k = 0; k < CMP_IPT; k++) {
    const int64_t i = wbase + k * 64 + lane;
    const bool kp = i < n && (f[k] != 0) == (mode == 1);
    keep |= (kp ? 1u : 0u) << k;
    wc += __popcll(__ballot(kp));
  }
  if (lane == 0) wsum[w] = wc;
  __syncthreads();
  int pre = cnt[blockIdx.x];
  for (int q = 0; q < w; q++) pre += wsum[q];
  int bs[NBINS];
#pragma unroll
  for (int b = 0; b < NBINS; b++) bs[b] = bstart_in[b];
#pragma unroll
  for (int k = 0; k < CMP_IPT; k++) {
    const int64_t i = wbase + k * 64 + lane;
    const bool kp = (keep >> k) & 1u;
    const unsigned long long bal = __ballot(kp);
    const int pos = pre + __popcll(bal & below);  // kept items before i
    if (kp) out[pos] = vs[k];
#pragma unroll
    for (int b = 0; b < NBINS; b++)
      if (bs[b] == i) bstart[b] = pos;
    pre += __popcll(bal);
  }
}

void partition_bins(int mode, const int32_t* order, const int32_t* order_bstart, const int8_t* binid,
                    const uint8_t* active, int64_t n, int32_t* out, int32_t* bstart, int32_t* tmp, hipStream_t s,
                    const int32_t* n_dev) {
  if (mode != 0) {
    const int nblk = std::max<int>(1, (int)((n + CMP_TILE - 1) / CMP_TILE));
    k_cmp_count<<<nblk, CMP_T, 0, s>>>(mode, order, active, n, tmp, n_dev);
    k_cmp_top<<<1, 1024, 0, s>>>(tmp, nblk, order_bstart, n, bstart, n_dev);
    k_cmp_scatter<<<nblk, CMP_T, 0, s>>>(mode, order, active, n, tmp, order_bstart, out, bstart, n_dev);
    return;
  }
  int nblk = (int)((n + PART_TILE - 1) / PART_TILE);
  if (nblk == 0) nblk = 1;
  k_part_count<<<nblk, PART_T, 0, s>>>(mode, order, binid, active, n, tmp, nblk);
  k_part_top<<<1, 1024, 0, s>>>(tmp, nblk, bstart);
  k_part_scatter<<<nblk, PART_T, 0, s>>>(mode, order, binid, active, n, tmp, nblk, out);
}

// ---------------------------------------------------------------------------
// fills / copies
// ---------------------------------------------------------------------------
template <class V>
__global__ void k_fill(V* __restrict__ p, int64_t n, V x) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = x;
}

template <class V>
void fill(V* p, int64_t n, V x, hipStream_t s) {
  if (n <= 0) return;
  int nb = (int)std::min<int64_t>((n + 255) / 256, 8192);
  k_fill<V><<<nb, 256, 0, s>>>(p, n, x);
}
template void fill<int32_t>(int32_t*, int64_t, int32_t, hipStream_t);
template void fill<int64_t>(int64_t*, int64_t, int64_t, hipStream_t);
template void fill<uint8_t>(uint8_t*, int64_t, uint8_t, hipStream_t);
template void fill<double>(double*, int64_t, double, hipStream_t);
template void fill<float>(float*, int64_t, float, hipStream_t);

__global__ void k_copy_i32(const int32_t* __restrict__ a, int32_t* __restrict__ b, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}
__global__ void k_accept(const int32_t* __restrict__ cur, int32_t* __restrict__ prev, int2* __restrict__ cp,
                         int32_t* __restrict__ cm, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int c = cur[i];
    prev[i] = c;
    cp[i] = make_int2(c, c);
    if (cm) cm[i] = c;
  }
}
void accept_sweep(const int32_t* cur, int32_t* prev, int2* cp, int32_t* cm, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  if (!cp) { copy_i32(cur, prev, n, s); return; }
  int nb = (int)std::min<int64_t>((n + 255) / 256, 8192);
  k_accept<<<nb, 256, 0, s>>>(cur, prev, cp, cm, n);
}

__global__ void k_settle_pairs(int2* __restrict__ cp, int32_t* __restrict__ cm, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int2 p = cp[i];
    if (p.x != p.y) {
      cp[i] = make_int2(p.x, p.x);
      if (cm) cm[i] = p.x;
    }
  }
}
void settle_pairs(int2* cp, int32_t* cm, int64_t n, hipStream_t s) {
  if (n <= 0 || !cp) return;
  int nb = (int)std::min<int64_t>((n + 255) / 256, 8192);
  k_settle_pairs<<<nb, 256, 0, s>>>(cp, cm, n);
}

__global__ void k_copy_list(const int32_t* __restrict__ a, int32_t* __restrict__ b, const int32_t* __restrict__ cnt,
                            int64_t n) {
  const int64_t m = *cnt < n ? *cnt : n;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}
void copy_list(const int32_t* a, int32_t* b, const int32_t* cnt, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  int nb = (int)std::min<int64_t>((n + 255) / 256, 8192);
  k_copy_list<<<nb, 256, 0, s>>>(a, b, cnt, n);
}

void copy_i32(const int32_t* a, int32_t* b, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  int nb = (int)std::min<int64_t>((n + 255) / 256, 8192);
  k_copy_i32<<<nb, 256, 0, s>>>(a, b, n);
}

// Unsigned 128-bit integer to double, rounded to nearest even: the 64 bits
// from the leading one down, with every lower bit folded into bit 0 as a
// sticky bit, round exactly as the whole value does (the rounding position is
// 11 bits above bit 0); the scale by 2^sh is exact.
__device__ __forceinline__ double u128_to_double(unsigned long long hi, unsigned long long lo) {
  if (hi == 0) return (double)lo;
  const int lz = __clzll(hi);  // 0..63
  const int sh = 64 - lz;      // bits below the 64-bit window: 1..64
  unsigned long long top = sh == 64 ? hi : (hi << lz) | (lo >> sh);
  const unsigned long long rest = sh == 64 ? lo : lo & ((1ull << sh) - 1);
  top |= rest != 0 ? 1ull : 0ull;
  return ldexp((double)top, sh);
}

// Q = sum_in/m2 - sum_tot2/m2^2, each operation rounded once, in the order the
// host evaluates it
__device__ __forceinline__ double q_of_counters(const Counters* c, double m2) {
  const double t2 = u128_to_double(c->sum_tot2_hi, c->sum_tot2);
  return __dsub_rn(__ddiv_rn((double)c->sum_in, m2), __ddiv_rn(t2, __dmul_rn(m2, m2)));
}

__global__ void k_round_start(const Counters* __restrict__ c, RoundCtl* __restrict__ rc, double m2) {
  if (threadIdx.x == 0) {
    rc->qprev = q_of_counters(c, m2);
    rc->live = 1;
    rc->accept = 0;
  }
}
void round_start(const Counters* c, RoundCtl* rc, double m2, hipStream_t s) {
  k_round_start<<<1, 64, 0, s>>>(c, rc, m2);
}

__global__ void k_round_step(Counters* __restrict__ c, RoundCtl* __restrict__ rc, double m2, double min_mod, int it,
                             int max_it, int32_t* __restrict__ bnext) {
  __shared__ int s_acc;
  if (threadIdx.x == 0) {
    const double q = q_of_counters(c, m2);
    const int acc = rc->live && !(__dsub_rn(q, rc->qprev) < min_mod) && it < max_it ? 1 : 0;
    c->step_q = q;
    c->step_accept = acc;
    rc->accept = acc;
    if (acc) rc->qprev = q;
    else rc->live = 0;
    s_acc = acc;
  }
  __syncthreads();
  if (!s_acc && threadIdx.x <= NBINS) bnext[threadIdx.x] = 0;
}
void round_step(Counters* c, RoundCtl* rc, double m2, double min_mod, int it, int max_it, int32_t* bnext,
                hipStream_t s) {
  k_round_step<<<1, 64, 0, s>>>(c, rc, m2, min_mod, it, max_it, bnext);
}

// accept_sweep over the vertices of a list whose length is on the device:
// the vertices of a sweep's frontier are the only ones whose cur can differ
// from prev (every other vertex already has prev = cur, cp = {cur, cur} and
// cm = cur).  rc (if given): only when the last round_step accepted.
__global__ void k_accept_list(const int32_t* __restrict__ cur, int32_t* __restrict__ prev, int2* __restrict__ cp,
                              int32_t* __restrict__ cm, const int32_t* __restrict__ list,
                              const int32_t* __restrict__ cnt, int64_t nmax, const RoundCtl* __restrict__ rc) {
  if (rc && !rc->accept) return;
  const int64_t m = *cnt < nmax ? *cnt : nmax;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
    const int v = list[i];
    const int c = cur[v];
    prev[v] = c;
    cp[v] = make_int2(c, c);
    if (cm) cm[v] = c;
  }
}
void accept_list(const int32_t* cur, int32_t* prev, int2* cp, int32_t* cm, const int32_t* list, const int32_t* cnt,
                 int64_t nmax, const RoundCtl* rc, hipStream_t s) {
  if (nmax <= 0) return;
  int nb = (int)std::min<int64_t>((nmax + 255) / 256, 8192);
  k_accept_list<<<nb, 256, 0, s>>>(cur, prev, cp, cm, list, cnt, nmax, rc);
}


__global__ void k_reset_counters(Counters* c, long long min_init) {
  if (threadIdx.x == 0) {
    c->sum_in = 0; c->sum_tot2 = 0; c->sum_tot2_hi = 0; c->min_tot = min_init; c->moved = 0; c->swept = 0;
    c->swept_edges = 0; c->active = 0; c->nonempty = 0; c->overflow = 0;
    c->step_q = 0; c->step_accept = 0; c->step_pad = 0;
  }
  if (threadIdx.x < NBINS) { c->bin_swept[threadIdx.x] = 0; c->bin_edges[threadIdx.x] = 0; }
  unsigned long long* sl = (unsigned long long*)c->slot;
  for (int i = threadIdx.x; i < (int)(sizeof(c->slot) / 8); i += blockDim.x) sl[i] = 0;
}
void reset_counters(Counters* c, long long min_init, hipStream_t s) {
  k_reset_counters<<<1, 256, 0, s>>>(c, min_init);
}

}  // namespace gala
