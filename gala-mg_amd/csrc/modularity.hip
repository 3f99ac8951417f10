// Modularity of an arbitrary partition (phase-level entry points).
//
// Q = sum_v In[v] / m2 - sum_c Tot[c]^2 / m2^2 with In[v] the weight of v's
// row entries whose neighbour shares its community and Tot[c] the summed
// weighted degree of community c -- the modularity_op reduction of LinXi-lx/
// GALA-mg src/louvain_gpu/kernel_functions.cuh:77-87 as the reference's
// phase 1 evaluates it (src/louvain_gpu/louvain_main_process.cu:538-548,
// 990-999).  Here every sum is an exact integer (float32 weights scaled to
// integers by the smallest power of two that makes them integral, as the
// engine's wide path does) and Q is finished with three FP64 operations, so it
// equals the engine's per-iteration Q bit for bit.
//
// Two entry points share the kernels:
//   gala_phase_modularity   host arrays in; allocates its device copies and
//                           its scratch (Tot per community, workgroup partials)
//   gala_modularity_noalloc device arrays in; allocates nothing: its scratch is
//                           the file-scope array `scratch` below, one copy per
//                           GPU, held until the process ends.
#include <mutex>

#include "../../include/gala.h"
#include "gala_device.h"
#include "gala_kernels.h"

namespace gala {

namespace {

constexpr int MOD_T = 256;            // threads per workgroup
constexpr int MOD_G = 16;             // lanes per row
constexpr int MOD_PARTS = 2048;       // workgroups of the row and Tot^2 passes
constexpr int64_t TWO62 = (int64_t)1 << 62;

struct ModPart {
  unsigned long long in_lo, in_hi;  // sum In  (128-bit, exact)
  unsigned long long m2_lo, m2_hi;  // sum K = m2
  unsigned long long t2_lo, t2_hi;  // sum Tot^2
  unsigned long long pad[2];
};
struct ModScalars {
  int32_t lo_exp;      // float32 weights: smallest exponent of a weight's lowest set bit (<= 0)
  int32_t bad;         // 1: a community or neighbour id out of range, or a bad weight
  int32_t overflow;    // 1: a scaled weight or a sum reached 2^62
  int32_t pad0;
  unsigned long long in_lo, in_hi, m2_lo, m2_hi, t2_lo, t2_hi;
  unsigned long long pad[1];
};

// scratch layout for n vertices: Tot[n] int64 | ModPart[MOD_PARTS] | ModScalars
inline int64_t align256(int64_t b) { return (b + 255) & ~int64_t(255); }
inline int64_t off_parts(int64_t n) { return align256(8 * n); }
inline int64_t off_scal(int64_t n) { return off_parts(n) + align256(sizeof(ModPart) * MOD_PARTS); }
inline int64_t scratch_bytes(int64_t n) { return off_scal(n) + align256(sizeof(ModScalars)); }

// The no-allocation entry point's scratch: communities of graphs of up to
// GALA_NOALLOC_MAX_N vertices (32 MiB of Tot plus the partials).
constexpr int64_t NOALLOC_BYTES = ((8 * (int64_t)GALA_NOALLOC_MAX_N + 255) & ~int64_t(255)) +
                                  ((sizeof(ModPart) * MOD_PARTS + 255) & ~int64_t(255)) +
                                  ((sizeof(ModScalars) + 255) & ~int64_t(255));
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

// Device address of `scratch` on every GPU (each holds its own copy).  The
// array's name is a host-side stand-in: only these addresses are passed to
// kernels.
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

__device__ __forceinline__ unsigned __int128 u128(unsigned long long lo, unsigned long long hi) {
  return ((unsigned __int128)hi << 64) | lo;
}
__device__ __forceinline__ unsigned __int128 sum_xor_u128(unsigned __int128 x) {
#pragma unroll
  for (int k = 32; k >= 1; k >>= 1) x += (unsigned __int128)__shfl_xor((__int128)x, k, 64);
  return x;
}

// lowest-set-bit exponent of a positive finite float32 (binary32 layout)
__device__ __forceinline__ int lowbit_exp(float x, bool* ok) {
  const uint32_t b = __float_as_uint(x);
  const uint32_t ex = (b >> 23) & 0xff, man = b & 0x7fffff;
  *ok = !(b >> 31) && ex != 0xff && !(ex == 0 && man == 0);
  const uint32_t sig = ex ? (man | 0x800000u) : man;
  const int e = ex ? (int)ex - 150 : -149;
  return sig ? e + __ffs(sig) - 1 : 0;
}

__global__ __launch_bounds__(MOD_T) void k_mod_clear(int64_t* __restrict__ tot, int32_t n,
                                                      ModScalars* __restrict__ sc) {
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x)
    tot[c] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->lo_exp = 0;
    sc->bad = 0;
    sc->overflow = 0;
  }
}

__global__ __launch_bounds__(MOD_T) void k_mod_minexp(const float* __restrict__ w, int64_t m,
                                                       ModScalars* __restrict__ sc) {
  int lo = 0;
  bool all_ok = true;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x) {
    bool ok;
    const int e = lowbit_exp(w[j], &ok);
    all_ok = all_ok && ok;
    lo = min(lo, e);
  }
#pragma unroll
  for (int k = 32; k >= 1; k >>= 1) lo = min(lo, __shfl_xor(lo, k, 64));
  if ((threadIdx.x & 63) == 0 && lo < 0) atomicMin(&sc->lo_exp, lo);
  if (!all_ok) sc->bad = 1;
}

// MOD_G lanes per vertex: K and In of the row, K added to Tot[comm[v]];
// workgroup partials of sum In and sum K.
__global__ __launch_bounds__(MOD_T) void k_mod_rows(int32_t n, int64_t m, const int64_t* __restrict__ end_incl,
                                                     const int32_t* __restrict__ nbr, const void* __restrict__ w,
                                                     int wmode, const int32_t* __restrict__ comm,
                                                     int64_t* __restrict__ tot, ModPart* __restrict__ parts,
                                                     ModScalars* __restrict__ sc) {
  __shared__ unsigned __int128 s_in[MOD_T / 64], s_m2[MOD_T / 64];
  const int lane = threadIdx.x & (MOD_G - 1);
  const double scale = wmode == GALA_W_F32 ? ldexp(1.0, -sc->lo_exp) : 1.0;
  unsigned __int128 in_acc = 0, m2_acc = 0;
  bool bad = false, over = false;
  const int groups = gridDim.x * (MOD_T / MOD_G);
  for (int v = (blockIdx.x * MOD_T + threadIdx.x) / MOD_G; v < n; v += groups) {
    const int64_t a = v == 0 ? 0 : end_incl[v - 1], b = end_incl[v];
    const int c = comm[v];
    // uniform over the group: a row outside [0, m) or a community id outside [0, n)
    if (c < 0 || c >= n || a < 0 || a > b || b > m) { bad = true; continue; }
    unsigned long long k = 0, in = 0;
    for (int64_t j = a + lane; j < b; j += MOD_G) {
      const int u = nbr[j];
      if (u < 0 || u >= n) { bad = true; continue; }
      unsigned long long wj = 1;
      if (wmode == GALA_W_I32) {
        const int32_t x = ((const int32_t*)w)[j];
        if (x <= 0) bad = true;
        wj = (unsigned long long)(x > 0 ? x : 0);
      } else if (wmode == GALA_W_F32) {
        const double x = (double)((const float*)w)[j] * scale;  // exact: a power-of-two scale
        if (!(x < (double)TWO62)) { over = true; continue; }
        wj = (unsigned long long)x;
      }
      k += wj;  // at most 2^62 per entry, checked against 2^62 below
      if (comm[u] == c) in += wj;
      if (k >= (unsigned long long)TWO62) over = true;
    }
#pragma unroll
    for (int s = MOD_G / 2; s >= 1; s >>= 1) {
      const unsigned long long ko = __shfl_xor(k, s, MOD_G), io = __shfl_xor(in, s, MOD_G);
      if (k + ko < k || k + ko >= (unsigned long long)TWO62) over = true;
      k += ko;
      in += io;
    }
    if (lane == 0) {
      atomicAdd((unsigned long long*)&tot[c], k);
      in_acc += in;
      m2_acc += k;
    }
  }
  const int wl = threadIdx.x & 63, wid = threadIdx.x >> 6;
  in_acc = sum_xor_u128(in_acc);
  m2_acc = sum_xor_u128(m2_acc);
  if (wl == 0) { s_in[wid] = in_acc; s_m2[wid] = m2_acc; }
  if (bad) sc->bad = 1;
  if (over) sc->overflow = 1;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < MOD_T / 64; q++) { in_acc += s_in[q]; m2_acc += s_m2[q]; }
    ModPart p{};
    p.in_lo = (unsigned long long)in_acc;
    p.in_hi = (unsigned long long)(in_acc >> 64);
    p.m2_lo = (unsigned long long)m2_acc;
    p.m2_hi = (unsigned long long)(m2_acc >> 64);
    parts[blockIdx.x] = p;
  }
}

__global__ __launch_bounds__(MOD_T) void k_mod_tot2(const int64_t* __restrict__ tot, int32_t n,
                                                     ModPart* __restrict__ parts) {
  __shared__ unsigned __int128 s_t2[MOD_T / 64];
  unsigned __int128 t2 = 0;  // each Tot <= m2 < 2^62 (checked), so Tot^2 < 2^124
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    const unsigned long long t = (unsigned long long)tot[c];
    t2 += (unsigned __int128)t * t;
  }
  t2 = sum_xor_u128(t2);
  if ((threadIdx.x & 63) == 0) s_t2[threadIdx.x >> 6] = t2;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < MOD_T / 64; q++) t2 += s_t2[q];
    parts[blockIdx.x].t2_lo = (unsigned long long)t2;
    parts[blockIdx.x].t2_hi = (unsigned long long)(t2 >> 64);
  }
}

// one workgroup: folds the partials into the scalars
__global__ __launch_bounds__(MOD_T) void k_mod_fin(const ModPart* __restrict__ parts, int nb,
                                                    ModScalars* __restrict__ sc) {
  __shared__ unsigned __int128 s_in[MOD_T / 64], s_m2[MOD_T / 64], s_t2[MOD_T / 64];
  unsigned __int128 in = 0, m2 = 0, t2 = 0;
  for (int b = threadIdx.x; b < nb; b += MOD_T) {
    in += u128(parts[b].in_lo, parts[b].in_hi);
    m2 += u128(parts[b].m2_lo, parts[b].m2_hi);
    t2 += u128(parts[b].t2_lo, parts[b].t2_hi);
  }
  in = sum_xor_u128(in);
  m2 = sum_xor_u128(m2);
  t2 = sum_xor_u128(t2);
  if ((threadIdx.x & 63) == 0) {
    s_in[threadIdx.x >> 6] = in;
    s_m2[threadIdx.x >> 6] = m2;
    s_t2[threadIdx.x >> 6] = t2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < MOD_T / 64; q++) { in += s_in[q]; m2 += s_m2[q]; t2 += s_t2[q]; }
    if (m2 >= (unsigned __int128)TWO62) sc->overflow = 1;
    sc->in_lo = (unsigned long long)in;
    sc->in_hi = (unsigned long long)(in >> 64);
    sc->m2_lo = (unsigned long long)m2;
    sc->m2_hi = (unsigned long long)(m2 >> 64);
    sc->t2_lo = (unsigned long long)t2;
    sc->t2_hi = (unsigned long long)(t2 >> 64);
  }
}

// Runs the passes on stream s over the scratch at `base` (laid out for n
// vertices) and reads the 64-byte result block back into *out.
int modularity_on(int32_t n, int64_t m, const int64_t* d_end, const int32_t* d_nbr, const void* d_w, int wmode,
                  const int32_t* d_comm, unsigned char* base, hipStream_t s, ModScalars* out) {
  int64_t* tot = (int64_t*)base;
  ModPart* parts = (ModPart*)(base + off_parts(n));
  ModScalars* sc = (ModScalars*)(base + off_scal(n));
  const int nbc = (int)std::min<int64_t>(std::max<int64_t>((n + MOD_T - 1) / MOD_T, 1), MOD_PARTS);
  k_mod_clear<<<nbc, MOD_T, 0, s>>>(tot, n, sc);
  if (wmode == GALA_W_F32 && m > 0)
    k_mod_minexp<<<(int)std::min<int64_t>((m + MOD_T - 1) / MOD_T, 4096), MOD_T, 0, s>>>((const float*)d_w, m, sc);
  const int nbr_blocks =
      (int)std::min<int64_t>(std::max<int64_t>(((int64_t)n * MOD_G + MOD_T - 1) / MOD_T, 1), MOD_PARTS);
  k_mod_rows<<<nbr_blocks, MOD_T, 0, s>>>(n, m, d_end, d_nbr, d_w, wmode, d_comm, tot, parts, sc);
  // the Tot^2 pass writes the same partial slots: use as many workgroups
  k_mod_tot2<<<nbr_blocks, MOD_T, 0, s>>>(tot, n, parts);
  k_mod_fin<<<1, MOD_T, 0, s>>>(parts, nbr_blocks, sc);
  if (hipMemcpyAsync(out, sc, sizeof(ModScalars), hipMemcpyDeviceToHost, s) != hipSuccess) return GALA_E_HIP;
  if (hipStreamSynchronize(s) != hipSuccess) return GALA_E_HIP;
  return GALA_OK;
}

int finish_q(const ModScalars& r, double* q_out) {
  if (r.bad) return GALA_E_INVALID;
  if (r.overflow || r.m2_hi || r.in_hi) return GALA_E_OVERFLOW;
  const int64_t m2 = (int64_t)r.m2_lo;
  if (m2 == 0) { *q_out = 0.0; return GALA_OK; }
  const unsigned __int128 t2 = ((unsigned __int128)r.t2_hi << 64) | r.t2_lo;
  const double m2d = (double)m2;
  *q_out = (double)r.in_lo / m2d - (double)t2 / (m2d * m2d);
  return GALA_OK;
}

}  // namespace

// Records the address of `scratch` on `device` (the current device), once.
int modularity_scratch_setup(int device) {
  if (device < 0 || device >= MAX_DEVICES) return GALA_E_INVALID;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return GALA_OK;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return GALA_E_HIP;
  g_scratch_ptr[device] = p;
  return GALA_OK;
}

int modularity_host(int device, hipStream_t s, int32_t n, const int64_t* end_incl, const int32_t* nbr, const void* w,
                    int wmode, const int32_t* comm, double* q_out) {
  const int64_t m = n > 0 ? end_incl[n - 1] : 0;
  const size_t wb = wmode == GALA_W_NONE ? 0 : 4;
  unsigned char *d_buf = nullptr, *d_scr = nullptr;
  const size_t in_bytes = 8 * (size_t)n + 4 * (size_t)m + wb * (size_t)m + 4 * (size_t)n + 64;
  (void)device;
  if (hipMalloc((void**)&d_buf, in_bytes) != hipSuccess) return GALA_E_OOM;
  if (hipMalloc((void**)&d_scr, (size_t)scratch_bytes(n)) != hipSuccess) { (void)hipFree(d_buf); return GALA_E_OOM; }
  int64_t* d_end = (int64_t*)d_buf;
  int32_t* d_nbr = (int32_t*)(d_buf + 8 * (size_t)n);
  int32_t* d_comm = (int32_t*)(d_buf + 8 * (size_t)n + 4 * (size_t)m);
  void* d_w = wb ? (void*)(d_buf + 8 * (size_t)n + 4 * (size_t)m + 4 * (size_t)n) : nullptr;
  int rc = GALA_OK;
  ModScalars r{};
  if (hipMemcpyAsync(d_end, end_incl, 8 * (size_t)n, hipMemcpyHostToDevice, s) != hipSuccess ||
      (m && hipMemcpyAsync(d_nbr, nbr, 4 * (size_t)m, hipMemcpyHostToDevice, s) != hipSuccess) ||
      hipMemcpyAsync(d_comm, comm, 4 * (size_t)n, hipMemcpyHostToDevice, s) != hipSuccess ||
      (wb && m && hipMemcpyAsync(d_w, w, wb * (size_t)m, hipMemcpyHostToDevice, s) != hipSuccess))
    rc = GALA_E_HIP;
  if (!rc) rc = modularity_on(n, m, d_end, d_nbr, d_w, wmode, d_comm, d_scr, s, &r);
  (void)hipFree(d_buf);
  (void)hipFree(d_scr);
  if (rc) return rc;
  return finish_q(r, q_out);
}

int modularity_noalloc(int device, hipStream_t s, int32_t n, const int64_t* d_end_incl, const int32_t* d_nbr,
                       const void* d_w, int wmode, const int32_t* d_comm, double* q_out) {
  if (n > GALA_NOALLOC_MAX_N) return GALA_E_INVALID;
  if (device < 0 || device >= MAX_DEVICES) return GALA_E_INVALID;
  void* base;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    base = g_scratch_ptr[device];
  }
  if (!base) return GALA_E_INVALID;  // gala_create sets it up
  // m: the last inclusive row end, read from the caller's device array
  int64_t m = 0;
  if (n > 0) {
    if (hipMemcpyAsync(&m, d_end_incl + (n - 1), 8, hipMemcpyDeviceToHost, s) != hipSuccess) return GALA_E_HIP;
    if (hipStreamSynchronize(s) != hipSuccess) return GALA_E_HIP;
  }
  ModScalars r{};
  const int rc = modularity_on(n, m, d_end_incl, d_nbr, d_w, wmode, d_comm, (unsigned char*)base, s, &r);
  if (rc) return rc;
  return finish_q(r, q_out);
}

}  // namespace gala
