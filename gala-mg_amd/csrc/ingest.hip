// Text edge list -> CSR on the device.
//
// The ingest of LinXi-lx/GALA-mg src/graph/graph.cpp:18-151 (Graph::Graph):
// one edge per line, '#' lines skipped, both directions pushed, each row
// deduplicated in ascending neighbour order with the weights of repeated
// entries summed, and the linked vertices renumbered densely in ascending
// original id.  The host loop there builds one std::map per vertex; here the
// same output comes from data-parallel passes over the bytes in HBM:
//
//   1. newline counts per 4 KiB tile (16-byte loads), scanned into each
//      tile's first line index;
//   2. each workgroup stages its tile (and the kilobyte before it) in LDS,
//      ranks the tile's newlines there, and parses the lines ending in the
//      tile, one thread per line (integers, and weights through the exact
//      decimal conversion of decimal.h); an error anywhere fails the call;
//   3. linked flags per original id, scanned into the dense renumbering;
//   4. two 64-bit keys (row << bits | nbr) per edge line, radix-sorted with a
//      stable sort (rocPRIM) over 2*bits key bits, weights as the payload, so
//      repeated entries keep their line order;
//   5. run heads scanned into output slots; the head of each run sums the
//      run's weights in line order (float64, then rounded once to float32,
//      or int64 range-checked to int32) and the last run of a row writes the
//      row's inclusive end.
// The line grammar and the output are those of gala_ingest_text on the host.
#include <rocprim/device/device_radix_sort.hpp>

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <thread>

#include "../../include/gala.h"
#include "decimal.h"
#include "gala_device.h"
#include "gala_kernels.h"

namespace gala {

namespace {

constexpr int NL_T = 256;               // threads per workgroup of the text passes (16 bytes each)
constexpr int NL_TILE = NL_T * 16;      // text bytes per workgroup
constexpr int PT_OV = 1024;             // bytes before a tile staged with it (starts of lines that end in it)
constexpr int PT_BUF = PT_OV + NL_TILE;
constexpr int64_t CHUNK = 64ll << 20;   // host read / copy pipeline chunk
constexpr int READ_THREADS = 8;         // parallel reads of one chunk

enum : int { ING_BAD_ID = 1, ING_BAD_WEIGHT = 2, ING_SUM_RANGE = 4 };

__device__ __forceinline__ int zero_bytes(uint32_t y) {
  // bit 7 of each byte of t is set iff that byte of y is nonzero (no carries cross bytes)
  const uint32_t t = ((y & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | y;
  return 4 - __popc(t & 0x80808080u);
}
__device__ __forceinline__ int nl_in(uint32_t x) { return zero_bytes(x ^ 0x0A0A0A0Au); }

__global__ __launch_bounds__(NL_T) void k_nl_count(const uint4* __restrict__ txt, int64_t nvec,
                                                   int64_t* __restrict__ cnt) {
  __shared__ int64_t sh[17];
  const int64_t i = (int64_t)blockIdx.x * NL_T + threadIdx.x;
  int64_t c = 0;
  if (i < nvec) {
    const uint4 v = txt[i];
    c = nl_in(v.x) + nl_in(v.y) + nl_in(v.z) + nl_in(v.w);
  }
  int64_t tot;
  block_excl_scan(c, sh, &tot);
  if (threadIdx.x == 0) cnt[blockIdx.x] = tot;
}

__device__ __forceinline__ int wave_max_i32(int x) {
#pragma unroll
  for (int k = 32; k >= 1; k >>= 1) x = max(x, __shfl_xor(x, k, 64));
  return x;
}

// One workgroup per 4 KiB tile of text, which it stages in LDS together with
// the PT_OV bytes before it.  It parses the lines whose terminating newline
// lies in the tile (the last tile also the final line, which ends at F):
// line i = base[tile] + rank of its newline.  A line starts after the previous
// newline: in the tile, in the staged overlap, or -- for a line longer than
// the overlap -- found by a backward scan of the text in global memory.
// valid[i] = 1 for an edge line; ea/eb its two ids; wd / wi its weight (1 when
// the token is absent).  Ids above INT32_MAX or negative, and malformed weight
// tokens, set bits of *err.
template <int WMODE>
__global__ __launch_bounds__(NL_T) void k_parse_tile(const uint8_t* __restrict__ txt, int64_t F, int64_t nvec,
                                                     int64_t ntiles, const int64_t* __restrict__ base,
                                                     int32_t* __restrict__ ea, int32_t* __restrict__ eb,
                                                     double* __restrict__ wd, int32_t* __restrict__ wi,
                                                     uint8_t* __restrict__ valid, int* __restrict__ tmax,
                                                     int* __restrict__ err) {
  __shared__ __align__(16) uint8_t buf[PT_BUF];  // text [T0 - PT_OV, T0 + NL_TILE)
  __shared__ int16_t nl[NL_TILE];                // offsets of the tile's newlines, ascending
  __shared__ int64_t sh[17];
  __shared__ int s_prev;                          // last newline of the overlap (buf offset), -1: none
  __shared__ int s_max;                           // largest id of the tile's edge lines
  const int64_t T0 = (int64_t)blockIdx.x * NL_TILE;
  const int64_t lo = T0 - PT_OV;
  for (int v = threadIdx.x; v < PT_BUF / 16; v += NL_T) {
    const int64_t g = lo / 16 + v;
    uint4 x = make_uint4(0, 0, 0, 0);
    if (g < 0) x = make_uint4(0x0A0A0A0Au, 0x0A0A0A0Au, 0x0A0A0A0Au, 0x0A0A0A0Au);  // before the file
    else if (g < nvec) x = ((const uint4*)txt)[g];
    *(uint4*)&buf[16 * v] = x;
  }
  if (threadIdx.x == 0) s_prev = s_max = -1;
  __syncthreads();
  const uint4 mine = *(const uint4*)&buf[PT_OV + 16 * threadIdx.x];
  const uint32_t w[4] = {mine.x, mine.y, mine.z, mine.w};
  int64_t tot;
  int64_t r = block_excl_scan((int64_t)(nl_in(mine.x) + nl_in(mine.y) + nl_in(mine.z) + nl_in(mine.w)), sh, &tot);
#pragma unroll
  for (int k = 0; k < 16; k++)
    if (((w[k >> 2] >> ((k & 3) * 8)) & 0xFF) == '\n') nl[r++] = (int16_t)(16 * threadIdx.x + k);
  if (threadIdx.x < PT_OV / 16) {
    int last = -1;
    for (int k = 0; k < 16; k++)
      if (buf[16 * threadIdx.x + k] == '\n') last = 16 * threadIdx.x + k;
    if (last >= 0) atomicMax(&s_prev, last);
  }
  __syncthreads();
  const int C = (int)tot;
  const bool final_tile = (int64_t)blockIdx.x == ntiles - 1;
  const int nlines = C + (final_tile ? 1 : 0);
  const int64_t b0 = base[blockIdx.x];
  auto get = [&](int64_t p) -> uint8_t { return p >= lo ? buf[p - lo] : txt[p]; };
  int mx = -1;
  for (int j = threadIdx.x; j < nlines; j += NL_T) {
    const int64_t e = j < C ? T0 + nl[j] : F;
    int64_t s;
    if (j > 0) {
      s = T0 + nl[j - 1] + 1;
    } else if (s_prev >= 0) {
      s = lo + s_prev + 1;
    } else {  // a line longer than the overlap: its start lies further back
      int64_t p = lo - 1;
      while (p >= 0 && txt[p] != '\n') p--;
      s = p + 1;
    }
    const int64_t i = b0 + j;
    uint8_t ok = 0;
    if (s < e && get(s) != '#') {
      int64_t p = s;
      while (p < e && dec::is_ws(get(p))) p++;
      const int64_t t1 = p;
      while (p < e && !dec::is_ws(get(p))) p++;
      const int64_t t1e = p;
      while (p < e && dec::is_ws(get(p))) p++;
      const int64_t t2 = p;
      while (p < e && !dec::is_ws(get(p))) p++;
      const int64_t t2e = p;
      bool n1, n2;
      uint64_t m1, m2;
      if (t1 < t1e && t2 < t2e && dec::parse_int(get, t1, t1e, &n1, &m1) && dec::parse_int(get, t2, t2e, &n2, &m2)) {
        int bad = 0;
        if ((n1 && m1 != 0) || m1 > 2147483647ull || (n2 && m2 != 0) || m2 > 2147483647ull) bad |= ING_BAD_ID;
        if (WMODE != GALA_W_NONE) {
          while (p < e && dec::is_ws(get(p))) p++;
          const int64_t t3 = p;
          while (p < e && !dec::is_ws(get(p))) p++;
          if (WMODE == GALA_W_F32) {
            double x = 1.0;
            if (t3 < p && dec::parse_decimal(get, t3, p, &x) != dec::DEC_OK) bad |= ING_BAD_WEIGHT;
            wd[i] = x;
          } else {
            int32_t x = 1;
            bool nw;
            uint64_t mw;
            if (t3 < p) {
              if (!dec::parse_int(get, t3, p, &nw, &mw) || mw > (nw ? 2147483648ull : 2147483647ull))
                bad |= ING_BAD_WEIGHT;
              else
                x = nw ? (int32_t)(-(int64_t)mw) : (int32_t)mw;
            }
            wi[i] = x;
          }
        }
        if (bad) {
          atomicOr(err, bad);
        } else {
          ok = 1;
          ea[i] = (int32_t)m1;
          eb[i] = (int32_t)m2;
          mx = max(mx, (int)max(m1, m2));
        }
      }
    }
    valid[i] = ok;
  }
  mx = wave_max_i32(mx);
  if ((threadIdx.x & 63) == 0) atomicMax(&s_max, mx);
  __syncthreads();
  if (threadIdx.x == 0) tmax[blockIdx.x] = s_max;
}

// *out = max(tmax[0, n)), one workgroup of 1024
__global__ __launch_bounds__(1024) void k_max_reduce(const int* __restrict__ tmax, int64_t n, int* __restrict__ out) {
  __shared__ int s;
  if (threadIdx.x == 0) s = -1;
  __syncthreads();
  int mx = -1;
  for (int64_t i = threadIdx.x; i < n; i += 1024) mx = max(mx, tmax[i]);
  mx = wave_max_i32(mx);
  if ((threadIdx.x & 63) == 0) atomicMax(&s, mx);
  __syncthreads();
  if (threadIdx.x == 0) *out = s;
}

__global__ void k_mark(const int32_t* __restrict__ ea, const int32_t* __restrict__ eb,
                       const uint8_t* __restrict__ valid, int64_t L, uint8_t* __restrict__ linked) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < L; i += (int64_t)gridDim.x * blockDim.x)
    if (valid[i]) { linked[ea[i]] = 1; linked[eb[i]] = 1; }
}

template <class V>
__global__ void k_emit(const int32_t* __restrict__ ea, const int32_t* __restrict__ eb, const V* __restrict__ wv,
                       const uint8_t* __restrict__ valid, const int32_t* __restrict__ eidx,
                       const int32_t* __restrict__ renum, int64_t L, int bits, uint64_t* __restrict__ key,
                       V* __restrict__ val) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < L; i += (int64_t)gridDim.x * blockDim.x) {
    if (!valid[i]) continue;
    const int64_t e = eidx[i];
    const uint64_t ra = (uint32_t)renum[ea[i]], rb = (uint32_t)renum[eb[i]];
    key[2 * e] = (ra << bits) | rb;
    key[2 * e + 1] = (rb << bits) | ra;
    if (val) {
      const V x = wv[i];
      val[2 * e] = x;
      val[2 * e + 1] = x;
    }
  }
}

__global__ void k_heads(const uint64_t* __restrict__ key, int64_t K, uint8_t* __restrict__ head) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < K; j += (int64_t)gridDim.x * blockDim.x)
    head[j] = j == 0 || key[j] != key[j - 1];
}

// WMODE: output for each run head j (slot u = uidx[j]): neighbour, summed
// weight, and the row end when the run is the row's last.
template <int WMODE, class V>
__global__ void k_dedup(const uint64_t* __restrict__ key, const V* __restrict__ val, int64_t K,
                        const uint8_t* __restrict__ head, const int32_t* __restrict__ uidx, int bits,
                        int32_t* __restrict__ nbr, void* __restrict__ wout, int64_t* __restrict__ end_incl,
                        int* __restrict__ err) {
  const uint64_t mask = (1ull << bits) - 1;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < K; j += (int64_t)gridDim.x * blockDim.x) {
    if (!head[j]) continue;
    const uint64_t k = key[j];
    const int64_t u = uidx[j];
    nbr[u] = (int32_t)(k & mask);
    int64_t r = j + 1;
    if (WMODE == GALA_W_F32) {
      double sum = (double)val[j];
      for (; r < K && key[r] == k; r++) sum += (double)val[r];
      const float f = (float)sum;
      if (!isfinite(f)) atomicOr(err, ING_SUM_RANGE);
      ((float*)wout)[u] = f;
    } else if (WMODE == GALA_W_I32) {
      int64_t sum = (int64_t)val[j];
      for (; r < K && key[r] == k; r++) sum += (int64_t)val[r];
      if (sum < INT32_MIN || sum > INT32_MAX) atomicOr(err, ING_SUM_RANGE);
      ((int32_t*)wout)[u] = (int32_t)sum;
    } else {
      while (r < K && key[r] == k) r++;
    }
    if (r == K || (key[r] >> bits) != (k >> bits)) end_incl[k >> bits] = u + 1;
  }
}

// device buffers released on every return path
struct Bufs {
  std::vector<void*> p;
  template <class T>
  hipError_t get(T** x, int64_t elems) {
    *x = nullptr;
    hipError_t e = hipMalloc((void**)x, (size_t)std::max<int64_t>(elems, 1) * sizeof(T));
    if (e == hipSuccess) p.push_back((void*)*x);
    return e;
  }
  void release(void* x) {
    if (!x) return;
    for (auto& q : p)
      if (q == x) { (void)hipFree(q); q = nullptr; }
  }
  ~Bufs() {
    for (void* q : p)
      if (q) (void)hipFree(q);
  }
};

struct Pinned {
  void* p[2] = {nullptr, nullptr};
  ~Pinned() {
    for (void* q : p)
      if (q) (void)hipHostFree(q);
  }
};

struct File {
  int fd = -1;
  ~File() {
    if (fd >= 0) ::close(fd);
  }
};

// dst[0, len) = file bytes [off, off + len), read by up to READ_THREADS
// threads in parallel (page-cache copies run at a few GB/s per thread)
bool pread_all(int fd, uint8_t* dst, int64_t off, int64_t len) {
  const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(READ_THREADS, len >> 22));
  std::atomic<bool> ok{true};
  auto part = [&](int t) {
    const int64_t a = len * t / nt, b = len * (t + 1) / nt;
    for (int64_t p = a; p < b;) {
      const ssize_t r = ::pread(fd, dst + p, (size_t)(b - p), (off_t)(off + p));
      if (r <= 0) { ok = false; return; }
      p += r;
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nt; t++) th.emplace_back(part, t);
  part(0);
  for (auto& x : th) x.join();
  return ok;
}

int grid_for(int64_t n, int t = 256) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + t - 1) / t, 65536)); }

#define CK(x)                                   \
  do {                                          \
    hipError_t e_ = (x);                        \
    if (e_ != hipSuccess) return e_ == hipErrorOutOfMemory ? GALA_E_OOM : GALA_E_HIP; \
  } while (0)

template <class V>
hipError_t sort_pairs(uint64_t* k_in, uint64_t* k_out, V* v_in, V* v_out, int64_t K, int end_bit, hipStream_t s,
                      Bufs& B) {
  size_t tb = 0;
  hipError_t e = rocprim::radix_sort_pairs(nullptr, tb, k_in, k_out, v_in, v_out, (size_t)K, 0u, (unsigned)end_bit, s);
  if (e != hipSuccess) return e;
  uint8_t* tmp = nullptr;
  if ((e = B.get(&tmp, (int64_t)tb)) != hipSuccess) return e;
  return rocprim::radix_sort_pairs((void*)tmp, tb, k_in, k_out, v_in, v_out, (size_t)K, 0u, (unsigned)end_bit, s);
}

hipError_t sort_keys(uint64_t* k_in, uint64_t* k_out, int64_t K, int end_bit, hipStream_t s, Bufs& B) {
  size_t tb = 0;
  hipError_t e = rocprim::radix_sort_keys(nullptr, tb, k_in, k_out, (size_t)K, 0u, (unsigned)end_bit, s);
  if (e != hipSuccess) return e;
  uint8_t* tmp = nullptr;
  if ((e = B.get(&tmp, (int64_t)tb)) != hipSuccess) return e;
  return rocprim::radix_sort_keys((void*)tmp, tb, k_in, k_out, (size_t)K, 0u, (unsigned)end_bit, s);
}

// two pinned buffers of cap bytes and an event each: the file -> HBM and
// HBM -> host pipelines move CHUNK pieces through them
struct Staging {
  void* pin[2];
  hipEvent_t done[2];
  int64_t cap;
};

// dst[0, len) = src[0, len) on up to READ_THREADS host threads
void par_memcpy(uint8_t* dst, const uint8_t* src, int64_t len) {
  const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(READ_THREADS, len >> 22));
  auto part = [&](int t) {
    const int64_t a = len * t / nt, b = len * (t + 1) / nt;
    std::memcpy(dst + a, src + a, (size_t)(b - a));
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nt; t++) th.emplace_back(part, t);
  part(0);
  for (auto& x : th) x.join();
}

// host dst[0, bytes) = device src: the device copy of piece k+1 into one
// pinned buffer runs while host threads move piece k out of the other
hipError_t d2h_staged(void* dst, const void* src, int64_t bytes, const Staging& g, hipStream_t s) {
  const int64_t np = (bytes + g.cap - 1) / g.cap;
  auto issue = [&](int64_t k) {
    const int64_t off = k * g.cap, len = std::min(g.cap, bytes - off);
    hipError_t e = hipMemcpyAsync(g.pin[k & 1], (const uint8_t*)src + off, (size_t)len, hipMemcpyDeviceToHost, s);
    return e == hipSuccess ? hipEventRecord(g.done[k & 1], s) : e;
  };
  hipError_t e = hipSuccess;
  if (np > 0 && (e = issue(0)) != hipSuccess) return e;
  for (int64_t k = 0; k < np; k++) {
    if (k + 1 < np && (e = issue(k + 1)) != hipSuccess) return e;  // its buffer was drained at step k - 1
    if ((e = hipEventSynchronize(g.done[k & 1])) != hipSuccess) return e;
    const int64_t off = k * g.cap;
    par_memcpy((uint8_t*)dst + off, (const uint8_t*)g.pin[k & 1], std::min(g.cap, bytes - off));
  }
  return hipSuccess;
}

template <int WMODE, class V>
int ingest_bytes(hipStream_t s, const uint8_t* d_txt, int64_t F, int64_t nvec, IngestResult* out, Bufs& B,
                 hipEvent_t* ev, const Staging& stg) {
  const int64_t tiles = std::max<int64_t>(1, (nvec + NL_T - 1) / NL_T);  // an empty file is one empty tile
  int64_t *cnt, *base, *tmp, *scal;  // scal: [0] newline count
  int* flags;                        // [0] max id, [1] error bits, [2] vertex count, [3] edge lines, [4] entries
  CK(B.get(&cnt, tiles));
  CK(B.get(&base, tiles));
  CK(B.get(&scal, 4));
  CK(B.get(&flags, 8));
  CK(B.get(&tmp, scan_tmp_elems(tiles) + 64));
  k_nl_count<<<(unsigned)tiles, NL_T, 0, s>>>((const uint4*)d_txt, nvec, cnt);
  scan_exclusive<int64_t, int64_t>(cnt, base, tiles, tmp, scal, s);
  int64_t NL = 0;
  CK(hipMemcpyAsync(&NL, scal, 8, hipMemcpyDeviceToHost, s));
  CK(hipStreamSynchronize(s));
  const int64_t L = NL + 1;  // the text after the last newline is a line too
  if (L > INT32_MAX) return GALA_E_OVERFLOW;  // line and entry indices are int32
  int32_t *ea, *eb;
  uint8_t* valid;
  double* wd = nullptr;
  int32_t* wi = nullptr;
  CK(B.get(&ea, L));
  CK(B.get(&eb, L));
  CK(B.get(&valid, L));
  if (WMODE == GALA_W_F32) CK(B.get(&wd, L));
  if (WMODE == GALA_W_I32) CK(B.get(&wi, L));
  const int init[8] = {-1, 0, 0, 0, 0, 0, 0, 0};
  CK(hipMemcpyAsync(flags, init, sizeof init, hipMemcpyHostToDevice, s));
  int* tmax;
  CK(B.get(&tmax, tiles));
  k_parse_tile<WMODE><<<(unsigned)tiles, NL_T, 0, s>>>(d_txt, F, nvec, tiles, base, ea, eb, wd, wi, valid, tmax,
                                                        flags + 1);
  k_max_reduce<<<1, 1024, 0, s>>>(tmax, tiles, flags);
  int32_t *eidx, *tmpL;
  CK(B.get(&eidx, L));
  CK(B.get(&tmpL, scan_tmp_elems(L) + 64));
  scan_exclusive<uint8_t, int32_t>(valid, eidx, L, tmpL, flags + 3, s);
  int hflags[8];
  CK(hipMemcpyAsync(hflags, flags, sizeof hflags, hipMemcpyDeviceToHost, s));
  CK(hipStreamSynchronize(s));
  if (hflags[1]) return GALA_E_FORMAT;
  const int64_t E = hflags[3];
  out->end.clear();
  out->nbr.clear();
  out->w.clear();
  out->n = 0;
  if (E == 0) return GALA_OK;
  if (E >= (1ll << 30)) return GALA_E_OVERFLOW;  // 2^31 directed entries and vertex ids stay int32
  const int64_t nid = (int64_t)hflags[0] + 1;
  uint8_t* linked;
  int32_t* renum;
  CK(B.get(&linked, nid));
  CK(B.get(&renum, nid));
  fill<uint8_t>(linked, nid, 0, s);
  k_mark<<<grid_for(L), 256, 0, s>>>(ea, eb, valid, L, linked);
  int32_t* tmpN;
  CK(B.get(&tmpN, scan_tmp_elems(nid) + 64));
  scan_exclusive<uint8_t, int32_t>(linked, renum, nid, tmpN, flags + 2, s);
  int32_t n = 0;
  CK(hipMemcpyAsync(&n, flags + 2, 4, hipMemcpyDeviceToHost, s));
  CK(hipStreamSynchronize(s));
  B.release(linked);
  B.release(tmpL);
  int bits = 1;
  while (bits < 31 && ((int64_t)1 << bits) < n) bits++;
  const int64_t K = 2 * E;
  uint64_t *k0, *k1;
  V *v0 = nullptr, *v1 = nullptr;
  CK(B.get(&k0, K));
  CK(B.get(&k1, K));
  if (WMODE != GALA_W_NONE) {
    CK(B.get(&v0, K));
    CK(B.get(&v1, K));
  }
  const V* wsrc = WMODE == GALA_W_F32 ? (const V*)(const void*)wd : (const V*)(const void*)wi;
  k_emit<V><<<grid_for(L), 256, 0, s>>>(ea, eb, wsrc, valid, eidx, renum, L, bits, k0, v0);
  B.release(renum);
  if (WMODE != GALA_W_NONE) CK(sort_pairs<V>(k0, k1, v0, v1, K, 2 * bits, s, B));
  else CK(sort_keys(k0, k1, K, 2 * bits, s, B));
  B.release(k0);
  B.release(v0);
  uint8_t* head;
  int32_t *uidx, *tmpK;
  CK(B.get(&head, K));
  CK(B.get(&uidx, K));
  CK(B.get(&tmpK, scan_tmp_elems(K) + 64));
  k_heads<<<grid_for(K), 256, 0, s>>>(k1, K, head);
  scan_exclusive<uint8_t, int32_t>(head, uidx, K, tmpK, flags + 4, s);
  int32_t M32 = 0;
  CK(hipMemcpyAsync(&M32, flags + 4, 4, hipMemcpyDeviceToHost, s));
  CK(hipStreamSynchronize(s));
  const int64_t M = M32;
  int64_t* d_end;
  int32_t* d_nbr;
  uint32_t* d_w = nullptr;
  CK(B.get(&d_end, n));
  CK(B.get(&d_nbr, M));
  if (WMODE != GALA_W_NONE) CK(B.get(&d_w, M));
  k_dedup<WMODE, V><<<grid_for(K), 256, 0, s>>>(k1, v1, K, head, uidx, bits, d_nbr, d_w, d_end, flags + 1);
  CK(hipEventRecord(ev[2], s));
  try {
    out->end.resize(n);
    out->nbr.resize(M);
    if (WMODE != GALA_W_NONE) out->w.resize(M);
  } catch (const std::exception&) {
    return GALA_E_OOM;
  }
  CK(hipMemcpyAsync(hflags, flags, sizeof hflags, hipMemcpyDeviceToHost, s));
  CK(hipStreamSynchronize(s));
  if (!hflags[1]) {
    const auto c0 = std::chrono::steady_clock::now();
    CK(d2h_staged(out->end.data(), d_end, 8 * (int64_t)n, stg, s));
    CK(d2h_staged(out->nbr.data(), d_nbr, 4 * M, stg, s));
    if (WMODE != GALA_W_NONE) CK(d2h_staged(out->w.data(), d_w, 4 * M, stg, s));
    out->d2h_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - c0).count();
  }
  if (hflags[1]) {
    out->end.clear();
    out->nbr.clear();
    out->w.clear();
    return GALA_E_FORMAT;
  }
  out->n = n;
  return GALA_OK;
}

}  // namespace

int ingest_text_device(int device, hipStream_t s, const char* path, int weight_mode, IngestResult* out) {
  if (!path || !out || weight_mode < GALA_W_NONE || weight_mode > GALA_W_F32) return GALA_E_INVALID;
  (void)hipSetDevice(device);
  const auto t0 = std::chrono::steady_clock::now();
  File fh;
  fh.fd = ::open(path, O_RDONLY);
  if (fh.fd < 0) return GALA_E_IO;
  struct stat st;
  if (::fstat(fh.fd, &st) != 0 || !S_ISREG(st.st_mode)) return GALA_E_IO;
  const int64_t F = st.st_size;
  const int64_t nvec = (F + 15) / 16;  // 16-byte words holding the text; padding bytes are zero
  Bufs B;
  uint8_t* d_txt;
  CK(B.get(&d_txt, 16 * (nvec + 2)));
  hipEvent_t ev[3] = {};
  struct Evs {
    hipEvent_t* e;
    ~Evs() {
      for (int k = 0; k < 3; k++)
        if (e[k]) (void)hipEventDestroy(e[k]);
    }
  } evs{ev};
  for (int k = 0; k < 3; k++) CK(hipEventCreate(&ev[k]));
  CK(hipMemsetAsync(d_txt + 16 * std::max<int64_t>(nvec - 1, 0), 0, 32, s));
  // read and copy in CHUNK pieces through two pinned buffers
  Pinned pin;
  hipEvent_t done[2] = {};
  struct Evs2 {
    hipEvent_t* e;
    ~Evs2() {
      for (int k = 0; k < 2; k++)
        if (e[k]) (void)hipEventDestroy(e[k]);
    }
  } evs2{done};
  const int64_t cb = std::min<int64_t>(std::max<int64_t>(F, 1ll << 20), CHUNK);
  for (int k = 0; k < 2; k++) {
    CK(hipHostMalloc(&pin.p[k], (size_t)cb, hipHostMallocDefault));
    CK(hipEventCreateWithFlags(&done[k], hipEventDisableTiming));
  }
  double read_ms = 0;
  CK(hipEventRecord(ev[0], s));
  for (int64_t off = 0, k = 0; off < F; off += CHUNK, k++) {
    const int64_t len = std::min<int64_t>(CHUNK, F - off);
    if (k >= 2) CK(hipEventSynchronize(done[k & 1]));
    const auto r0 = std::chrono::steady_clock::now();
    if (!pread_all(fh.fd, (uint8_t*)pin.p[k & 1], off, len)) return GALA_E_IO;
    read_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - r0).count();
    CK(hipMemcpyAsync(d_txt + off, pin.p[k & 1], (size_t)len, hipMemcpyHostToDevice, s));
    CK(hipEventRecord(done[k & 1], s));
  }
  CK(hipEventRecord(ev[1], s));
  const Staging stg{{pin.p[0], pin.p[1]}, {done[0], done[1]}, cb};
  out->d2h_ms = 0;
  int rc;
  if (weight_mode == GALA_W_F32) rc = ingest_bytes<GALA_W_F32, double>(s, d_txt, F, nvec, out, B, ev, stg);
  else if (weight_mode == GALA_W_I32) rc = ingest_bytes<GALA_W_I32, int32_t>(s, d_txt, F, nvec, out, B, ev, stg);
  else rc = ingest_bytes<GALA_W_NONE, int32_t>(s, d_txt, F, nvec, out, B, ev, stg);
  if (rc) return rc;
  float ms = 0;
  out->read_ms = read_ms;
  out->h2d_ms = (hipEventElapsedTime(&ms, ev[0], ev[1]) == hipSuccess) ? ms : 0;
  // an edgeless file ends before the device passes' end event is recorded
  out->device_ms = out->n > 0 && hipEventElapsedTime(&ms, ev[1], ev[2]) == hipSuccess ? ms : 0;
  out->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  out->bytes = F;
  return GALA_OK;
}

}  // namespace gala
