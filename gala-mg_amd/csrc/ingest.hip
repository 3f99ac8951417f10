// Text edge list -> CSR on the device.
//
// The ingest of LinXi-lx/GALA-mg src/graph/graph.cpp:18-151 (Graph::Graph):
// one edge per line, '#' lines skipped, both directions pushed, each row
// deduplicated in ascending neighbour order with the weights of repeated
// entries summed, and the linked vertices renumbered densely in ascending
// original id.  The host loop there builds one std::map per vertex; here the
// same output comes from data-parallel passes over the bytes in HBM:
//
//   1. newline positions: per-workgroup counts of '\n' in 4 KiB tiles (16-byte
//      loads), a scan, then each thread writes the positions in its 16 bytes;
//   2. one thread per line tokenises and parses it (integers, and weights
//      through the exact decimal conversion of decimal.h); an error anywhere
//      fails the whole call;
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

#include <chrono>
#include <cstdio>

#include "../../include/gala.h"
#include "decimal.h"
#include "gala_device.h"
#include "gala_kernels.h"

namespace gala {

namespace {

constexpr int NL_T = 256;               // threads per workgroup of the newline passes (16 bytes each)
constexpr int64_t CHUNK = 64ll << 20;   // host read / copy pipeline chunk

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

__global__ __launch_bounds__(NL_T) void k_nl_write(const uint4* __restrict__ txt, int64_t nvec,
                                                   const int64_t* __restrict__ base, int64_t* __restrict__ nlpos) {
  __shared__ int64_t sh[17];
  const int64_t i = (int64_t)blockIdx.x * NL_T + threadIdx.x;
  uint32_t w[4] = {0, 0, 0, 0};
  int64_t c = 0;
  if (i < nvec) {
    const uint4 v = txt[i];
    w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
    c = nl_in(v.x) + nl_in(v.y) + nl_in(v.z) + nl_in(v.w);
  }
  int64_t tot;
  int64_t o = block_excl_scan(c, sh, &tot) + base[blockIdx.x];
  if (c == 0) return;
#pragma unroll
  for (int k = 0; k < 16; k++)
    if (((w[k >> 2] >> ((k & 3) * 8)) & 0xFF) == '\n') nlpos[o++] = i * 16 + k;
}

// byte reader over the padded text with a one-word cache
struct Rd {
  const uint32_t* w;
  int64_t cw;
  uint32_t v;
  __device__ uint8_t operator()(int64_t p) {
    const int64_t k = p >> 2;
    if (k != cw) { cw = k; v = w[k]; }
    return (uint8_t)(v >> ((p & 3) * 8));
  }
};

__device__ __forceinline__ int wave_max_i32(int x) {
#pragma unroll
  for (int k = 32; k >= 1; k >>= 1) x = max(x, __shfl_xor(x, k, 64));
  return x;
}

// One thread per line.  valid[i] = 1 for an edge line; a/b its two ids; wd /
// wi its weight (1 when the token is absent).  Ids above INT32_MAX or negative,
// and malformed weight tokens, set bits of *err.
template <int WMODE>
__global__ __launch_bounds__(256) void k_parse(const uint32_t* __restrict__ txt, int64_t F,
                                               const int64_t* __restrict__ nlpos, int64_t NL,
                                               int32_t* __restrict__ ea, int32_t* __restrict__ eb,
                                               double* __restrict__ wd, int32_t* __restrict__ wi,
                                               uint8_t* __restrict__ valid, int* __restrict__ maxid,
                                               int* __restrict__ err) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int mx = -1;
  if (i <= NL) {
    const int64_t s = i == 0 ? 0 : nlpos[i - 1] + 1;
    const int64_t e = i < NL ? nlpos[i] : F;
    Rd rd{txt, -1, 0};
    auto get = [&](int64_t p) { return rd(p); };
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
          mx = (int)max(m1, m2);
        }
      }
    }
    valid[i] = ok;
  }
  mx = wave_max_i32(mx);
  if ((threadIdx.x & 63) == 0 && mx >= 0) atomicMax(maxid, mx);
}

__global__ void k_mark(const int32_t* __restrict__ ea, const int32_t* __restrict__ eb,
                       const uint8_t* __restrict__ valid, int64_t L, uint8_t* __restrict__ linked) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < L; i += (int64_t)gridDim.x * blockDim.x)
    if (valid[i]) { linked[ea[i]] = 1; linked[eb[i]] = 1; }
}

template <class V>
__global__ void k_emit(const int32_t* __restrict__ ea, const int32_t* __restrict__ eb, const V* __restrict__ wv,
                       const uint8_t* __restrict__ valid, const int64_t* __restrict__ eidx,
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
                        const uint8_t* __restrict__ head, const int64_t* __restrict__ uidx, int bits,
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
  FILE* f = nullptr;
  ~File() {
    if (f) std::fclose(f);
  }
};

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

template <int WMODE, class V>
int ingest_bytes(hipStream_t s, const uint8_t* d_txt, int64_t F, int64_t nvec, IngestResult* out, Bufs& B,
                 hipEvent_t* ev) {
  const int64_t tiles = (nvec + NL_T - 1) / NL_T;
  int64_t *cnt, *base, *tmp, *scal;  // scal: [0] newline count, [1] edge lines, [2] entries
  int* flags;                        // [0] max id, [1] error bits, [2] vertex count
  CK(B.get(&cnt, tiles));
  CK(B.get(&base, tiles));
  CK(B.get(&scal, 4));
  CK(B.get(&flags, 4));
  CK(B.get(&tmp, scan_tmp_elems(std::max<int64_t>(tiles, 1)) + 64));
  k_nl_count<<<(unsigned)std::max<int64_t>(tiles, 1), NL_T, 0, s>>>((const uint4*)d_txt, nvec, cnt);
  scan_exclusive<int64_t, int64_t>(cnt, base, tiles, tmp, scal, s);
  int64_t NL = 0;
  CK(hipMemcpyAsync(&NL, scal, 8, hipMemcpyDeviceToHost, s));
  CK(hipStreamSynchronize(s));
  const int64_t L = NL + 1;  // the text after the last newline is a line too
  int64_t* nlpos;
  int32_t *ea, *eb;
  uint8_t* valid;
  double* wd = nullptr;
  int32_t* wi = nullptr;
  CK(B.get(&nlpos, NL));
  CK(B.get(&ea, L));
  CK(B.get(&eb, L));
  CK(B.get(&valid, L));
  if (WMODE == GALA_W_F32) CK(B.get(&wd, L));
  if (WMODE == GALA_W_I32) CK(B.get(&wi, L));
  if (NL > 0) k_nl_write<<<(unsigned)tiles, NL_T, 0, s>>>((const uint4*)d_txt, nvec, base, nlpos);
  const int init[4] = {-1, 0, 0, 0};
  CK(hipMemcpyAsync(flags, init, sizeof init, hipMemcpyHostToDevice, s));
  k_parse<WMODE><<<(unsigned)((L + 255) / 256), 256, 0, s>>>((const uint32_t*)d_txt, F, nlpos, NL, ea, eb, wd, wi,
                                                              valid, flags, flags + 1);
  int64_t* eidx;
  CK(B.get(&eidx, L));
  int64_t* tmpL;
  CK(B.get(&tmpL, scan_tmp_elems(L) + 64));
  scan_exclusive<uint8_t, int64_t>(valid, eidx, L, tmpL, scal + 1, s);
  int hflags[4];
  int64_t E = 0;
  CK(hipMemcpyAsync(hflags, flags, sizeof hflags, hipMemcpyDeviceToHost, s));
  CK(hipMemcpyAsync(&E, scal + 1, 8, hipMemcpyDeviceToHost, s));
  CK(hipStreamSynchronize(s));
  B.release(nlpos);
  if (hflags[1]) return GALA_E_FORMAT;
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
  int64_t *uidx, *tmpK;
  CK(B.get(&head, K));
  CK(B.get(&uidx, K));
  CK(B.get(&tmpK, scan_tmp_elems(K) + 64));
  k_heads<<<grid_for(K), 256, 0, s>>>(k1, K, head);
  scan_exclusive<uint8_t, int64_t>(head, uidx, K, tmpK, scal + 2, s);
  int64_t M = 0;
  CK(hipMemcpyAsync(&M, scal + 2, 8, hipMemcpyDeviceToHost, s));
  CK(hipStreamSynchronize(s));
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
  CK(hipMemcpyAsync(out->end.data(), d_end, 8 * n, hipMemcpyDeviceToHost, s));
  CK(hipMemcpyAsync(out->nbr.data(), d_nbr, 4 * M, hipMemcpyDeviceToHost, s));
  if (WMODE != GALA_W_NONE) CK(hipMemcpyAsync(out->w.data(), d_w, 4 * M, hipMemcpyDeviceToHost, s));
  CK(hipEventRecord(ev[3], s));
  CK(hipStreamSynchronize(s));
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
  fh.f = std::fopen(path, "rb");
  if (!fh.f) return GALA_E_IO;
  if (std::fseek(fh.f, 0, SEEK_END) != 0) return GALA_E_IO;
  const int64_t F = std::ftell(fh.f);
  if (F < 0 || std::fseek(fh.f, 0, SEEK_SET) != 0) return GALA_E_IO;
  const int64_t nvec = (F + 15) / 16;  // 16-byte words holding the text; padding bytes are zero
  Bufs B;
  uint8_t* d_txt;
  CK(B.get(&d_txt, 16 * (nvec + 2)));
  hipEvent_t ev[4] = {};
  struct Evs {
    hipEvent_t* e;
    ~Evs() {
      for (int k = 0; k < 4; k++)
        if (e[k]) (void)hipEventDestroy(e[k]);
    }
  } evs{ev};
  for (int k = 0; k < 4; k++) CK(hipEventCreate(&ev[k]));
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
  const int64_t cb = std::min<int64_t>(std::max<int64_t>(F, 1), CHUNK);
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
    if ((int64_t)std::fread(pin.p[k & 1], 1, (size_t)len, fh.f) != len) return GALA_E_IO;
    read_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - r0).count();
    CK(hipMemcpyAsync(d_txt + off, pin.p[k & 1], (size_t)len, hipMemcpyHostToDevice, s));
    CK(hipEventRecord(done[k & 1], s));
  }
  CK(hipEventRecord(ev[1], s));
  int rc;
  if (weight_mode == GALA_W_F32) rc = ingest_bytes<GALA_W_F32, double>(s, d_txt, F, nvec, out, B, ev);
  else if (weight_mode == GALA_W_I32) rc = ingest_bytes<GALA_W_I32, int32_t>(s, d_txt, F, nvec, out, B, ev);
  else rc = ingest_bytes<GALA_W_NONE, int32_t>(s, d_txt, F, nvec, out, B, ev);
  if (rc) return rc;
  float ms = 0;
  out->read_ms = read_ms;
  out->h2d_ms = (hipEventElapsedTime(&ms, ev[0], ev[1]) == hipSuccess) ? ms : 0;
  // an edgeless file ends before the output events are recorded
  if (out->n > 0) {
    out->device_ms = (hipEventElapsedTime(&ms, ev[1], ev[2]) == hipSuccess) ? ms : 0;
    out->d2h_ms = (hipEventElapsedTime(&ms, ev[2], ev[3]) == hipSuccess) ? ms : 0;
  } else {
    out->device_ms = out->d2h_ms = 0;
  }
  out->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  out->bytes = F;
  return GALA_OK;
}

}  // namespace gala
