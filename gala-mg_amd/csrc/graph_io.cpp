// Host-side graph IO and synthetic generators (no device code).
//
// .bin layout (LinXi-lx/GALA-mg src/graph/graph.cpp:153-219): little-endian
//   int32 n; int64 end[n] (inclusive row ends, no leading 0);
//   int32 nbr[end[n-1]]; optional weight section of end[n-1] 4-byte values.
// The weight section is selected out of band (the reference's -w flag); this
// build adds float32 in the same slot.
//
// Generators (deterministic per seed, xoshiro256** seeded by splitmix64):
//   SBM       : every pair i<j drawn once, Bernoulli(p_in | p_out), geometric skipping
//   Chung-Lu  : expected degrees w_i = min(cap, c (i+1)^(-1/(gamma-1))), c set by
//               bisection for the mean; edges by the Miller-Hagberg O(n+m) skip method
//   R-MAT     : scale/edge factor/a,b,c, fp32 weights U[lo,hi]; symmetrised,
//               duplicates and self-loops merged by summing weights
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <string>
#include <thread>
#include <atomic>
#include <vector>

#include "../../include/gala.h"

namespace {

struct Rng {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto& v : s) v = splitmix(x);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
  double uniform_pos() {  // (0,1]
    return 1.0 - uniform();
  }
};

// Symmetric CSR from an undirected edge list (u,v) (u != v); rows sorted, unique.
struct Csr {
  int32_t n = 0;
  std::vector<int64_t> end;  // inclusive ends
  std::vector<int32_t> nbr;
  std::vector<float> w;     // float32 weights
  std::vector<int32_t> wi;  // int32 weights (text ingest, GALA_W_I32)
};

void csr_from_pairs(int32_t n, const std::vector<std::pair<int32_t, int32_t>>& e, Csr& out) {
  std::vector<int64_t> deg(n + 1, 0);
  for (auto& p : e) { deg[p.first + 1]++; deg[p.second + 1]++; }
  for (int32_t i = 0; i < n; i++) deg[i + 1] += deg[i];
  std::vector<int32_t> nb(deg[n]);
  std::vector<int64_t> cur(deg.begin(), deg.end() - 1);
  for (auto& p : e) { nb[cur[p.first]++] = p.second; nb[cur[p.second]++] = p.first; }
  out.n = n;
  out.end.assign(n, 0);
  out.nbr.clear();
  out.nbr.reserve(nb.size());
  int64_t pos = 0;
  for (int32_t i = 0; i < n; i++) {
    auto b = nb.begin() + deg[i], en = nb.begin() + deg[i + 1];
    std::sort(b, en);
    auto last = std::unique(b, en);
    for (auto it = b; it != last; ++it) out.nbr.push_back(*it);
    pos += last - b;
    out.end[i] = pos;
  }
}

int emit(const Csr& c, int32_t* out_n, int64_t* out_m, int64_t* end_incl, int32_t* nbr, float* w) {
  if (out_n) *out_n = c.n;
  if (out_m) *out_m = c.n ? c.end[c.n - 1] : 0;
  if (end_incl) std::memcpy(end_incl, c.end.data(), sizeof(int64_t) * c.n);
  if (nbr) std::memcpy(nbr, c.nbr.data(), sizeof(int32_t) * c.nbr.size());
  if (w && !c.w.empty()) std::memcpy(w, c.w.data(), sizeof(float) * c.w.size());
  return GALA_OK;
}

// Bernoulli(p) over a run of `len` consecutive candidates: calls f(offset) for hits.
template <class F>
void bernoulli_run(Rng& r, int64_t len, double p, F f) {
  if (p <= 0 || len <= 0) return;
  if (p >= 1) { for (int64_t k = 0; k < len; k++) f(k); return; }
  const double lq = std::log1p(-p);
  int64_t k = -1;
  for (;;) {
    double u = r.uniform_pos();
    int64_t skip = (int64_t)std::floor(std::log(u) / lq);
    k += 1 + skip;
    if (k >= len) break;
    f(k);
  }
}

// Planted-partition SBM: vertex v in block floor(v*blocks/n); every pair i<j
// drawn once, p_in inside a block, p_out across.
void sbm_csr(int32_t n, int32_t blocks, double p_in, double p_out, uint64_t seed, Csr& out) {
  Rng r(seed);
  auto blk = [&](int64_t v) { return (int32_t)((v * blocks) / n); };
  auto bstart = [&](int32_t b) {  // first vertex of block b
    return (int32_t)(((int64_t)b * n + blocks - 1) / blocks);
  };
  std::vector<std::pair<int32_t, int32_t>> e;
  for (int32_t i = 0; i < n; i++) {
    int32_t bi = blk(i);
    // runs of constant probability over j > i: rest of own block, then other blocks
    int32_t j0 = i + 1;
    for (int32_t b = bi; b < blocks && j0 < n; b++) {
      int32_t j1 = (b + 1 < blocks) ? bstart(b + 1) : n;
      if (j1 <= j0) continue;
      double p = (b == bi) ? p_in : p_out;
      bernoulli_run(r, j1 - j0, p, [&](int64_t k) { e.push_back({i, (int32_t)(j0 + k)}); });
      j0 = j1;
    }
  }
  csr_from_pairs(n, e, out);
}

int host_threads() {
  return (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
}

// f(i) for i in [0, count) over nth threads (dynamic assignment).
template <class F>
void parallel_for(int64_t count, int nth, F f) {
  std::atomic<int64_t> next{0};
  auto work = [&]() {
    for (int64_t i; (i = next.fetch_add(1)) < count;) f(i);
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nth && t < count; t++) pool.emplace_back(work);
  work();
  for (auto& t : pool) t.join();
}

// cache of the last generated graph so the two-call protocol does not regenerate
struct GenCache {
  std::string key;
  Csr csr;
} g_cache;

// union of a batch of graphs: graph g owns vertices [vstart[g], vstart[g+1])
struct BatchCsr {
  std::vector<int32_t> vstart;
  std::vector<int64_t> end;  // inclusive ends
  std::vector<int32_t> nbr;
};
struct BatchCache {
  std::string key;
  BatchCsr u;
} g_batch;

// ---- text edge list ingest (Graph::Graph, graph.cpp:18-151) ----
// Line rules: split at '\n'; empty and '#' lines skipped; tokens separated by
// ' ' '\t' '\r' '\v' '\f'; an edge line starts with two integer tokens
// [+-]?[0-9]+ (other lines are skipped); ids outside [0, INT32_MAX] are a
// format error; the optional third token is the weight (int32 integer, or a
// decimal of at most 40 significant digits rounded by strtod).  Rows are
// bucketed by source in line order, each bucket stably sorted by neighbour,
// and repeated entries summed in that order (float64 then float32, or int64
// range-checked to int32) -- the per-row std::map of graph.cpp:54-75.
bool tok_ws(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

bool int_token(const char* p, const char* q, bool* neg, uint64_t* mag) {
  *neg = false;
  if (p < q && (*p == '+' || *p == '-')) *neg = *p++ == '-';
  if (p == q) return false;
  uint64_t v = 0;
  for (; p < q; p++) {
    if (*p < '0' || *p > '9') return false;
    if (v < (1ull << 40)) v = v * 10 + (uint64_t)(*p - '0');
  }
  *mag = v;
  return true;
}

// decimal token grammar and the 40-significant-digit limit
bool decimal_token(const char* p, const char* q) {
  if (p < q && (*p == '+' || *p == '-')) p++;
  int64_t digits = 0, first = -1, last = -1;
  bool point = false;
  for (; p < q && ((*p >= '0' && *p <= '9') || *p == '.'); p++) {
    if (*p == '.') {
      if (point) return false;
      point = true;
      continue;
    }
    if (*p != '0') {
      if (first < 0) first = digits;
      last = digits;
    }
    digits++;
  }
  if (digits == 0 || (first >= 0 && last - first + 1 > 40)) return false;
  if (p == q) return true;
  if (*p != 'e' && *p != 'E') return false;
  if (++p < q && (*p == '+' || *p == '-')) p++;
  if (p == q) return false;
  for (; p < q; p++)
    if (*p < '0' || *p > '9') return false;
  return true;
}

int ingest_text_host(const char* path, int wmode, Csr& out) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return GALA_E_IO;
  std::string txt;
  char buf[1 << 16];
  for (size_t k; (k = std::fread(buf, 1, sizeof buf, f)) > 0;) txt.append(buf, k);
  const bool bad_read = std::ferror(f) != 0;
  std::fclose(f);
  if (bad_read) return GALA_E_IO;
  std::vector<int32_t> ea, eb, wi;
  std::vector<double> wd;
  int32_t maxid = -1;
  const char* t = txt.data();
  const char* tend = t + txt.size();
  std::string tok;
  for (size_t ls = 0; ls <= txt.size();) {
    const char* s = t + ls;
    const char* e = static_cast<const char*>(std::memchr(s, '\n', (size_t)(tend - s)));
    if (!e) e = tend;
    ls = (size_t)(e - t) + 1;
    if (s < e && *s != '#') {
      const char* p = s;
      const char* tk[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
      for (int k = 0; k < 3; k++) {
        while (p < e && tok_ws(*p)) p++;
        tk[k][0] = p;
        while (p < e && !tok_ws(*p)) p++;
        tk[k][1] = p;
      }
      bool n1, n2;
      uint64_t a, b;
      if (tk[0][0] < tk[0][1] && tk[1][0] < tk[1][1] && int_token(tk[0][0], tk[0][1], &n1, &a) &&
          int_token(tk[1][0], tk[1][1], &n2, &b)) {
        if ((n1 && a) || (n2 && b) || a > INT32_MAX || b > INT32_MAX) return GALA_E_FORMAT;
        const bool has_w = tk[2][0] < tk[2][1];
        if (wmode == GALA_W_I32) {
          bool nw = false;
          uint64_t mw = 1;
          if (has_w && (!int_token(tk[2][0], tk[2][1], &nw, &mw) || mw > (nw ? 2147483648ull : 2147483647ull)))
            return GALA_E_FORMAT;
          wi.push_back(nw ? (int32_t)(-(int64_t)mw) : (int32_t)mw);
        } else if (wmode == GALA_W_F32) {
          double x = 1.0;
          if (has_w) {
            if (!decimal_token(tk[2][0], tk[2][1])) return GALA_E_FORMAT;
            tok.assign(tk[2][0], tk[2][1]);
            x = std::strtod(tok.c_str(), nullptr);
            if (std::isinf(x)) return GALA_E_FORMAT;
          }
          wd.push_back(x);
        }
        ea.push_back((int32_t)a);
        eb.push_back((int32_t)b);
        maxid = std::max(maxid, (int32_t)std::max(a, b));
      }
    }
  }
  const int64_t E = (int64_t)ea.size();
  out = Csr();
  if (E == 0) return GALA_OK;
  // renumbering of the linked ids, ascending: a table over [0, maxid] when the
  // ids are dense enough, else the sorted distinct ids searched per endpoint
  int32_t nn = 0;
  if ((int64_t)maxid + 1 <= 8 * E + (1 << 20)) {
    std::vector<int32_t> renum((size_t)maxid + 1, -1);
    for (int64_t i = 0; i < E; i++) renum[ea[i]] = renum[eb[i]] = 0;
    for (auto& r : renum)
      if (r == 0) r = nn++;
    for (int64_t i = 0; i < E; i++) { ea[i] = renum[ea[i]]; eb[i] = renum[eb[i]]; }
  } else {
    std::vector<int32_t> ids(ea);
    ids.insert(ids.end(), eb.begin(), eb.end());
    std::sort(ids.begin(), ids.end());
    ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
    nn = (int32_t)ids.size();
    auto rank = [&](int32_t x) { return (int32_t)(std::lower_bound(ids.begin(), ids.end(), x) - ids.begin()); };
    for (int64_t i = 0; i < E; i++) { ea[i] = rank(ea[i]); eb[i] = rank(eb[i]); }
  }
  // bucket entries by row in line order (a->b, then b->a)
  std::vector<int64_t> start((size_t)nn + 1, 0);
  for (int64_t i = 0; i < E; i++) { start[ea[i] + 1]++; start[eb[i] + 1]++; }
  for (int32_t v = 0; v < nn; v++) start[v + 1] += start[v];
  std::vector<int64_t> cur(start.begin(), start.end() - 1);
  std::vector<std::pair<int32_t, int64_t>> ent((size_t)(2 * E));  // (neighbour, line)
  for (int64_t i = 0; i < E; i++) {
    const int32_t ra = ea[i], rb = eb[i];
    ent[cur[ra]++] = {rb, i};
    ent[cur[rb]++] = {ra, i};
  }
  out.n = nn;
  out.end.assign(nn, 0);
  for (int32_t v = 0; v < nn; v++) {
    auto b0 = ent.begin() + start[v], b1 = ent.begin() + start[v + 1];
    std::stable_sort(b0, b1, [](const std::pair<int32_t, int64_t>& x, const std::pair<int32_t, int64_t>& y) {
      return x.first < y.first;
    });
    for (auto it = b0; it != b1;) {
      auto jt = it;
      double sd = 0;
      int64_t si = 0;
      for (; jt != b1 && jt->first == it->first; ++jt) {
        if (wmode == GALA_W_F32) sd = jt == it ? wd[jt->second] : sd + wd[jt->second];
        if (wmode == GALA_W_I32) si += wi[jt->second];
      }
      out.nbr.push_back(it->first);
      if (wmode == GALA_W_F32) {
        const float fw = (float)sd;
        if (!std::isfinite(fw)) return GALA_E_FORMAT;
        out.w.push_back(fw);
      }
      if (wmode == GALA_W_I32) {
        if (si < INT32_MIN || si > INT32_MAX) return GALA_E_FORMAT;
        out.wi.push_back((int32_t)si);
      }
      it = jt;
    }
    out.end[v] = (int64_t)out.nbr.size();
  }
  return GALA_OK;
}

}  // namespace

extern "C" {

int gala_read_bin(const char* path, int weight_mode, int32_t* n, int64_t* m, int64_t* end_incl,
                  int32_t* nbr, void* w) {
  if (!path || !n || !m) return GALA_E_INVALID;
  FILE* f = std::fopen(path, "rb");
  if (!f) return GALA_E_IO;
  std::fseek(f, 0, SEEK_END);
  long long fsz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  int32_t nn = 0;
  if (std::fread(&nn, 4, 1, f) != 1) { std::fclose(f); return GALA_E_FORMAT; }
  if (nn < 0 || 4 + 8ll * nn > fsz) { std::fclose(f); return GALA_E_FORMAT; }
  int64_t last = 0;
  if (nn > 0) {
    std::fseek(f, 4 + 8ll * (nn - 1), SEEK_SET);
    if (std::fread(&last, 8, 1, f) != 1) { std::fclose(f); return GALA_E_FORMAT; }
  }
  // the entry count comes from the file: bound it by the bytes present before
  // multiplying, so a corrupt count cannot overflow the size test
  const long long per_entry = weight_mode != GALA_W_NONE ? 8 : 4;
  if (last < 0 || last > (fsz - 4 - 8ll * nn) / per_entry) { std::fclose(f); return GALA_E_FORMAT; }
  *n = nn;
  *m = last;
  if (!end_incl) { std::fclose(f); return GALA_OK; }
  std::fseek(f, 4, SEEK_SET);
  if (nn > 0 && std::fread(end_incl, 8, nn, f) != (size_t)nn) { std::fclose(f); return GALA_E_IO; }
  int64_t prev = 0;
  for (int32_t i = 0; i < nn; i++) {
    if (end_incl[i] < prev) { std::fclose(f); return GALA_E_FORMAT; }
    prev = end_incl[i];
  }
  if (nbr && last > 0 && std::fread(nbr, 4, last, f) != (size_t)last) { std::fclose(f); return GALA_E_IO; }
  if (nbr)
    for (int64_t j = 0; j < last; j++)
      if (nbr[j] < 0 || nbr[j] >= nn) { std::fclose(f); return GALA_E_FORMAT; }
  if (w && weight_mode != GALA_W_NONE && last > 0 &&
      std::fread(w, 4, last, f) != (size_t)last) { std::fclose(f); return GALA_E_IO; }
  std::fclose(f);
  return GALA_OK;
}

int gala_write_bin(const char* path, int32_t n, const int64_t* end_incl, const int32_t* nbr,
                   const void* w, int weight_mode) {
  if (!path || n < 0 || (n > 0 && (!end_incl || !nbr))) return GALA_E_INVALID;
  FILE* f = std::fopen(path, "wb");
  if (!f) return GALA_E_IO;
  int64_t m = n ? end_incl[n - 1] : 0;
  bool ok = std::fwrite(&n, 4, 1, f) == 1;
  if (n) ok = ok && std::fwrite(end_incl, 8, n, f) == (size_t)n;
  if (m) ok = ok && std::fwrite(nbr, 4, m, f) == (size_t)m;
  if (w && weight_mode != GALA_W_NONE && m) ok = ok && std::fwrite(w, 4, m, f) == (size_t)m;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? GALA_OK : GALA_E_IO;
}

int gala_ingest_text(const char* path, int weight_mode, int32_t* n, int64_t* m, int64_t* end_incl,
                     int32_t* nbr, void* w) {
  if (!path || !n || !m || weight_mode < GALA_W_NONE || weight_mode > GALA_W_F32) return GALA_E_INVALID;
  std::string key = std::string("txt:") + path + ":" + std::to_string(weight_mode);
  if (g_cache.key != key) {
    try {
      Csr c;
      int rc = ingest_text_host(path, weight_mode, c);
      if (rc) return rc;
      g_cache.key = key;
      g_cache.csr = std::move(c);
    } catch (const std::exception&) {
      return GALA_E_OOM;
    }
  }
  const Csr& c = g_cache.csr;
  *n = c.n;
  *m = c.n ? c.end[c.n - 1] : 0;
  if (end_incl) std::memcpy(end_incl, c.end.data(), 8 * c.n);
  if (nbr) std::memcpy(nbr, c.nbr.data(), 4 * c.nbr.size());
  if (w && weight_mode == GALA_W_F32) std::memcpy(w, c.w.data(), 4 * c.w.size());
  if (w && weight_mode == GALA_W_I32) std::memcpy(w, c.wi.data(), 4 * c.wi.size());
  return GALA_OK;
}

int gala_gen_sbm(int32_t n, int32_t blocks, double p_in, double p_out, uint64_t seed, int32_t* out_n,
                 int64_t* out_m, int64_t* end_incl, int32_t* nbr) {
  if (n < 0 || blocks < 1 || blocks > std::max(n, 1)) return GALA_E_INVALID;
  char kb[256];
  std::snprintf(kb, sizeof kb, "sbm:%d:%d:%.17g:%.17g:%llu", n, blocks, p_in, p_out,
                (unsigned long long)seed);
  if (g_cache.key != kb) {
    Csr c;
    sbm_csr(n, blocks, p_in, p_out, seed, c);
    g_cache.key = kb;
    g_cache.csr = std::move(c);
  }
  return emit(g_cache.csr, out_n, out_m, end_incl, nbr, nullptr);
}

int gala_gen_sbm_batch(int32_t n_graphs, const int32_t* n, const int32_t* blocks, const double* p_in,
                       const double* p_out, const uint64_t* seeds, int32_t threads, int32_t* vstart,
                       int64_t* out_m, int64_t* end_incl, int32_t* nbr) {
  if (n_graphs < 0 || (n_graphs > 0 && (!n || !blocks || !p_in || !p_out || !seeds))) return GALA_E_INVALID;
  int64_t nt = 0;
  for (int32_t g = 0; g < n_graphs; g++) {
    if (n[g] < 0 || blocks[g] < 1 || blocks[g] > std::max(n[g], 1)) return GALA_E_INVALID;
    nt += n[g];
  }
  if (nt > INT32_MAX) return GALA_E_INVALID;
  // key: FNV-1a over the parameter arrays
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](const void* p, size_t len) {
    const unsigned char* b = (const unsigned char*)p;
    for (size_t i = 0; i < len; i++) { h ^= b[i]; h *= 1099511628211ull; }
  };
  mix(&n_graphs, sizeof n_graphs);
  mix(n, sizeof(int32_t) * n_graphs); mix(blocks, sizeof(int32_t) * n_graphs);
  mix(p_in, sizeof(double) * n_graphs); mix(p_out, sizeof(double) * n_graphs);
  mix(seeds, sizeof(uint64_t) * n_graphs);
  char kb[64];
  std::snprintf(kb, sizeof kb, "sbmbatch:%d:%016llx", n_graphs, (unsigned long long)h);
  if (g_batch.key != kb) {
    std::vector<Csr> parts(n_graphs);
    int nth = threads > 0 ? threads : host_threads();
    // largest graphs first so the threads finish together
    std::vector<int32_t> ord(n_graphs);
    for (int32_t g = 0; g < n_graphs; g++) ord[g] = g;
    std::sort(ord.begin(), ord.end(), [&](int32_t a, int32_t b) { return n[a] > n[b]; });
    std::atomic<int32_t> next{0};
    auto work = [&]() {
      for (;;) {
        int32_t k = next.fetch_add(1);
        if (k >= n_graphs) break;
        int32_t g = ord[k];
        sbm_csr(n[g], blocks[g], p_in[g], p_out[g], seeds[g], parts[g]);
      }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nth; t++) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    BatchCsr u;
    u.vstart.assign(n_graphs + 1, 0);
    int64_t mt = 0;
    for (int32_t g = 0; g < n_graphs; g++) {
      u.vstart[g + 1] = u.vstart[g] + n[g];
      mt += parts[g].nbr.size();
    }
    u.end.resize(nt);
    u.nbr.resize(mt);
    int64_t e0 = 0;
    for (int32_t g = 0; g < n_graphs; g++) {
      const Csr& c = parts[g];
      const int32_t v0 = u.vstart[g];
      for (int32_t i = 0; i < c.n; i++) u.end[v0 + i] = e0 + c.end[i];
      for (size_t j = 0; j < c.nbr.size(); j++) u.nbr[e0 + j] = v0 + c.nbr[j];
      e0 += c.nbr.size();
      parts[g] = Csr();
    }
    g_batch.key = kb;
    g_batch.u = std::move(u);
  }
  const BatchCsr& u = g_batch.u;
  if (vstart) std::memcpy(vstart, u.vstart.data(), sizeof(int32_t) * (n_graphs + 1));
  if (out_m) *out_m = (int64_t)u.nbr.size();
  if (end_incl) std::memcpy(end_incl, u.end.data(), sizeof(int64_t) * u.end.size());
  if (nbr) std::memcpy(nbr, u.nbr.data(), sizeof(int32_t) * u.nbr.size());
  return GALA_OK;
}

int gala_gen_chung_lu(int32_t n, double gamma, double mean_deg, double max_deg, uint64_t seed,
                      int32_t* out_n, int64_t* out_m, int64_t* end_incl, int32_t* nbr) {
  if (n < 2 || gamma <= 1 || mean_deg <= 0 || max_deg <= 0) return GALA_E_INVALID;
  char kb[256];
  std::snprintf(kb, sizeof kb, "cl:%d:%.17g:%.17g:%.17g:%llu", n, gamma, mean_deg, max_deg,
                (unsigned long long)seed);
  if (g_cache.key != kb) {
    const double ex = -1.0 / (gamma - 1.0);
    std::vector<double> base(n);
    for (int32_t i = 0; i < n; i++) base[i] = std::pow((double)i + 1.0, ex);
    auto mean_for = [&](double c) {
      double s = 0;
      for (int32_t i = 0; i < n; i++) s += std::min(max_deg, c * base[i]);
      return s / n;
    };
    double lo = 0, hi = 1;
    while (mean_for(hi) < mean_deg && hi < 1e300) hi *= 2;
    for (int it = 0; it < 200; it++) {
      double mid = 0.5 * (lo + hi);
      (mean_for(mid) < mean_deg ? lo : hi) = mid;
    }
    std::vector<double> wv(n);
    double S = 0;
    for (int32_t i = 0; i < n; i++) { wv[i] = std::min(max_deg, hi * base[i]); S += wv[i]; }
    Rng r(seed);
    std::vector<std::pair<int32_t, int32_t>> e;
    e.reserve((size_t)(S / 2 * 1.02));
    for (int32_t u = 0; u < n - 1; u++) {
      int64_t v = u + 1;
      double p = std::min(wv[u] * wv[v] / S, 1.0);
      while (v < n && p > 0) {
        if (p != 1.0) {
          double rr = r.uniform_pos();
          v += (int64_t)std::floor(std::log(rr) / std::log1p(-p));
        }
        if (v < n) {
          double q = std::min(wv[u] * wv[v] / S, 1.0);
          if (r.uniform() < q / p) e.push_back({u, (int32_t)v});
          p = q;
          v += 1;
        }
      }
    }
    Csr c;
    csr_from_pairs(n, e, c);
    g_cache.key = kb;
    g_cache.csr = std::move(c);
  }
  return emit(g_cache.csr, out_n, out_m, end_incl, nbr, nullptr);
}

int gala_gen_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                  float w_lo, float w_hi, int32_t* out_n, int64_t* out_m, int64_t* end_incl,
                  int32_t* nbr, float* w) {
  if (scale < 1 || scale > 30 || edge_factor < 1 || a < 0 || b < 0 || c < 0 || a + b + c > 1)
    return GALA_E_INVALID;
  char kb[256];
  std::snprintf(kb, sizeof kb, "rmat:%d:%d:%.17g:%.17g:%.17g:%llu:%g:%g", scale, edge_factor, a, b, c,
                (unsigned long long)seed, (double)w_lo, (double)w_hi);
  if (g_cache.key != kb) {
    const int32_t n = 1 << scale;
    const int64_t ne = (int64_t)edge_factor << scale;
    const int nth = host_threads();
    // edges in fixed chunks, each from its own stream (output independent of
    // the thread count)
    constexpr int64_t CHUNK = 1 << 20;
    const int64_t nchunks = (ne + CHUNK - 1) / CHUNK;
    std::vector<int32_t> src(ne), dst(ne);
    std::vector<float> wt(ne);
    parallel_for(nchunks, nth, [&](int64_t ch) {
      uint64_t x = seed ^ (0xA0761D6478BD642Full * (uint64_t)(ch + 1));
      Rng r(Rng::splitmix(x));
      for (int64_t k = ch * CHUNK; k < std::min(ne, (ch + 1) * CHUNK); k++) {
        int32_t u = 0, v = 0;
        for (int lvl = 0; lvl < scale; lvl++) {
          double q = r.uniform();
          int bu = 0, bv = 0;
          if (q < a) {}
          else if (q < a + b) bv = 1;
          else if (q < a + b + c) bu = 1;
          else { bu = 1; bv = 1; }
          u = (u << 1) | bu;
          v = (v << 1) | bv;
        }
        src[k] = u; dst[k] = v;
        wt[k] = (float)(w_lo + (w_hi - w_lo) * r.uniform());
      }
    });
    // symmetrise into rows; a self-loop contributes one entry
    std::vector<int64_t> deg(n + 1, 0);
    parallel_for(nchunks, nth, [&](int64_t ch) {
      for (int64_t k = ch * CHUNK; k < std::min(ne, (ch + 1) * CHUNK); k++) {
        __atomic_fetch_add(&deg[src[k] + 1], 1, __ATOMIC_RELAXED);
        if (src[k] != dst[k]) __atomic_fetch_add(&deg[dst[k] + 1], 1, __ATOMIC_RELAXED);
      }
    });
    for (int32_t i = 0; i < n; i++) deg[i + 1] += deg[i];
    std::vector<std::pair<int32_t, float>> ent(deg[n]);
    std::vector<int64_t> cur(deg.begin(), deg.end() - 1);
    parallel_for(nchunks, nth, [&](int64_t ch) {
      for (int64_t k = ch * CHUNK; k < std::min(ne, (ch + 1) * CHUNK); k++) {
        ent[__atomic_fetch_add(&cur[src[k]], 1, __ATOMIC_RELAXED)] = {dst[k], wt[k]};
        if (src[k] != dst[k]) ent[__atomic_fetch_add(&cur[dst[k]], 1, __ATOMIC_RELAXED)] = {src[k], wt[k]};
      }
    });
    src.clear(); src.shrink_to_fit(); dst.clear(); dst.shrink_to_fit(); wt.clear(); wt.shrink_to_fit();
    // rows sorted by (neighbour, weight) so the merge order is fixed; parallel
    // entries summed in FP64
    constexpr int64_t RCH = 4096;
    const int64_t nrch = (n + RCH - 1) / RCH;
    std::vector<int64_t> uniq(n + 1, 0);
    parallel_for(nrch, nth, [&](int64_t ch) {
      for (int64_t i = ch * RCH; i < std::min<int64_t>(n, (ch + 1) * RCH); i++) {
        auto bgn = ent.begin() + deg[i], en = ent.begin() + deg[i + 1];
        std::sort(bgn, en);
        int64_t u = 0;
        for (auto it = bgn; it != en; ++it)
          if (it == bgn || it->first != (it - 1)->first) u++;
        uniq[i + 1] = u;
      }
    });
    for (int32_t i = 0; i < n; i++) uniq[i + 1] += uniq[i];
    Csr cs;
    cs.n = n;
    cs.end.assign(n, 0);
    cs.nbr.resize(uniq[n]);
    cs.w.resize(uniq[n]);
    parallel_for(nrch, nth, [&](int64_t ch) {
      for (int64_t i = ch * RCH; i < std::min<int64_t>(n, (ch + 1) * RCH); i++) {
        auto bgn = ent.begin() + deg[i], en = ent.begin() + deg[i + 1];
        int64_t pos = uniq[i];
        for (auto it = bgn; it != en;) {
          int32_t k = it->first;
          double sum = 0;
          while (it != en && it->first == k) { sum += it->second; ++it; }
          cs.nbr[pos] = k;
          cs.w[pos] = (float)sum;
          pos++;
        }
        cs.end[i] = uniq[i + 1];
      }
    });
    g_cache.key = kb;
    g_cache.csr = std::move(cs);
  }
  return emit(g_cache.csr, out_n, out_m, end_incl, nbr, w);
}

}  // extern "C"
