// Host driver of the gfx950 Louvain engine and the C ABI of include/gala.h.
//
// Replaces the reference's round driver (louvain_gpu, LinXi-lx/GALA-mg
// src/louvain_gpu/louvain.cu:136-229, with the multi-round loop of :193-226
// made live), its phase-1 driver (louvain_main_process,
// src/louvain_gpu/louvain_main_process.cu:237-1060) and phase-2 driver
// (build_compressed_graph, src/louvain_gpu/build_compressed_graph.cu:5-173).
//
// Device residency: the input CSR is uploaded once; every round's graph,
// state and frontier stay in HBM.  Per iteration the host reads back one
// 128-byte counter block (exact sums for Q, move/sweep counts) -- the only
// synchronisation, needed for the reference's convergence test.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gala.h"
#include "gala_kernels.h"

using namespace gala;

#define HIPOK(x)                                    \
  do {                                              \
    hipError_t e_ = (x);                            \
    if (e_ != hipSuccess) {                         \
      ctx->last_hip = e_;                           \
      return GALA_E_HIP;                            \
    }                                               \
  } while (0)

namespace {

constexpr int64_t M2_LIMIT = 1000000000LL;  // int32 path: 4*m2^2 < 2^63 for the exact int64 gains

template <class P>
hipError_t dalloc(P** p, int64_t elems) {
  *p = nullptr;
  size_t bytes = (size_t)std::max<int64_t>(elems, 1) * sizeof(P);
  return hipMalloc((void**)p, bytes);
}

}  // namespace

struct gala_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;  // long-row degree bins beside the others (over_bins)
  hipEvent_t evf[2] = {};         // fork / join of the two streams
  hipEvent_t evr[2] = {};         // batched runs: readback of the live-graph count, one per iteration in flight
  gala_opts opts{};
  hipError_t last_hip = hipSuccess;

  // last device text ingest (two-call protocol) and its key: path + weight mode
  gala::IngestResult ingest;
  std::string ingest_key;

  // input graph (sizes) and mode
  int32_t n0 = 0;
  int64_t m0 = 0;
  int wmode = GALA_W_NONE;
  // exact wide path (int64 sums, int128 gains): float32 weights, or m2 > 1e9
  bool wide = false;
  int32_t wshift = 0;   // float32 weights: w * 2^wshift is an integer for every weight
  int64_t m2_wide = 0;  // m2 of the wide path (scaled integer), computed on upload
  bool have_graph = false;
  double t_h2d_ms = 0;

  // CSR buffers: [0] = input (read-only, kept for re-runs), [1]/[2] = work
  // buffers the contraction ping-pongs between.  Weights are int32 (int path)
  // or int64 sums (wide path; its input round may hold float32 weights);
  // the input may have implicit unit weights.
  int64_t* off[3] = {nullptr, nullptr, nullptr};
  // row ends of the work graphs when they are gapped (contracted inside a run)
  int64_t* end[3] = {nullptr, nullptr, nullptr};
  bool gapped[3] = {false, false, false};
  int32_t* nbr[3] = {nullptr, nullptr, nullptr};
  void* w[3] = {nullptr, nullptr, nullptr};
  bool unit0 = true;
  bool f32in = false;  // input weights kept as float32 in w[0] (wide path, scaled on load)
  bool unique0 = false;  // input rows strictly ascending (duplicate-free)
  bool noself0 = false;  // input rows hold no self-loop
  int src = 0;  // buffer holding the current round graph

  // per-vertex state (sized n0)
  void *K = nullptr, *Self = nullptr, *Tot = nullptr, *In = nullptr;
  int32_t *size = nullptr, *prev = nullptr, *cur = nullptr;
  uint8_t* active = nullptr;
  uint8_t* dirty = nullptr;
  int8_t* binid = nullptr;
  int32_t *order = nullptr, *front = nullptr, *fin = nullptr;
  void* cv = nullptr;  // community view, 8 (int) or 16 (wide) bytes per vertex
  int2* cp = nullptr;  // {cur, prev} per vertex (integer path, incremental In)
  int32_t* cm = nullptr;  // cur with bit 31 set for a vertex that moved in the sweep in progress
  unsigned long long* tsd = nullptr;  // packed Tot/size deltas of a sweep (integer path; zero between sweeps)
  BlockPart* part = nullptr;  // workgroup partials of the O(V) reductions
  int32_t *bstart_full = nullptr, *bstart_front = nullptr;
  int64_t* hub_off = nullptr;
  int64_t* hub_tmp = nullptr;
  int32_t* hub_keys = nullptr;
  void* hub_vals = nullptr;
  int32_t* hub_rc = nullptr;  // resolved split-hub rows (hub_cap slots each)
  void* hub_rw = nullptr;
  int64_t hub_cap = 0;
  int64_t hub_dmax = -1;  // longest row of the round graph's hub bins (round_setup; -1: unknown)
  int32_t* hub_plan = nullptr;  // split hub decide scratch
  void* hub_part = nullptr;
  uint8_t* hub_fail = nullptr;
  int32_t* part_tmp = nullptr;
  int64_t* scan64_tmp = nullptr;
  Counters* ctr = nullptr;
  Counters* ctr_h = nullptr;  // pinned
  RoundCtl* rctl = nullptr;   // the device-side stopping rule of queued rounds
  char* hpin = nullptr;       // pinned scratch of the small readbacks (HPIN_BYTES)

  // contraction buffers
  int32_t *c_cnt = nullptr, *c_nid = nullptr, *c_label = nullptr, *c_skey = nullptr, *c_dcnt = nullptr,
          *c_big_list = nullptr, *c_nbig = nullptr, *c_i32_tmp = nullptr, *c_i32_scalar = nullptr;
  int64_t *c_ecnt = nullptr, *c_eoff = nullptr, *c_ecur = nullptr, *c_big_tab = nullptr,
          *c_i64_tmp = nullptr, *c_i64_scalar = nullptr, *c_mtotal = nullptr, *c_voff = nullptr;
  void* c_sval = nullptr;
  int8_t* c_tier = nullptr;
  int32_t *c_tlist = nullptr, *c_tstart = nullptr;
  int32_t* arena_keys = nullptr;
  void* arena_vals = nullptr;
  int64_t arena_slots = 0;
  DenseChunk* chunks = nullptr;  // dense-row chunk table
  int64_t chunks_cap = 0;
  char* dscratch = nullptr;      // scratch of the dense-row sort
  int64_t dscratch_bytes = 0;

  // learned frontier
  float* last_gain = nullptr;
  int32_t* moved_it = nullptr;
  float* nbr_moved = nullptr;
  int32_t* front2 = nullptr;   // predicted movers of a learned sweep, then its complement
  int32_t* bstart2 = nullptr;
  int32_t* front3 = nullptr;   // second frontier buffer (the pre-sweep frontier survives a sweep)
  int32_t* bstart3 = nullptr;
  void* tot_snap = nullptr;    // Tot and size before a learned sweep
  int32_t* size_snap = nullptr;
  uint8_t* keep = nullptr;
  float* feat = nullptr;        // logged features of the frontier
  uint8_t* lab = nullptr;
  float* mlp_prm = nullptr;     // MLP_NPARAM fp32 (persistent across graphs)
  short* mlp_w1b = nullptr;     // BF16 copy of W1
  bool have_mlp = false;
  gala_log_cb logger = nullptr;
  void* logger_user = nullptr;

  // batched graphs (gala_batch_*): union layout and per-graph state
  int32_t b_G = 0;
  std::vector<int32_t> b_vstart;       // host copy of vstart[G+1]
  int32_t* b_gid[3] = {nullptr, nullptr, nullptr};  // [0] input vertices, [1]/[2] round graphs
  int32_t* b_m2g = nullptr;
  uint8_t *b_live = nullptr, *b_accept = nullptr, *b_done = nullptr;
  double *b_qprev = nullptr, *b_cur_mod = nullptr;
  int32_t *b_iters = nullptr, *b_tot_iters = nullptr, *b_rounds = nullptr, *b_ncomm = nullptr;
  int32_t* b_scalars = nullptr;        // [0] nlive, [1] nopen, [2] iteration counter
  unsigned long long *b_gin = nullptr, *b_gtot2 = nullptr;
  int32_t* b_host = nullptr;           // pinned readback of the scalars
  // vertices (of positive degree) of the graphs still iterating, grouped by
  // degree bin: the per-sweep O(n) passes visit only these; [1] is scratch
  int32_t* b_vl[2] = {nullptr, nullptr};
  int32_t* b_lbs[2] = {nullptr, nullptr};  // their bin starts (NBINS + 1)
  uint8_t* b_vlive = nullptr;
  int32_t b_m2max = 0;
  std::vector<void*> bowned;
  bool have_batch = false, have_batch_result = false;
  std::vector<int32_t> b_comm_local;   // per-vertex community within its graph
  bool b_comm_ready = false;           // b_comm_local read back from fin (gala_batch_communities)
  std::vector<double> b_q;
  std::vector<int32_t> b_nc, b_rd, b_it;

  // decide kernel variant per degree bin, general and first sweeps; the
  // defaults are what gala_autotune picks on the 4M-vertex power-law graph
  // (one thread per vertex below degree 32, the leader-hash kernels with
  // resident workgroups for 32..511, leader evaluation in the workgroup kernels
  // for 512..8191, every hub of 8192+ split by community over workgroups)
  int dvar[NBINS] = {3, 3, 3, 3, 8, 5, 1, 1, 6};
  int fvar[NBINS] = {1, 2, 2, 2, 3, 2, 0, 0, 0};
  // weight update: the incremental integer path with narrow lane groups
  // (variant 2); the dirty-row path (tracked runs) reads variant 2 as one
  // thread per vertex for bins 0-3
  int wvar[NBINS] = {2, 2, 2, 2, 2, 2, 0, 0, 0};
  // degree-bin cut points (lowest degree of every bin); gala_autotune may
  // lower those of bins 4..8
  BinCuts cuts = {{BIN_LO[0], BIN_LO[1], BIN_LO[2], BIN_LO[3], BIN_LO[4], BIN_LO[5], BIN_LO[6], BIN_LO[7],
                   BIN_LO[8]}};

  std::vector<void*> owned;
  int32_t cap_n = -1;  // sizes the owned buffers were allocated for
  int64_t cap_m = -1;
  bool cap_wide = false;
  gala_result res{};
  bool have_result = false;

  hipEvent_t ev[8] = {};
  hipEvent_t bev[NBINS + 1] = {};  // per-bin decide timing (opts.profile_bins)
  hipEvent_t kev[2] = {};          // the bin-8 part kernel's own time (opts.profile_bins)
  hipEvent_t sev[2][3] = {};       // queued rounds: decide start / decide end / update end, per sweep parity
};

namespace {

bool is_wide(const gala_ctx* ctx) { return ctx->wide; }

void free_batch(gala_ctx* ctx) {
  for (void* p : ctx->bowned) (void)hipFree(p);
  ctx->bowned.clear();
  ctx->have_batch = false;
  ctx->have_batch_result = false;
  ctx->b_G = 0;
}

void free_all(gala_ctx* ctx) {
  for (void* p : ctx->owned) (void)hipFree(p);
  ctx->owned.clear();
  ctx->cap_n = -1;
  ctx->cap_m = -1;
  free_batch(ctx);
  if (ctx->hub_keys) (void)hipFree(ctx->hub_keys);
  if (ctx->hub_vals) (void)hipFree(ctx->hub_vals);
  if (ctx->hub_rc) (void)hipFree(ctx->hub_rc);
  if (ctx->hub_rw) (void)hipFree(ctx->hub_rw);
  ctx->hub_keys = nullptr; ctx->hub_vals = nullptr; ctx->hub_cap = 0;
  ctx->hub_rc = nullptr; ctx->hub_rw = nullptr;
  if (ctx->arena_keys) (void)hipFree(ctx->arena_keys);
  if (ctx->arena_vals) (void)hipFree(ctx->arena_vals);
  ctx->arena_keys = nullptr; ctx->arena_vals = nullptr;
  ctx->arena_slots = 0;
  for (void* p : {(void*)ctx->chunks, (void*)ctx->dscratch})
    if (p) (void)hipFree(p);
  ctx->chunks = nullptr; ctx->dscratch = nullptr; ctx->chunks_cap = 0; ctx->dscratch_bytes = 0;
  ctx->have_graph = false;
}

template <class P>
int own(gala_ctx* ctx, P** p, int64_t elems) {
  hipError_t e = dalloc(p, elems);
  if (e != hipSuccess) { ctx->last_hip = e; return GALA_E_OOM; }
  ctx->owned.push_back((void*)*p);
  return GALA_OK;
}

#define OWN(p, n)                                   \
  do {                                              \
    int rc_ = own(ctx, &(p), (n));                  \
    if (rc_) { free_all(ctx); return rc_; }         \
  } while (0)

// Allocate every device buffer for an input of n vertices / m entries (kept
// when the buffers of the previous graph are large enough).
int alloc_buffers(gala_ctx* ctx, int32_t n, int64_t m, bool wide) {
  free_batch(ctx);
  if (!ctx->owned.empty() && n <= ctx->cap_n && m <= ctx->cap_m && wide == ctx->cap_wide) return GALA_OK;
  free_all(ctx);
  const int64_t N = std::max<int32_t>(n, 1), M = std::max<int64_t>(m, 1);
  const size_t tsz = wide ? 8 : 4;  // tot_t / w_t / acc_t size
  for (int k = 0; k < 3; k++) {
    OWN(ctx->off[k], N + 1);
    OWN(ctx->end[k], N);
    ctx->gapped[k] = false;
    // 16 entries of padding: row loads with 16-byte vectors may read up to 3
    // entries past the end of the last row
    OWN(ctx->nbr[k], M + CSR_PAD);
    char* p;
    OWN(p, (M + CSR_PAD) * tsz);
    ctx->w[k] = p;
  }
  { char* p;
    OWN(p, N * tsz); ctx->K = p; OWN(p, N * tsz); ctx->Self = p; OWN(p, N * tsz); ctx->Tot = p;
    OWN(p, N * tsz); ctx->In = p; }
  OWN(ctx->size, N); OWN(ctx->prev, N); OWN(ctx->cur, N);
  OWN(ctx->active, N); OWN(ctx->dirty, N); OWN(ctx->binid, N);
  OWN(ctx->order, N); OWN(ctx->front, N); OWN(ctx->fin, N);
  { char* p; OWN(p, N * 16); ctx->cv = p; }
  OWN(ctx->cp, N);
  OWN(ctx->cm, N);
  OWN(ctx->tsd, N);
  fill<int64_t>((int64_t*)ctx->tsd, N, 0, ctx->stream);
  OWN(ctx->last_gain, N); OWN(ctx->moved_it, N); OWN(ctx->nbr_moved, N);
  OWN(ctx->front2, N); OWN(ctx->bstart2, NBINS + 1); OWN(ctx->keep, N);
  OWN(ctx->front3, N); OWN(ctx->bstart3, NBINS + 1); OWN(ctx->size_snap, N);
  { char* p; OWN(p, N * tsz); ctx->tot_snap = p; }
  OWN(ctx->feat, N * MLP_IN); OWN(ctx->lab, N);
  OWN(ctx->bstart_full, NBINS + 1); OWN(ctx->bstart_front, NBINS + 1);
  OWN(ctx->hub_off, N + 1); OWN(ctx->hub_tmp, 2 * (N + 1));
  OWN(ctx->hub_plan, hub_plan_elems(M));
  OWN(ctx->hub_fail, hub_plan_elems(M));
  { char* p; OWN(p, hub_part_elems(M) * (int64_t)sizeof(HubPart<WideW>)); ctx->hub_part = p; }
  OWN(ctx->part_tmp, part_tmp_elems(N));
  OWN(ctx->scan64_tmp, scan_tmp_elems(N) + 2);
  OWN(ctx->ctr, 2);  // [1] holds a deferred iteration-0 copy
  OWN(ctx->rctl, 1);
  OWN(ctx->part, MAX_PARTS);
  // contraction
  OWN(ctx->c_cnt, N); OWN(ctx->c_nid, N); OWN(ctx->c_label, N); OWN(ctx->c_skey, M);
  { char* p; OWN(p, M * tsz); ctx->c_sval = p; }
  OWN(ctx->c_dcnt, N + 1); OWN(ctx->c_big_list, 2 * N); OWN(ctx->c_nbig, 2);
  OWN(ctx->c_i32_tmp, N / 4 + 2 + scan_tmp_elems(N));
  OWN(ctx->c_i32_scalar, 4);
  OWN(ctx->c_tier, N); OWN(ctx->c_tlist, N); OWN(ctx->c_tstart, NBINS + 1);
  OWN(ctx->c_ecnt, N + 1); OWN(ctx->c_eoff, N + 1); OWN(ctx->c_ecur, N);
  OWN(ctx->c_big_tab, N + 1); OWN(ctx->c_i64_tmp, 2 * (N + 1) + scan_tmp_elems(N));
  OWN(ctx->c_i64_scalar, 16);
  OWN(ctx->c_mtotal, 16 * MTOTAL_SLOTS);
  OWN(ctx->c_voff, N);
  if (!ctx->ctr_h) {
    hipError_t e = hipHostMalloc((void**)&ctx->ctr_h, 2 * sizeof(Counters), hipHostMallocDefault);
    if (e != hipSuccess) { ctx->last_hip = e; free_all(ctx); return GALA_E_OOM; }
  }
  if (!ctx->hpin) {
    hipError_t e = hipHostMalloc((void**)&ctx->hpin, HPIN_BYTES, hipHostMallocDefault);
    if (e != hipSuccess) { ctx->last_hip = e; free_all(ctx); return GALA_E_OOM; }
  }
  ctx->cap_n = n;
  ctx->cap_m = m;
  ctx->cap_wide = wide;
  return GALA_OK;
}

int grow_hub(gala_ctx* ctx, int64_t slots, bool wide) {
  if (slots <= ctx->hub_cap) return GALA_OK;
  if (ctx->hub_keys) (void)hipFree(ctx->hub_keys);
  if (ctx->hub_vals) (void)hipFree(ctx->hub_vals);
  if (ctx->hub_rc) (void)hipFree(ctx->hub_rc);
  if (ctx->hub_rw) (void)hipFree(ctx->hub_rw);
  ctx->hub_keys = nullptr; ctx->hub_vals = nullptr; ctx->hub_rc = nullptr; ctx->hub_rw = nullptr;
  int64_t cap = slots + slots / 4;
  hipError_t e = dalloc(&ctx->hub_keys, cap);
  if (e == hipSuccess) {
    char* p;
    e = dalloc(&p, cap * (wide ? 8 : 4));
    ctx->hub_vals = p;
  }
  if (e == hipSuccess) e = dalloc(&ctx->hub_rc, cap);
  if (e == hipSuccess) {
    char* p;
    e = dalloc(&p, cap * (wide ? 8 : 4));
    ctx->hub_rw = p;
  }
  if (e != hipSuccess) { ctx->last_hip = e; ctx->hub_cap = 0; return GALA_E_OOM; }
  // the hub kernels keep overflow slices all-zero between uses
  e = hipMemset(ctx->hub_keys, 0, sizeof(int32_t) * cap);
  if (e == hipSuccess) e = hipMemset(ctx->hub_vals, 0, (size_t)cap * (wide ? 8 : 4));
  if (e != hipSuccess) { ctx->last_hip = e; return GALA_E_HIP; }
  ctx->hub_cap = cap;
  return GALA_OK;
}

int grow_arena(gala_ctx* ctx, int64_t slots, bool wide) {
  if (slots > ctx->arena_slots) {
    if (ctx->arena_keys) (void)hipFree(ctx->arena_keys);
    if (ctx->arena_vals) (void)hipFree(ctx->arena_vals);
    ctx->arena_keys = nullptr; ctx->arena_vals = nullptr;
    int64_t cap = slots + slots / 4;  // slack so later rounds rarely reallocate
    hipError_t e = dalloc(&ctx->arena_keys, cap);
    if (e == hipSuccess) { char* p; e = dalloc(&p, cap * (wide ? 8 : 4)); ctx->arena_vals = p; }
    if (e != hipSuccess) { ctx->last_hip = e; ctx->arena_slots = 0; return GALA_E_OOM; }
    // the row-merge kernels keep arena slices all-zero between uses
    e = hipMemset(ctx->arena_keys, 0, sizeof(int32_t) * cap);
    if (e == hipSuccess) e = hipMemset(ctx->arena_vals, 0, (size_t)cap * (wide ? 8 : 4));
    if (e != hipSuccess) { ctx->last_hip = e; return GALA_E_HIP; }
    ctx->arena_slots = cap;
  }
  return GALA_OK;
}

// Dense-row buffers for a contraction plan: the chunk table and the scratch of
// the segmented sort (kept for later rounds when large enough).
int grow_dense(gala_ctx* ctx, const ContractPlan& plan, bool wide) {
  hipError_t e = hipSuccess;
  const int64_t nch = (int64_t)plan.chunks.size();
  if (nch > ctx->chunks_cap) {
    if (ctx->chunks) (void)hipFree(ctx->chunks);
    ctx->chunks = nullptr;
    e = dalloc(&ctx->chunks, nch);
    if (e != hipSuccess) { ctx->last_hip = e; ctx->chunks_cap = 0; return GALA_E_OOM; }
    ctx->chunks_cap = nch;
  }
  const int64_t bytes = dense_scratch_bytes(nch, plan.ndense, plan.dense_entries, wide);
  if (bytes > ctx->dscratch_bytes) {
    if (ctx->dscratch) (void)hipFree(ctx->dscratch);
    ctx->dscratch = nullptr;
    e = hipMalloc((void**)&ctx->dscratch, (size_t)bytes);
    if (e != hipSuccess) { ctx->last_hip = e; ctx->dscratch_bytes = 0; return GALA_E_OOM; }
    ctx->dscratch_bytes = bytes;
  }
  return GALA_OK;
}

// --------------------------------------------------------------------------
// per-vertex hub-table offsets for the bins that use the global arena
// --------------------------------------------------------------------------
// per hub row: its table size next_pow2(2d); *dmax (zero before) = the longest row
__global__ void k_hub_sizes(const int32_t* __restrict__ order, const int32_t* __restrict__ bstart,
                            int gbin, const int64_t* __restrict__ off, const int64_t* __restrict__ end,
                            int64_t* __restrict__ tsz, int64_t* __restrict__ dmax) {
  const int lo = bstart[gbin], hi = bstart[NBINS];
  int64_t mx = 0;
  for (int i = lo + blockIdx.x * blockDim.x + threadIdx.x; i < hi; i += gridDim.x * blockDim.x) {
    int v = order[i];
    int64_t d = end[v] - off[v];
    int64_t t = 1;
    while (t < 2 * d) t <<= 1;
    tsz[i - lo] = t;
    mx = d > mx ? d : mx;
  }
  if (mx > 0) atomicMax((unsigned long long*)dmax, (unsigned long long)mx);
}
__global__ void k_hub_scatter(const int32_t* __restrict__ order, const int32_t* __restrict__ bstart,
                              int gbin, const int64_t* __restrict__ pos, int64_t* __restrict__ hub_off) {
  const int lo = bstart[gbin], hi = bstart[NBINS];
  for (int i = lo + blockIdx.x * blockDim.x + threadIdx.x; i < hi; i += gridDim.x * blockDim.x)
    hub_off[order[i]] = pos[i - lo];
}
__global__ void k_offsets_from_incl(const int64_t* __restrict__ end_incl, int32_t n, int64_t* __restrict__ off) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += gridDim.x * blockDim.x)
    off[i] = i == 0 ? 0 : end_incl[i - 1];
}
// gathers the round set-up readback into one 80-byte block (see round_setup)
__global__ void k_setup_stage(const int32_t* __restrict__ bstart, const int64_t* __restrict__ slots,
                              const int64_t* __restrict__ dmax, int64_t* __restrict__ stage) {
  const int t = threadIdx.x;
  if (t <= NBINS) ((int32_t*)stage)[t] = bstart[t];
  if (t == 0) stage[8] = *slots;
  if (t == 1) stage[9] = *dmax;
}

__global__ void k_i32_to_i64(const int32_t* __restrict__ a, int64_t* __restrict__ b, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}
// integer-valued float32 weights below 2^31 (checked on the host): exact
__global__ void k_f32_to_i32(const float* __restrict__ a, int32_t* __restrict__ b, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (int32_t)a[i];
}

// Exact integer image of positive finite float32 weights: the smallest shift
// k >= 0 such that w * 2^k is an integer for every weight, the exact sum of the
// scaled weights (< 2^62, else GALA_E_OVERFLOW) and whether every scaled
// weight fits 31 bits.  From the binary32 layout: a normal value is
// (2^23 + mantissa) * 2^(exponent - 150), a subnormal one mantissa * 2^-149;
// with the significand made odd its exponent is that of the lowest set bit.
int f32_exact_scale(const float* w, int64_t m, int32_t* shift, int64_t* m2, bool* fits32) {
  auto odd = [](float x, uint64_t* sig, int32_t* e) {
    uint32_t b;
    std::memcpy(&b, &x, 4);
    const uint32_t ex = (b >> 23) & 0xff, man = b & 0x7fffff;
    *sig = ex ? (man | 0x800000u) : man;
    *e = ex ? (int32_t)ex - 150 : -149;
    const int tz = __builtin_ctzll(*sig);
    *sig >>= tz;
    *e += tz;
  };
  int32_t lo = 0;
  for (int64_t j = 0; j < m; j++) {
    uint64_t sig;
    int32_t e;
    odd(w[j], &sig, &e);
    lo = std::min(lo, e);
  }
  const int32_t k = -lo;
  __int128 sum = 0;
  bool f32 = true;
  for (int64_t j = 0; j < m; j++) {
    uint64_t sig;
    int32_t e;
    odd(w[j], &sig, &e);
    e += k;  // >= 0 by the choice of k
    if (e > 62) return GALA_E_OVERFLOW;
    const __int128 W = (__int128)sig << e;
    sum += W;
    if (W >= WIDE_M2_LIMIT || sum >= WIDE_M2_LIMIT) return GALA_E_OVERFLOW;
    f32 = f32 && W < ((__int128)1 << 31);
  }
  *shift = k;
  *m2 = (int64_t)sum;
  *fits32 = f32;
  return GALA_OK;
}

__global__ void k_iota(int32_t* __restrict__ p, int32_t n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = i;
}
// c->sum_in += sum of w: one atomic per workgroup of 256
__global__ __launch_bounds__(256) void k_sum_w64(const int32_t* __restrict__ w, int64_t m, Counters* c) {
  __shared__ long long s_w[4];
  long long s = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
    s += w[i];
  for (int k = 32; k >= 1; k >>= 1) s += __shfl_xor(s, k, 64);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    s = s_w[0] + s_w[1] + s_w[2] + s_w[3];
    if (s) atomicAdd(&c->sum_in, (unsigned long long)s);
  }
}

struct RoundOut {
  double q;
  int iters;
  double t_dm, t_upd;
  int32_t caps[NBINS];  // degree-bin sizes of the round graph
};

template <class T>
DevGraph<T> graph_at(gala_ctx* ctx, int k, int32_t n, int64_t m) {
  DevGraph<T> g;
  g.n = n;
  g.m = m;
  g.off = ctx->off[k];
  g.end = ctx->gapped[k] ? ctx->end[k] : ctx->off[k] + 1;
  g.nbr = ctx->nbr[k];
  g.w = (k == 0 && (ctx->unit0 || ctx->f32in)) ? nullptr : (const typename T::w_t*)ctx->w[k];
  if (T::WIDE && k == 0 && ctx->f32in) {
    g.wf = (const float*)ctx->w[0];
    g.wsc = std::ldexp(1.0, ctx->wshift);
  }
  return g;
}
inline int dst_of(int src) { return src == 1 ? 2 : 1; }

template <class T>
DevState<T> dstate(gala_ctx* ctx) {
  using tt = typename T::tot_t;
  DevState<T> st;
  st.K = (tt*)ctx->K; st.Self = (tt*)ctx->Self; st.Tot = (tt*)ctx->Tot; st.In = (tt*)ctx->In;
  st.size = ctx->size; st.prev = ctx->prev; st.cur = ctx->cur;
  st.active = ctx->active; st.dirty = ctx->dirty; st.binid = ctx->binid; st.hub_off = ctx->hub_off;
  st.cuts = ctx->cuts;
  st.hub_keys = ctx->hub_keys; st.hub_vals = (typename T::acc_t*)ctx->hub_vals;
  st.hub_rc = ctx->hub_rc; st.hub_rw = (typename T::acc_t*)ctx->hub_rw;
  st.hub_plan = ctx->hub_plan; st.hub_part = ctx->hub_part; st.hub_fail = ctx->hub_fail;
  st.hub_dmax = ctx->hub_dmax;
  st.ctr = ctx->ctr;
  st.track = 0;
  st.iter = 0;
  st.last_gain = ctx->last_gain;
  st.moved_it = ctx->moved_it;
  st.nbr_moved = ctx->nbr_moved;
  st.cv = (typename T::cv_t*)ctx->cv;
  st.part = ctx->part;
  st.cp = ctx->cp;
  st.cm = ctx->cm;
  st.tsd = T::WIDE ? nullptr : ctx->tsd;
  return st;
}

// Q from the exact sums, finished with three FP64 operations
double q_from(const Counters& c, double m2) {
  const unsigned __int128 t2 = ((unsigned __int128)c.sum_tot2_hi << 64) | c.sum_tot2;
  return (double)c.sum_in / m2 - (double)t2 / (m2 * m2);
}

// Bins, order, hub offsets and community init for the round graph.
// caps[b] = vertices of bin b; *nitems = vertices in any bin (degree > 0),
// i.e. the valid prefix of ctx->order.
template <class T>
int round_setup(gala_ctx* ctx, const DevGraph<T>& g, int32_t* caps, int32_t* nitems, int track = 0,
                bool degree_init = false) {
  hipStream_t s = ctx->stream;
  DevState<T> st = dstate<T>(ctx);
  st.track = track;
  launch_init<T>(g, st, s, degree_init);
  partition_bins(0, nullptr, nullptr, ctx->binid, nullptr, g.n, ctx->order, ctx->bstart_full, ctx->part_tmp, s);
  // hub-table sizes and their prefix sums before the one readback: the scan
  // runs over n slots (entries past the hub count are zero)
  const int gb = global_bin_min<T>();
  const int64_t nn = std::max<int32_t>(g.n, 1);
  // hub_tmp[0, nn): table sizes; hub_tmp[nn]: the longest hub row
  fill<int64_t>(ctx->hub_tmp, nn + 1, 0, s);
  k_hub_sizes<<<std::min<int64_t>((nn + 255) / 256, 1024), 256, 0, s>>>(ctx->order, ctx->bstart_full, gb, g.off,
                                                                         g.end, ctx->hub_tmp, ctx->hub_tmp + nn);
  int64_t* pos = ctx->hub_tmp + nn + 1;
  int64_t* tot = ctx->scan64_tmp;
  scan_exclusive<int64_t, int64_t>(ctx->hub_tmp, pos, nn, ctx->c_i64_tmp, tot, s);
  // one transfer: bytes [0, 40) the NBINS + 1 bin starts, [64, 72) the hub-table
  // slots, [72, 80) the longest hub row
  k_setup_stage<<<1, 64, 0, s>>>(ctx->bstart_full, tot, ctx->hub_tmp + nn, ctx->c_i64_scalar);
  HIPOK(hipMemcpyAsync(ctx->hpin, ctx->c_i64_scalar, 80, hipMemcpyDeviceToHost, s));
  HIPOK(hipStreamSynchronize(s));
  const int32_t* hb = (const int32_t*)ctx->hpin;
  const int64_t* hs = (const int64_t*)(ctx->hpin + 64);
  ctx->hub_dmax = hs[1];
  for (int b = 0; b < NBINS; b++) caps[b] = hb[b + 1] - hb[b];
  *nitems = hb[NBINS];
  if (!degree_init) launch_init_rows<T>(g, st, ctx->order, ctx->bstart_full, caps, s);
  const int nh = hb[NBINS] - hb[gb];
  if (nh > 0) {
    int rc = grow_hub(ctx, hs[0], T::WIDE);
    if (rc) return rc;
    k_hub_scatter<<<std::min((nh + 255) / 256, 4096), 256, 0, s>>>(ctx->order, ctx->bstart_full, gb, pos,
                                                                   ctx->hub_off);
  }
  return GALA_OK;
}

// The decide-and-move totals of the counters: the sums of their slots.
void fold_slots(Counters* c) {
  c->moved = c->swept = c->swept_edges = 0;
  for (int b = 0; b < NBINS; b++) c->bin_swept[b] = c->bin_edges[b] = 0;
  for (const DecideSlot& sl : c->slot) {
    c->moved += sl.moved;
    c->swept += sl.swept;
    c->swept_edges += sl.swept_edges;
    for (int b = 0; b < NBINS; b++) {
      c->bin_swept[b] += sl.bin_swept[b];
      c->bin_edges[b] += sl.bin_edges[b];
    }
  }
}

// Reads ctr[0] into out and, when saved is given, the copy kept in ctr[1] as
// well, in one transfer and one synchronisation.
template <class T>
int read_counters(gala_ctx* ctx, Counters* out, Counters* saved = nullptr) {
  const int k = saved ? 2 : 1;
  HIPOK(hipMemcpyAsync(ctx->ctr_h, ctx->ctr, k * sizeof(Counters), hipMemcpyDeviceToHost, ctx->stream));
  HIPOK(hipStreamSynchronize(ctx->stream));
  *out = ctx->ctr_h[0];
  fold_slots(out);
  if (saved) {
    *saved = ctx->ctr_h[1];
    fold_slots(saved);
    if (saved->overflow) return GALA_E_OVERFLOW;
  }
  if (out->overflow) return GALA_E_OVERFLOW;
  return GALA_OK;
}

// Launches one kernel per degree bin.  concurrent: bins CONC_BIN.. (few long
// rows, one LDS-heavy workgroup per vertex, long tails) go on the second
// stream beside the short-row bins, so the two halves share the CUs instead
// of draining one after the other; both halves write disjoint vertices (and
// add exact integer deltas where they meet).  Joined back before returning.
// pre2(stream): independent work queued on the second stream ahead of the
// long-row bins (on the one stream ahead of every bin when not concurrent).
constexpr int CONC_BIN = 6;
struct NoPre {
  void operator()(hipStream_t) const {}
};
template <class F, class P = NoPre>
int over_bins(gala_ctx* ctx, bool concurrent, F&& launch, P&& pre2 = P()) {
  hipStream_t s = ctx->stream;
  if (!concurrent) {
    pre2(s);
    for (int b = 0; b < NBINS; b++) launch(b, s);
    return GALA_OK;
  }
  HIPOK(hipEventRecord(ctx->evf[0], s));
  HIPOK(hipStreamWaitEvent(ctx->stream2, ctx->evf[0], 0));
  pre2(ctx->stream2);
  for (int b = CONC_BIN; b < NBINS; b++) launch(b, ctx->stream2);
  for (int b = 0; b < CONC_BIN; b++) launch(b, s);
  HIPOK(hipEventRecord(ctx->evf[1], ctx->stream2));
  HIPOK(hipStreamWaitEvent(s, ctx->evf[1], 0));
  return GALA_OK;
}

// One round of phase 1 (louvain_main_process).  On return ctx->prev holds the
// accepted assignment of the round graph.
template <class T>
int run_round(gala_ctx* ctx, const DevGraph<T>& g, typename T::tot_t m2, int round, bool unique_rows,
              gala_iter_cb icb, void* user, RoundOut* ro, std::vector<int32_t>& hostvec) {
  using tt = typename T::tot_t;
  hipStream_t s = ctx->stream;
  const double m2d = (double)m2;
  // learned sweeps in the rounds whose graph is large enough for the filter
  // to save more than it costs
  const bool learned = ctx->opts.frontier_mode == 1 && ctx->have_mlp && g.n >= ctx->opts.learned_min_vertices;
  const bool logging = ctx->logger != nullptr;
  const int track = (learned || logging) ? 1 : 0;
  int32_t caps[NBINS], nitems = 0;
  // the input graph with unit weights and no self-loop: K = degree, Self = 0
  const bool degree_init = round == 0 && ctx->unit0 && ctx->noself0;
  int rc = round_setup<T>(ctx, g, caps, &nitems, track, degree_init);
  if (rc) return rc;
  DevState<T> st = dstate<T>(ctx);
  st.track = track;
  // Iteration 0: Q on the singleton partition (louvain_main_process.cu:538-548)
  reset_counters(ctx->ctr, (long long)m2, s);
  launch_apply<T>(st, g.n, m2, s, nullptr, nullptr, nullptr, nullptr, false);  // nobody has moved yet
  if (unique_rows) {
    // singleton partition, at most one self entry per row: In = Self
    HIPOK(hipMemcpyAsync(st.In, st.Self, sizeof(tt) * g.n, hipMemcpyDeviceToDevice, s));
  } else {
    for (int b = 0; b < NBINS; b++)
      launch_weight<T>(b, g, st, ctx->order, ctx->bstart_full, caps[b], false, s, ctx->wvar[b]);
  }
  launch_prune<T>(st, g.n, m2, s);
  Counters c;
  double q = 0, qprev = 0;
  // Q0 is first needed by iteration 1's stopping test, so its counters are
  // parked in ctr[1] and read with iteration 1's, saving a synchronisation;
  // a per-iteration dump needs prev as it is now and reads at once.
  const bool dump0 = icb && ctx->opts.dump_iterations;
  bool q0_pending = !dump0;
  if (q0_pending) {
    HIPOK(hipMemcpyAsync(ctx->ctr + 1, ctx->ctr, sizeof(Counters), hipMemcpyDeviceToDevice, s));
  } else {
    if ((rc = read_counters<T>(ctx, &c))) return rc;
    q = qprev = q_from(c, m2d);
    hostvec.resize(g.n);
    HIPOK(hipMemcpy(hostvec.data(), ctx->prev, sizeof(int32_t) * g.n, hipMemcpyDeviceToHost));
    icb(user, round, 0, q, 0, hostvec.data(), g.n);
  }
  // first frontier: every vertex (fill(active,1), louvain_main_process.cu:574).
  // Two frontier buffers alternate so that the frontier a sweep started from
  // is still at hand when a stalled learned sweep must be certified.
  int32_t* fcur = ctx->front;
  int32_t* bcur = ctx->bstart_front;
  int32_t* fnext = ctx->front3;
  int32_t* bnext = ctx->bstart3;
  copy_i32(ctx->order, fcur, nitems, s);
  copy_i32(ctx->bstart_full, bcur, NBINS + 1, s);
  int32_t nfront = nitems;  // vertices in the plain frontier
  const int max_it = ctx->opts.max_iterations > 0 ? ctx->opts.max_iterations : 100000;
  const float tau = ctx->opts.learned_tau > 0 ? ctx->opts.learned_tau : 0.5f;
  const bool prof = ctx->opts.profile_bins != 0;
  const bool conc = !prof && !ctx->opts.serial_bins;  // long-row bins on the second stream
  const size_t tsz = sizeof(tt);
  double t_dm = 0, t_upd = 0;
  std::vector<float> hfeat;
  std::vector<uint8_t> hlab;
  // decide-and-move over one frontier, timed into t_dm (and per bin)
  // build_cv: refresh the community view first (a certification sweep reads
  // the view of the pre-sweep state, which is still in place)
  // In is updated from the movers' rows alone (no dirty marks)
  const bool incr = true;
  // ev: the events recorded at its start and end (ctx->ev unless given)
  auto decide = [&](DevState<T>& sd0, const int32_t* fr, const int32_t* bs, bool first, bool build_cv,
                    hipEvent_t* ev = nullptr) -> int {
    if (!ev) ev = ctx->ev;
    DevState<T> sd = sd0;
    if (incr) sd.dirty = nullptr;
    HIPOK(hipEventRecord(ev[0], s));
    if (build_cv) launch_build_cv<T>(sd, g.n, s);
    if (prof) {  // per-bin events: one stream
      sd.kev = ctx->kev;
      for (int b = 0; b < NBINS; b++) {
        HIPOK(hipEventRecord(ctx->bev[b], s));
        if (first) launch_decide_first<T>(b, g, sd, fr, bs, caps[b], m2, s, ctx->fvar[b]);
        else launch_decide<T>(b, g, sd, fr, bs, caps[b], m2, s, ctx->dvar[b]);
      }
      HIPOK(hipEventRecord(ctx->bev[NBINS], s));
    } else {
      int r = over_bins(ctx, conc, [&](int b, hipStream_t sb) {
        if (first) launch_decide_first<T>(b, g, sd, fr, bs, caps[b], m2, sb, ctx->fvar[b]);
        else launch_decide<T>(b, g, sd, fr, bs, caps[b], m2, sb, ctx->dvar[b]);
      });
      if (r) return r;
    }
    HIPOK(hipEventRecord(ev[1], s));
    return GALA_OK;
  };
  // the decide time of the sweep (events ev[0] -> ev[1]) and its work, from
  // the sweep's counters
  auto account_decide = [&](bool first, const Counters& cs, hipEvent_t* ev = nullptr) {
    if (!ev) ev = ctx->ev;
    const Counters* out = &cs;
    float a = 0;
    (void)hipEventElapsedTime(&a, ev[0], ev[1]);
    t_dm += a;
    const int hv = ctx->dvar[NBINS - 1];
    if (prof && !first && caps[NBINS - 1] > 0 && hv >= 4 && hv <= 6) {  // the split's part kernel ran
      float kp = 0;
      (void)hipEventElapsedTime(&kp, ctx->kev[0], ctx->kev[1]);
      ctx->res.hub_part_ms += kp;
      ctx->res.hub_part_launches++;
    }
    ctx->res.swept_vertices += (int64_t)out->swept;
    ctx->res.swept_edges += (int64_t)out->swept_edges;
    if (!g.weighted()) {
      ctx->res.swept_vertices_unit += (int64_t)out->swept;
      ctx->res.swept_edges_unit += (int64_t)out->swept_edges;
    }
    // per entry: 4 neighbour id + the weight + 4 prev[u] (not in a first
    // sweep) + the community total (4 bytes, 8 on the wide path); the weight
    // is 4 bytes (int32, or float32 of a wide input round), 8 for the int64
    // sums of a wide contracted round, 0 when implicit
    const int64_t wbytes = !g.weighted() ? 0 : (T::WIDE && g.w) ? 8 : 4;
    const int64_t per_edge = 4 + wbytes + (first ? 0 : 4) + (T::WIDE ? 8 : 4);
    ctx->res.decide_bytes += per_edge * (int64_t)out->swept_edges + 40 * (int64_t)out->swept;
    for (int b = 0; b < NBINS; b++) {
      const int64_t by = per_edge * (int64_t)out->bin_edges[b] + 40 * (int64_t)out->bin_swept[b];
      float ms = 0;
      if (prof) (void)hipEventElapsedTime(&ms, ctx->bev[b], ctx->bev[b + 1]);
      const int launched = caps[b] > 0 ? 1 : 0;  // the decide launchers skip empty bins
      if (first) {
        ctx->res.bin_first_bytes[b] += by;
        ctx->res.bin_first_edges[b] += (int64_t)out->bin_edges[b];
        ctx->res.bin_first_ms[b] += ms;
        ctx->res.bin_first_launches[b] += launched;
      } else {
        ctx->res.bin_decide_bytes[b] += by;
        ctx->res.bin_swept_edges[b] += (int64_t)out->bin_edges[b];
        ctx->res.bin_decide_ms[b] += ms;
        ctx->res.bin_launches[b] += launched;
      }
    }
  };
  // Tot/size update, In of the invalidated rows, prune, next frontier and the
  // counters; accumulates the sweep's work and timings
  // fr / bs: the vertices the sweep decided (only they can have moved)
  // the update's launches: Tot/size from the movers (on the second stream
  // when concurrent) beside In: In changes only where a vertex or one of its
  // neighbours moved, so only the movers' rows are read.  (Recomputing every
  // In after the first sweep of a round, when most vertices move, measured
  // slower.)  Then the prune and the next frontier (fnext, bnext).
  auto update_launch = [&](DevState<T>& sp, const int32_t* fr, const int32_t* bs) -> int {
    if (incr) {
      int r = over_bins(
          ctx, conc,
          [&](int b, hipStream_t sb) { launch_weight_moved<T>(b, g, st, fr, bs, caps[b], sb, ctx->wvar[b]); },
          [&](hipStream_t s2) { launch_apply<T>(sp, g.n, m2, s2, nullptr, nullptr, fr, bs + NBINS); });
      if (r) return r;
    } else {
      launch_apply<T>(sp, g.n, m2, s, nullptr, nullptr, fr, bs + NBINS);
    }
    if (!incr)
      for (int b = 0; b < NBINS; b++)
        launch_weight<T>(b, g, st, ctx->order, ctx->bstart_full, caps[b], true, s, ctx->wvar[b]);
    launch_prune<T>(sp, g.n, m2, s);
    partition_bins(1, ctx->order, ctx->bstart_full, ctx->binid, ctx->active, nitems, fnext, bnext, ctx->part_tmp,
                   s);
    return GALA_OK;
  };
  auto update = [&](DevState<T>& sp, bool first, const int32_t* fr, const int32_t* bs, Counters* out) -> int {
    int r = update_launch(sp, fr, bs);
    if (r) return r;
    HIPOK(hipEventRecord(ctx->ev[2], s));
    Counters c0;
    r = read_counters<T>(ctx, out, q0_pending ? &c0 : nullptr);
    if (r) return r;
    if (q0_pending) {
      q0_pending = false;
      qprev = q_from(c0, m2d);
      if (icb) icb(user, round, 0, qprev, 0, nullptr, g.n);
    }
    float b2 = 0;
    (void)hipEventElapsedTime(&b2, ctx->ev[1], ctx->ev[2]);
    t_upd += b2;
    account_decide(first, *out);
    std::swap(fcur, fnext);
    std::swap(bcur, bnext);
    return GALA_OK;
  };
  // opts.queue_sweeps: plain rounds queue sweep k+1 before sweep k's counters
  // are read, so the GPU does not wait on the host between sweeps.  The
  // stopping rule is then applied on the device: a round_step after each
  // update accepts the sweep (prev = cur at the start of the next one) or
  // ends the round and empties the next frontier, so the one sweep queued
  // behind the last finds no vertex.  The host applies the same rule to the
  // same counters one sweep late and checks that both agree.  (On the 4M
  // power-law graph this is 1-2% slower than the loop after this one, which
  // reads every sweep before the next: the readback gaps it hides are shorter
  // than the launches of the one empty sweep it leaves per round.)
  const bool queued = ctx->opts.queue_sweeps && !learned && !logging && !dump0 && !prof;
  if (queued) {
    ctx->res.queued_rounds++;
    const double min_mod = ctx->opts.min_modularity;
    // the result of a failed step surfaces after the queued work has drained
    auto fail = [&](int r) -> int {
      (void)hipStreamSynchronize(s);
      return r;
    };
    round_start(ctx->ctr, ctx->rctl, m2d, s);  // qprev = Q0 (iteration-0 counters)
    int it = 1;
    st.iter = 1;
    reset_counters(ctx->ctr, (long long)m2, s);
    if ((rc = decide(st, fcur, bcur, unique_rows, !unique_rows, ctx->sev[1]))) return fail(rc);
    for (;; it++) {
      hipEvent_t* ev = ctx->sev[it & 1];
      const bool first = it == 1 && unique_rows;
      if ((rc = update_launch(st, fcur, bcur))) return fail(rc);
      round_step(ctx->ctr, ctx->rctl, m2d, min_mod, it, max_it, bnext, s);
      // sweep it's counters (with the iteration-0 copy after sweep 1)
      HIPOK(hipMemcpyAsync(ctx->ctr_h, ctx->ctr, (it == 1 ? 2 : 1) * sizeof(Counters), hipMemcpyDeviceToHost, s));
      HIPOK(hipEventRecord(ev[2], s));
      std::swap(fcur, fnext);
      std::swap(bcur, bnext);
      const bool more = it < max_it;
      if (more) {  // sweep it + 1
        // the swept frontier (now in fnext / bnext) holds every vertex that moved
        accept_list(ctx->cur, ctx->prev, st.cp, st.cm, fnext, bnext + NBINS, nitems, ctx->rctl, s);
        st.iter = it + 1;
        reset_counters(ctx->ctr, (long long)m2, s);
        if ((rc = decide(st, fcur, bcur, false, true, ctx->sev[(it + 1) & 1]))) return fail(rc);
      }
      HIPOK(hipEventSynchronize(ev[2]));
      c = ctx->ctr_h[0];
      fold_slots(&c);
      if (it == 1) {
        Counters c0 = ctx->ctr_h[1];
        fold_slots(&c0);
        if (c0.overflow) return fail(GALA_E_OVERFLOW);
        qprev = q_from(c0, m2d);
        if (icb) icb(user, round, 0, qprev, 0, nullptr, g.n);
      }
      if (c.overflow) return fail(GALA_E_OVERFLOW);
      float b2 = 0;
      (void)hipEventElapsedTime(&b2, ev[1], ev[2]);
      t_upd += b2;
      account_decide(first, c, ev);
      q = q_from(c, m2d);
      const bool accept = !(q - qprev < min_mod) && it < max_it;
      if (c.step_accept != (accept ? 1 : 0) || !(c.step_q == q)) return fail(GALA_E_INTERNAL);
      if (icb) icb(user, round, it, q, (int64_t)c.moved, nullptr, g.n);
      if (!accept) {
        if (more) ctx->res.inert_sweeps++;
        break;
      }
      qprev = q;
    }
    ro->q = qprev;
    ro->iters = it;
    for (int b = 0; b < NBINS; b++) ro->caps[b] = caps[b];
    ro->t_dm = t_dm;
    ro->t_upd = t_upd;
    return GALA_OK;
  }
  int it = 1;
  for (;; it++) {
    st.iter = it;
    reset_counters(ctx->ctr, (long long)m2, s);
    // A learned sweep visits only the predicted movers P of the plain frontier
    // F.  The first sweep of a round is plain.
    const bool lsweep = learned && it >= 2;
    const bool first = it == 1 && unique_rows;  // singleton partition: no aggregation needed
    const int32_t nf0 = nfront;
    if (lsweep) {
      launch_mlp_filter<T>(g, st, fcur, nfront, it, ctx->mlp_prm, ctx->mlp_w1b, tau, ctx->opts.learned_min_degree,
                           ctx->keep, s);
      partition_bins(1, fcur, bcur, ctx->binid, ctx->keep, nfront, ctx->front2, ctx->bstart2, ctx->part_tmp, s);
      // pre-sweep Tot and size, read again if the sweep has to be certified
      HIPOK(hipMemcpyAsync(ctx->tot_snap, st.Tot, tsz * g.n, hipMemcpyDeviceToDevice, s));
      HIPOK(hipMemcpyAsync(ctx->size_snap, st.size, sizeof(int32_t) * g.n, hipMemcpyDeviceToDevice, s));
    }
    const bool log_it = logging && !lsweep && it >= 2;
    if (log_it) launch_features<T>(g, st, fcur, nfront, it, ctx->feat, s);
    // the first-sweep kernels read Tot directly: no community view needed
    if ((rc = decide(st, lsweep ? ctx->front2 : fcur, lsweep ? ctx->bstart2 : bcur, first, !first))) return rc;
    if (log_it) {
      launch_labels(fcur, nfront, ctx->moved_it, it, ctx->lab, s);
      hfeat.resize((size_t)nfront * MLP_IN);
      hlab.resize(nfront);
      HIPOK(hipMemcpyAsync(hfeat.data(), ctx->feat, sizeof(float) * hfeat.size(), hipMemcpyDeviceToHost, s));
      HIPOK(hipMemcpyAsync(hlab.data(), ctx->lab, hlab.size(), hipMemcpyDeviceToHost, s));
      HIPOK(hipStreamSynchronize(s));
      ctx->logger(ctx->logger_user, round, it, nfront, hfeat.data(), hlab.data());
    }
    if ((rc = update(st, first, lsweep ? ctx->front2 : fcur, lsweep ? ctx->bstart2 : bcur, &c))) return rc;
    nfront = (int32_t)c.active;
    int64_t moved = (int64_t)c.moved;
    q = q_from(c, m2d);
    if (lsweep) {
      ctx->res.learned_sweeps++;
      ctx->res.learned_frontier += nf0;
      ctx->res.learned_kept += (int64_t)c.swept;
    }
    if (lsweep && q - qprev < ctx->opts.min_modularity) {
      // Certification.  The sweep is synchronous, so the plain sweep from the
      // same state is the learned one plus the decisions of F \ P, taken
      // against the pre-sweep Tot and size (prev is still the pre-sweep
      // assignment).  Their moves land on the updated state; rows none of
      // them touches keep the In of the learned sweep, which is exact.
      ctx->res.certify_sweeps++;
      reset_counters(ctx->ctr, (long long)m2, s);
      settle_pairs(ctx->cp, ctx->cm, g.n, s);  // the learned sweep's moves are in Tot and size already
      partition_bins(2, fnext, bnext, ctx->binid, ctx->keep, nf0, ctx->front2, ctx->bstart2, ctx->part_tmp, s);
      DevState<T> sc = st;
      sc.Tot = (tt*)ctx->tot_snap;
      sc.size = ctx->size_snap;
      if ((rc = decide(sc, ctx->front2, ctx->bstart2, false, false))) return rc;
      Counters cc;
      if ((rc = read_counters<T>(ctx, &cc))) return rc;
      if (cc.moved == 0) {
        // the rest of the frontier moved nobody: the plain sweep from this
        // state is the learned one, whose Q already stalled (its update stands;
        // the round ends on it, reverted)
        ctx->res.certify_empty++;
        account_decide(false, cc);
      } else {
        if ((rc = update(st, false, ctx->front2, ctx->bstart2, &c))) return rc;
        nfront = (int32_t)c.active;
        moved += (int64_t)c.moved;
        q = q_from(c, m2d);
      }
    }
    if (icb) {
      const int32_t* vec = nullptr;
      if (ctx->opts.dump_iterations) {
        hostvec.resize(g.n);
        HIPOK(hipMemcpy(hostvec.data(), ctx->cur, sizeof(int32_t) * g.n, hipMemcpyDeviceToHost));
        vec = hostvec.data();
      }
      icb(user, round, it, q, moved, vec, g.n);
    }
    // plain rule, with revert (louvain_main_process.cu:1004-1007,1026-1031)
    if (q - qprev < ctx->opts.min_modularity || it >= max_it) break;
    // prev = cur (louvain_main_process.cu:1029); after a plain sweep only its
    // frontier (now in fnext / bnext) can hold movers: a short frontier is
    // visited through its list, a long one (its gathers and scattered writes
    // cost more than streaming every vertex) by the pass over all n
    if (lsweep || 4 * (int64_t)c.swept > (int64_t)g.n) accept_sweep(ctx->cur, ctx->prev, st.cp, st.cm, g.n, s);
    else accept_list(ctx->cur, ctx->prev, st.cp, st.cm, fnext, bnext + NBINS, nitems, nullptr, s);
    qprev = q;
  }
  ro->q = qprev;
  ro->iters = it;
  for (int b = 0; b < NBINS; b++) ro->caps[b] = caps[b];
  ro->t_dm = t_dm;
  ro->t_upd = t_upd;
  (void)sizeof(tt);
  return GALA_OK;
}

template <class T>
ContractBufs<T> cbufs(gala_ctx* ctx) {
  ContractBufs<T> cb;
  cb.cnt = ctx->c_cnt; cb.nid = ctx->c_nid; cb.label = ctx->c_label; cb.ecnt = ctx->c_ecnt;
  cb.eoff = ctx->c_eoff;
  cb.ecur = ctx->c_ecur; cb.skey = ctx->c_skey; cb.sval = (typename T::acc_t*)ctx->c_sval;
  cb.dcnt = ctx->c_dcnt; cb.noff = nullptr; cb.big_list = ctx->c_big_list; cb.nbig = ctx->c_nbig;
  cb.big_tab = ctx->c_big_tab; cb.arena_keys = ctx->arena_keys;
  cb.arena_vals = (typename T::acc_t*)ctx->arena_vals; cb.arena_slots = ctx->arena_slots;
  cb.order = ctx->order;  // degree-bin order of the round graph (round_setup)
  cb.bstart = ctx->bstart_full;
  cb.chunks = ctx->chunks;
  cb.dscratch = ctx->dscratch;
  cb.i32_tmp = ctx->c_i32_tmp; cb.i64_tmp = ctx->c_i64_tmp; cb.i32_scalar = ctx->c_i32_scalar;
  cb.tier = ctx->c_tier; cb.tlist = ctx->c_tlist; cb.tstart = ctx->c_tstart; cb.part_tmp = ctx->part_tmp;
  cb.i64_scalar = ctx->c_i64_scalar;
  cb.hpin = ctx->hpin;
  cb.hpin_bytes = HPIN_BYTES - 8 * 16 * MTOTAL_SLOTS;  // the tail: the round's entry-count slots
  cb.mtotal = ctx->c_mtotal;
  cb.voff = ctx->c_voff;
  return cb;
}

// Contract the round graph under comm into buffer `dst`.  Returns C and new m;
// ctx->c_label then holds new_id(comm[v]) for the composition.
// canonical: rows in ascending neighbour order (when the CSR leaves the
// engine); otherwise rows of up to the LDS table size keep hash order.
template <class T>
int contract(gala_ctx* ctx, const DevGraph<T>& g, const int32_t* comm, int dst, const int32_t* caps,
             int32_t* C_out, int64_t* m_out, bool canonical) {
  hipStream_t s = ctx->stream;
  ContractBufs<T> cb = cbufs<T>(ctx);
  // non-canonical: stage straight into the output graph's arrays and keep
  // the merged rows there (gapped CSR); canonical: stage into the scratch
  // arrays and compact into a sorted, gap-free CSR
  auto gap_bufs = [&](ContractBufs<T>& b) {
    if (canonical) return;
    b.skey = ctx->nbr[dst];
    b.sval = (typename T::acc_t*)ctx->w[dst];
    b.eoff = ctx->off[dst];
  };
  gap_bufs(cb);
  ContractPlan plan;
  plan.canonical = canonical;
  for (int b = 0; b < NBINS; b++) plan.caps[b] = caps[b];
  if (contract_prepare<T>(g, comm, cb, &plan, s)) { ctx->last_hip = hipGetLastError(); return GALA_E_HIP; }
  int rc = grow_arena(ctx, plan.slots, T::WIDE);
  if (rc) return rc;
  if (plan.ndense > 0 && (rc = grow_dense(ctx, plan, T::WIDE))) return rc;
  ContractOut<T> out;
  out.off = ctx->off[dst];
  out.nbr = ctx->nbr[dst];
  out.w = (typename T::w_t*)ctx->w[dst];
  int64_t m = 0;
  cb = cbufs<T>(ctx);
  gap_bufs(cb);
  if (!canonical) out.end = ctx->end[dst];
  if (contract_build<T>(g, cb, plan, out, &m, s)) { ctx->last_hip = hipGetLastError(); return GALA_E_HIP; }
  ctx->gapped[dst] = !canonical;
  *C_out = plan.C;
  *m_out = canonical ? m : g.m;  // gapped: the extent of the staged arrays
  return GALA_OK;
}

template <class T>
int run_all(gala_ctx* ctx, gala_iter_cb icb, gala_round_cb rcb, void* user, gala_result* out) {
  using tt = typename T::tot_t;
  hipStream_t s = ctx->stream;
  ctx->have_result = false;  // fin is overwritten
  ctx->have_batch_result = false;
  gala_result& R = ctx->res;
  std::memset(&R, 0, sizeof R);
  R.n_vertices = ctx->n0;
  R.input_edges = ctx->m0 / 2;
  R.t_h2d_ms = ctx->t_h2d_ms;
  R.wide = T::WIDE ? 1 : 0;
  R.weight_shift = ctx->wshift;
  // m2 = sum of weights (louvain.cu:165-167): summed exactly on upload for the
  // wide path, on the device for int32 weights
  reset_counters(ctx->ctr, 0, s);
  DevGraph<T> g = graph_at<T>(ctx, 0, ctx->n0, ctx->m0);
  int src = 0;
  int64_t m2i = 0;
  if (T::WIDE) {
    m2i = ctx->m2_wide;
  } else if (g.w) {
    int nb = (int)std::min<int64_t>((g.m + 255) / 256, 4096);
    k_sum_w64<<<nb, 256, 0, s>>>((const int32_t*)g.w, g.m, ctx->ctr);
    Counters c;
    int rc = read_counters<T>(ctx, &c);
    if (rc) return rc;
    m2i = (int64_t)c.sum_in;
  } else {
    m2i = g.m;
  }
  if (m2i <= 0) return GALA_E_INVALID;
  if (m2i > (T::WIDE ? WIDE_M2_LIMIT - 1 : M2_LIMIT)) return GALA_E_OVERFLOW;
  const tt m2 = (tt)m2i;

  HIPOK(hipEventRecord(ctx->ev[4], s));
  k_iota<<<std::min((ctx->n0 + 255) / 256, 8192), 256, 0, s>>>(ctx->fin, ctx->n0);
  int64_t m_in = ctx->m0;  // directed entries of the round graph
  double prev_mod = -1, cur_mod = -1;
  int round = 0;
  std::vector<int32_t> hostvec;
  const int max_rounds = ctx->opts.max_rounds > 0 ? std::min(ctx->opts.max_rounds, 64) : 64;
  for (;;) {
    if (round > 0) prev_mod = cur_mod;
    HIPOK(hipEventRecord(ctx->ev[5], s));
    RoundOut ro;
    const bool unique_rows = src != 0 || ctx->unique0;  // contracted rows are duplicate-free
    int rc = run_round<T>(ctx, g, m2, round, unique_rows, icb, user, &ro, hostvec);
    if (rc) return rc;
    cur_mod = ro.q;
    HIPOK(hipEventRecord(ctx->ev[6], s));
    int32_t C = 0;
    int64_t m = 0;
    const int dst = dst_of(src);
    const bool canonical = ctx->opts.dump_rounds != 0;
    // A round whose first sweep was refused keeps the singleton partition:
    // its contraction is the identity on duplicate-free rows (C = n, the
    // same rows under the same ids) and composes nothing, so it is skipped
    // and the round graph stays.  (Q then equals the last round's: the run
    // ends here unless min_modularity <= 0.)
    const bool identity = ro.iters == 1 && unique_rows && !rcb;
    int64_t m_out = m_in;
    if (identity) {
      C = g.n;
      HIPOK(hipEventRecord(ctx->ev[7], s));
      HIPOK(hipStreamSynchronize(s));
    } else {
      rc = contract<T>(ctx, g, ctx->prev, dst, ro.caps, &C, &m, canonical);
      if (rc) return rc;
      compose_map(ctx->fin, ctx->n0, ctx->c_label, s);
      HIPOK(hipEventRecord(ctx->ev[7], s));
      // entries of the contracted graph: the sum of the slots of c_mtotal
      int64_t* m_rd = (int64_t*)(ctx->hpin + HPIN_BYTES - 8 * 16 * MTOTAL_SLOTS);
      if (!canonical)
        HIPOK(hipMemcpyAsync(m_rd, ctx->c_mtotal, 8 * 16 * MTOTAL_SLOTS, hipMemcpyDeviceToHost, s));
      HIPOK(hipStreamSynchronize(s));
      int64_t m_slots = 0;
      if (!canonical)
        for (int k = 0; k < MTOTAL_SLOTS; k++) m_slots += m_rd[16 * k];
      m_out = canonical ? m : m_slots;
    }
    if (round < 64) {
      R.round_vertices[round] = g.n;
      R.round_entries[round] = m_in;
      R.round_contracted_entries[round] = m_out;
    }
    m_in = m_out;
    float t_round = 0, t_build = 0;
    (void)hipEventElapsedTime(&t_round, ctx->ev[5], ctx->ev[6]);
    (void)hipEventElapsedTime(&t_build, ctx->ev[6], ctx->ev[7]);
    R.t_decide_ms += ro.t_dm;
    R.t_update_ms += ro.t_upd;
    R.t_contract_ms += t_build;
    R.total_iterations += ro.iters;
    if (round < 64) {
      R.round_q[round] = cur_mod;
      R.round_nc[round] = C;
      R.round_iters[round] = ro.iters;
      R.round_ms[round] = t_round;
    }
    if (round < 64) {
      R.round_decide_ms[round] = ro.t_dm;
      R.round_update_ms[round] = ro.t_upd;
      R.round_build_ms[round] = t_build;
    }
    R.rounds = round + 1;
    R.n_communities = C;
    if (rcb) {
      std::vector<int32_t> cm(g.n);
      HIPOK(hipMemcpy(cm.data(), ctx->prev, sizeof(int32_t) * g.n, hipMemcpyDeviceToHost));
      if (ctx->opts.dump_rounds) {
        std::vector<int64_t> off(C + 1), endi(C);
        std::vector<int32_t> nb(m);
        std::vector<int64_t> wi;
        std::vector<double> wf;
        HIPOK(hipMemcpy(off.data(), ctx->off[dst], sizeof(int64_t) * (C + 1), hipMemcpyDeviceToHost));
        HIPOK(hipMemcpy(nb.data(), ctx->nbr[dst], sizeof(int32_t) * m, hipMemcpyDeviceToHost));
        for (int32_t i = 0; i < C; i++) endi[i] = off[i + 1];
        if (T::WIDE) {
          // the exact integer sums, and their value on the input's scale
          wi.resize(m);
          HIPOK(hipMemcpy(wi.data(), ctx->w[dst], sizeof(int64_t) * m, hipMemcpyDeviceToHost));
          wf.resize(m);
          for (int64_t j = 0; j < m; j++) wf[j] = std::ldexp((double)wi[j], -ctx->wshift);
        } else {
          std::vector<int32_t> w32(m);
          HIPOK(hipMemcpy(w32.data(), ctx->w[dst], sizeof(int32_t) * m, hipMemcpyDeviceToHost));
          wi.assign(w32.begin(), w32.end());
        }
        rcb(user, round, cur_mod, g.n, cm.data(), C, endi.data(), nb.data(), wi.data(),
            T::WIDE ? wf.data() : nullptr);
      } else {
        rcb(user, round, cur_mod, g.n, cm.data(), C, nullptr, nullptr, nullptr, nullptr);
      }
    }
    // next round's graph = contracted graph
    if (!identity) {
      src = dst;
      g = graph_at<T>(ctx, src, C, m);
    }
    round++;
    R.rounds = round;
    R.n_communities = C;
    if (!(cur_mod - prev_mod >= ctx->opts.min_modularity)) break;
    if (round >= max_rounds) break;
  }
  HIPOK(hipEventRecord(ctx->ev[3], s));
  HIPOK(hipStreamSynchronize(s));
  float tt_ms = 0;
  (void)hipEventElapsedTime(&tt_ms, ctx->ev[4], ctx->ev[3]);
  R.t_total_ms = tt_ms;
  R.modularity = cur_mod;
  ctx->have_result = true;  // the composed partition stays in fin until gala_get_communities
  if (out) *out = R;
  return GALA_OK;
}

// Upload a graph given as inclusive ends into buffer A (+ weights).
int upload(gala_ctx* ctx, int32_t n, const int64_t* end_incl, const int32_t* nbr, const void* w,
           int weight_mode) {
  if (n < 0 || (n > 0 && (!end_incl || !nbr))) return GALA_E_INVALID;
  int64_t m = n ? end_incl[n - 1] : 0;
  int64_t prev = 0;
  for (int32_t i = 0; i < n; i++) {
    if (end_incl[i] < prev) return GALA_E_FORMAT;
    prev = end_incl[i];
  }
  for (int64_t j = 0; j < m; j++)
    if (nbr[j] < 0 || nbr[j] >= n) return GALA_E_FORMAT;
  bool unique = true;  // strictly ascending rows (what preprocess and the generators write)
  bool noself = true;  // no row holds its own vertex
  for (int32_t i = 0; i < n; i++) {
    const int64_t a = i == 0 ? 0 : end_incl[i - 1];
    for (int64_t j = a; j < end_incl[i]; j++) {
      if (j > a && nbr[j] <= nbr[j - 1]) unique = false;
      if (nbr[j] == i) noself = false;
    }
  }
  if (weight_mode != GALA_W_NONE && m > 0 && !w) return GALA_E_INVALID;
  if (weight_mode == GALA_W_I32)
    for (int64_t j = 0; j < m; j++)
      if (((const int32_t*)w)[j] <= 0) return GALA_E_FORMAT;  // Louvain needs positive weights
  if (weight_mode == GALA_W_F32)
    for (int64_t j = 0; j < m; j++)
      if (!(((const float*)w)[j] > 0) || !std::isfinite(((const float*)w)[j])) return GALA_E_FORMAT;
  // The path: int32 sums while m2 <= 1e9 and every weight is a 31-bit
  // integer, else the wide path on the exact integer image of the weights.
  bool wide = false, f32_as_i32 = false;
  int32_t shift = 0;
  int64_t m2x = m;
  if (weight_mode == GALA_W_I32) {
    __int128 sum = 0;
    for (int64_t j = 0; j < m; j++) sum += ((const int32_t*)w)[j];
    if (sum >= WIDE_M2_LIMIT) return GALA_E_OVERFLOW;
    m2x = (int64_t)sum;
    wide = m2x > M2_LIMIT;
  } else if (weight_mode == GALA_W_F32) {
    bool fits32 = false;
    const int rs = f32_exact_scale((const float*)w, m, &shift, &m2x, &fits32);
    if (rs) return rs;
    f32_as_i32 = shift == 0 && fits32 && m2x <= M2_LIMIT;
    wide = !f32_as_i32;
  } else {
    wide = m > M2_LIMIT;
  }
  ctx->wmode = weight_mode;
  int rc = alloc_buffers(ctx, n, m, wide);
  if (rc) return rc;
  ctx->wide = wide;
  ctx->wshift = shift;
  ctx->m2_wide = m2x;
  hipStream_t s = ctx->stream;
  HIPOK(hipEventRecord(ctx->ev[0], s));
  int64_t* dend = ctx->off[1];  // staging of the inclusive ends
  if (n > 0) HIPOK(hipMemcpyAsync(dend, end_incl, sizeof(int64_t) * n, hipMemcpyHostToDevice, s));
  k_offsets_from_incl<<<std::min((n + 256) / 256 + 1, 8192), 256, 0, s>>>(dend, n, ctx->off[0]);
  if (m > 0) HIPOK(hipMemcpyAsync(ctx->nbr[0], nbr, sizeof(int32_t) * m, hipMemcpyHostToDevice, s));
  ctx->unit0 = weight_mode == GALA_W_NONE;
  ctx->f32in = weight_mode == GALA_W_F32 && wide;
  const int nbw = (int)std::min<int64_t>((m + 255) / 256, 8192);
  if (weight_mode == GALA_W_I32 && m > 0) {
    if (!wide) {
      HIPOK(hipMemcpyAsync(ctx->w[0], w, sizeof(int32_t) * m, hipMemcpyHostToDevice, s));
    } else {  // widened to the int64 sums of the wide path on the device
      int32_t* tmp = (int32_t*)ctx->c_sval;  // staging (8 bytes per entry available)
      HIPOK(hipMemcpyAsync(tmp, w, sizeof(int32_t) * m, hipMemcpyHostToDevice, s));
      k_i32_to_i64<<<nbw, 256, 0, s>>>(tmp, (int64_t*)ctx->w[0], m);
    }
  }
  if (weight_mode == GALA_W_F32 && m > 0) {
    if (wide) {  // kept as float32 in HBM; the kernels scale each weight by 2^shift on load
      HIPOK(hipMemcpyAsync(ctx->w[0], w, sizeof(float) * m, hipMemcpyHostToDevice, s));
    } else {  // integer-valued floats: the int32 path
      float* tmp = (float*)ctx->c_sval;
      HIPOK(hipMemcpyAsync(tmp, w, sizeof(float) * m, hipMemcpyHostToDevice, s));
      k_f32_to_i32<<<nbw, 256, 0, s>>>(tmp, (int32_t*)ctx->w[0], m);
    }
  }
  HIPOK(hipEventRecord(ctx->ev[1], s));
  HIPOK(hipEventSynchronize(ctx->ev[1]));
  float ms = 0;
  (void)hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
  ctx->t_h2d_ms = ms;
  ctx->n0 = n;
  ctx->m0 = m;
  ctx->unique0 = unique;
  ctx->noself0 = noself;
  ctx->have_graph = true;
  ctx->have_result = false;
  return GALA_OK;
}

// ---------------------------------------------------------------------------
// batched graphs
// ---------------------------------------------------------------------------
template <class P>
int bown(gala_ctx* ctx, P** p, int64_t elems) {
  hipError_t e = dalloc(p, std::max<int64_t>(elems, 1));
  if (e != hipSuccess) { ctx->last_hip = e; return GALA_E_OOM; }
  ctx->bowned.push_back((void*)*p);
  return GALA_OK;
}
#define BOWN(p, n)                                  \
  do {                                              \
    int rc_ = bown(ctx, &(p), (n));                 \
    if (rc_) { free_batch(ctx); return rc_; }       \
  } while (0)

BatchDev bdev(gala_ctx* ctx) {
  BatchDev b;
  b.G = ctx->b_G;
  b.m2g = ctx->b_m2g;
  b.live = ctx->b_live; b.accept = ctx->b_accept; b.done = ctx->b_done;
  b.qprev = ctx->b_qprev; b.cur_mod = ctx->b_cur_mod;
  b.iters = ctx->b_iters; b.tot_iters = ctx->b_tot_iters; b.rounds = ctx->b_rounds;
  b.g_in = ctx->b_gin; b.g_tot2 = ctx->b_gtot2;
  b.nlive = ctx->b_scalars;
  return b;
}

// Union upload plus per-graph m2 and the graph of every vertex.
int batch_upload(gala_ctx* ctx, int32_t G, const int32_t* vstart, int32_t n, const int64_t* end_incl,
                 const int32_t* nbr, const void* w, int weight_mode) {
  if (G <= 0 || !vstart || weight_mode == GALA_W_F32 || weight_mode < 0 || weight_mode > 1) return GALA_E_INVALID;
  if (vstart[0] != 0 || vstart[G] != n) return GALA_E_INVALID;
  for (int32_t g = 0; g < G; g++)
    if (vstart[g + 1] < vstart[g]) return GALA_E_INVALID;
  if (n <= 0 || !end_incl || !nbr) return GALA_E_INVALID;
  if (weight_mode == GALA_W_I32 && !w) return GALA_E_INVALID;
  // the CSR checks upload() makes, before any row is walked here
  for (int32_t v = 0; v < n; v++)
    if (end_incl[v] < (v == 0 ? 0 : end_incl[v - 1])) return GALA_E_FORMAT;
  if (weight_mode == GALA_W_I32)
    for (int64_t j = 0; j < end_incl[n - 1]; j++)
      if (((const int32_t*)w)[j] <= 0) return GALA_E_FORMAT;
  // rows stay inside their graph; per-graph m2 (exact, integer)
  std::vector<int32_t> m2g(G), gid0(n);
  int32_t m2max = 0;
  for (int32_t g = 0; g < G; g++) {
    int64_t m2 = 0;
    for (int32_t v = vstart[g]; v < vstart[g + 1]; v++) {
      gid0[v] = g;
      const int64_t a = v == 0 ? 0 : end_incl[v - 1];
      for (int64_t j = a; j < end_incl[v]; j++) {
        if (nbr[j] < vstart[g] || nbr[j] >= vstart[g + 1]) return GALA_E_FORMAT;
        m2 += weight_mode == GALA_W_I32 ? ((const int32_t*)w)[j] : 1;
      }
    }
    if (m2 > M2_LIMIT) return GALA_E_OVERFLOW;
    m2g[g] = (int32_t)m2;
    m2max = std::max(m2max, (int32_t)m2);
  }
  int rc = upload(ctx, n, end_incl, nbr, w, weight_mode);
  if (rc) return rc;
  const int64_t N = n;
  for (int k = 0; k < 3; k++) BOWN(ctx->b_gid[k], N);
  BOWN(ctx->b_m2g, G); BOWN(ctx->b_live, G); BOWN(ctx->b_accept, G); BOWN(ctx->b_done, G);
  BOWN(ctx->b_qprev, G); BOWN(ctx->b_cur_mod, G); BOWN(ctx->b_iters, G); BOWN(ctx->b_tot_iters, G);
  BOWN(ctx->b_rounds, G); BOWN(ctx->b_ncomm, G); BOWN(ctx->b_scalars, 4); BOWN(ctx->b_gin, G);
  BOWN(ctx->b_gtot2, G);
  for (int k = 0; k < 2; k++) { BOWN(ctx->b_vl[k], N); BOWN(ctx->b_lbs[k], NBINS + 1); }
  BOWN(ctx->b_vlive, N);
  if (!ctx->b_host) {
    hipError_t e = hipHostMalloc((void**)&ctx->b_host, 4 * sizeof(int32_t), hipHostMallocDefault);  // [0] [1] scalars, [2] [3] per-iteration live counts
    if (e != hipSuccess) { ctx->last_hip = e; free_batch(ctx); return GALA_E_OOM; }
  }
  hipStream_t s = ctx->stream;
  HIPOK(hipMemcpyAsync(ctx->b_gid[0], gid0.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
  HIPOK(hipMemcpyAsync(ctx->b_m2g, m2g.data(), sizeof(int32_t) * G, hipMemcpyHostToDevice, s));
  HIPOK(hipStreamSynchronize(s));
  ctx->b_G = G;
  ctx->b_vstart.assign(vstart, vstart + G + 1);
  ctx->b_m2max = m2max;
  ctx->have_batch = true;
  ctx->have_batch_result = false;
  return GALA_OK;
}

// All graphs of the batch, each with the semantics of its own run_all: the
// union's sweeps, updates and contractions are shared, every per-graph
// decision (iteration break with revert, round stop) is taken on the device
// from that graph's own exact sums.
int run_batch(gala_ctx* ctx, gala_batch_result* out) {
  using T = IntW;
  hipStream_t s = ctx->stream;
  ctx->have_result = false;  // fin is overwritten
  ctx->have_batch_result = false;
  gala_batch_result R{};
  R.n_graphs = ctx->b_G;
  const int32_t G = ctx->b_G;
  const int32_t m2max = std::max(ctx->b_m2max, 1);
  const double min_mod = ctx->opts.min_modularity;
  const int max_it = ctx->opts.max_iterations > 0 ? ctx->opts.max_iterations : 100000;
  const int max_rounds = ctx->opts.max_rounds > 0 ? std::min(ctx->opts.max_rounds, 64) : 64;
  const bool graphs = ctx->opts.no_hip_graphs == 0;
  if (ctx->m0 == 0) {  // no graph has an edge: every vertex stays alone, Q = 0
    ctx->b_q.assign(G, 0.0);
    ctx->b_rd.assign(G, 0);
    ctx->b_it.assign(G, 0);
    ctx->b_nc.resize(G);
    ctx->b_comm_local.resize(ctx->n0);
    for (int32_t gg = 0; gg < G; gg++) {
      ctx->b_nc[gg] = ctx->b_vstart[gg + 1] - ctx->b_vstart[gg];
      for (int32_t v = ctx->b_vstart[gg]; v < ctx->b_vstart[gg + 1]; v++)
        ctx->b_comm_local[v] = v - ctx->b_vstart[gg];
    }
    ctx->b_comm_ready = true;
    ctx->have_batch_result = true;
    if (out) *out = R;
    return GALA_OK;
  }
  BatchDev b = bdev(ctx);
  int32_t* nlive = ctx->b_scalars;
  int32_t* nopen = ctx->b_scalars + 1;
  int32_t* it_dev = ctx->b_scalars + 2;
  int src = 0;
  DevGraph<T> g = graph_at<T>(ctx, 0, ctx->n0, ctx->m0);
  fill<uint8_t>(ctx->b_done, G, 0, s);
  fill<int32_t>(ctx->b_tot_iters, G, 0, s);
  fill<int32_t>(ctx->b_rounds, G, 0, s);
  fill<double>(ctx->b_cur_mod, G, 0.0, s);
  copy_i32(ctx->b_gid[0], ctx->b_gid[1], ctx->n0, s);
  reset_counters(ctx->ctr, m2max, s);  // the decide totals start at zero
  HIPOK(hipEventRecord(ctx->ev[4], s));
  k_iota<<<std::min((ctx->n0 + 255) / 256, 8192), 256, 0, s>>>(ctx->fin, ctx->n0);
  int gcur = 1;  // b_gid index of the round graph
  int round = 0;
  for (;;) {
    HIPOK(hipEventRecord(ctx->ev[5], s));
    int32_t* gid = ctx->b_gid[gcur];
    int32_t caps[NBINS], nitems = 0;
    int rc = round_setup<T>(ctx, g, caps, &nitems, 0);
    if (rc) return rc;
    DevState<T> st = dstate<T>(ctx);
    st.gid = gid;
    st.m2g = ctx->b_m2g;
    launch_batch_live(b, it_dev, s);
    // iteration 0: the singleton partition's Q of every running graph
    launch_batch_reset(ctx->ctr, m2max, nlive, s);
    const bool unique_rows = src != 0 || ctx->unique0;
    launch_apply<T>(st, g.n, m2max, s);
    if (unique_rows) {
      // singleton partition, at most one self entry per row: In = Self
      HIPOK(hipMemcpyAsync(st.In, st.Self, sizeof(*st.In) * g.n, hipMemcpyDeviceToDevice, s));
    } else {
      for (int k = 0; k < NBINS; k++)
        launch_weight<T>(k, g, st, ctx->order, ctx->bstart_full, caps[k], false, s, ctx->wvar[k]);
    }
    launch_prune<T>(st, g.n, m2max, s);
    launch_graph_sums(gid, ctx->binid, (const int32_t*)ctx->In, (const int32_t*)ctx->Tot, g.n, b.g_in, b.g_tot2, s);
    launch_batch_step(b, it_dev, min_mod, max_it, s);
    // the round's first frontier, and the list of the live graphs' vertices
    int32_t *LA = ctx->b_vl[0], *LB = ctx->b_vl[1], *BA = ctx->b_lbs[0], *BB = ctx->b_lbs[1];
    const int32_t* nA = BA + NBINS;  // length of the live list (device)
    launch_batch_mask(b, gid, ctx->cur, ctx->prev, ctx->active, g.n, true, it_dev, nullptr, nullptr, s, ctx->order,
                      ctx->bstart_full + NBINS, ctx->b_vlive);
    partition_bins(1, ctx->order, ctx->bstart_full, ctx->binid, ctx->b_vlive, nitems, LA, BA, ctx->part_tmp, s);
    partition_bins(1, LA, BA, ctx->binid, ctx->active, nitems, ctx->front, ctx->bstart_front, ctx->part_tmp, s, nA);
    // one iteration for every running graph; a graph that stopped is inert
    // integer path: In from the movers' rows ({cur, prev} pairs, reset by the
    // mask after each sweep), so decide-and-move writes no dirty marks
    constexpr bool incr = true;
    DevState<T> sdec = st;
    if (incr) sdec.dirty = nullptr;
    // Every O(n) pass of a sweep visits only the live list (a graph that
    // stopped is inert), which shrinks as graphs converge; the list is
    // re-compacted after every sweep (fixed buffers, so a captured body
    // replays correctly).
    // The long-row bins' decide and In update run on the second stream
    // beside the others (forked and joined inside a captured body too).
    const bool conc = !ctx->opts.serial_bins;
    auto body = [&](bool first) -> int {
      launch_batch_reset(ctx->ctr, m2max, nlive, s);
      launch_build_cv<T>(st, g.n, s, LA, nA);
      int r = over_bins(ctx, conc, [&](int k, hipStream_t sb) {
        if (first) launch_decide_first<T>(k, g, sdec, ctx->front, ctx->bstart_front, caps[k], m2max, sb, ctx->fvar[k]);
        else launch_decide<T>(k, g, sdec, ctx->front, ctx->bstart_front, caps[k], m2max, sb, ctx->dvar[k]);
      });
      if (r) return r;
      if constexpr (incr) {
        r = over_bins(
            ctx, conc, [&](int k, hipStream_t sb) { launch_weight_moved<T>(k, g, st, LA, BA, caps[k], sb, ctx->wvar[k]); },
            [&](hipStream_t s2) { launch_apply<T>(st, g.n, m2max, s2, LA, nA); });
        if (r) return r;
      } else {
        launch_apply<T>(st, g.n, m2max, s, LA, nA);
        for (int k = 0; k < NBINS; k++) launch_weight<T>(k, g, st, LA, BA, caps[k], true, s, ctx->wvar[k]);
      }
      launch_prune<T>(st, g.n, m2max, s, LA, nA);
      launch_graph_sums(gid, ctx->binid, (const int32_t*)ctx->In, (const int32_t*)ctx->Tot, g.n, b.g_in, b.g_tot2,
                        s, LA, nA);
      launch_batch_step(b, it_dev, min_mod, max_it, s);
      launch_batch_mask(b, gid, ctx->cur, ctx->prev, ctx->active, g.n, false, it_dev, incr ? st.cp : nullptr,
                        incr ? st.cm : nullptr, s, LA, nA, ctx->b_vlive);
      partition_bins(1, LA, BA, ctx->binid, ctx->b_vlive, nitems, LB, BB, ctx->part_tmp, s, nA);
      copy_list(LB, LA, BB + NBINS, nitems, s);
      copy_i32(BB, BA, NBINS + 1, s);
      partition_bins(1, LA, BA, ctx->binid, ctx->active, nitems, ctx->front, ctx->bstart_front, ctx->part_tmp, s,
                     nA);
      return GALA_OK;
    };
    auto live_after = [&](int32_t* out_live) -> int {
      HIPOK(hipMemcpyAsync(ctx->b_host, nlive, sizeof(int32_t), hipMemcpyDeviceToHost, s));
      HIPOK(hipStreamSynchronize(s));
      *out_live = ctx->b_host[0];
      return GALA_OK;
    };
    int32_t live = 0;
    if ((rc = live_after(&live))) return rc;
    if (round < 64) R.round_open[round] = live;
    hipGraphExec_t exec = nullptr;
    // Iteration k + 1 is queued before the live count of iteration k is read
    // (a graph that stopped is inert, so the one iteration queued past the
    // last one changes nothing): the host waits one iteration behind and the
    // GPU does not idle on the round trip.  Each queued iteration is followed
    // by the copy of its live count into its own pinned slot and an event.
    auto launch_it = [&](int k) -> int {
      if (k == 1 || !graphs) {
        if (int r = body(k == 1 && unique_rows)) return r;
      } else {
        if (!exec) {
          hipGraph_t gr = nullptr;
          HIPOK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
          const int rb = body(false);
          hipError_t e = hipStreamEndCapture(s, &gr);
          if (rb) { if (gr) (void)hipGraphDestroy(gr); return rb; }
          if (e != hipSuccess) { ctx->last_hip = e; return GALA_E_HIP; }
          e = hipGraphInstantiate(&exec, gr, nullptr, nullptr, 0);
          (void)hipGraphDestroy(gr);
          if (e != hipSuccess) { ctx->last_hip = e; return GALA_E_HIP; }
        }
        HIPOK(hipGraphLaunch(exec, s));
      }
      HIPOK(hipMemcpyAsync(ctx->b_host + 2 + (k & 1), nlive, sizeof(int32_t), hipMemcpyDeviceToHost, s));
      HIPOK(hipEventRecord(ctx->evr[k & 1], s));
      return GALA_OK;
    };
    int it = 0;     // iterations queued
    int known = 0;  // iterations whose live count has been read (the live count of iteration `known` is `live`)
    if (live > 0 && max_it > 0) {
      rc = launch_it(++it);
      while (!rc) {
        if (it < max_it) rc = launch_it(++it);  // the live count of iteration it - 2 was positive
        if (rc) break;
        known++;
        const hipError_t e = hipEventSynchronize(ctx->evr[known & 1]);
        if (e != hipSuccess) {
          ctx->last_hip = e;
          rc = GALA_E_HIP;
          break;
        }
        live = ctx->b_host[2 + (known & 1)];
        if (live == 0 || known >= max_it) break;
      }
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (rc) return rc;
    it = known;  // iterations that did work (a queued one past them is inert)
    HIPOK(hipEventRecord(ctx->ev[6], s));
    // contraction of the whole union under the accepted assignments
    int32_t C = 0;
    int64_t m = 0;
    const int dst = dst_of(src);
    rc = contract<T>(ctx, g, ctx->prev, dst, caps, &C, &m, false);
    if (rc) return rc;
    compose_map(ctx->fin, ctx->n0, ctx->c_label, s);
    const int gnext = gcur == 1 ? 2 : 1;
    launch_batch_gid_next(gid, ctx->c_cnt, ctx->c_nid, g.n, ctx->b_gid[gnext], s);
    fill<int32_t>(nopen, 1, 0, s);
    launch_batch_round(b, round, min_mod, max_rounds, nopen, s);
    HIPOK(hipEventRecord(ctx->ev[7], s));
    HIPOK(hipMemcpyAsync(ctx->b_host + 1, nopen, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPOK(hipStreamSynchronize(s));
    const int32_t open = ctx->b_host[1];
    float t_p1 = 0, t_build = 0;
    (void)hipEventElapsedTime(&t_p1, ctx->ev[5], ctx->ev[6]);
    (void)hipEventElapsedTime(&t_build, ctx->ev[6], ctx->ev[7]);
    R.t_phase1_ms += t_p1;
    R.t_contract_ms += t_build;
    R.iterations += it;
    if (round < 64) {
      R.round_iters[round] = it;
      R.round_ms[round] = t_p1 + t_build;
    }
    src = dst;
    gcur = gnext;
    g = graph_at<T>(ctx, src, C, m);
    round++;
    R.rounds = round;
    if (open == 0 || round >= max_rounds) break;
  }
  HIPOK(hipEventRecord(ctx->ev[3], s));
  HIPOK(hipStreamSynchronize(s));
  float tt_ms = 0;
  (void)hipEventElapsedTime(&tt_ms, ctx->ev[4], ctx->ev[3]);
  R.t_total_ms = tt_ms;
  Counters c;
  int rc = read_counters<T>(ctx, &c);
  if (rc) return rc;
  R.swept_vertices = (int64_t)c.swept;
  R.swept_edges = (int64_t)c.swept_edges;
  // per-graph results (the communities stay in fin, renumbered within each
  // graph when they are requested)
  fill<int32_t>(ctx->b_ncomm, G, 0, s);
  launch_batch_count(ctx->b_gid[gcur], g.n, ctx->b_ncomm, s);
  ctx->b_q.resize(G); ctx->b_nc.resize(G); ctx->b_rd.resize(G); ctx->b_it.resize(G);
  HIPOK(hipMemcpyAsync(ctx->b_q.data(), ctx->b_cur_mod, sizeof(double) * G, hipMemcpyDeviceToHost, s));
  HIPOK(hipMemcpyAsync(ctx->b_nc.data(), ctx->b_ncomm, sizeof(int32_t) * G, hipMemcpyDeviceToHost, s));
  HIPOK(hipMemcpyAsync(ctx->b_rd.data(), ctx->b_rounds, sizeof(int32_t) * G, hipMemcpyDeviceToHost, s));
  HIPOK(hipMemcpyAsync(ctx->b_it.data(), ctx->b_tot_iters, sizeof(int32_t) * G, hipMemcpyDeviceToHost, s));
  HIPOK(hipStreamSynchronize(s));
  ctx->b_comm_ready = false;  // read back on request (gala_batch_communities)
  ctx->have_batch_result = true;
  if (out) *out = R;
  return GALA_OK;
}

const char* kErr[] = {"ok", "I/O error", "malformed graph", "out of memory", "HIP runtime error",
                      "exact-integer range exceeded or hash overflow", "invalid argument or state",
                      "internal consistency check failed"};

}  // namespace

extern "C" {

const char* gala_version(void) { return "gala-mi355x 0.1 (gfx950)"; }

const char* gala_strerror(int code) {
  int k = -code;
  if (k < 0 || k > 7) return "unknown error";
  return kErr[k];
}

void gala_default_opts(gala_opts* o) {
  std::memset(o, 0, sizeof *o);
  o->min_modularity = 1e-6;
  o->learned_tau = 0.5f;
}

int gala_create(gala_ctx** out, int device, const gala_opts* opts) {
  if (!out) return GALA_E_INVALID;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return GALA_E_HIP;
  if (device < 0 || device >= ndev) return GALA_E_INVALID;
  if (hipSetDevice(device) != hipSuccess) return GALA_E_HIP;
  gala_ctx* ctx = new gala_ctx();
  ctx->device = device;
  if (opts) ctx->opts = *opts;
  else gala_default_opts(&ctx->opts);
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) { delete ctx; return GALA_E_HIP; }
  if (hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking) != hipSuccess) {
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return GALA_E_HIP;
  }
  // device address of the no-allocation entry point's static scratch on this GPU
  if (int rs = modularity_scratch_setup(device)) {
    (void)hipStreamDestroy(ctx->stream);
    (void)hipStreamDestroy(ctx->stream2);
    delete ctx;
    return rs;
  }
  for (auto& e : ctx->evf)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) { gala_destroy(ctx); return GALA_E_HIP; }
  for (auto& e : ctx->evr)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) { gala_destroy(ctx); return GALA_E_HIP; }
  for (auto& e : ctx->ev)
    if (hipEventCreate(&e) != hipSuccess) { gala_destroy(ctx); return GALA_E_HIP; }
  for (auto& e : ctx->bev)
    if (hipEventCreate(&e) != hipSuccess) { gala_destroy(ctx); return GALA_E_HIP; }
  for (auto& e : ctx->kev)
    if (hipEventCreate(&e) != hipSuccess) { gala_destroy(ctx); return GALA_E_HIP; }
  for (auto& p : ctx->sev)
    for (auto& e : p)
      if (hipEventCreate(&e) != hipSuccess) { gala_destroy(ctx); return GALA_E_HIP; }
  *out = ctx;
  return GALA_OK;
}

void gala_destroy(gala_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
  free_all(ctx);
  if (ctx->mlp_prm) (void)hipFree(ctx->mlp_prm);
  if (ctx->mlp_w1b) (void)hipFree(ctx->mlp_w1b);
  if (ctx->ctr_h) (void)hipHostFree(ctx->ctr_h);
  if (ctx->hpin) (void)hipHostFree(ctx->hpin);
  if (ctx->b_host) (void)hipHostFree(ctx->b_host);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : ctx->bev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : ctx->kev)
    if (e) (void)hipEventDestroy(e);
  for (auto& p : ctx->sev)
    for (auto& e : p)
      if (e) (void)hipEventDestroy(e);
  for (auto& e : ctx->evf)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : ctx->evr)
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

int gala_set_graph(gala_ctx* ctx, int32_t n, const int64_t* end_incl, const int32_t* nbr, const void* w,
                   int weight_mode) {
  if (!ctx || weight_mode < 0 || weight_mode > 2) return GALA_E_INVALID;
  (void)hipSetDevice(ctx->device);
  return upload(ctx, n, end_incl, nbr, w, weight_mode);
}

int gala_load_bin(gala_ctx* ctx, const char* path, int weight_mode) {
  if (!ctx || weight_mode < 0 || weight_mode > 2) return GALA_E_INVALID;
  int32_t n = 0;
  int64_t m = 0;
  int rc = gala_read_bin(path, weight_mode, &n, &m, nullptr, nullptr, nullptr);
  if (rc) return rc;
  std::vector<int64_t> end;
  std::vector<int32_t> nbr, w;
  try {  // host staging of the file: no C++ exception may cross the C ABI
    end.resize(n);
    nbr.resize(m);
    w.resize(weight_mode != GALA_W_NONE ? m : 0);
  } catch (const std::exception&) {
    return GALA_E_OOM;
  }
  rc = gala_read_bin(path, weight_mode, &n, &m, end.data(), nbr.data(), w.empty() ? nullptr : w.data());
  if (rc) return rc;
  (void)hipSetDevice(ctx->device);
  return upload(ctx, n, end.data(), nbr.data(), w.empty() ? nullptr : w.data(), weight_mode);
}

int gala_run(gala_ctx* ctx, gala_iter_cb icb, gala_round_cb rcb, void* user, gala_result* out) {
  if (!ctx || !ctx->have_graph) return GALA_E_INVALID;
  if (ctx->n0 == 0) return GALA_E_INVALID;
  (void)hipSetDevice(ctx->device);
  // the input buffers are read-only: a context can run the same graph again
  return is_wide(ctx) ? run_all<WideW>(ctx, icb, rcb, user, out) : run_all<IntW>(ctx, icb, rcb, user, out);
}

// weight-update kernel variants per bin.  Dirty-row path (tracked runs):
// 1 = pipelined (bins 0-4), 2 = one thread per vertex (bins 0-3).  Incremental
// integer path: narrower lane groups for bins 1-5 (launch_weight_moved).
static int weight_variants(int bin) { return bin <= 5 ? 3 : 1; }

int gala_set_variants(gala_ctx* ctx, const int32_t* variants) {
  if (!ctx || !variants) return GALA_E_INVALID;
  for (int b = 0; b < NBINS; b++) {
    if (variants[b] < 0 || variants[b] >= decide_variants(b)) return GALA_E_INVALID;
    if (variants[NBINS + b] < 0 || variants[NBINS + b] >= decide_first_variants(b)) return GALA_E_INVALID;
    if (variants[2 * NBINS + b] < 0 || variants[2 * NBINS + b] >= weight_variants(b)) return GALA_E_INVALID;
  }
  for (int b = 0; b < NBINS; b++) {
    ctx->dvar[b] = variants[b];
    ctx->fvar[b] = variants[NBINS + b];
    ctx->wvar[b] = variants[2 * NBINS + b];
  }
  return GALA_OK;
}

int gala_get_variants(gala_ctx* ctx, int32_t* variants) {
  if (!ctx || !variants) return GALA_E_INVALID;
  for (int b = 0; b < NBINS; b++) {
    variants[b] = ctx->dvar[b];
    variants[NBINS + b] = ctx->fvar[b];
    variants[2 * NBINS + b] = ctx->wvar[b];
  }
  return GALA_OK;
}

int gala_variant_counts(int32_t* counts_out) {
  if (!counts_out) return GALA_E_INVALID;
  for (int b = 0; b < NBINS; b++) {
    counts_out[b] = decide_variants(b);
    counts_out[NBINS + b] = decide_first_variants(b);
    counts_out[2 * NBINS + b] = weight_variants(b);
  }
  return GALA_OK;
}

int gala_set_bin_cuts(gala_ctx* ctx, const int32_t* lo) {
  if (!ctx || !lo || !bin_cuts_valid(lo)) return GALA_E_INVALID;
  for (int b = 0; b < NBINS; b++) ctx->cuts.lo[b] = lo[b];
  ctx->have_result = false;
  return GALA_OK;
}

int gala_get_bin_cuts(gala_ctx* ctx, int32_t* lo) {
  if (!ctx || !lo) return GALA_E_INVALID;
  for (int b = 0; b < NBINS; b++) lo[b] = ctx->cuts.lo[b];
  return GALA_OK;
}

// Cut points by coordinate descent: for bins 8 down to 4, the lower cut is
// tried at 1, 7/8, 3/4, 5/8 and 1/2 of its default (never at or below the
// previous bin's default cut) and the one with the least decide + update time
// over the best of two timed plain runs is kept.  Binning changes which kernel
// evaluates a vertex, never the decision, so every run is identical but for time.
int tune_bin_cuts(gala_ctx* ctx) {
  const gala_opts saved = ctx->opts;
  ctx->opts.profile_bins = 0;
  ctx->opts.frontier_mode = 0;
  auto timed = [&](double* ms) -> int {
    gala_result r;
    double best = 1e300;
    for (int rep = 0; rep < 3; rep++) {  // the first run warms up
      int rc = is_wide(ctx) ? run_all<WideW>(ctx, nullptr, nullptr, nullptr, &r)
                            : run_all<IntW>(ctx, nullptr, nullptr, nullptr, &r);
      if (rc) return rc;
      if (rep > 0) best = std::min(best, r.t_decide_ms + r.t_update_ms);
    }
    *ms = best;
    return GALA_OK;
  };
  double cur_ms = 0;
  int rc = timed(&cur_ms);
  for (int b = NBINS - 1; b >= BIN_TUNE_FIRST && rc == GALA_OK; b--) {
    const int keep = ctx->cuts.lo[b];
    int best_lo = keep;
    for (int num : {8, 7, 6, 5, 4}) {
      const int lo = std::max(BIN_LO[b - 1] + 1, BIN_LO[b] * num / 8);
      if (lo == keep) continue;
      ctx->cuts.lo[b] = lo;
      double ms = 0;
      if ((rc = timed(&ms)) != GALA_OK) break;
      if (ms < cur_ms) { cur_ms = ms; best_lo = lo; }
    }
    ctx->cuts.lo[b] = best_lo;
  }
  ctx->opts = saved;
  ctx->have_result = false;
  return rc;
}

int gala_autotune_bins(gala_ctx* ctx, int32_t* lo_out) {
  if (!ctx || !ctx->have_graph || ctx->n0 == 0) return GALA_E_INVALID;
  (void)hipSetDevice(ctx->device);
  int rc = tune_bin_cuts(ctx);
  if (rc) return rc;
  if (lo_out) gala_get_bin_cuts(ctx, lo_out);
  return GALA_OK;
}

int gala_autotune(gala_ctx* ctx, int32_t* variants_out) {
  if (!ctx || !ctx->have_graph || ctx->n0 == 0) return GALA_E_INVALID;
  (void)hipSetDevice(ctx->device);
  // one full plain run per variant index with per-bin HIP events; every
  // variant takes the same decisions, so the runs are identical but for time.
  // Decide kernels are chosen per bin; the weight-update kernels, timed as a
  // phase, are chosen all together.
  const gala_opts saved = ctx->opts;
  ctx->opts.profile_bins = 1;
  ctx->opts.frontier_mode = 0;
  int nmax = 1;
  for (int b = 0; b < NBINS; b++)
    nmax = std::max(nmax, std::max(decide_variants(b), std::max(decide_first_variants(b), weight_variants(b))));
  double best_d[NBINS], best_f[NBINS], best_w = 1e300;
  int bd[NBINS], bf[NBINS], bw = 0;
  for (int b = 0; b < NBINS; b++) { best_d[b] = best_f[b] = 1e300; bd[b] = bf[b] = 0; }
  int rc = GALA_OK;
  for (int k = 0; k < nmax && rc == GALA_OK; k++) {
    for (int b = 0; b < NBINS; b++) {
      ctx->dvar[b] = std::min(k, decide_variants(b) - 1);
      ctx->fvar[b] = std::min(k, decide_first_variants(b) - 1);
      ctx->wvar[b] = std::min(k, weight_variants(b) - 1);
    }
    gala_result r;
    for (int rep = 0; rep < 2 && rc == GALA_OK; rep++)  // the second run is timed
      rc = is_wide(ctx) ? run_all<WideW>(ctx, nullptr, nullptr, nullptr, &r) : run_all<IntW>(ctx, nullptr, nullptr, nullptr, &r);
    if (rc) break;
    for (int b = 0; b < NBINS; b++) {
      if (k < decide_variants(b) && r.bin_decide_ms[b] < best_d[b]) { best_d[b] = r.bin_decide_ms[b]; bd[b] = k; }
      if (k < decide_first_variants(b) && r.bin_first_ms[b] < best_f[b]) { best_f[b] = r.bin_first_ms[b]; bf[b] = k; }
    }
    if (k < weight_variants(0) && r.t_update_ms < best_w) { best_w = r.t_update_ms; bw = k; }
  }
  ctx->opts = saved;
  for (int b = 0; b < NBINS; b++) {
    ctx->dvar[b] = bd[b];
    ctx->fvar[b] = bf[b];
    ctx->wvar[b] = std::min(bw, weight_variants(b) - 1);
  }
  ctx->have_result = false;
  if (rc) return rc;
  if ((rc = tune_bin_cuts(ctx)) != GALA_OK) return rc;  // then the cut points, with those kernels
  if (variants_out) gala_get_variants(ctx, variants_out);
  return GALA_OK;
}

int gala_batch_set(gala_ctx* ctx, int32_t n_graphs, const int32_t* vstart, int32_t n, const int64_t* end_incl,
                   const int32_t* nbr, const void* w, int weight_mode) {
  if (!ctx) return GALA_E_INVALID;
  (void)hipSetDevice(ctx->device);
  return batch_upload(ctx, n_graphs, vstart, n, end_incl, nbr, w, weight_mode);
}

int gala_batch_run(gala_ctx* ctx, gala_batch_result* out) {
  if (!ctx || !ctx->have_batch || !ctx->have_graph) return GALA_E_INVALID;
  (void)hipSetDevice(ctx->device);
  return run_batch(ctx, out);
}

int gala_batch_graph_results(gala_ctx* ctx, double* modularity, int32_t* n_communities, int32_t* rounds,
                             int32_t* iterations) {
  if (!ctx || !ctx->have_batch_result) return GALA_E_INVALID;
  const size_t G = ctx->b_q.size();
  if (modularity) std::memcpy(modularity, ctx->b_q.data(), sizeof(double) * G);
  if (n_communities) std::memcpy(n_communities, ctx->b_nc.data(), sizeof(int32_t) * G);
  if (rounds) std::memcpy(rounds, ctx->b_rd.data(), sizeof(int32_t) * G);
  if (iterations) std::memcpy(iterations, ctx->b_it.data(), sizeof(int32_t) * G);
  return GALA_OK;
}

int gala_batch_communities(gala_ctx* ctx, int32_t* comm) {
  if (!ctx || !comm || !ctx->have_batch_result) return GALA_E_INVALID;
  if (!ctx->b_comm_ready) {
    // the composed union partition, renumbered within each graph (the final
    // union's vertices are grouped by graph in ascending order)
    (void)hipSetDevice(ctx->device);
    ctx->b_comm_local.resize(ctx->n0);
    HIPOK(hipMemcpy(ctx->b_comm_local.data(), ctx->fin, sizeof(int32_t) * ctx->n0, hipMemcpyDeviceToHost));
    int64_t first = 0;
    for (int32_t gg = 0; gg < ctx->b_G; gg++) {
      for (int32_t v = ctx->b_vstart[gg]; v < ctx->b_vstart[gg + 1]; v++) ctx->b_comm_local[v] -= (int32_t)first;
      first += ctx->b_nc[gg];
    }
    ctx->b_comm_ready = true;
  }
  std::memcpy(comm, ctx->b_comm_local.data(), sizeof(int32_t) * ctx->b_comm_local.size());
  return GALA_OK;
}

int gala_partial_result(gala_ctx* ctx, gala_result* out) {
  if (!ctx || !out) return GALA_E_INVALID;
  *out = ctx->res;
  return GALA_OK;
}

int gala_get_communities(gala_ctx* ctx, int32_t* out_n0) {
  if (!ctx || !out_n0 || !ctx->have_result) return GALA_E_INVALID;
  (void)hipSetDevice(ctx->device);
  HIPOK(hipMemcpy(out_n0, ctx->fin, sizeof(int32_t) * ctx->n0, hipMemcpyDeviceToHost));
  return GALA_OK;
}

// ---------------------------------------------------------------------------
// phase-level entry points
// ---------------------------------------------------------------------------
int gala_phase_decide(gala_ctx* ctx, int32_t n, const int64_t* end_incl, const int32_t* nbr,
                      const int32_t* w, int64_t m2, const int32_t* K, const int32_t* Self,
                      const int32_t* Tot, const int32_t* size, const int32_t* prev,
                      const uint8_t* frontier, int32_t* cur_out) {
  if (!ctx || n <= 0 || !K || !Self || !Tot || !size || !prev || !cur_out) return GALA_E_INVALID;
  if (m2 <= 0 || m2 > M2_LIMIT) return GALA_E_OVERFLOW;
  (void)hipSetDevice(ctx->device);
  int rc = upload(ctx, n, end_incl, nbr, w, w ? GALA_W_I32 : GALA_W_NONE);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  DevGraph<IntW> g = graph_at<IntW>(ctx, 0, n, ctx->m0);
  int32_t caps[NBINS], nitems = 0;
  if ((rc = round_setup<IntW>(ctx, g, caps, &nitems))) return rc;  // bins, order, hub tables
  HIPOK(hipMemcpyAsync(ctx->K, K, 4ll * n, hipMemcpyHostToDevice, s));
  HIPOK(hipMemcpyAsync(ctx->Self, Self, 4ll * n, hipMemcpyHostToDevice, s));
  HIPOK(hipMemcpyAsync(ctx->Tot, Tot, 4ll * n, hipMemcpyHostToDevice, s));
  HIPOK(hipMemcpyAsync(ctx->size, size, 4ll * n, hipMemcpyHostToDevice, s));
  HIPOK(hipMemcpyAsync(ctx->prev, prev, 4ll * n, hipMemcpyHostToDevice, s));
  HIPOK(hipMemcpyAsync(ctx->cur, prev, 4ll * n, hipMemcpyHostToDevice, s));
  if (frontier) HIPOK(hipMemcpyAsync(ctx->active, frontier, n, hipMemcpyHostToDevice, s));
  else fill<uint8_t>(ctx->active, n, 1, s);
  partition_bins(1, ctx->order, ctx->bstart_full, ctx->binid, ctx->active, nitems, ctx->front, ctx->bstart_front,
                 ctx->part_tmp, s);
  DevState<IntW> st = dstate<IntW>(ctx);
  reset_counters(ctx->ctr, m2, s);
  launch_build_cv<IntW>(st, n, s);
  for (int b = 0; b < NBINS; b++)
    launch_decide<IntW>(b, g, st, ctx->front, ctx->bstart_front, caps[b], (int32_t)m2, s, ctx->dvar[b]);
  Counters c;
  if ((rc = read_counters<IntW>(ctx, &c))) return rc;
  HIPOK(hipMemcpy(cur_out, ctx->cur, 4ll * n, hipMemcpyDeviceToHost));
  ctx->have_graph = false;
  return GALA_OK;
}

int gala_phase_weight(gala_ctx* ctx, int32_t n, const int64_t* end_incl, const int32_t* nbr,
                      const int32_t* w, int64_t m2, const int32_t* K, const int32_t* Self,
                      const int32_t* Tot, const int32_t* cur, int64_t min_tot, int32_t* In_out,
                      uint8_t* active_out) {
  if (!ctx || n <= 0 || !K || !Self || !Tot || !cur || !In_out || !active_out) return GALA_E_INVALID;
  if (m2 <= 0 || m2 > M2_LIMIT) return GALA_E_OVERFLOW;
  (void)hipSetDevice(ctx->device);
  int rc = upload(ctx, n, end_incl, nbr, w, w ? GALA_W_I32 : GALA_W_NONE);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  DevGraph<IntW> g = graph_at<IntW>(ctx, 0, n, ctx->m0);
  int32_t caps[NBINS], nitems = 0;
  if ((rc = round_setup<IntW>(ctx, g, caps, &nitems))) return rc;
  HIPOK(hipMemcpyAsync(ctx->K, K, 4ll * n, hipMemcpyHostToDevice, s));
  HIPOK(hipMemcpyAsync(ctx->Self, Self, 4ll * n, hipMemcpyHostToDevice, s));
  HIPOK(hipMemcpyAsync(ctx->Tot, Tot, 4ll * n, hipMemcpyHostToDevice, s));
  HIPOK(hipMemcpyAsync(ctx->cur, cur, 4ll * n, hipMemcpyHostToDevice, s));
  DevState<IntW> st = dstate<IntW>(ctx);
  reset_counters(ctx->ctr, min_tot, s);
  // degree-0 vertices are in no bin: In stays 0 and active 1, as the
  // reference's zero-initialised In and fill(active, 1) leave them
  fill<int32_t>((int32_t*)ctx->In, n, 0, s);
  fill<uint8_t>(ctx->active, n, 1, s);
  for (int b = 0; b < NBINS; b++) launch_weight<IntW>(b, g, st, ctx->order, ctx->bstart_full, caps[b], false, s, ctx->wvar[b]);
  launch_prune<IntW>(st, n, (int32_t)m2, s);
  Counters c;
  if ((rc = read_counters<IntW>(ctx, &c))) return rc;
  HIPOK(hipMemcpy(In_out, ctx->In, 4ll * n, hipMemcpyDeviceToHost));
  HIPOK(hipMemcpy(active_out, ctx->active, n, hipMemcpyDeviceToHost));
  ctx->have_graph = false;
  return GALA_OK;
}

// Device text ingest: a first call (NULL arrays) parses the file on the GPU
// and keeps the CSR in the context; a call with arrays copies it out (and
// ingests first when the context holds another file or mode).
int gala_ingest_text_device(gala_ctx* ctx, const char* path, int weight_mode, int32_t* n, int64_t* m,
                            int64_t* end_incl, int32_t* nbr, void* w) {
  if (!ctx || !path || !n || !m || weight_mode < GALA_W_NONE || weight_mode > GALA_W_F32) return GALA_E_INVALID;
  std::string key;
  try {
    key = std::string(path) + "\n" + std::to_string(weight_mode);
  } catch (const std::exception&) {
    return GALA_E_OOM;
  }
  if (!end_incl || ctx->ingest_key != key) {
    ctx->ingest_key.clear();
    int rc;
    try {
      rc = ingest_text_device(ctx->device, ctx->stream, path, weight_mode, &ctx->ingest);
    } catch (const std::exception&) {
      rc = GALA_E_OOM;
    }
    if (rc) return rc;
    ctx->ingest_key = key;
  }
  const IngestResult& r = ctx->ingest;
  *n = r.n;
  *m = (int64_t)r.nbr.size();
  if (end_incl && r.n) std::memcpy(end_incl, r.end.data(), 8 * (size_t)r.n);
  if (nbr && !r.nbr.empty()) std::memcpy(nbr, r.nbr.data(), 4 * r.nbr.size());
  if (w && weight_mode != GALA_W_NONE && !r.w.empty()) std::memcpy(w, r.w.data(), 4 * r.w.size());
  return GALA_OK;
}

int gala_load_text(gala_ctx* ctx, const char* path, int weight_mode) {
  int32_t n = 0;
  int64_t m = 0;
  int rc = gala_ingest_text_device(ctx, path, weight_mode, &n, &m, nullptr, nullptr, nullptr);
  if (rc) return rc;
  (void)hipSetDevice(ctx->device);
  const IngestResult& r = ctx->ingest;
  return upload(ctx, n, r.end.data(), r.nbr.data(), r.w.empty() ? nullptr : r.w.data(), weight_mode);
}

int gala_ingest_times(gala_ctx* ctx, double* read_ms, double* load_ms, double* device_ms, double* d2h_ms,
                      double* wall_ms) {
  if (!ctx || ctx->ingest_key.empty()) return GALA_E_INVALID;
  const IngestResult& r = ctx->ingest;
  if (read_ms) *read_ms = r.read_ms;
  if (load_ms) *load_ms = r.h2d_ms;
  if (device_ms) *device_ms = r.device_ms;
  if (d2h_ms) *d2h_ms = r.d2h_ms;
  if (wall_ms) *wall_ms = r.wall_ms;
  return GALA_OK;
}

int gala_phase_modularity(gala_ctx* ctx, int32_t n, const int64_t* end_incl, const int32_t* nbr, const void* w,
                          int weight_mode, const int32_t* comm, double* q_out) {
  if (!ctx || n < 0 || !q_out || weight_mode < 0 || weight_mode > 2) return GALA_E_INVALID;
  if (n > 0 && (!end_incl || !nbr || !comm)) return GALA_E_INVALID;
  for (int32_t i = 0; i < n; i++)
    if (end_incl[i] < (i == 0 ? 0 : end_incl[i - 1])) return GALA_E_FORMAT;
  if (weight_mode != GALA_W_NONE && n > 0 && end_incl[n - 1] > 0 && !w) return GALA_E_INVALID;
  (void)hipSetDevice(ctx->device);
  return modularity_host(ctx->device, ctx->stream, n, end_incl, nbr, w, weight_mode, comm, q_out);
}

int gala_modularity_noalloc(gala_ctx* ctx, int32_t n, const int64_t* d_end_incl, const int32_t* d_nbr,
                            const void* d_w, int weight_mode, const int32_t* d_comm, double* q_out) {
  if (!ctx || n < 0 || !q_out || weight_mode < 0 || weight_mode > 2) return GALA_E_INVALID;
  if (n > 0 && (!d_end_incl || !d_nbr || !d_comm || (weight_mode != GALA_W_NONE && !d_w))) return GALA_E_INVALID;
  (void)hipSetDevice(ctx->device);
  return modularity_noalloc(ctx->device, ctx->stream, n, d_end_incl, d_nbr, d_w, weight_mode, d_comm, q_out);
}

int gala_phase_contract(gala_ctx* ctx, int32_t n, const int64_t* end_incl, const int32_t* nbr,
                        const int32_t* w, const int32_t* comm, int32_t* out_C, int64_t* out_m,
                        int32_t* new_id, int64_t* out_end_incl, int32_t* out_nbr, int64_t* out_w) {
  if (!ctx || n <= 0 || !comm || !out_C || !out_m) return GALA_E_INVALID;
  for (int32_t v = 0; v < n; v++)
    if (comm[v] < 0 || comm[v] >= n) return GALA_E_INVALID;
  (void)hipSetDevice(ctx->device);
  int rc = upload(ctx, n, end_incl, nbr, w, w ? GALA_W_I32 : GALA_W_NONE);
  if (rc) return rc;
  DevGraph<IntW> g = graph_at<IntW>(ctx, 0, n, ctx->m0);
  int32_t caps[NBINS], nitems = 0;
  if ((rc = round_setup<IntW>(ctx, g, caps, &nitems))) return rc;  // degree bins for the staging kernels
  HIPOK(hipMemcpy(ctx->prev, comm, 4ll * n, hipMemcpyHostToDevice));
  reset_counters(ctx->ctr, 0, ctx->stream);
  int32_t C = 0;
  int64_t m = 0;
  if ((rc = contract<IntW>(ctx, g, ctx->prev, 1, caps, &C, &m, true))) return rc;
  *out_C = C;
  *out_m = m;
  if (new_id) {
    std::vector<int32_t> cnt(n), nid(n);
    HIPOK(hipMemcpy(cnt.data(), ctx->c_cnt, 4ll * n, hipMemcpyDeviceToHost));
    HIPOK(hipMemcpy(nid.data(), ctx->c_nid, 4ll * n, hipMemcpyDeviceToHost));
    for (int32_t c = 0; c < n; c++) new_id[c] = cnt[c] > 0 ? nid[c] : -1;
  }
  if (out_end_incl && out_nbr && out_w) {
    std::vector<int64_t> off(C + 1);
    HIPOK(hipMemcpy(off.data(), ctx->off[1], 8ll * (C + 1), hipMemcpyDeviceToHost));
    for (int32_t c = 0; c < C; c++) out_end_incl[c] = off[c + 1];
    HIPOK(hipMemcpy(out_nbr, ctx->nbr[1], 4ll * m, hipMemcpyDeviceToHost));
    std::vector<int32_t> w32(m);
    HIPOK(hipMemcpy(w32.data(), ctx->w[1], 4ll * m, hipMemcpyDeviceToHost));
    for (int64_t j = 0; j < m; j++) out_w[j] = w32[j];
  }
  ctx->have_graph = false;
  return GALA_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// learned frontier: MLP parameters, training, move logs
// ---------------------------------------------------------------------------
namespace {

int ensure_mlp(gala_ctx* ctx) {
  if (ctx->mlp_prm) return GALA_OK;
  hipError_t e = hipMalloc((void**)&ctx->mlp_prm, sizeof(float) * MLP_NPARAM);
  if (e == hipSuccess) e = hipMalloc((void**)&ctx->mlp_w1b, sizeof(short) * MLP_B1);
  if (e != hipSuccess) { ctx->last_hip = e; return GALA_E_OOM; }
  return GALA_OK;
}

int upload_mlp(gala_ctx* ctx, const float* params) {
  int rc = ensure_mlp(ctx);
  if (rc) return rc;
  std::vector<float> p(params, params + MLP_NPARAM);
  for (int k = MLP_IN; k < 16; k++)  // padded feature rows of W1 stay zero
    for (int h = 0; h < MLP_H; h++) p[MLP_W1 + k * MLP_H + h] = 0.f;
  HIPOK(hipMemcpy(ctx->mlp_prm, p.data(), sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice));
  launch_w1_bf16(ctx->mlp_prm, ctx->mlp_w1b, ctx->stream);
  HIPOK(hipStreamSynchronize(ctx->stream));
  return GALA_OK;
}

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() { if (p) (void)hipFree(p); }
};

}  // namespace

extern "C" {

int gala_set_logger(gala_ctx* ctx, gala_log_cb cb, void* user) {
  if (!ctx) return GALA_E_INVALID;
  ctx->logger = cb;
  ctx->logger_user = user;
  return GALA_OK;
}

int gala_set_mlp(gala_ctx* ctx, const float* params) {
  if (!ctx || !params) return GALA_E_INVALID;
  (void)hipSetDevice(ctx->device);
  int rc = upload_mlp(ctx, params);
  if (rc) return rc;
  ctx->have_mlp = true;
  return GALA_OK;
}

int gala_get_mlp(gala_ctx* ctx, float* params) {
  if (!ctx || !params || !ctx->have_mlp) return GALA_E_INVALID;
  (void)hipSetDevice(ctx->device);
  HIPOK(hipMemcpy(params, ctx->mlp_prm, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost));
  return GALA_OK;
}

int gala_mlp_forward(gala_ctx* ctx, const float* params, const float* X, int64_t n, float* p_out) {
  if (!ctx || !params || (n > 0 && (!X || !p_out)) || n < 0 || n > INT32_MAX) return GALA_E_INVALID;
  (void)hipSetDevice(ctx->device);
  std::vector<float> keep_prm;
  int rc = upload_mlp(ctx, params);
  if (rc) return rc;
  ctx->have_mlp = true;
  DevBuf dx, dp;
  HIPOK(hipMalloc(&dx.p, sizeof(float) * std::max<int64_t>(n, 1) * MLP_IN));
  HIPOK(hipMalloc(&dp.p, sizeof(float) * std::max<int64_t>(n, 1)));
  HIPOK(hipMemcpy(dx.p, X, sizeof(float) * n * MLP_IN, hipMemcpyHostToDevice));
  launch_mlp_forward((const float*)dx.p, (int)n, ctx->mlp_prm, ctx->mlp_w1b, (float*)dp.p, ctx->stream);
  HIPOK(hipStreamSynchronize(ctx->stream));
  HIPOK(hipMemcpy(p_out, dp.p, sizeof(float) * n, hipMemcpyDeviceToHost));
  return GALA_OK;
}

int gala_mlp_grad(gala_ctx* ctx, const float* params, const float* X, const uint8_t* y, int64_t n, float pos_weight,
                  float* grad_out, float* loss_out) {
  if (!ctx || !params || !X || !y || !grad_out || n <= 0 || n > INT32_MAX) return GALA_E_INVALID;
  (void)hipSetDevice(ctx->device);
  int rc = upload_mlp(ctx, params);
  if (rc) return rc;
  ctx->have_mlp = true;
  hipStream_t s = ctx->stream;
  double wsum = 0;
  for (int64_t i = 0; i < n; i++) wsum += y[i] ? pos_weight : 1.0;
  std::vector<int32_t> idx(n);
  for (int64_t i = 0; i < n; i++) idx[i] = (int32_t)i;
  DevBuf dx, dy, di, dg, dl;
  HIPOK(hipMalloc(&dx.p, sizeof(float) * n * MLP_IN));
  HIPOK(hipMalloc(&dy.p, n));
  HIPOK(hipMalloc(&di.p, sizeof(int32_t) * n));
  HIPOK(hipMalloc(&dg.p, sizeof(float) * MLP_NPARAM));
  HIPOK(hipMalloc(&dl.p, sizeof(float)));
  HIPOK(hipMemcpy(dx.p, X, sizeof(float) * n * MLP_IN, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(dy.p, y, n, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(di.p, idx.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
  fill<float>((float*)dg.p, MLP_NPARAM, 0.f, s);
  fill<float>((float*)dl.p, 1, 0.f, s);
  launch_mlp_grad((const float*)dx.p, (const uint8_t*)dy.p, (const int32_t*)di.p, (int)n, ctx->mlp_prm, ctx->mlp_w1b,
                  pos_weight, (float)(1.0 / wsum), (float*)dg.p, (float*)dl.p, s);
  HIPOK(hipStreamSynchronize(s));
  HIPOK(hipMemcpy(grad_out, dg.p, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost));
  if (loss_out) HIPOK(hipMemcpy(loss_out, dl.p, sizeof(float), hipMemcpyDeviceToHost));
  return GALA_OK;
}

int gala_mlp_train(gala_ctx* ctx, const float* X, const uint8_t* y, int64_t n, int epochs, int batch, float lr,
                   float pos_weight, uint64_t seed, int init_given, float* params_io, float* loss_out) {
  if (!ctx || !X || !y || !params_io || n <= 0 || n > INT32_MAX || epochs <= 0 || batch <= 0) return GALA_E_INVALID;
  (void)hipSetDevice(ctx->device);
  hipStream_t s = ctx->stream;
  uint64_t st = seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
  auto rnd = [&]() {  // splitmix64 -> [0, 1)
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  std::vector<float> prm(MLP_NPARAM, 0.f);
  if (init_given) {
    std::copy(params_io, params_io + MLP_NPARAM, prm.begin());
  } else {  // Glorot-uniform weights, zero biases
    const double a1 = std::sqrt(6.0 / (MLP_IN + MLP_H)), a2 = std::sqrt(6.0 / (MLP_H + 1));
    for (int k = 0; k < MLP_IN; k++)
      for (int h = 0; h < MLP_H; h++) prm[MLP_W1 + k * MLP_H + h] = (float)((2 * rnd() - 1) * a1);
    for (int h = 0; h < MLP_H; h++) prm[MLP_W2 + h] = (float)((2 * rnd() - 1) * a2);
  }
  int rc = upload_mlp(ctx, prm.data());
  if (rc) return rc;
  DevBuf dx, dy, di, dg, dl, dm, dv;
  HIPOK(hipMalloc(&dx.p, sizeof(float) * n * MLP_IN));
  HIPOK(hipMalloc(&dy.p, n));
  HIPOK(hipMalloc(&di.p, sizeof(int32_t) * n));
  HIPOK(hipMalloc(&dg.p, sizeof(float) * MLP_NPARAM));
  HIPOK(hipMalloc(&dl.p, sizeof(float)));
  HIPOK(hipMalloc(&dm.p, sizeof(float) * MLP_NPARAM));
  HIPOK(hipMalloc(&dv.p, sizeof(float) * MLP_NPARAM));
  HIPOK(hipMemcpy(dx.p, X, sizeof(float) * n * MLP_IN, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(dy.p, y, n, hipMemcpyHostToDevice));
  fill<float>((float*)dm.p, MLP_NPARAM, 0.f, s);
  fill<float>((float*)dv.p, MLP_NPARAM, 0.f, s);
  std::vector<int32_t> idx(n);
  for (int64_t i = 0; i < n; i++) idx[i] = (int32_t)i;
  int step = 0;
  for (int ep = 0; ep < epochs; ep++) {
    for (int64_t i = n - 1; i > 0; i--) {  // Fisher-Yates
      int64_t j = (int64_t)(rnd() * (double)(i + 1));
      std::swap(idx[i], idx[j]);
    }
    HIPOK(hipMemcpy(di.p, idx.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    fill<float>((float*)dl.p, 1, 0.f, s);
    int nb = 0;
    for (int64_t b0 = 0; b0 < n; b0 += batch, nb++) {
      const int bn = (int)std::min<int64_t>(batch, n - b0);
      double wsum = 0;
      for (int k = 0; k < bn; k++) wsum += y[idx[b0 + k]] ? pos_weight : 1.0;
      fill<float>((float*)dg.p, MLP_NPARAM, 0.f, s);
      launch_mlp_grad((const float*)dx.p, (const uint8_t*)dy.p, (const int32_t*)di.p + b0, bn, ctx->mlp_prm,
                      ctx->mlp_w1b, pos_weight, (float)(1.0 / wsum), (float*)dg.p, (float*)dl.p, s);
      launch_adam(ctx->mlp_prm, (float*)dm.p, (float*)dv.p, (const float*)dg.p, lr, ++step, ctx->mlp_w1b, s);
    }
    float ls = 0;
    HIPOK(hipMemcpyAsync(&ls, dl.p, sizeof(float), hipMemcpyDeviceToHost, s));
    HIPOK(hipStreamSynchronize(s));
    if (loss_out) loss_out[ep] = ls / (float)nb;
  }
  HIPOK(hipMemcpy(params_io, ctx->mlp_prm, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost));
  ctx->have_mlp = true;
  return GALA_OK;
}

}  // extern "C"
