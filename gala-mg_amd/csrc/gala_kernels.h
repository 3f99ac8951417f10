// Launch wrappers for the gfx950 kernels (host-callable, stream-ordered).
#pragma once
#include "gala_internal.h"
#include <algorithm>
#include <vector>

namespace gala {

// ---- common (kernels_common.hip) ----
int64_t scan_tmp_elems(int64_t n);
template <class In, class V>
void scan_exclusive(const In* in, V* out, int64_t n, V* tmp, V* total, hipStream_t s);

int64_t part_tmp_elems(int64_t n);
// Stable partition into degree bins.  mode 0: every vertex 0..n-1 by binid;
// mode 1: the vertices of order[0..n) with active[v] != 0; mode 2: those with
// active[v] == 0.  bstart receives NBINS+1 bin starts.  Modes 1 and 2 take
// order already grouped by bin, with its NBINS+1 bin starts in order_bstart
// (device), and keep that grouping: a stable compaction.
void partition_bins(int mode, const int32_t* order, const int32_t* order_bstart, const int8_t* binid,
                    const uint8_t* active, int64_t n, int32_t* out, int32_t* bstart, int32_t* tmp, hipStream_t s,
                    const int32_t* n_dev = nullptr);

template <class V>
void fill(V* p, int64_t n, V x, hipStream_t s);
void copy_i32(const int32_t* a, int32_t* b, int64_t n, hipStream_t s);
// b[0 .. min(*cnt, n)) = a[...]: a list whose length is on the device
void copy_list(const int32_t* a, int32_t* b, const int32_t* cnt, int64_t n, hipStream_t s);
// prev = cur after an accepted sweep; cp (if not null) = {cur, cur}, cm = cur
void accept_sweep(const int32_t* cur, int32_t* prev, int2* cp, int32_t* cm, int64_t n, hipStream_t s);
// cp = {cur, cur} and cm = cur without touching prev: the moves of a sweep already applied
// to Tot and size are not applied again by the next update
void settle_pairs(int2* cp, int32_t* cm, int64_t n, hipStream_t s);
void reset_counters(Counters* c, long long min_init, hipStream_t s);
// Stopping rule of a round on the device (louvain_main_process.cu:1004-1007,
// 1026-1031).  round_start: qprev = Q of the counters (the singleton
// partition), live = 1.  round_step, after sweep `it`'s update: the sweep is
// accepted iff live, Q - qprev >= min_mod and it < max_it; then qprev = Q,
// else live = 0 and the next frontier's bin starts (NBINS + 1) are zeroed, so
// a sweep already queued behind it finds no vertex.  The Q and the decision
// are also written into the counters for the host.
void round_start(const Counters* c, RoundCtl* rc, double m2, hipStream_t s);
void round_step(Counters* c, RoundCtl* rc, double m2, double min_mod, int it, int max_it, int32_t* bnext,
                hipStream_t s);
// accept_sweep over the vertices list[0, *cnt) (the swept frontier; *cnt <=
// nmax), when rc is null or the last round_step accepted its sweep
void accept_list(const int32_t* cur, int32_t* prev, int2* cp, int32_t* cm, const int32_t* list, const int32_t* cnt,
                 int64_t nmax, const RoundCtl* rc, hipStream_t s);

// ---- Louvain phase kernels (kernels_louvain.hip) ----
// Graph view on the device.  off has n+1 exclusive offsets.
template <class T>
struct DevGraph {
  int32_t n;
  int64_t m;  // directed entries (compact CSR) or the extent of the arrays (gapped)
  // row v is nbr[off[v] .. end[v]); a compact CSR has end = off + 1, a graph
  // contracted inside a run keeps each merged row at the start of the row's
  // staging segment, with a gap up to the next row's off
  const int64_t* off;
  const int64_t* end;
  const int32_t* nbr;
  const typename T::w_t* w;  // nullptr (and wf nullptr) => every weight is 1
  // wide path, input round of a float32-weighted graph: the weights as read
  // from the file, each scaled exactly to an integer by wsc = 2^k on load
  const float* wf = nullptr;
  double wsc = 1.0;
  __host__ __device__ bool weighted() const { return w != nullptr || wf != nullptr; }
};

// Weight of entry j as the trait's integer.  A float32 weight times 2^k is an
// integer below 2^62 (checked on upload), so both conversions are exact.
template <class T>
__device__ __forceinline__ typename T::w_t load_w(const DevGraph<T>& g, int64_t j) {
  if constexpr (T::WIDE) {
    if (g.wf) return (typename T::w_t)((double)g.wf[j] * g.wsc);
  }
  return g.w ? g.w[j] : (typename T::w_t)1;
}

// entries of padding after every CSR neighbour / weight array (vector row loads)
constexpr int64_t CSR_PAD = 16;

// split hub decide (bin 8): row chunk of one aggregation workgroup, table
// chunk of one scan workgroup, and the scan's per-chunk partial result
constexpr int HUB_CH = 4096, HUB_TCH = 4096;
template <class T>
struct HubPart {
  typename T::gain_t h;
  typename T::acc_t wc;
  int32_t d;
};
// capacities for a graph of m directed entries: bin-8 hubs have >= 8192
// entries and a next_pow2(2d) < 4d slot table
inline int64_t hub_plan_elems(int64_t m) { return 2 * (m / 8192 + 2); }
inline int64_t hub_part_elems(int64_t m) { return 4 * m / HUB_TCH + m / 8192 + 2; }

template <class T>
struct DevState {
  typename T::tot_t *K, *Self, *Tot, *In;
  int32_t *size, *prev, *cur;
  uint8_t* active;
  uint8_t* dirty;    // In must be recomputed (the vertex or a neighbour moved)
  int8_t* binid;
  BinCuts cuts = {{BIN_LO[0], BIN_LO[1], BIN_LO[2], BIN_LO[3], BIN_LO[4], BIN_LO[5], BIN_LO[6], BIN_LO[7],
                   BIN_LO[8]}};  // degree bin cut points (binid)
  int64_t* hub_off;  // per-vertex offset into the global hash arena (bin 8), in slots
  int32_t* hub_keys;
  typename T::acc_t* hub_vals;
  // split hubs: each row's (community of the neighbour, weight) pairs, resolved
  // once per sweep into the hub's arena-sized slice (not kept zero)
  int32_t* hub_rc = nullptr;
  typename T::acc_t* hub_rw = nullptr;
  int32_t* hub_plan = nullptr;  // split hub decide: chunk starts (hub_plan_elems)
  void* hub_part = nullptr;     // split hub decide: HubPart<T>[hub_part_elems]
  uint8_t* hub_fail = nullptr;  // community-split hub decide: per-hub overflow flags
  // longest row of the round graph's hub bins (host-side bound from the round
  // set-up; -1: unknown): lets the launcher skip the split-hub kernels that no
  // row of the round can need
  int64_t hub_dmax = -1;
  Counters* ctr;
  // learned-frontier features (tracked only when track != 0)
  int track;
  int iter;          // sweep number within the round (1 = first sweep)
  float* last_gain;  // normalised net gain of the best move at the vertex's last sweep
  int32_t* moved_it; // sweep in which the vertex last moved (-1: never this round)
  float* nbr_moved;  // fraction of neighbours that moved (dirty-row In updates of tracked runs)
  // batched graphs (block-diagonal union): graph of every vertex and the m2 of
  // every graph; null for a single graph
  const int32_t* gid = nullptr;
  const typename T::tot_t* m2g = nullptr;
  // community view {prev[u], size[prev[u]], Tot[prev[u]]} of every vertex,
  // rebuilt before each sweep (launch_build_cv); the ring, per-group hash and
  // pipelined kernels read it
  typename T::cv_t* cv = nullptr;
  BlockPart* part = nullptr;  // MAX_PARTS per-workgroup partials of k_apply / k_prune
  // {cur, prev} of every vertex for the sweep in progress (null: not kept):
  // the commit of a swept vertex writes it, accept_sweep resets it to
  // {cur, cur}, so the incremental In update reads both with one gather
  int2* cp = nullptr;
  // the same in 4 bytes for the neighbour gathers of the incremental In
  // update: cur, with bit 31 set while the vertex has moved in the sweep in
  // progress (half the bytes of cp, so twice the vertices per cached line)
  int32_t* cm = nullptr;
  // integer path: each community's Tot and size deltas of a sweep packed in
  // one word (Tot delta * 2^32 + size delta, two's complement), so a
  // (community, tile) sum is one global atomic instead of two; k_apply adds
  // them to Tot and size and leaves the words zero (null: separate atomics)
  unsigned long long* tsd = nullptr;
  // profiling (per-bin timing): two events recorded around the bin-8 split's
  // part kernel, which does all of that bin's row work (null: not recorded)
  hipEvent_t* kev = nullptr;
};

// Per-graph state of a batch (one entry per graph; integer path).
struct BatchDev {
  int32_t G;
  const int32_t* m2g;            // m2 of every graph
  uint8_t* live;                 // still iterating in this round
  uint8_t* accept;               // latest sweep accepted (prev = cur)
  uint8_t* done;                 // round loop finished
  double* qprev;                 // Q of the accepted assignment of the round
  double* cur_mod;               // Q at the end of the latest round
  int32_t* iters;                // iterations of the current round
  int32_t* tot_iters;            // iterations over all rounds
  int32_t* rounds;               // rounds run
  unsigned long long* g_in;      // per-graph sum In   (zero between sweeps)
  unsigned long long* g_tot2;    // per-graph sum Tot^2 (zero between sweeps)
  int32_t* nlive;                // graphs still iterating after the latest sweep
};

// K, Self, Tot = K, size = 1, prev = cur = v, degree bins; K and Self of the
// rows of 8 entries and more come from launch_init_rows, once the bins are
// partitioned (order, bstart: the bin-grouped vertices; caps: host bin sizes).
template <class T>
void launch_init(const DevGraph<T>& g, DevState<T>& st, hipStream_t s, bool degree_only = false);
template <class T>
void launch_init_rows(const DevGraph<T>& g, DevState<T>& st, const int32_t* order, const int32_t* bstart,
                      const int32_t* caps, hipStream_t s);

// decide-and-move for one bin over front[bstart[b] .. bstart[b+1]) (device counts);
// `cap` = host upper bound of that range (used only to size the grid).
// cv[u] = {prev[u], size[prev[u]], Tot[prev[u]]} for u < n, before a sweep.
// Vertex lists: the O(n) passes below take an optional device list of the
// vertices to visit (vlist[0 .. *vcnt)); without one they visit 0..n-1.  The
// batched path passes the vertices of the graphs still iterating.
template <class T>
void launch_build_cv(DevState<T>& st, int32_t n, hipStream_t s, const int32_t* vlist = nullptr,
                     const int32_t* vcnt = nullptr);

// variant: kernel variant of the bin (0 = default); see decide_variants().
template <class T>
void launch_decide(int bin, const DevGraph<T>& g, DevState<T>& st, const int32_t* front,
                   const int32_t* bstart, int cap, typename T::tot_t m2, hipStream_t s, int variant = 0);
int decide_variants(int bin);
int decide_first_variants(int bin);

// First sweep of a round (singleton partition, duplicate-free rows): no
// aggregation needed, every entry is its own candidate community.
template <class T>
void launch_decide_first(int bin, const DevGraph<T>& g, DevState<T>& st, const int32_t* front,
                         const int32_t* bstart, int cap, typename T::tot_t m2, hipStream_t s, int variant = 0);

// Tot and size take the sweep's moves (from the {cur, prev} pairs, tile sums
// per community), then sums of Tot^2, min Tot, non-empty count.
// Tot/size take the deltas of the movers among dlist[0 .. *dcnt) (the swept
// vertices: only they can have moved; every vertex when null, else vlist),
// then sum Tot^2, min Tot and the non-empty count over the communities of
// vlist[0 .. *vcnt) (all n when null).  moves: false when no vertex has moved
// since Tot and size were set (skips the delta pass).
template <class T>
void launch_apply(DevState<T>& st, int32_t n, typename T::tot_t m2, hipStream_t s, const int32_t* vlist = nullptr,
                  const int32_t* vcnt = nullptr, const int32_t* dlist = nullptr, const int32_t* dcnt = nullptr,
                  bool moves = true);

// In[v] over order[bstart[b] .. bstart[b+1]) (only the dirty vertices when only_dirty).
// variant 1: pipelined kernels for bins 0-4 (same result).
template <class T>
void launch_weight(int bin, const DevGraph<T>& g, DevState<T>& st, const int32_t* order,
                   const int32_t* bstart, int cap, bool only_dirty, hipStream_t s, int variant = 0);
// Incremental In update of the integer path: the rows of the vertices that
// moved (cur != prev) over order[bstart[b] .. bstart[b+1]); a mover rewrites
// its own In and adds the change of every shared entry to the In of each
// neighbour that stayed.  Decide-and-move needs no dirty marks for it.
template <class T>
void launch_weight_moved(int bin, const DevGraph<T>& g, DevState<T>& st, const int32_t* order,
                         const int32_t* bstart, int cap, hipStream_t s, int variant = 0);
// prune flags, sum In, active count; clears the dirty marks.
template <class T>
void launch_prune(DevState<T>& st, int32_t n, typename T::tot_t m2, hipStream_t s, const int32_t* vlist = nullptr,
                  const int32_t* vcnt = nullptr);

// ---- batched graphs (kernels_batch.hip) ----
void launch_graph_sums(const int32_t* gid, const int8_t* binid, const int32_t* In, const int32_t* Tot, int32_t n,
                       unsigned long long* g_in, unsigned long long* g_tot2, hipStream_t s,
                       const int32_t* vlist = nullptr, const int32_t* vcnt = nullptr);
void launch_batch_step(const BatchDev& b, const int32_t* it_dev, double min_mod, int max_it, hipStream_t s);
void launch_batch_mask(const BatchDev& b, const int32_t* gid, const int32_t* cur, int32_t* prev, uint8_t* active,
                       int32_t n, bool first, int32_t* it_dev, int2* cp, int32_t* cm, hipStream_t s,
                       const int32_t* vlist = nullptr,
                       const int32_t* vcnt = nullptr, uint8_t* vlive = nullptr);
void launch_batch_live(const BatchDev& b, int32_t* it_dev, hipStream_t s);
void launch_batch_round(const BatchDev& b, int round, double min_mod, int max_rounds, int32_t* nopen, hipStream_t s);
void launch_batch_gid_next(const int32_t* gid, const int32_t* cnt, const int32_t* nid, int32_t n, int32_t* gid_next,
                           hipStream_t s);
void launch_batch_count(const int32_t* gid, int32_t n, int32_t* per_graph, hipStream_t s);
void launch_batch_reset(Counters* c, long long min_init, int32_t* nlive, hipStream_t s);

// ---- learned frontier (kernels_learned.hip) ----
// Features of vertex v before sweep it (all in [-1, 1]):
//   f0 log2(1 + degree)/16, f1 In/K (weight fraction inside its community),
//   f2 normalised net gain of its best move at its last sweep,
//   f3 moved in sweep it-1, f4 log(1 + size of its community) / log(1 + n),
//   f5 min(it, 16)/16.
// All of them are O(1) per vertex, so tracked runs keep the incremental In
// update (no row rescans).
template <class T>
void launch_mlp_filter(const DevGraph<T>& g, DevState<T>& st, const int32_t* front, int nfront, int it,
                       const float* prm, const short* w1b, float tau, int dmin, uint8_t* keep, hipStream_t s);
template <class T>
void launch_features(const DevGraph<T>& g, DevState<T>& st, const int32_t* front, int nfront, int it, float* out,
                     hipStream_t s);
void launch_labels(const int32_t* front, int nfront, const int32_t* moved_it, int it, uint8_t* lab, hipStream_t s);
void launch_mlp_forward(const float* X, int n, const float* prm, const short* w1b, float* p, hipStream_t s);
void launch_mlp_grad(const float* X, const uint8_t* y, const int32_t* idx, int n, const float* prm, const short* w1b,
                     float pos_w, float inv_wsum, float* grad, float* loss, hipStream_t s);
void launch_adam(float* prm, float* m, float* v, const float* grad, float lr, int step, short* w1b, hipStream_t s);
void launch_w1_bf16(const float* prm, short* w1b, hipStream_t s);

// ---- contraction (kernels_contract.hip) ----
// One chunk of a dense row's staged segment (see kernels_contract.hip).
struct DenseChunk {
  int32_t slot, c;  // dense row index (ascending community), community
  int64_t start;    // offset inside the community's staged segment
  int32_t len;      // staged entries of the chunk (<= LT/2)
  int32_t first;    // index of the row's first chunk
  int32_t nch;      // chunks of the row
  int32_t pad;
};

struct ContractPlan {
  int32_t caps[NBINS];  // vertices per degree bin of the graph being contracted
  int32_t tstart[NBINS + 1];  // host copy of the row length tier starts
  int32_t C;      // communities (rows of the contracted graph)
  int32_t nbig;   // rows with 4096 < L <= LT staged entries
  int64_t slots;  // global overflow slots those rows need
  int32_t ndense; // rows with L > LT (dense accumulation)
  std::vector<int32_t> dense_c;     // their communities, ascending
  std::vector<DenseChunk> chunks;   // their segments cut into LT/2-entry chunks
  int64_t dense_entries = 0;        // staged entries of all dense rows
  // rows of up to LT entries leave the merge in ascending neighbour order
  // (the order preprocess writes); otherwise in hash-table order.  Louvain's
  // decisions do not depend on the order of a row's entries.
  bool canonical = true;
};

// Bytes of device scratch the dense-row sort of a plan needs.
int64_t dense_scratch_bytes(int64_t nchunks, int64_t nrows, int64_t entries, bool fp);

template <class T>
struct ContractBufs {
  int32_t* cnt;       // [n]   community sizes
  int32_t* nid;       // [n]   new id of community (exclusive scan of cnt > 0)
  int32_t* label;     // [n]   new id of every vertex's community
  const int32_t* order;   // the graph's vertices grouped by degree bin
  const int32_t* bstart;  // [NBINS+1] bin offsets into order (device)
  int64_t* ecnt;      // [n+1] entries per new community
  int64_t* eoff;      // [n+1]
  int64_t* ecur;      // [n]
  int32_t* skey;      // [m]   staging keys
  typename T::acc_t* sval;  // [m]   staging values
  int32_t* dcnt;      // [n+1] distinct neighbours per new community
  int64_t* noff;      // [n+1]
  int32_t* big_list;  // [2n]  big rows in [0, n), dense rows from n
  int32_t* nbig;      // [2]   counts of big and dense rows
  int64_t* big_tab;   // [n+1] arena offsets (slots) of each large community's table
  int32_t* arena_keys;            // big-row overflow slices (all-zero between uses)
  typename T::acc_t* arena_vals;
  int64_t arena_slots;
  DenseChunk* chunks;             // chunk table of the dense rows
  char* dscratch;                 // dense_scratch_bytes() of scratch for their sort
  int32_t* i32_tmp;
  int64_t* i64_tmp;
  int32_t* i32_scalar;  // [4]
  int64_t* i64_scalar;  // [16] staging of the plan readback
  // communities grouped by row length tier (L <= 64, <= 512, <= 2048, <= 4096)
  int8_t* tier;       // [n]
  int32_t* tlist;     // [n]
  int32_t* tstart;    // [NBINS+1] (device)
  int32_t* part_tmp;  // partition scratch
  // pinned host scratch for the small readbacks (>= 64 bytes); dense-row
  // lists that do not fit go through pageable copies
  char* hpin;
  int64_t hpin_bytes;
  int64_t* mtotal;  // [16 * MTOTAL_SLOTS] entries of a gapped contracted graph: the sum of mtotal[16 k]
  int64_t* voff;    // [n] offset of every vertex's row inside its community's staging segment
};

template <class T>
struct ContractOut {
  int64_t* off;  // [C+1]
  int32_t* nbr;
  typename T::w_t* w;
  // non-null: gapped output.  The merged rows stay where they were staged (the
  // caller pointed the staging arrays and the segment offsets at the output
  // graph's nbr / w / off), end[c] = off[c] + distinct count, no compaction.
  int64_t* end = nullptr;
};

// Renumber (nid, label), per-community entry counts and offsets, long-row
// classification and their hash-table offsets.  Fills *plan (one sync, two
// with dense rows).
template <class T>
int contract_prepare(const DevGraph<T>& g, const int32_t* comm, ContractBufs<T>& cb, ContractPlan* plan,
                     hipStream_t s);
// Stage, merge rows in canonical order, write the new CSR into out.  Needs
// arena_slots >= plan.slots, dense_slots >= 1 (x C accumulators) when
// plan.ndense > 0, chunk/tile buffers sized for the plan.  Returns 0 or -1.
template <class T>
int contract_build(const DevGraph<T>& g, ContractBufs<T>& cb, const ContractPlan& plan,
                   ContractOut<T>& out, int64_t* m_out, hipStream_t s);

// final[v] = label[final[v]]
void compose_map(int32_t* fin, int32_t n0, const int32_t* label, hipStream_t s);

// ---- text edge list -> CSR on the device (ingest.hip) ----
struct IngestResult {
  int32_t n = 0;
  std::vector<int64_t> end;  // inclusive row ends
  std::vector<int32_t> nbr;
  std::vector<uint32_t> w;   // float32 or int32 bits, empty when unweighted
  double read_ms = 0, h2d_ms = 0, device_ms = 0, d2h_ms = 0, wall_ms = 0;
  int64_t bytes = 0;
};
int ingest_text_device(int device, hipStream_t s, const char* path, int weight_mode, IngestResult* out);

// ---- modularity of a given partition (modularity.hip) ----
int modularity_scratch_setup(int device);
int modularity_host(int device, hipStream_t s, int32_t n, const int64_t* end_incl, const int32_t* nbr, const void* w,
                    int wmode, const int32_t* comm, double* q_out);
int modularity_noalloc(int device, hipStream_t s, int32_t n, const int64_t* d_end_incl, const int32_t* d_nbr,
                       const void* d_w, int wmode, const int32_t* d_comm, double* q_out);

}  // namespace gala
