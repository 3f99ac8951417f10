// Internal declarations shared by the gfx950 kernels and the host engine.
// Not part of the C ABI (include/gala.h is).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gala {

// ---------------------------------------------------------------------------
// Degree bins for decide-and-move and weight update (wave64 re-cut of the
// reference's eight bins, src/louvain_gpu/louvain_main_process.cu:276-288).
//   bin 0: 1..3      4-lane sub-wave group, ring exchange
//   bin 1: 4..7      8-lane group
//   bin 2: 8..15     16-lane group
//   bin 3: 16..31    32-lane group
//   bin 4: 32..127   one wave per vertex, 256-slot LDS hash per wave
//   bin 5: 128..511  one wave per vertex, 1024-slot LDS hash per wave
//   bin 6: 512..4095 one 256-thread workgroup per vertex, 8192-slot LDS hash
//   bin 7: 4096..8191 one 1024-thread workgroup per vertex, 16384-slot LDS hash
//   bin 8: >= 8192   one 1024-thread workgroup per vertex, global hash table
// Degree-0 vertices belong to no bin (they are never swept).
// ---------------------------------------------------------------------------
constexpr int NBINS = 9;
constexpr int BIN_LO[NBINS] = {1, 4, 8, 16, 32, 128, 512, 4096, 8192};

// Lowest degree of every bin.  The autotuner may lower the cut points of bins
// 4..8 (never raise them, never below the previous default cut + 1): every
// kernel of a bin then receives rows no longer than its default range allows
// (the LDS tables and per-thread row buffers are sized for that maximum) and
// handles shorter rows like any others; bins 0..3 keep their cuts, so rows
// below 8 entries still get K and Self from the init kernel.
struct BinCuts {
  int lo[NBINS];
};
constexpr int BIN_TUNE_FIRST = 4;  // first bin whose lower cut may move

__host__ __device__ inline int bin_of_degree(int64_t d, const BinCuts& c) {
  if (d <= 0) return -1;
  int b = NBINS - 1;
  while (b > 0 && d < c.lo[b]) b--;
  return b;
}

// cuts valid: bins 0..3 at their defaults; bin b >= 4 in (BIN_LO[b-1], BIN_LO[b]]
inline bool bin_cuts_valid(const int32_t* lo) {
  for (int b = 0; b < NBINS; b++) {
    if (b < BIN_TUNE_FIRST && lo[b] != BIN_LO[b]) return false;
    if (b >= BIN_TUNE_FIRST && (lo[b] <= BIN_LO[b - 1] || lo[b] > BIN_LO[b])) return false;
  }
  return true;
}

// First bin that uses the global hash arena (tables of 8-byte sums do not fit
// 16384 slots in LDS, so the wide path sends bin 7 to the arena as well).
template <class T>
constexpr int global_bin_min() { return T::WIDE ? 7 : 8; }

// Learned-frontier MLP: 6 features -> 64 hidden (ReLU) -> 1 logit.
// Parameter vector layout (fp32): W1[16][64] (feature rows 6..15 are zero
// padding for the 16-deep MFMA), b1[64], w2[64], b2.
constexpr int MLP_IN = 6, MLP_H = 64;
constexpr int MLP_W1 = 0, MLP_B1 = 16 * MLP_H, MLP_W2 = MLP_B1 + MLP_H, MLP_B2 = MLP_W2 + MLP_H;
constexpr int MLP_NPARAM = MLP_B2 + 1;

// Per-iteration device counters (one small struct read back per iteration).
// Decide-and-move counts, spread over CTR_SLOTS slots of 256 bytes: the
// workgroups of a launch add to slot (workgroup id % CTR_SLOTS), so their
// end-of-launch atomics land on different L2 lines instead of queueing on one.
constexpr int CTR_SLOTS = 32;
constexpr int MTOTAL_SLOTS = 32;  // slots of the contracted-entry count (one 128-byte line each)
struct alignas(256) DecideSlot {
  unsigned long long moved, swept, swept_edges, pad;
  unsigned long long bin_swept[NBINS];
  unsigned long long bin_edges[NBINS];
};

struct Counters {
  unsigned long long sum_in;     // sum over v of In[v]            (exact, <= m2 < 2^62)
  unsigned long long sum_tot2;   // sum over c of Tot[c]^2, low 64 bits (exact)
  unsigned long long sum_tot2_hi;// ... high 64 bits (wide path: Tot^2 up to 2^124)
  long long min_tot;             // min non-zero Tot (zero -> m2)
  unsigned long long moved;      // vertices that changed community this sweep
  unsigned long long swept;      // vertices evaluated by decide-and-move
  unsigned long long swept_edges;// row entries read by decide-and-move
  unsigned long long active;     // vertices kept in the next frontier
  int nonempty;                  // communities with size > 0
  int overflow;                  // hash-table overflow flag (must stay 0)
  double step_q;                 // Q of the sweep as the device's stopping rule saw it
  int step_accept;               // ... and its decision (1: accepted, prev = cur)
  int step_pad;
  unsigned long long bin_swept[NBINS];  // decide-and-move per degree bin
  unsigned long long bin_edges[NBINS];
  // written by the decide kernels; moved / swept / swept_edges / bin_* above
  // are their sums, formed on the host when the counters are read
  DecideSlot slot[CTR_SLOTS];
};

// Device state of a round's stopping rule when the host reads each sweep's
// counters one sweep late (the sweep after it is already queued).
struct RoundCtl {
  double qprev;  // Q of the last accepted sweep (of the singleton partition before sweep 1)
  int live;      // 1 while the round goes on; 0 once a sweep was refused
  int accept;    // the last sweep's decision: 1 = accepted (prev = cur before the next sweep)
};

// Per-workgroup partial results of the O(V) reductions (k_apply, k_prune).
// One workgroup combines them afterwards, so thousands of workgroups do not
// serialise on the same counter addresses.  Every sum is an exact integer.
struct BlockPart {
  unsigned long long s0;  // apply: sum Tot^2 (low 64 bits)  prune: sum In
  unsigned long long s1;  // apply: min non-zero Tot          prune: vertices kept active
  unsigned long long s0hi;// apply: sum Tot^2 (high 64 bits)
  unsigned long long pad0;
  long long i0;           // apply: non-empty communities
  long long pad[3];
};
constexpr int MAX_PARTS = 4096;  // workgroups of k_apply / k_prune
constexpr int64_t HPIN_BYTES = 1 << 20;  // pinned host scratch of the engine's small readbacks

// Community view of a vertex for one sweep: its community, that community's
// Tot and whether the community is a singleton, so one gather gives all
// three.  Integer path: 8 bytes, Tot in the low 31 bits of ts (Tot <= m2 <
// 2^31) and the singleton flag in its sign bit.  Wide path: 16 bytes.
struct alignas(8) CvI {
  int32_t c;
  int32_t ts;
};
struct alignas(16) CvW {
  int32_t c;
  int32_t size;
  int64_t tot;
};
__host__ __device__ inline int32_t cv_tot(const CvI& e) { return e.ts & 0x7fffffff; }
__host__ __device__ inline int64_t cv_tot(const CvW& e) { return e.tot; }
// 1 for a singleton community, some other positive size otherwise
__host__ __device__ inline int32_t cv_sz(const CvI& e) { return e.ts < 0 ? 1 : 2; }
__host__ __device__ inline int32_t cv_sz(const CvW& e) { return e.size; }
__host__ __device__ inline void cv_set(CvI& e, int32_t c, int32_t size, int32_t tot) {
  e.c = c;
  e.ts = size == 1 ? (int32_t)((uint32_t)tot | 0x80000000u) : tot;
}
__host__ __device__ inline void cv_set(CvW& e, int32_t c, int32_t size, int64_t tot) {
  e.c = c;
  e.size = size;
  e.tot = tot;
}

// Weight traits.  Both are exact integer arithmetic:
//  IntW  -- int32 weights and sums, int64 gains; m2 <= 1e9 (the fast path for
//           unweighted and int32-weighted graphs of up to 1e9 directed entries).
//  WideW -- int64 weights and sums, int128 gains; m2 < 2^62.  Float32 weights
//           run here as exact integers W = w * 2^k (k the smallest shift that
//           makes every weight integral; decisions and Q are invariant under
//           the common scale), as do integer graphs with m2 > 1e9.  The input
//           round may keep its float32 weights in HBM (DevGraph::wf) and scale
//           them in registers; contracted rounds carry int64 sums.
struct IntW {
  using w_t = int32_t;     // edge weight as stored
  using acc_t = int32_t;   // per-community neighbour weight sum (<= K < 2^31)
  using tot_t = int32_t;   // K, Self, Tot, In (all <= m2 < 2^31)
  using gain_t = long long;
  using cv_t = CvI;
  static constexpr bool WIDE = false;
};
struct WideW {
  using w_t = int64_t;
  using acc_t = int64_t;
  using tot_t = int64_t;
  using gain_t = __int128;
  using cv_t = CvW;
  static constexpr bool WIDE = true;
};
constexpr int64_t WIDE_M2_LIMIT = (int64_t)1 << 62;  // every sum <= m2 fits int64, every product int128

}  // namespace gala
