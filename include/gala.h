/*
 * gala.h -- C ABI of the gfx950 Louvain engine (libgala.so).
 *
 * Drop-in boundary for the single-GPU path of LinXi-lx/GALA-mg.  The reference
 * has no plugin/FFI layer; its in-process interface is C++ over thrust vectors:
 *   double louvain_gpu(Graph&, vertex_t* h_community, double min_modularity,
 *                      int gpu_id, int gpu_num)          src/louvain_gpu/louvain.cuh:8
 *   louvain_main_process(...)  (phase 1, one round)       src/louvain_gpu/louvain.cuh:10-17
 *   build_compressed_graph(...) (phase 2, contraction)    src/louvain_gpu/louvain.cuh:19-22
 *   Graph::load_bin_graph / store_bin_graph (.bin format) src/graph/graph.cpp:153-219
 * Each entry point below names the reference interface it replaces.
 *
 * Conventions: every function returns 0 on success or a negative GALA_E_* code;
 * the caller owns all host buffers, the context owns device memory; a context
 * is single-threaded (independent contexts may run concurrently).  No C++ or
 * HIP types cross this boundary.
 */
#ifndef GALA_H
#define GALA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GALA_OK 0
#define GALA_E_IO -1        /* file missing / unreadable / short read */
#define GALA_E_FORMAT -2    /* malformed .bin or CSR (offsets, ids, sizes) */
#define GALA_E_OOM -3       /* device or host allocation failed */
#define GALA_E_HIP -4       /* HIP runtime error / no device */
#define GALA_E_OVERFLOW -5  /* exact-integer range exceeded (m2 > 1e9) or hash overflow */
#define GALA_E_INVALID -6   /* bad argument or state (no graph, m2 == 0, ...) */
#define GALA_E_INTERNAL -7  /* internal consistency check failed (a bug: please report) */

/* weight modes of the .bin weight section (out of band, as the reference's -w flag) */
#define GALA_W_NONE 0 /* no weight section: every weight is 1 (graph.cpp:205-212) */
#define GALA_W_I32 1  /* int32 weights (graph.cpp:201-204) */
#define GALA_W_F32 2  /* IEEE-754 float32 weights in the same slot, run in exact
                         arithmetic on their values (scaled to integers) */

typedef struct gala_ctx gala_ctx;

typedef struct {
  double min_modularity; /* convergence threshold, 1e-6 in gala_main.cu:48 */
  int max_iterations;    /* per-round sweep cap, 0 = 100000 */
  int max_rounds;        /* 0 = until the round's dQ < min_modularity */
  int dump_iterations;   /* 1: iteration callback receives the community vector */
  int dump_rounds;       /* 1: round callback receives the contracted CSR */
  int verbose;           /* 1: print the reference's stdout lines from gala_run */
  int frontier_mode;     /* 0 plain (sound prune), 1 learned frontier + certification */
  float learned_tau;     /* learned frontier threshold on the move probability */
  int profile_bins;      /* 1: time every degree bin's decide launch with HIP events */
  int no_hip_graphs;     /* batched runs replay each round's iteration body from a
                            captured hipGraph; 1 = launch kernel by kernel instead */
  int serial_bins;       /* single-graph runs launch the long-row degree bins (512+)
                            on a second stream beside the others; 1 = one stream */
  int learned_min_degree;/* learned sweeps keep every vertex with fewer row entries
                            than this (0: the network filters every vertex) */
  int learned_min_vertices; /* rounds whose graph has fewer vertices run plain
                               sweeps only (0: every round is filtered) */
  int queue_sweeps;      /* 1: plain rounds queue each sweep before the previous
                            one's counters are read, the stopping rule applied on
                            the device (0: the host reads every sweep first) */
  int reserved[2];
} gala_opts;

/* Called after every iteration (Iteration:0 is Q before any move).
 * cur is a host copy of the proposed assignment when dump_iterations=1, else NULL. */
typedef void (*gala_iter_cb)(void* user, int round, int iteration, double q, int64_t n_moved,
                             const int32_t* cur, int32_t n);

/* Called after each round's contraction.  comm = accepted (pre-renumbering)
 * assignment of the round graph (n entries); the contracted CSR (C rows,
 * inclusive ends, int64 weights) is given when dump_rounds=1, else NULLs. */
typedef void (*gala_round_cb)(void* user, int round, double q, int32_t n, const int32_t* comm,
                              int32_t C, const int64_t* end_incl, const int32_t* nbr,
                              const int64_t* w_i64, const double* w_f64);

typedef struct {
  double modularity;       /* final Q */
  int32_t n_vertices;      /* input vertices */
  int32_t n_communities;   /* final communities */
  int32_t rounds;
  int32_t total_iterations;/* sweeps over all rounds (including reverted ones) */
  double t_total_ms;       /* graph resident in HBM -> final composed map on device */
  double t_decide_ms;      /* decide-and-move */
  double t_update_ms;      /* Tot/size apply + weight update + modularity + frontier */
  double t_contract_ms;    /* contraction incl. composition */
  double t_h2d_ms;         /* upload of the input graph (outside t_total_ms) */
  int64_t swept_vertices;  /* vertices evaluated by decide-and-move */
  int64_t swept_edges;     /* row entries read by decide-and-move */
  int64_t input_edges;     /* undirected input edges (directed entries / 2) */
  int64_t swept_vertices_unit; /* ...of which on graphs with implicit unit weights */
  int64_t swept_edges_unit;
  int64_t decide_bytes;    /* algorithmic bytes of all decide-and-move sweeps:
                              16d+40 per swept vertex of degree d (12d+40 with implicit
                              unit weights); first sweeps of a round need no prev[]
                              gathers: 12d+40 (8d+40 unit) */
  double round_q[64];
  int32_t round_nc[64];
  int32_t round_iters[64];
  double round_ms[64];        /* phase 1 of the round (louvain_main_process) */
  double round_decide_ms[64];
  double round_update_ms[64];
  double round_build_ms[64];  /* contraction of the round */
  int32_t learned_sweeps;      /* sweeps run on the learned frontier */
  int32_t certify_sweeps;      /* stalled learned sweeps certified by sweeping the rest of the frontier */
  int64_t bin_decide_bytes[9]; /* decide-and-move algorithmic bytes per degree bin */
  int64_t bin_swept_edges[9];
  double bin_decide_ms[9];     /* per-bin decide time (profile_bins=1), HIP events */
  int64_t learned_frontier;    /* plain-frontier vertices at the learned sweeps */
  int64_t learned_kept;        /* of those, vertices the network kept (swept) */
  /* The bin_* arrays above count the general sweeps; these the first sweep of
   * each round (singleton partition, streaming-argmax kernels). */
  int64_t bin_first_bytes[9];
  int64_t bin_first_edges[9];
  double bin_first_ms[9];      /* profile_bins=1 */
  int32_t bin_launches[9];     /* decide launches per bin, general sweeps */
  int32_t bin_first_launches[9];
  /* Arithmetic of the run.  wide = 0: int32 sums, int64 gains (m2 <= 1e9).
   * wide = 1: int64 sums, int128 gains (m2 < 2^62): float32 weights run as the
   * exact integers w * 2^weight_shift, as do integer graphs with m2 > 1e9.
   * Both are exact: Q and every decision are those of exact arithmetic. */
  int32_t wide;
  int32_t weight_shift;
  /* Per round: vertices and directed entries of the round graph, entries of
   * the graph its contraction produced (the contraction's byte count). */
  int32_t round_vertices[64];
  int64_t round_entries[64];
  int64_t round_contracted_entries[64];
  int32_t certify_empty;  /* of the certify_sweeps, those in which the rest of the
                             frontier moved nobody (their update pass is skipped) */
  int32_t hub_part_launches; /* profile_bins=1: general sweeps that ran the bin-8
                                split's part kernel, and its own time */
  double hub_part_ms;
  int32_t queued_rounds;  /* rounds run with the device-side stopping rule (sweep k+1
                             queued before sweep k's counters are read) */
  int32_t inert_sweeps;   /* sweeps queued after the round's last one: their frontier
                             is emptied on the device, so they move nobody */
} gala_result;

/* Version string of the library. */
const char* gala_version(void);
const char* gala_strerror(int code);
void gala_default_opts(gala_opts* o);

/* Create a context on HIP device `device` (replaces cudaSetDevice + thrust
 * state set-up, louvain.cu:149-173).  opts may be NULL for defaults. */
int gala_create(gala_ctx** out, int device, const gala_opts* opts);
void gala_destroy(gala_ctx* ctx);

/* Load a reference .bin file (Graph::load_bin_graph, graph.cpp:177-219) and
 * upload it (GraphGPU, graph_gpu.h:13-19).  weight_mode = GALA_W_*. */
int gala_load_bin(gala_ctx* ctx, const char* path, int weight_mode);

/* Same from host arrays: n vertices, end_incl[n] inclusive row ends (the .bin
 * layout), nbr[end_incl[n-1]], w = NULL or int32/float32 per weight_mode. */
int gala_set_graph(gala_ctx* ctx, int32_t n, const int64_t* end_incl, const int32_t* nbr,
                   const void* w, int weight_mode);

/* Multi-round Louvain (louvain_gpu, louvain.cu:136-229 with the intended round
 * loop of :193-226).  Callbacks may be NULL. */
int gala_run(gala_ctx* ctx, gala_iter_cb icb, gala_round_cb rcb, void* user, gala_result* out);

/* Result so far (valid inside a round callback for the rounds completed). */
int gala_partial_result(gala_ctx* ctx, gala_result* out);

/* Composed final community of every input vertex (the reference's
 * h_community output, louvain.cu:226), renumbered 0..C-1. */
int gala_get_communities(gala_ctx* ctx, int32_t* out_n0);

/* ---- phase-level entry points (host arrays in and out), for parity tests ---- */

/* One synchronous decide-and-move sweep (the build_and_select* family,
 * kernel_functions.cu:289-887) over the vertices with frontier[v] != 0
 * (frontier NULL = all), on a given state.  int32 weights (NULL = unit). */
int gala_phase_decide(gala_ctx* ctx, int32_t n, const int64_t* end_incl, const int32_t* nbr,
                      const int32_t* w, int64_t m2, const int32_t* K, const int32_t* Self,
                      const int32_t* Tot, const int32_t* size, const int32_t* prev,
                      const uint8_t* frontier, int32_t* cur_out);

/* Weight update + prune (compute_In / compute_In_blk, kernel_functions.cu:889-1034)
 * for all vertices: In_out[v], active_out[v]. */
int gala_phase_weight(gala_ctx* ctx, int32_t n, const int64_t* end_incl, const int32_t* nbr,
                      const int32_t* w, int64_t m2, const int32_t* K, const int32_t* Self,
                      const int32_t* Tot, const int32_t* cur, int64_t min_tot, int32_t* In_out,
                      uint8_t* active_out);

/* Contraction (build_compressed_graph, build_compressed_graph.cu:5-173).
 * Two calls: with out arrays NULL it returns C and the entry count through
 * out_C / out_m; then with arrays of those sizes it fills them.  new_id[n]
 * receives the renumbering (-1 for empty communities). */
int gala_phase_contract(gala_ctx* ctx, int32_t n, const int64_t* end_incl, const int32_t* nbr,
                        const int32_t* w, const int32_t* comm, int32_t* out_C, int64_t* out_m,
                        int32_t* new_id, int64_t* out_end_incl, int32_t* out_nbr, int64_t* out_w);

/* Text edge list -> CSR on the GPU, with the line rules and output of
 * gala_ingest_text (host IO section below); replaces Graph::Graph(file, is_wt),
 * graph.cpp:18-151, the step the preprocess CLI runs (preprocess.cpp:43): the file is streamed
 * into HBM, newlines are found, lines parsed, ids renumbered, entries sorted
 * and deduplicated there; the output equals gala_ingest_text's byte for byte.
 * Two-call protocol: a call with NULL arrays ingests and keeps the CSR in the
 * context; a call with arrays copies it out.  Up to 2^30 edge lines. */
int gala_ingest_text_device(gala_ctx* ctx, const char* path, int weight_mode, int32_t* n, int64_t* m,
                            int64_t* end_incl, int32_t* nbr, void* w);
/* Device ingest of a text edge list, then set as the context's graph (as gala_load_bin). */
int gala_load_text(gala_ctx* ctx, const char* path, int weight_mode);
/* Timings of the last device ingest (any pointer may be NULL): host file read,
 * file-to-HBM pipeline (read + copy), device passes (parse to dedup, including
 * their allocations), copy-out, and the whole call. */
int gala_ingest_times(gala_ctx* ctx, double* read_ms, double* load_ms, double* device_ms, double* d2h_ms,
                      double* wall_ms);

/* Modularity of the partition comm[n] (community ids in [0, n)) of a graph
 * (modularity_op, kernel_functions.cuh:77-87, as louvain_main_process.cu:
 * 538-548,990-999 evaluate it): Q = sum In / m2 - sum Tot^2 / m2^2 from exact
 * integer sums (float32 weights scaled to integers as gala_run does), so it
 * equals gala_run's Q for the same partition bit for bit.  Host arrays;
 * GALA_E_INVALID for an id out of range or a non-positive weight,
 * GALA_E_OVERFLOW when the scaled m2 reaches 2^62; an edgeless graph gives 0. */
int gala_phase_modularity(gala_ctx* ctx, int32_t n, const int64_t* end_incl, const int32_t* nbr, const void* w,
                          int weight_mode, const int32_t* comm, double* q_out);

/* The same from DEVICE arrays (on ctx's device, read on ctx's stream), with
 * no allocation of any kind: no counterpart in the reference; for callers
 * that cannot allocate device memory (serving loops, code under stream
 * capture, harnesses that forbid hipMalloc).  Its scratch is a static device
 * array of the library, one copy per GPU, sized for n <= GALA_NOALLOC_MAX_N
 * (32 MiB of per-community totals plus 128 KiB of partials), held until the
 * process ends; larger n gives GALA_E_INVALID.  Every call on one GPU uses
 * that GPU's copy: calls that overlap on one GPU (several threads, several
 * contexts or streams) race and must be serialised by the caller.  Blocks
 * until q_out is written. */
#define GALA_NOALLOC_MAX_N (1 << 22)
int gala_modularity_noalloc(gala_ctx* ctx, int32_t n, const int64_t* d_end_incl, const int32_t* d_nbr,
                            const void* d_w, int weight_mode, const int32_t* d_comm, double* q_out);

/* ---- learned frontier (no counterpart in the reference) ----
 * A 6 -> 64 -> 1 MLP (BF16 MFMA, FP32 accumulation, FP32 master weights)
 * predicts which vertices the next sweep moves; with opts.frontier_mode = 1 a
 * sweep only visits the predicted movers among the prune-active vertices.  A
 * learned sweep that does not raise Q by min_modularity is certified: the rest
 * of the frontier is swept against the same pre-sweep state, which completes
 * it into the plain sweep, and the plain break/revert rule decides.  A round
 * therefore only ends on the outcome of a plain sweep.  Parameter
 * vector: GALA_MLP_NPARAM floats, W1[16][64] (rows 6..15 zero), b1[64],
 * w2[64], b2.  Features (6 per vertex) are defined in gala_kernels.h. */
#define GALA_MLP_NFEAT 6
#define GALA_MLP_NPARAM 1153

/* Move logs of a plain run: before every sweep after the first, the features
 * of the swept vertices and, after it, whether each one moved. */
typedef void (*gala_log_cb)(void* user, int round, int iteration, int32_t n, const float* features,
                            const uint8_t* moved);
int gala_set_logger(gala_ctx* ctx, gala_log_cb cb, void* user);
int gala_set_mlp(gala_ctx* ctx, const float* params);
int gala_get_mlp(gala_ctx* ctx, float* params);
/* Train on (X[n][6], y[n]) with the device's own forward/backward kernels and
 * Adam; params_io holds the initial weights when init_given, else they are
 * drawn from seed; receives the trained weights, which the context keeps.
 * loss_out[epochs] (optional) receives the mean weighted BCE per epoch. */
int gala_mlp_train(gala_ctx* ctx, const float* X, const uint8_t* y, int64_t n, int epochs, int batch, float lr,
                   float pos_weight, uint64_t seed, int init_given, float* params_io, float* loss_out);
/* Move probabilities for X[n][6] under params. */
int gala_mlp_forward(gala_ctx* ctx, const float* params, const float* X, int64_t n, float* p_out);
/* Gradient of the weighted BCE over (X, y) at params (no update). */
int gala_mlp_grad(gala_ctx* ctx, const float* params, const float* X, const uint8_t* y, int64_t n, float pos_weight,
                  float* grad_out, float* loss_out);

/* ---- decide kernel variants and the autotuner (no counterpart in the
 * reference, whose bins and block sizes are fixed) ----
 * Every degree bin has one or more decide-and-move kernels that take exactly
 * the same decisions and differ in how lanes map to vertices.  variants[0..8]
 * select the kernel of bins 0..8 for general sweeps, variants[9..17] for the
 * first sweep of a round, variants[18..26] the weight-update kernel.
 * gala_autotune times every variant on the loaded graph (full plain runs with
 * per-bin HIP events) and keeps the fastest for later runs of the context;
 * results are unchanged by construction. */
#define GALA_NVARIANTS 27
int gala_set_variants(gala_ctx* ctx, const int32_t* variants);
int gala_get_variants(gala_ctx* ctx, int32_t* variants);
/* Chooses the variants, then the degree-bin cut points (gala_autotune_bins). */
int gala_autotune(gala_ctx* ctx, int32_t* variants_out);
/* Degree-bin cut points: lo[b] is the lowest degree of bin b (GALA_NBINS
 * entries; defaults 1, 4, 8, 16, 32, 128, 512, 4096, 8192 -- the reference
 * cuts its bins at fixed degrees, louvain_main_process.cu:276-288).  Bins 0-3
 * keep their cuts; the cut of bin b >= 4 may be lowered down to one above the
 * default cut of bin b - 1 (each kernel's tables are sized for its bin's
 * default range, so no bin may receive longer rows): GALA_E_INVALID
 * otherwise.  gala_autotune_bins picks them by timed plain runs on the loaded
 * graph (coordinate descent, bins 8 down to 4, cuts at 1, 7/8 .. 1/2 of the
 * default); results are unchanged by construction. */
#define GALA_NBINS 9
int gala_set_bin_cuts(gala_ctx* ctx, const int32_t* lo);
int gala_get_bin_cuts(gala_ctx* ctx, int32_t* lo);
int gala_autotune_bins(gala_ctx* ctx, int32_t* lo_out);
/* Number of kernel variants of every slot (GALA_NVARIANTS entries, same order
 * as gala_set_variants): a valid variant of slot i is 0 .. counts[i]-1.  Needs
 * no device and no context. */
int gala_variant_counts(int32_t* counts_out);

/* ---- batched graphs (no counterpart in the reference, which runs one graph
 * per process) ----
 * Many independent graphs run as one block-diagonal union: graph g owns input
 * vertices [vstart[g], vstart[g+1]) and its rows may only name vertices of the
 * same range.  Every graph gets exactly the result of its own gala_run (same
 * per-iteration decisions, break/revert, round loop and Q), while the sweeps,
 * updates and contractions of all graphs share launches.  Integer weights
 * only (GALA_W_NONE or GALA_W_I32); each graph's m2 must be <= 1e9.  A graph
 * with no edge keeps every vertex alone and reports Q = 0 and 0 rounds. */
typedef struct {
  int32_t n_graphs;
  int32_t rounds;            /* rounds of the union (the longest graph) */
  int32_t iterations;        /* iterations of the union, summed over rounds */
  int32_t reserved0;
  double t_total_ms;         /* device time of the run, HIP events */
  double t_phase1_ms;        /* sweeps + updates, all rounds */
  double t_contract_ms;      /* contractions, all rounds */
  int64_t swept_vertices;    /* decide-and-move work over all graphs */
  int64_t swept_edges;
  int32_t round_open[64];    /* graphs still running in each round */
  int32_t round_iters[64];   /* iterations of the union in each round */
  double round_ms[64];
} gala_batch_result;

/* Upload a batch (the union in the gala_set_graph layout plus vstart[n_graphs+1]). */
int gala_batch_set(gala_ctx* ctx, int32_t n_graphs, const int32_t* vstart, int32_t n, const int64_t* end_incl,
                   const int32_t* nbr, const void* w, int weight_mode);
int gala_batch_run(gala_ctx* ctx, gala_batch_result* out);
/* Per graph: final Q, communities, rounds and iterations (any may be NULL). */
int gala_batch_graph_results(gala_ctx* ctx, double* modularity, int32_t* n_communities, int32_t* rounds,
                             int32_t* iterations);
/* Final community of every union vertex, numbered 0..C_g-1 within its graph
 * (the numbering gala_get_communities gives for that graph alone). */
int gala_batch_communities(gala_ctx* ctx, int32_t* comm);

/* ---- .bin IO and generators (host only; no device needed) ---- */

/* Read a .bin: first call with NULL arrays returns n and m (directed entries). */
int gala_read_bin(const char* path, int weight_mode, int32_t* n, int64_t* m, int64_t* end_incl,
                  int32_t* nbr, void* w);
/* Write a .bin (Graph::store_bin_graph, graph.cpp:153-175); w may be NULL. */
int gala_write_bin(const char* path, int32_t n, const int64_t* end_incl, const int32_t* nbr,
                   const void* w, int weight_mode);
/* Text edge list -> CSR with the reference preprocess semantics (graph.cpp:18-151):
 * both directions pushed, per-row dedup in ascending order (weights summed when
 * weighted), dense renumbering of the linked vertices in ascending original id.
 * Self-loops are kept as one entry (the reference truncates them).
 * Line rules (shared with gala_ingest_text_device):
 *   - the file is split at '\n'; an empty line or one starting with '#' is skipped;
 *   - tokens are separated by ' ', '\t', '\r', '\v', '\f';
 *   - an edge line's first two tokens are integers [+-]?[0-9]+; a line whose
 *     first two tokens are not both integers is skipped (a header, a '%' line);
 *   - an id below 0 or above INT32_MAX is GALA_E_FORMAT;
 *   - weighted: the third token is the weight (1 when absent), further tokens
 *     are ignored.  GALA_W_I32: an int32 integer token, summed in int64 per
 *     entry and range-checked to int32.  GALA_W_F32: a decimal
 *     [+-]?(D+(.D*)?|.D+)([eE][+-]?D+)? of at most 40 significant digits,
 *     rounded to the nearest double (ties to even); repeated entries are summed
 *     in float64 in line order and rounded once to float32.  A malformed weight,
 *     one that rounds to an infinity, or a sum out of range is GALA_E_FORMAT.
 * Two-call protocol as gala_read_bin. */
int gala_ingest_text(const char* path, int weight_mode, int32_t* n, int64_t* m, int64_t* end_incl,
                     int32_t* nbr, void* w);

/* Synthetic generators (deterministic for a seed).  Two-call protocol: with
 * NULL arrays they return n and m; the arrays are then filled on a second call
 * with the same arguments.  Rows are sorted and duplicate-free. */
int gala_gen_sbm(int32_t n, int32_t blocks, double p_in, double p_out, uint64_t seed, int32_t* out_n,
                 int64_t* out_m, int64_t* end_incl, int32_t* nbr);
int gala_gen_chung_lu(int32_t n, double gamma, double mean_deg, double max_deg, uint64_t seed,
                      int32_t* out_n, int64_t* out_m, int64_t* end_incl, int32_t* nbr);
/* A batch of SBMs (graph g as gala_gen_sbm(n[g], blocks[g], p_in[g], p_out[g],
 * seeds[g])) laid out as one union: vstart[n_graphs+1], vertex ids of graph g
 * offset by vstart[g].  threads = 0: up to 16 host threads.  Two-call protocol:
 * the first call fills vstart and out_m. */
int gala_gen_sbm_batch(int32_t n_graphs, const int32_t* n, const int32_t* blocks, const double* p_in,
                       const double* p_out, const uint64_t* seeds, int32_t threads, int32_t* vstart,
                       int64_t* out_m, int64_t* end_incl, int32_t* nbr);
int gala_gen_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                  float w_lo, float w_hi, int32_t* out_n, int64_t* out_m, int64_t* end_incl,
                  int32_t* nbr, float* w);

#ifdef __cplusplus
}
#endif

#endif /* GALA_H */
