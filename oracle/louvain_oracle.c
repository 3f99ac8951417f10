/*
 * louvain_oracle.c -- CPU restatement of GALA's single-GPU Louvain path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP path: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it,
 * and never as the product path.  It is plain C (gcc, optional OpenMP) and does
 * all gain / modularity arithmetic in exact integers so that a bit-exact
 * comparison with the GPU is meaningful.
 *
 * Semantics followed (reference = LinXi-lx/GALA-mg, paths relative to its root):
 *   init          src/louvain_gpu/kernel_functions.cu:34-73   K, Tot=K, Self, prev=v
 *   decide/move   src/louvain_gpu/kernel_functions.cu:289-441 (sub-warp),
 *                 :473-593 (warp hash), :640-813 (block hash), :102-221 (hashing)
 *     dq(d) = 2 w_d - 2 K (Tot_d - Tot_c + K) / m2 for d != c, argmax with ties to
 *     the lowest community id (:372,:524,:705,:733); gain = dq(best) - 2 w_c + 2 Self
 *     (:381,:532,:748); move iff gain > 0 unless size[best]==size[c]==1 and best > c
 *     (:386,:536,:754); Tot/size deltas (:397-400).  Reads are all pre-iteration
 *     (synchronous).  Here dq is scaled by m2/2:  H(d) = w_d m2 - K (Tot_d - Tot_c + K),
 *     and the move test is H(best) - m2 (w_c - Self) > 0  -- exact in int128 with
 *     every weight sum in int64 while m2 < 2^62 (checked).
 *   Tot update    src/louvain_gpu/louvain_main_process.cu:949-956 and Tot_op
 *                 kernel_functions.cuh:112-121 (min over non-zero Tot, zero -> m2)
 *   weight update src/louvain_gpu/kernel_functions.cu:889-1034
 *     In[v] = sum w(v,u) over cur[u]==cur[v]; prune (active=0) iff
 *     m2 (2 In - S - K) + K (minTot - (Tot_c - K)) > 0  (:944, :1028, scaled by m2)
 *   modularity    kernel_functions.cuh:77-87, louvain_main_process.cu:541-545,990-999
 *     Q = sum In / m2 - sum Tot^2 / m2^2, here from exact integer sums (Tot^2 in int128):
 *     Q = (double)sumIn / (double)m2 - (double)sumTot2 / ((double)m2 * (double)m2)
 *   iteration     louvain_main_process.cu:550-1059: Iteration:0 is Q before any move,
 *                 break when Q_i - Q_{i-1} < min_modularity, return the previous
 *                 assignment and Q_{i-1} (revert)
 *   frontier      louvain_main_process.cu:1009-1018 (active vertices only); the prune
 *                 is sound so it never changes results -- prune=0 sweeps everyone
 *   contraction   src/louvain_gpu/build_compressed_graph.cu:5-173: new id = rank of
 *                 non-empty community (:18-23), rows merged by new id, internal
 *                 weight becomes a self-loop; rows here are canonical (ascending
 *                 neighbour id), the order graph.cpp:58-70 produces
 *   round loop    src/louvain_gpu/louvain.cu:176-224 (the commented-out multi-round
 *                 loop), final composed map final[v] = new_id_R(...new_id_0(v))
 *
 * Weights: int32 (the reference's weight_t), int64 (contracted sums), or float32
 * scaled exactly to int64 by orc_scale_f32 -- one exact-integer implementation
 * for all of them.  A separate FP64 variant (orc_louvain_f: float32 weights,
 * double sums in row order) is the FP64 CPU oracle the weighted path's final Q
 * is compared with (1e-9).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef __int128 i128;

#define ORC_OK 0
#define ORC_E_INVALID -1
#define ORC_E_OVERFLOW -2
#define ORC_E_OOM -3

int cmp_i32(const void *a, const void *b);

typedef void (*orc_iter_cb)(void *user, int round, int iteration, double q,
                            const int32_t *cur, int32_t n, int64_t n_moved);

/* ------------------------------------------------------------------ */
/* graph held as exclusive offsets off[n+1]                            */
/* ------------------------------------------------------------------ */
typedef struct {
    int32_t n;
    int64_t *off;  /* n+1 */
    int32_t *nbr;
    int64_t *w;    /* integer weights held in int64 (contracted sums) */
    double *wf;    /* fp weights (weighted-FP mode) */
} og_t;

static int g_threads = 1;

void orc_set_threads(int t) { g_threads = t > 0 ? t : 1; }

/* per-thread sparse accumulator: dense slot array + touched list */
typedef struct {
    int64_t *acc;
    double *accf;
    int32_t *touched;
    int32_t nt;
    uint8_t *seen;
} spa_t;

static int spa_init(spa_t *s, int32_t n, int fp) {
    s->acc = fp ? NULL : (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    s->accf = fp ? (double *)calloc((size_t)n + 1, sizeof(double)) : NULL;
    s->touched = (int32_t *)malloc(((size_t)n + 1) * sizeof(int32_t));
    s->seen = (uint8_t *)calloc((size_t)n + 1, 1);
    s->nt = 0;
    if ((!s->acc && !s->accf) || !s->touched || !s->seen) return ORC_E_OOM;
    return ORC_OK;
}
static void spa_free(spa_t *s) {
    free(s->acc); free(s->accf); free(s->touched); free(s->seen);
}

/* ------------------------------------------------------------------ */
/* One Louvain round (phase 1) -- integer weights, exact arithmetic    */
/* ------------------------------------------------------------------ */
typedef struct {
    int64_t m2;
    int32_t n;
    int64_t *K, *Self, *Tot, *In;
    int32_t *size;
    int32_t *prev, *cur;
    int64_t *dTot;
    int32_t *dsize;
    uint8_t *active;
} st_t;

static void st_free(st_t *s) {
    free(s->K); free(s->Self); free(s->Tot); free(s->In); free(s->size);
    free(s->prev); free(s->cur); free(s->dTot); free(s->dsize); free(s->active);
}

static int st_alloc(st_t *s, int32_t n) {
    memset(s, 0, sizeof(*s));
    s->n = n;
    size_t N = (size_t)(n > 0 ? n : 1);
    s->K = (int64_t *)calloc(N, 8); s->Self = (int64_t *)calloc(N, 8);
    s->Tot = (int64_t *)calloc(N, 8); s->In = (int64_t *)calloc(N, 8);
    s->size = (int32_t *)calloc(N, 4); s->prev = (int32_t *)calloc(N, 4);
    s->cur = (int32_t *)calloc(N, 4); s->dTot = (int64_t *)calloc(N, 8);
    s->dsize = (int32_t *)calloc(N, 4); s->active = (uint8_t *)calloc(N, 1);
    if (!s->K || !s->Self || !s->Tot || !s->In || !s->size || !s->prev || !s->cur ||
        !s->dTot || !s->dsize || !s->active) { st_free(s); return ORC_E_OOM; }
    return ORC_OK;
}

/* init_communities  (kernel_functions.cu:34-73) */
static void st_init(st_t *s, const og_t *g) {
    for (int32_t v = 0; v < g->n; v++) {
        int64_t k = 0, self = 0;
        for (int64_t j = g->off[v]; j < g->off[v + 1]; j++) {
            k += g->w[j];
            if (g->nbr[j] == v) self = g->w[j];
        }
        s->K[v] = k; s->Tot[v] = k; s->Self[v] = self;
        s->prev[v] = v; s->cur[v] = v; s->size[v] = 1; s->active[v] = 1;
    }
}

/* Decide for one vertex; returns the chosen community (== prev[v] when staying). */
static int32_t decide_vertex(const og_t *g, const st_t *s, int32_t v, spa_t *sp) {
    int32_t c = s->prev[v];
    int64_t deg = g->off[v + 1] - g->off[v];
    if (deg == 0) return c; /* degree bins start at 1 (louvain_main_process.cu:279) */
    int64_t K = s->K[v], S = s->Self[v], Tc = s->Tot[c], m2 = s->m2;
    sp->nt = 0;
    for (int64_t j = g->off[v]; j < g->off[v + 1]; j++) {
        int32_t d = s->prev[g->nbr[j]];
        if (!sp->seen[d]) { sp->seen[d] = 1; sp->acc[d] = 0; sp->touched[sp->nt++] = d; }
        sp->acc[d] += g->w[j];
    }
    int64_t wc = sp->seen[c] ? sp->acc[c] : 0;
    int32_t best = -1;
    i128 bestH = 0;
    for (int32_t t = 0; t < sp->nt; t++) {
        int32_t d = sp->touched[t];
        sp->seen[d] = 0;
        if (d == c) continue;
        i128 H = (i128)sp->acc[d] * m2 - (i128)K * (s->Tot[d] - Tc + K);
        if (H > bestH || (H == bestH && H > 0 && d < best)) { bestH = H; best = d; }
    }
    if (best < 0) return c;
    if (bestH - (i128)m2 * (wc - S) <= 0) return c;
    if (s->size[best] == 1 && s->size[c] == 1 && best > c) return c; /* singleton guard */
    return best;
}

/* In and prune for one vertex (compute_In, kernel_functions.cu:889-955) */
static void weight_vertex(const og_t *g, st_t *s, int32_t v, int64_t minTot) {
    int32_t c = s->cur[v];
    int64_t in = 0;
    for (int64_t j = g->off[v]; j < g->off[v + 1]; j++)
        if (s->cur[g->nbr[j]] == c) in += g->w[j];
    s->In[v] = in;
    int64_t K = s->K[v], S = s->Self[v];
    i128 P = (i128)s->m2 * (2 * in - S - K) + (i128)K * (minTot - (s->Tot[c] - K));
    s->active[v] = (P > 0) ? 0 : 1;
}

/* Q from exact sums, finished with three FP64 operations (the engine's order) */
static double q_of(int64_t sumIn, i128 sumTot2, int64_t m2) {
    return (double)sumIn / (double)m2 - (double)sumTot2 / ((double)m2 * (double)m2);
}

/* Weight update over all vertices + modularity sums. */
static double st_weight_and_q(const og_t *g, st_t *s, int64_t minTot) {
    int32_t n = g->n;
    int64_t sumIn = 0;
    i128 sumT2 = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : sumIn) num_threads(g_threads)
    for (int32_t v = 0; v < n; v++) { weight_vertex(g, s, v, minTot); sumIn += s->In[v]; }
    for (int32_t c = 0; c < n; c++) sumT2 += (i128)s->Tot[c] * s->Tot[c];
    return q_of(sumIn, sumT2, s->m2);
}

typedef struct {
    int32_t iterations;     /* number of sweeps run (the last one is reverted) */
    int64_t swept;          /* vertices evaluated by decide (over all sweeps) */
    int64_t swept_edges;    /* row entries read by decide */
} round_stats_t;

/*
 * One round of phase 1 (louvain_main_process).  On return s->prev holds the
 * accepted assignment and *q_out its modularity.
 */
static int run_round(const og_t *g, st_t *s, double min_mod, int prune, int max_iter,
                     int round, orc_iter_cb cb, void *user, double *q_out,
                     round_stats_t *rs) {
    int32_t n = g->n;
    int nth = g_threads;
    spa_t *sp = (spa_t *)calloc((size_t)nth, sizeof(spa_t));
    if (!sp) return ORC_E_OOM;
    for (int t = 0; t < nth; t++)
        if (spa_init(&sp[t], n, 0)) { for (int u = 0; u <= t; u++) spa_free(&sp[u]); free(sp); return ORC_E_OOM; }
    st_init(s, g);
    /* Iteration 0: In on the singleton partition, min_Tot = 1 (louvain_main_process.cu:538-540) */
    double q = st_weight_and_q(g, s, 1);
    if (cb) cb(user, round, 0, q, s->prev, n, 0);
    for (int32_t v = 0; v < n; v++) s->active[v] = 1; /* fill(active,1) (:574) */
    double qprev = q;
    int it = 1;
    if (rs) memset(rs, 0, sizeof(*rs));
    for (;; it++) {
        /* decide-and-move, synchronous */
        int64_t moved = 0, swept = 0, sedges = 0;
#pragma omp parallel num_threads(nth) reduction(+ : moved, swept, sedges)
        {
#ifdef _OPENMP
            spa_t *my = &sp[omp_get_thread_num()];
#else
            spa_t *my = &sp[0];
#endif
#pragma omp for schedule(dynamic, 1024)
            for (int32_t v = 0; v < n; v++) {
                if (prune && !s->active[v]) { s->cur[v] = s->prev[v]; continue; }
                swept++;
                sedges += g->off[v + 1] - g->off[v];
                int32_t b = decide_vertex(g, s, v, my);
                s->cur[v] = b;
                if (b != s->prev[v]) moved++;
            }
        }
        if (rs) { rs->swept += swept; rs->swept_edges += sedges; }
        /* deltas (sequential: exact and order-free) */
        for (int32_t v = 0; v < n; v++) {
            int32_t c = s->prev[v], b = s->cur[v];
            if (b != c) {
                s->dTot[b] += s->K[v]; s->dTot[c] -= s->K[v];
                s->dsize[b] += 1; s->dsize[c] -= 1;
            }
        }
        int64_t minTot = s->m2;
        for (int32_t c = 0; c < n; c++) {
            s->Tot[c] += s->dTot[c]; s->size[c] += s->dsize[c];
            s->dTot[c] = 0; s->dsize[c] = 0;
            if (s->Tot[c] != 0 && s->Tot[c] < minTot) minTot = s->Tot[c];
        }
        q = st_weight_and_q(g, s, minTot);
        if (cb) cb(user, round, it, q, s->cur, n, moved);
        if (q - qprev < min_mod || (max_iter > 0 && it >= max_iter)) break;
        if (!prune) for (int32_t v = 0; v < n; v++) s->active[v] = 1;
        memcpy(s->prev, s->cur, (size_t)n * sizeof(int32_t));
        qprev = q;
    }
    if (rs) rs->iterations = it;
    *q_out = qprev;
    for (int t = 0; t < nth; t++) spa_free(&sp[t]);
    free(sp);
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Contraction (build_compressed_graph.cu:5-173), canonical rows       */
/* ------------------------------------------------------------------ */
static int contract(const og_t *g, const int32_t *comm, int32_t *new_id_out, og_t *out) {
    int32_t n = g->n;
    int32_t *cnt = (int32_t *)calloc((size_t)n + 1, 4);
    int32_t *nid = new_id_out;
    if (!cnt) return ORC_E_OOM;
    for (int32_t v = 0; v < n; v++) cnt[comm[v]]++;
    int32_t C = 0;
    for (int32_t c = 0; c < n; c++) nid[c] = cnt[c] > 0 ? C++ : -1;
    /* members grouped by new id, ascending vertex order */
    int32_t *moff = (int32_t *)calloc((size_t)C + 1, 4);
    int32_t *mem = (int32_t *)malloc((size_t)(n > 0 ? n : 1) * 4);
    if (!moff || !mem) { free(cnt); free(moff); free(mem); return ORC_E_OOM; }
    for (int32_t v = 0; v < n; v++) moff[nid[comm[v]] + 1]++;
    for (int32_t c = 0; c < C; c++) moff[c + 1] += moff[c];
    {
        int32_t *cursor = (int32_t *)malloc((size_t)(C > 0 ? C : 1) * 4);
        if (!cursor) { free(cnt); free(moff); free(mem); return ORC_E_OOM; }
        memcpy(cursor, moff, (size_t)C * 4);
        for (int32_t v = 0; v < n; v++) mem[cursor[nid[comm[v]]]++] = v;
        free(cursor);
    }
    /* two passes: row lengths, then fill */
    int64_t *acc = (int64_t *)calloc((size_t)(C > 0 ? C : 1), 8);
    uint8_t *seen = (uint8_t *)calloc((size_t)(C > 0 ? C : 1), 1);
    int32_t *touched = (int32_t *)malloc((size_t)(C > 0 ? C : 1) * 4);
    int64_t *off = (int64_t *)calloc((size_t)C + 1, 8);
    if (!acc || !seen || !touched || !off) {
        free(cnt); free(moff); free(mem); free(acc); free(seen); free(touched); free(off);
        return ORC_E_OOM;
    }
    int64_t cap = g->off[n] > 0 ? g->off[n] : 1;
    int32_t *nbr = (int32_t *)malloc((size_t)cap * 4);
    int64_t *w = (int64_t *)malloc((size_t)cap * 8);
    if (!nbr || !w) {
        free(cnt); free(moff); free(mem); free(acc); free(seen); free(touched); free(off);
        free(nbr); free(w); return ORC_E_OOM;
    }
    int64_t pos = 0;
    for (int32_t c = 0; c < C; c++) {
        int32_t nt = 0;
        for (int32_t i = moff[c]; i < moff[c + 1]; i++) {
            int32_t v = mem[i];
            for (int64_t j = g->off[v]; j < g->off[v + 1]; j++) {
                int32_t d = nid[comm[g->nbr[j]]];
                if (!seen[d]) { seen[d] = 1; acc[d] = 0; touched[nt++] = d; }
                acc[d] += g->w[j];
            }
        }
        /* canonical ascending order: the touched list is small, sort it */
        /* insertion sort for short rows, qsort otherwise */
        if (nt > 32) {
            qsort(touched, (size_t)nt, 4, cmp_i32);
        } else {
            for (int32_t a = 1; a < nt; a++) {
                int32_t x = touched[a], b = a - 1;
                while (b >= 0 && touched[b] > x) { touched[b + 1] = touched[b]; b--; }
                touched[b + 1] = x;
            }
        }
        for (int32_t t = 0; t < nt; t++) {
            int32_t d = touched[t];
            nbr[pos] = d; w[pos] = acc[d]; pos++;
            seen[d] = 0;
        }
        off[c + 1] = pos;
    }
    out->n = C; out->off = off; out->nbr = nbr; out->w = w; out->wf = NULL;
    free(cnt); free(moff); free(mem); free(acc); free(seen); free(touched);
    return ORC_OK;
}

int cmp_i32(const void *a, const void *b) {
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

static void og_free(og_t *g) { free(g->off); free(g->nbr); free(g->w); free(g->wf); }

static int og_from_incl(og_t *g, int32_t n, const int64_t *end_incl, const int32_t *nbr,
                        const int32_t *w, const int64_t *w64) {
    memset(g, 0, sizeof(*g));
    if (n < 0) return ORC_E_INVALID;
    g->n = n;
    g->off = (int64_t *)malloc(((size_t)n + 1) * 8);
    if (!g->off) return ORC_E_OOM;
    g->off[0] = 0;
    for (int32_t i = 0; i < n; i++) g->off[i + 1] = end_incl[i];
    int64_t e = g->off[n];
    for (int32_t i = 0; i < n; i++) if (g->off[i + 1] < g->off[i]) return ORC_E_INVALID;
    g->nbr = (int32_t *)malloc((size_t)(e > 0 ? e : 1) * 4);
    g->w = (int64_t *)malloc((size_t)(e > 0 ? e : 1) * 8);
    if (!g->nbr || !g->w) return ORC_E_OOM;
    for (int64_t j = 0; j < e; j++) {
        if (nbr[j] < 0 || nbr[j] >= n) return ORC_E_INVALID;
        g->nbr[j] = nbr[j];
        g->w[j] = w64 ? w64[j] : w ? (int64_t)w[j] : 1;
        if (g->w[j] <= 0) return ORC_E_INVALID; /* Louvain needs positive weights */
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Public (ctypes) entry points                                        */
/* ------------------------------------------------------------------ */

/* m2 and its range: every weight sum (K, Tot, In, w_d) is <= m2 < 2^62, so
 * they are exact in int64 and every gain / prune product fits int128. */
#define ORC_M2_MAX ((int64_t)1 << 62)
static int check_m2(int64_t m2) {
    if (m2 <= 0) return ORC_E_INVALID;
    if (m2 >= ORC_M2_MAX) return ORC_E_OVERFLOW;
    return ORC_OK;
}

/* sum of positive int64 weights with the 2^62 range check (no wraparound) */
static int sum_m2(const og_t *g, int64_t *m2) {
    i128 s = 0;
    for (int64_t j = 0; j < g->off[g->n]; j++) {
        s += g->w[j];
        if (s >= ORC_M2_MAX) return ORC_E_OVERFLOW;
    }
    *m2 = (int64_t)s;
    return check_m2(*m2);
}

/*
 * Full multi-round Louvain (louvain.cu:136-229 with the intended loop).
 * final_comm[n0] receives the composed community of every input vertex.
 * stats (optional, int64[4]): rounds, total sweeps, swept vertices, swept edges.
 */
static int louvain_exact(int32_t n, const int64_t *end_incl, const int32_t *nbr, const int32_t *w,
                         const int64_t *w64, double min_mod, int prune, int max_rounds, orc_iter_cb cb,
                         void *user, int32_t *final_comm, double *final_q, int32_t *final_nc,
                         double *round_q, int32_t *round_nc, int64_t *stats) {
    og_t g;
    int rc = og_from_incl(&g, n, end_incl, nbr, w, w64);
    if (rc) { og_free(&g); return rc; }
    int64_t m2 = 0;
    if ((rc = sum_m2(&g, &m2))) { og_free(&g); return rc; }
    for (int32_t v = 0; v < n; v++) final_comm[v] = v;
    double prev_mod = -1, cur_mod = -1;
    int round = 0;
    int64_t st_sweeps = 0, st_sv = 0, st_se = 0;
    for (;;) {
        st_t s;
        if ((rc = st_alloc(&s, g.n))) { og_free(&g); return rc; }
        s.m2 = m2;
        round_stats_t rs;
        double q;
        if (round > 0) prev_mod = cur_mod;
        rc = run_round(&g, &s, min_mod, prune, 0, round, cb, user, &q, &rs);
        if (rc) { st_free(&s); og_free(&g); return rc; }
        st_sweeps += rs.iterations; st_sv += rs.swept; st_se += rs.swept_edges;
        cur_mod = q;
        int32_t *nid = (int32_t *)malloc((size_t)(g.n > 0 ? g.n : 1) * 4);
        og_t ng;
        rc = nid ? contract(&g, s.prev, nid, &ng) : ORC_E_OOM;
        if (rc) { free(nid); st_free(&s); og_free(&g); return rc; }
        for (int32_t v = 0; v < n; v++) final_comm[v] = nid[s.prev[final_comm[v]]];
        if (round_q) round_q[round] = cur_mod;
        if (round_nc) round_nc[round] = ng.n;
        free(nid);
        st_free(&s);
        og_free(&g);
        g = ng;
        round++;
        if (!(cur_mod - prev_mod >= min_mod)) break;
        if (max_rounds > 0 && round >= max_rounds) break;
    }
    *final_q = cur_mod;
    *final_nc = g.n;
    if (stats) { stats[0] = round; stats[1] = st_sweeps; stats[2] = st_sv; stats[3] = st_se; }
    og_free(&g);
    return ORC_OK;
}

int orc_louvain(int32_t n, const int64_t *end_incl, const int32_t *nbr, const int32_t *w,
                double min_mod, int prune, int max_rounds, orc_iter_cb cb, void *user,
                int32_t *final_comm, double *final_q, int32_t *final_nc,
                double *round_q, int32_t *round_nc, int64_t *stats) {
    return louvain_exact(n, end_incl, nbr, w, NULL, min_mod, prune, max_rounds, cb, user, final_comm,
                         final_q, final_nc, round_q, round_nc, stats);
}

/*
 * The same on int64 weights: contracted graphs, and float32 weights scaled to
 * exact integers by orc_scale_f32 (Q and every decision are invariant under a
 * common scale of the weights, so the run is the exact-arithmetic run on the
 * float values).  m2 must stay below 2^62.
 */
int orc_louvain_w64(int32_t n, const int64_t *end_incl, const int32_t *nbr, const int64_t *w64,
                    double min_mod, int prune, int max_rounds, orc_iter_cb cb, void *user,
                    int32_t *final_comm, double *final_q, int32_t *final_nc,
                    double *round_q, int32_t *round_nc, int64_t *stats) {
    return louvain_exact(n, end_incl, nbr, NULL, w64, min_mod, prune, max_rounds, cb, user, final_comm,
                         final_q, final_nc, round_q, round_nc, stats);
}

/*
 * Exact integer image of positive finite float32 weights: w_j = W_j / 2^k with
 * W_j integer and k >= 0 the smallest shift that makes every weight integral
 * (k = 0 when all weights are integers).  Written from the IEEE-754 binary32
 * layout: a normal value is (2^23 + mantissa) * 2^(exponent - 150), a
 * subnormal one mantissa * 2^-149.  Returns ORC_E_INVALID on a non-positive or
 * non-finite weight, ORC_E_OVERFLOW when the scaled sum reaches 2^62.
 */
int orc_scale_f32(int64_t m, const float *wf, int64_t *out, int32_t *k_out) {
    int32_t lo = 0; /* smallest exponent of a weight's lowest set bit (<= 0) */
    for (int64_t j = 0; j < m; j++) {
        uint32_t b;
        memcpy(&b, &wf[j], 4);
        uint32_t ex = (b >> 23) & 0xff, man = b & 0x7fffff;
        if ((b >> 31) || ex == 0xff || (ex == 0 && man == 0)) return ORC_E_INVALID;
        uint64_t sig = ex ? (man | 0x800000u) : man;
        int32_t e = ex ? (int32_t)ex - 150 : -149;
        while (!(sig & 1)) { sig >>= 1; e++; }
        if (e < lo) lo = e;
    }
    int32_t k = -lo;
    i128 sum = 0;
    for (int64_t j = 0; j < m; j++) {
        uint32_t b;
        memcpy(&b, &wf[j], 4);
        uint32_t ex = (b >> 23) & 0xff, man = b & 0x7fffff;
        uint64_t sig = ex ? (man | 0x800000u) : man;
        int32_t e = (ex ? (int32_t)ex - 150 : -149) + k;
        while (!(sig & 1)) { sig >>= 1; e++; } /* odd significand: e >= 0 by the choice of k */
        if (e > 62) return ORC_E_OVERFLOW;
        i128 W = (i128)sig << e;
        sum += W;
        if (W >= ORC_M2_MAX || sum >= ORC_M2_MAX) return ORC_E_OVERFLOW;
        out[j] = (int64_t)W;
    }
    *k_out = k;
    return ORC_OK;
}

/*
 * One round on a graph given with int64 weights (contracted graphs carry sums).
 * comm_out[n] = accepted assignment (pre-renumbering), q_out its modularity.
 */
int orc_round(int32_t n, const int64_t *end_incl, const int32_t *nbr, const int64_t *w,
              int64_t m2, double min_mod, int prune, int round, orc_iter_cb cb, void *user,
              int32_t *comm_out, double *q_out, int32_t *iters_out) {
    og_t g;
    int32_t *w32 = NULL;
    (void)w32;
    memset(&g, 0, sizeof(g));
    g.n = n;
    g.off = (int64_t *)malloc(((size_t)n + 1) * 8);
    if (!g.off) return ORC_E_OOM;
    g.off[0] = 0;
    for (int32_t i = 0; i < n; i++) g.off[i + 1] = end_incl[i];
    int64_t e = g.off[n];
    g.nbr = (int32_t *)malloc((size_t)(e > 0 ? e : 1) * 4);
    g.w = (int64_t *)malloc((size_t)(e > 0 ? e : 1) * 8);
    if (!g.nbr || !g.w) { og_free(&g); return ORC_E_OOM; }
    memcpy(g.nbr, nbr, (size_t)e * 4);
    memcpy(g.w, w, (size_t)e * 8);
    int rc = check_m2(m2);
    if (rc) { og_free(&g); return rc; }
    st_t s;
    if ((rc = st_alloc(&s, n))) { og_free(&g); return rc; }
    s.m2 = m2;
    round_stats_t rs;
    rc = run_round(&g, &s, min_mod, prune, 0, round, cb, user, q_out, &rs);
    if (!rc) { memcpy(comm_out, s.prev, (size_t)n * 4); *iters_out = rs.iterations; }
    st_free(&s);
    og_free(&g);
    return rc;
}

/*
 * Contraction of (n, end_incl, nbr, w) under comm.  Output arrays are allocated
 * here (release with orc_free); new_id[n] receives the renumbering map.
 */
int orc_contract(int32_t n, const int64_t *end_incl, const int32_t *nbr, const int64_t *w,
                 const int32_t *comm, int32_t *new_id, int32_t *out_n, int64_t **out_end,
                 int32_t **out_nbr, int64_t **out_w) {
    og_t g;
    memset(&g, 0, sizeof(g));
    g.n = n;
    g.off = (int64_t *)malloc(((size_t)n + 1) * 8);
    if (!g.off) return ORC_E_OOM;
    g.off[0] = 0;
    for (int32_t i = 0; i < n; i++) g.off[i + 1] = end_incl[i];
    int64_t e = g.off[n];
    g.nbr = (int32_t *)malloc((size_t)(e > 0 ? e : 1) * 4);
    g.w = (int64_t *)malloc((size_t)(e > 0 ? e : 1) * 8);
    if (!g.nbr || !g.w) { og_free(&g); return ORC_E_OOM; }
    memcpy(g.nbr, nbr, (size_t)e * 4);
    memcpy(g.w, w, (size_t)e * 8);
    for (int32_t v = 0; v < n; v++)
        if (comm[v] < 0 || comm[v] >= n) { og_free(&g); return ORC_E_INVALID; }
    og_t o;
    int rc = contract(&g, comm, new_id, &o);
    og_free(&g);
    if (rc) return rc;
    int64_t *end = (int64_t *)malloc((size_t)(o.n > 0 ? o.n : 1) * 8);
    if (!end) { og_free(&o); return ORC_E_OOM; }
    for (int32_t c = 0; c < o.n; c++) end[c] = o.off[c + 1];
    *out_n = o.n; *out_end = end; *out_nbr = o.nbr; *out_w = o.w;
    free(o.off);
    return ORC_OK;
}

void orc_free(void *p) { free(p); }

/*
 * Single synchronous decide-and-move step on given state (phase-level parity):
 * reads prev/Tot/size, writes cur; only vertices with frontier[v]!=0 are swept
 * (frontier may be NULL for all).
 */
int orc_decide_step(int32_t n, const int64_t *end_incl, const int32_t *nbr, const int64_t *w,
                    int64_t m2, const int64_t *K, const int64_t *Self, const int64_t *Tot,
                    const int32_t *size, const int32_t *prev, const uint8_t *frontier,
                    int32_t *cur) {
    og_t g;
    memset(&g, 0, sizeof(g));
    g.n = n;
    g.off = (int64_t *)malloc(((size_t)n + 1) * 8);
    if (!g.off) return ORC_E_OOM;
    g.off[0] = 0;
    for (int32_t i = 0; i < n; i++) g.off[i + 1] = end_incl[i];
    g.nbr = (int32_t *)nbr;
    g.w = (int64_t *)w;
    st_t s;
    memset(&s, 0, sizeof(s));
    s.m2 = m2; s.n = n;
    s.K = (int64_t *)K; s.Self = (int64_t *)Self; s.Tot = (int64_t *)Tot;
    s.size = (int32_t *)size; s.prev = (int32_t *)prev;
    spa_t sp;
    if (spa_init(&sp, n, 0)) { free(g.off); return ORC_E_OOM; }
    for (int32_t v = 0; v < n; v++)
        cur[v] = (frontier && !frontier[v]) ? prev[v] : decide_vertex(&g, &s, v, &sp);
    spa_free(&sp);
    free(g.off);
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Weighted FP64 variant (fp32 weights on input, FP64 sums)            */
/* ------------------------------------------------------------------ */
typedef struct {
    double m2;
    double *K, *Self, *Tot, *In;
    int32_t *size, *prev, *cur;
} stf_t;

static int32_t decide_vertex_f(const og_t *g, const stf_t *s, int32_t v, spa_t *sp) {
    int32_t c = s->prev[v];
    if (g->off[v + 1] == g->off[v]) return c;
    double K = s->K[v], S = s->Self[v], Tc = s->Tot[c], m2 = s->m2;
    sp->nt = 0;
    for (int64_t j = g->off[v]; j < g->off[v + 1]; j++) {
        int32_t d = s->prev[g->nbr[j]];
        if (!sp->seen[d]) { sp->seen[d] = 1; sp->accf[d] = 0; sp->touched[sp->nt++] = d; }
        sp->accf[d] += g->wf[j];
    }
    double wc = sp->seen[c] ? sp->accf[c] : 0;
    int32_t best = -1;
    double bestH = 0;
    for (int32_t t = 0; t < sp->nt; t++) {
        int32_t d = sp->touched[t];
        sp->seen[d] = 0;
        if (d == c) continue;
        double H = sp->accf[d] * m2 - K * (s->Tot[d] - Tc + K);
        if (H > bestH || (H == bestH && H > 0 && d < best)) { bestH = H; best = d; }
    }
    if (best < 0) return c;
    if (!(bestH - m2 * (wc - S) > 0)) return c;
    if (s->size[best] == 1 && s->size[c] == 1 && best > c) return c;
    return best;
}

/*
 * Full multi-round Louvain on FP weights (weights given as float32).  Same
 * control flow as orc_louvain; used for the weighted-path Q tolerance.
 */
int orc_louvain_f(int32_t n, const int64_t *end_incl, const int32_t *nbr, const float *wf,
                  double min_mod, int max_rounds, int32_t *final_comm, double *final_q,
                  int32_t *final_nc, double *round_q, int32_t *round_nc) {
    og_t g;
    memset(&g, 0, sizeof(g));
    g.n = n;
    g.off = (int64_t *)malloc(((size_t)n + 1) * 8);
    if (!g.off) return ORC_E_OOM;
    g.off[0] = 0;
    for (int32_t i = 0; i < n; i++) g.off[i + 1] = end_incl[i];
    int64_t e = g.off[n];
    g.nbr = (int32_t *)malloc((size_t)(e > 0 ? e : 1) * 4);
    g.wf = (double *)malloc((size_t)(e > 0 ? e : 1) * 8);
    if (!g.nbr || !g.wf) { og_free(&g); return ORC_E_OOM; }
    double m2 = 0;
    for (int64_t j = 0; j < e; j++) { g.nbr[j] = nbr[j]; g.wf[j] = (double)wf[j]; m2 += g.wf[j]; }
    if (!(m2 > 0)) { og_free(&g); return ORC_E_INVALID; }
    for (int32_t v = 0; v < n; v++) final_comm[v] = v;
    double prev_mod = -1, cur_mod = -1;
    int round = 0;
    for (;;) {
        int32_t N = g.n;
        size_t NN = (size_t)(N > 0 ? N : 1);
        stf_t s;
        s.m2 = m2;
        s.K = (double *)calloc(NN, 8); s.Self = (double *)calloc(NN, 8);
        s.Tot = (double *)calloc(NN, 8); s.In = (double *)calloc(NN, 8);
        s.size = (int32_t *)calloc(NN, 4); s.prev = (int32_t *)calloc(NN, 4);
        s.cur = (int32_t *)calloc(NN, 4);
        double *dTot = (double *)calloc(NN, 8);
        int32_t *dsize = (int32_t *)calloc(NN, 4);
        spa_t sp;
        int bad = !s.K || !s.Self || !s.Tot || !s.In || !s.size || !s.prev || !s.cur || !dTot ||
                  !dsize || spa_init(&sp, N, 1);
        if (bad) { og_free(&g); return ORC_E_OOM; }
        for (int32_t v = 0; v < N; v++) {
            double k = 0, self = 0;
            for (int64_t j = g.off[v]; j < g.off[v + 1]; j++) {
                k += g.wf[j];
                if (g.nbr[j] == v) self = g.wf[j];
            }
            s.K[v] = k; s.Tot[v] = k; s.Self[v] = self; s.prev[v] = v; s.cur[v] = v; s.size[v] = 1;
        }
        if (round > 0) prev_mod = cur_mod;
        /* Q0 */
        double sumIn = 0, sumT2 = 0;
        for (int32_t v = 0; v < N; v++) sumIn += s.Self[v];
        for (int32_t c = 0; c < N; c++) sumT2 += s.Tot[c] * s.Tot[c];
        double q = sumIn / m2 - sumT2 / (m2 * m2), qprev = q;
        for (int it = 1;; it++) {
            for (int32_t v = 0; v < N; v++) s.cur[v] = decide_vertex_f(&g, &s, v, &sp);
            for (int32_t v = 0; v < N; v++) {
                int32_t c = s.prev[v], b = s.cur[v];
                if (b != c) { dTot[b] += s.K[v]; dTot[c] -= s.K[v]; dsize[b]++; dsize[c]--; }
            }
            for (int32_t c = 0; c < N; c++) {
                s.Tot[c] += dTot[c]; s.size[c] += dsize[c]; dTot[c] = 0; dsize[c] = 0;
            }
            sumIn = 0; sumT2 = 0;
            for (int32_t v = 0; v < N; v++)
                for (int64_t j = g.off[v]; j < g.off[v + 1]; j++)
                    if (s.cur[g.nbr[j]] == s.cur[v]) sumIn += g.wf[j];
            for (int32_t c = 0; c < N; c++) sumT2 += s.Tot[c] * s.Tot[c];
            q = sumIn / m2 - sumT2 / (m2 * m2);
            if (q - qprev < min_mod) break;
            memcpy(s.prev, s.cur, (size_t)N * 4);
            qprev = q;
        }
        cur_mod = qprev;
        /* contract with FP sums */
        int32_t *nid = (int32_t *)malloc(NN * 4);
        int32_t C = 0;
        int32_t *cnt = (int32_t *)calloc(NN, 4);
        for (int32_t v = 0; v < N; v++) cnt[s.prev[v]]++;
        for (int32_t c = 0; c < N; c++) nid[c] = cnt[c] > 0 ? C++ : -1;
        int32_t *moff = (int32_t *)calloc((size_t)C + 1, 4);
        int32_t *mem = (int32_t *)malloc(NN * 4);
        for (int32_t v = 0; v < N; v++) moff[nid[s.prev[v]] + 1]++;
        for (int32_t c = 0; c < C; c++) moff[c + 1] += moff[c];
        int32_t *cursor = (int32_t *)malloc(((size_t)C + 1) * 4);
        memcpy(cursor, moff, (size_t)C * 4);
        for (int32_t v = 0; v < N; v++) mem[cursor[nid[s.prev[v]]]++] = v;
        og_t ng;
        memset(&ng, 0, sizeof(ng));
        ng.n = C;
        ng.off = (int64_t *)calloc((size_t)C + 1, 8);
        int64_t cap = g.off[N] > 0 ? g.off[N] : 1;
        ng.nbr = (int32_t *)malloc((size_t)cap * 4);
        ng.wf = (double *)malloc((size_t)cap * 8);
        spa_t sc;
        spa_init(&sc, C, 1);
        int64_t pos = 0;
        for (int32_t c = 0; c < C; c++) {
            sc.nt = 0;
            for (int32_t i = moff[c]; i < moff[c + 1]; i++) {
                int32_t v = mem[i];
                for (int64_t j = g.off[v]; j < g.off[v + 1]; j++) {
                    int32_t d = nid[s.prev[g.nbr[j]]];
                    if (!sc.seen[d]) { sc.seen[d] = 1; sc.accf[d] = 0; sc.touched[sc.nt++] = d; }
                    sc.accf[d] += g.wf[j];
                }
            }
            qsort(sc.touched, (size_t)sc.nt, 4, cmp_i32);
            for (int32_t t = 0; t < sc.nt; t++) {
                int32_t d = sc.touched[t];
                ng.nbr[pos] = d; ng.wf[pos] = sc.accf[d]; pos++; sc.seen[d] = 0;
            }
            ng.off[c + 1] = pos;
        }
        spa_free(&sc);
        for (int32_t v = 0; v < n; v++) final_comm[v] = nid[s.prev[final_comm[v]]];
        if (round_q) round_q[round] = cur_mod;
        if (round_nc) round_nc[round] = C;
        free(nid); free(cnt); free(moff); free(mem); free(cursor);
        free(s.K); free(s.Self); free(s.Tot); free(s.In); free(s.size); free(s.prev); free(s.cur);
        free(dTot); free(dsize); spa_free(&sp);
        og_free(&g);
        g = ng;
        round++;
        if (!(cur_mod - prev_mod >= min_mod)) break;
        if (max_rounds > 0 && round >= max_rounds) break;
    }
    *final_q = cur_mod;
    *final_nc = g.n;
    og_free(&g);
    return ORC_OK;
}

/* Modularity of an arbitrary partition (exact integer sums), for tests. */
double orc_modularity(int32_t n, const int64_t *end_incl, const int32_t *nbr, const int64_t *w,
                      const int32_t *comm) {
    int64_t m2 = 0, sumIn = 0;
    i128 sumT2 = 0;
    int64_t *tot = (int64_t *)calloc((size_t)(n > 0 ? n : 1), 8);
    int64_t prev = 0;
    for (int32_t v = 0; v < n; v++) {
        for (int64_t j = prev; j < end_incl[v]; j++) {
            m2 += w[j];
            tot[comm[v]] += w[j];
            if (comm[nbr[j]] == comm[v]) sumIn += w[j];
        }
        prev = end_incl[v];
    }
    for (int32_t c = 0; c < n; c++) sumT2 += (i128)tot[c] * tot[c];
    free(tot);
    if (m2 == 0) return 0.0;
    return q_of(sumIn, sumT2, m2);
}

int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
