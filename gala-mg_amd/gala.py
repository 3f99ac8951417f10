"""Python host mirror of the gfx950 Louvain engine (ctypes over libgala.so).

Mirrors the reference's in-process interface (LinXi-lx/GALA-mg):
  * ``Graph`` with ``load_bin_graph`` / ``store_bin_graph``   (src/graph/graph.h:7-19)
  * ``louvain_gpu(graph, min_modularity)`` -> (Q, communities) (src/louvain_gpu/louvain.cuh:8)
plus phase-level calls (decide, weight update, contraction) for parity tests
and the synthetic generators.  Every call goes through the C ABI declared in
include/gala.h; there is no CPU fallback -- if libgala.so is missing or has no
device, calls raise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgala.so")

W_NONE, W_I32, W_F32 = 0, 1, 2

EXPORTS = [
    "gala_version", "gala_strerror", "gala_default_opts", "gala_create", "gala_destroy",
    "gala_load_bin", "gala_set_graph", "gala_run", "gala_partial_result",
    "gala_get_communities", "gala_phase_decide", "gala_phase_weight", "gala_phase_contract",
    "gala_read_bin", "gala_write_bin", "gala_ingest_text", "gala_gen_sbm", "gala_gen_chung_lu",
    "gala_gen_rmat", "gala_set_logger", "gala_set_mlp", "gala_get_mlp", "gala_mlp_train",
    "gala_mlp_forward", "gala_mlp_grad", "gala_batch_set", "gala_batch_run", "gala_batch_graph_results",
    "gala_batch_communities", "gala_gen_sbm_batch", "gala_set_variants", "gala_get_variants", "gala_autotune",
    "gala_variant_counts", "gala_phase_modularity", "gala_modularity_noalloc",
    "gala_ingest_text_device", "gala_load_text", "gala_ingest_times",
    "gala_set_bin_cuts", "gala_get_bin_cuts", "gala_autotune_bins",
]
BIN_LO_DEFAULT = (1, 4, 8, 16, 32, 128, 512, 4096, 8192)
NOALLOC_MAX_N = 1 << 22  # GALA_NOALLOC_MAX_N
N_VARIANT_SLOTS = 27  # kernel variant per bin: 9 general sweeps, 9 first sweeps, 9 weight updates


def variant_counts():
    """Kernel variants of every slot (9 general sweeps, 9 first sweeps, 9
    weight updates), as the library defines them; needs no device."""
    import numpy as np
    c = np.zeros(N_VARIANT_SLOTS, np.int32)
    _check(lib().gala_variant_counts(_p(c, ctypes.c_int32)), "gala_variant_counts")
    return [int(x) for x in c]

MLP_NFEAT = 6
MLP_NPARAM = 1153  # W1[16][64] (rows 6..15 zero), b1[64], w2[64], b2


class GalaError(RuntimeError):
    pass


class Opts(ctypes.Structure):
    _fields_ = [("min_modularity", ctypes.c_double), ("max_iterations", ctypes.c_int),
                ("max_rounds", ctypes.c_int), ("dump_iterations", ctypes.c_int),
                ("dump_rounds", ctypes.c_int), ("verbose", ctypes.c_int),
                ("frontier_mode", ctypes.c_int), ("learned_tau", ctypes.c_float),
                ("profile_bins", ctypes.c_int), ("no_hip_graphs", ctypes.c_int),
                ("serial_bins", ctypes.c_int), ("learned_min_degree", ctypes.c_int),
                ("learned_min_vertices", ctypes.c_int), ("queue_sweeps", ctypes.c_int),
                ("reserved", ctypes.c_int * 2)]


class Result(ctypes.Structure):
    _fields_ = [("modularity", ctypes.c_double), ("n_vertices", ctypes.c_int32),
                ("n_communities", ctypes.c_int32), ("rounds", ctypes.c_int32),
                ("total_iterations", ctypes.c_int32), ("t_total_ms", ctypes.c_double),
                ("t_decide_ms", ctypes.c_double), ("t_update_ms", ctypes.c_double),
                ("t_contract_ms", ctypes.c_double), ("t_h2d_ms", ctypes.c_double),
                ("swept_vertices", ctypes.c_int64), ("swept_edges", ctypes.c_int64),
                ("input_edges", ctypes.c_int64), ("swept_vertices_unit", ctypes.c_int64),
                ("swept_edges_unit", ctypes.c_int64), ("decide_bytes", ctypes.c_int64),
                ("round_q", ctypes.c_double * 64),
                ("round_nc", ctypes.c_int32 * 64), ("round_iters", ctypes.c_int32 * 64),
                ("round_ms", ctypes.c_double * 64), ("round_decide_ms", ctypes.c_double * 64),
                ("round_update_ms", ctypes.c_double * 64), ("round_build_ms", ctypes.c_double * 64),
                ("learned_sweeps", ctypes.c_int32), ("certify_sweeps", ctypes.c_int32),
                ("bin_decide_bytes", ctypes.c_int64 * 9), ("bin_swept_edges", ctypes.c_int64 * 9),
                ("bin_decide_ms", ctypes.c_double * 9),
                ("learned_frontier", ctypes.c_int64), ("learned_kept", ctypes.c_int64),
                ("bin_first_bytes", ctypes.c_int64 * 9), ("bin_first_edges", ctypes.c_int64 * 9),
                ("bin_first_ms", ctypes.c_double * 9), ("bin_launches", ctypes.c_int32 * 9),
                ("bin_first_launches", ctypes.c_int32 * 9), ("wide", ctypes.c_int32),
                ("weight_shift", ctypes.c_int32), ("round_vertices", ctypes.c_int32 * 64),
                ("round_entries", ctypes.c_int64 * 64), ("round_contracted_entries", ctypes.c_int64 * 64),
                ("certify_empty", ctypes.c_int32), ("hub_part_launches", ctypes.c_int32),
                ("hub_part_ms", ctypes.c_double), ("queued_rounds", ctypes.c_int32),
                ("inert_sweeps", ctypes.c_int32)]

    def to_dict(self):
        r = self.rounds
        return dict(modularity=self.modularity, n_vertices=self.n_vertices,
                    n_communities=self.n_communities, rounds=r,
                    total_iterations=self.total_iterations, t_total_ms=self.t_total_ms,
                    t_decide_ms=self.t_decide_ms, t_update_ms=self.t_update_ms,
                    t_contract_ms=self.t_contract_ms, t_h2d_ms=self.t_h2d_ms,
                    swept_vertices=self.swept_vertices, swept_edges=self.swept_edges,
                    input_edges=self.input_edges, swept_vertices_unit=self.swept_vertices_unit,
                    swept_edges_unit=self.swept_edges_unit, decide_bytes=self.decide_bytes,
                    round_q=list(self.round_q[:r]),
                    round_nc=list(self.round_nc[:r]), round_iters=list(self.round_iters[:r]),
                    round_ms=list(self.round_ms[:r]),
                    round_decide_ms=list(self.round_decide_ms[:r]),
                    round_update_ms=list(self.round_update_ms[:r]),
                    round_build_ms=list(self.round_build_ms[:r]),
                    learned_sweeps=self.learned_sweeps, certify_sweeps=self.certify_sweeps,
                    bin_decide_bytes=list(self.bin_decide_bytes), bin_swept_edges=list(self.bin_swept_edges),
                    bin_decide_ms=list(self.bin_decide_ms),
                    learned_frontier=self.learned_frontier, learned_kept=self.learned_kept,
                    bin_first_bytes=list(self.bin_first_bytes), bin_first_edges=list(self.bin_first_edges),
                    bin_first_ms=list(self.bin_first_ms), bin_launches=list(self.bin_launches),
                    bin_first_launches=list(self.bin_first_launches), wide=self.wide,
                    weight_shift=self.weight_shift, round_vertices=list(self.round_vertices[:r]),
                    round_entries=list(self.round_entries[:r]),
                    round_contracted_entries=list(self.round_contracted_entries[:r]),
                    certify_empty=self.certify_empty, hub_part_launches=self.hub_part_launches,
                    hub_part_ms=self.hub_part_ms, queued_rounds=self.queued_rounds,
                    inert_sweeps=self.inert_sweeps)


ITER_CB = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                           ctypes.c_int64, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32)
ROUND_CB = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int32,
                            ctypes.POINTER(ctypes.c_int32), ctypes.c_int32,
                            ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32),
                            ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double))

class BatchResult(ctypes.Structure):
    _fields_ = [("n_graphs", ctypes.c_int32), ("rounds", ctypes.c_int32), ("iterations", ctypes.c_int32),
                ("reserved0", ctypes.c_int32), ("t_total_ms", ctypes.c_double), ("t_phase1_ms", ctypes.c_double),
                ("t_contract_ms", ctypes.c_double), ("swept_vertices", ctypes.c_int64),
                ("swept_edges", ctypes.c_int64), ("round_open", ctypes.c_int32 * 64),
                ("round_iters", ctypes.c_int32 * 64), ("round_ms", ctypes.c_double * 64)]

    def to_dict(self):
        r = self.rounds
        return dict(n_graphs=self.n_graphs, rounds=r, iterations=self.iterations, t_total_ms=self.t_total_ms,
                    t_phase1_ms=self.t_phase1_ms, t_contract_ms=self.t_contract_ms,
                    swept_vertices=self.swept_vertices, swept_edges=self.swept_edges,
                    round_open=list(self.round_open[:r]), round_iters=list(self.round_iters[:r]),
                    round_ms=list(self.round_ms[:r]))


LOG_CB = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int32,
                          ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_uint8))

_lib = None
I64P = ctypes.POINTER(ctypes.c_int64)
I32P = ctypes.POINTER(ctypes.c_int32)
U8P = ctypes.POINTER(ctypes.c_uint8)
VP = ctypes.c_void_p


def lib():
    """Load libgala.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GalaError(f"{LIB_PATH} not built: run `make -C gala-mg_amd`")
        L = ctypes.CDLL(LIB_PATH)
        L.gala_version.restype = ctypes.c_char_p
        L.gala_strerror.restype = ctypes.c_char_p
        L.gala_strerror.argtypes = [ctypes.c_int]
        L.gala_default_opts.argtypes = [ctypes.POINTER(Opts)]
        L.gala_create.argtypes = [ctypes.POINTER(VP), ctypes.c_int, ctypes.POINTER(Opts)]
        L.gala_destroy.argtypes = [VP]
        L.gala_load_bin.argtypes = [VP, ctypes.c_char_p, ctypes.c_int]
        L.gala_set_graph.argtypes = [VP, ctypes.c_int32, I64P, I32P, VP, ctypes.c_int]
        L.gala_run.argtypes = [VP, ITER_CB, ROUND_CB, VP, ctypes.POINTER(Result)]
        L.gala_partial_result.argtypes = [VP, ctypes.POINTER(Result)]
        L.gala_get_communities.argtypes = [VP, I32P]
        L.gala_phase_decide.argtypes = [VP, ctypes.c_int32, I64P, I32P, I32P, ctypes.c_int64, I32P,
                                        I32P, I32P, I32P, I32P, U8P, I32P]
        L.gala_phase_weight.argtypes = [VP, ctypes.c_int32, I64P, I32P, I32P, ctypes.c_int64, I32P,
                                        I32P, I32P, I32P, ctypes.c_int64, I32P, U8P]
        L.gala_phase_contract.argtypes = [VP, ctypes.c_int32, I64P, I32P, I32P, I32P, I32P, I64P,
                                          I32P, I64P, I32P, I64P]
        L.gala_phase_modularity.argtypes = [VP, ctypes.c_int32, I64P, I32P, VP, ctypes.c_int, I32P,
                                            ctypes.POINTER(ctypes.c_double)]
        L.gala_modularity_noalloc.argtypes = [VP, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                              ctypes.POINTER(ctypes.c_double)]
        L.gala_read_bin.argtypes = [ctypes.c_char_p, ctypes.c_int, I32P, I64P, I64P, I32P, VP]
        L.gala_write_bin.argtypes = [ctypes.c_char_p, ctypes.c_int32, I64P, I32P, VP, ctypes.c_int]
        L.gala_ingest_text.argtypes = [ctypes.c_char_p, ctypes.c_int, I32P, I64P, I64P, I32P, VP]
        L.gala_ingest_text_device.argtypes = [VP, ctypes.c_char_p, ctypes.c_int, I32P, I64P, I64P, I32P, VP]
        L.gala_load_text.argtypes = [VP, ctypes.c_char_p, ctypes.c_int]
        DP = ctypes.POINTER(ctypes.c_double)
        L.gala_ingest_times.argtypes = [VP, DP, DP, DP, DP, DP]
        L.gala_gen_sbm.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_uint64, I32P, I64P, I64P, I32P]
        L.gala_gen_chung_lu.argtypes = [ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_uint64, I32P, I64P, I64P, I32P]
        L.gala_gen_rmat.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_uint64, ctypes.c_float,
                                    ctypes.c_float, I32P, I64P, I64P, I32P,
                                    ctypes.POINTER(ctypes.c_float)]
        F32P = ctypes.POINTER(ctypes.c_float)
        L.gala_set_logger.argtypes = [VP, LOG_CB, VP]
        L.gala_set_mlp.argtypes = [VP, F32P]
        L.gala_get_mlp.argtypes = [VP, F32P]
        L.gala_mlp_train.argtypes = [VP, F32P, U8P, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                     ctypes.c_float, ctypes.c_uint64, ctypes.c_int, F32P, F32P]
        L.gala_mlp_forward.argtypes = [VP, F32P, F32P, ctypes.c_int64, F32P]
        L.gala_mlp_grad.argtypes = [VP, F32P, F32P, U8P, ctypes.c_int64, ctypes.c_float, F32P, F32P]
        L.gala_batch_set.argtypes = [VP, ctypes.c_int32, I32P, ctypes.c_int32, I64P, I32P, VP, ctypes.c_int]
        L.gala_batch_run.argtypes = [VP, ctypes.POINTER(BatchResult)]
        L.gala_batch_graph_results.argtypes = [VP, ctypes.POINTER(ctypes.c_double), I32P, I32P, I32P]
        L.gala_batch_communities.argtypes = [VP, I32P]
        L.gala_gen_sbm_batch.argtypes = [ctypes.c_int32, I32P, I32P, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint64),
                                         ctypes.c_int32, I32P, I64P, I64P, I32P]
        L.gala_set_variants.argtypes = [VP, I32P]
        L.gala_get_variants.argtypes = [VP, I32P]
        L.gala_autotune.argtypes = [VP, I32P]
        L.gala_set_bin_cuts.argtypes = [VP, I32P]
        L.gala_get_bin_cuts.argtypes = [VP, I32P]
        L.gala_autotune_bins.argtypes = [VP, I32P]
        L.gala_variant_counts.argtypes = [I32P]
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        raise GalaError(f"{what}: {lib().gala_strerror(rc).decode()} ({rc})")


def _p(a, ct):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ct))


# ---------------------------------------------------------------------------
# Graph (src/graph/graph.h:7-19)
# ---------------------------------------------------------------------------
class Graph:
    """CSR graph in the reference layout: ``degrees`` are inclusive row ends."""

    def __init__(self, degrees=None, neighbors=None, weights=None, weight_mode=W_NONE):
        self.degrees = np.zeros(0, np.int64) if degrees is None else np.ascontiguousarray(degrees, np.int64)
        self.neighbors = np.zeros(0, np.int32) if neighbors is None else np.ascontiguousarray(neighbors, np.int32)
        self.weight_mode = weight_mode
        if weights is None:
            self.weights = None
        else:
            self.weights = np.ascontiguousarray(weights, np.float32 if weight_mode == W_F32 else np.int32)

    @property
    def vertex_num(self):
        return len(self.degrees)

    @property
    def edge_num(self):
        return int(self.degrees[-1]) // 2 if len(self.degrees) else 0

    @property
    def is_weighted(self):
        return self.weight_mode != W_NONE

    def offsets(self):
        return np.concatenate([[0], self.degrees]).astype(np.int64)

    @staticmethod
    def load_bin_graph(path, weight_mode=W_NONE):
        """Graph::load_bin_graph (graph.cpp:177-219)."""
        L = lib()
        n, m = ctypes.c_int32(), ctypes.c_int64()
        _check(L.gala_read_bin(path.encode(), weight_mode, ctypes.byref(n), ctypes.byref(m), None, None,
                               None), f"read {path}")
        end = np.zeros(max(n.value, 1), np.int64)
        nbr = np.zeros(max(m.value, 1), np.int32)
        w = None
        if weight_mode != W_NONE:
            w = np.zeros(max(m.value, 1), np.float32 if weight_mode == W_F32 else np.int32)
        _check(L.gala_read_bin(path.encode(), weight_mode, ctypes.byref(n), ctypes.byref(m),
                               _p(end, ctypes.c_int64), _p(nbr, ctypes.c_int32),
                               None if w is None else w.ctypes.data_as(VP)), f"read {path}")
        return Graph(end[:n.value], nbr[:m.value], None if w is None else w[:m.value], weight_mode)

    def store_bin_graph(self, path):
        """Graph::store_bin_graph (graph.cpp:153-175)."""
        w = self.weights
        _check(lib().gala_write_bin(path.encode(), self.vertex_num, _p(self.degrees, ctypes.c_int64),
                                    _p(self.neighbors, ctypes.c_int32),
                                    None if w is None else w.ctypes.data_as(VP), self.weight_mode),
               f"write {path}")

    @staticmethod
    def from_text(path, weight_mode=W_NONE, engine=None):
        """Graph::Graph(file, is_wt) (graph.cpp:18-151), i.e. preprocess's ingest:
        on the host (gala_ingest_text), or on engine's GPU when an Engine is
        given (gala_ingest_text_device; same output byte for byte)."""
        L = lib()
        if engine is None:
            fn, pre = L.gala_ingest_text, ()
        else:
            fn, pre = L.gala_ingest_text_device, (engine.ctx,)
        n, m = ctypes.c_int32(), ctypes.c_int64()
        _check(fn(*pre, path.encode(), weight_mode, ctypes.byref(n), ctypes.byref(m), None, None, None),
               f"ingest {path}")
        end = np.zeros(max(n.value, 1), np.int64)
        nbr = np.zeros(max(m.value, 1), np.int32)
        w = None if weight_mode == W_NONE else np.zeros(max(m.value, 1),
                                                          np.float32 if weight_mode == W_F32 else np.int32)
        _check(fn(*pre, path.encode(), weight_mode, ctypes.byref(n), ctypes.byref(m),
                  _p(end, ctypes.c_int64), _p(nbr, ctypes.c_int32),
                  None if w is None else w.ctypes.data_as(VP)), f"ingest {path}")
        return Graph(end[:n.value], nbr[:m.value], None if w is None else w[:m.value], weight_mode)


# ---------------------------------------------------------------------------
# generators
# ---------------------------------------------------------------------------
def _two_call(fn, args, with_w=False):
    n, m = ctypes.c_int32(), ctypes.c_int64()
    if with_w:
        _check(fn(*args, ctypes.byref(n), ctypes.byref(m), None, None, None), fn.__name__)
    else:
        _check(fn(*args, ctypes.byref(n), ctypes.byref(m), None, None), fn.__name__)
    end = np.zeros(max(n.value, 1), np.int64)
    nbr = np.zeros(max(m.value, 1), np.int32)
    if with_w:
        w = np.zeros(max(m.value, 1), np.float32)
        _check(fn(*args, ctypes.byref(n), ctypes.byref(m), _p(end, ctypes.c_int64),
                  _p(nbr, ctypes.c_int32), _p(w, ctypes.c_float)), fn.__name__)
        return Graph(end[:n.value], nbr[:m.value], w[:m.value], W_F32)
    _check(fn(*args, ctypes.byref(n), ctypes.byref(m), _p(end, ctypes.c_int64), _p(nbr, ctypes.c_int32)),
           fn.__name__)
    return Graph(end[:n.value], nbr[:m.value])


def gen_sbm(n, blocks, p_in, p_out, seed):
    return _two_call(lib().gala_gen_sbm, (n, blocks, p_in, p_out, seed))


def gen_chung_lu(n, gamma=2.1, mean_deg=17.35, max_deg=15000.0, seed=0):
    return _two_call(lib().gala_gen_chung_lu, (n, gamma, mean_deg, max_deg, seed))


def gen_rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=0, w_lo=0.5, w_hi=2.0):
    return _two_call(lib().gala_gen_rmat, (scale, edge_factor, a, b, c, seed, w_lo, w_hi), with_w=True)


def gen_sbm_batch(n, blocks, p_in, p_out, seeds, threads=0):
    """Union of SBMs (graph g = gen_sbm(n[g], blocks[g], p_in[g], p_out[g], seeds[g]),
    vertex ids offset by vstart[g]).  Returns (vstart[G+1], union Graph)."""
    G = len(n)
    n = np.ascontiguousarray(n, np.int32)
    blocks = np.ascontiguousarray(blocks, np.int32)
    p_in = np.ascontiguousarray(p_in, np.float64)
    p_out = np.ascontiguousarray(p_out, np.float64)
    seeds = np.ascontiguousarray(seeds, np.uint64)
    vstart = np.zeros(G + 1, np.int32)
    m = ctypes.c_int64()
    D = ctypes.POINTER(ctypes.c_double)
    args = (G, _p(n, ctypes.c_int32), _p(blocks, ctypes.c_int32), _p(p_in, ctypes.c_double),
            _p(p_out, ctypes.c_double), _p(seeds, ctypes.c_uint64), threads, _p(vstart, ctypes.c_int32),
            ctypes.byref(m))
    _check(lib().gala_gen_sbm_batch(*args, None, None), "gala_gen_sbm_batch")
    nt = int(vstart[-1])
    end = np.zeros(max(nt, 1), np.int64)
    nbr = np.zeros(max(m.value, 1), np.int32)
    _check(lib().gala_gen_sbm_batch(*args, _p(end, ctypes.c_int64), _p(nbr, ctypes.c_int32)), "gala_gen_sbm_batch")
    del D
    return vstart, Graph(end[:nt], nbr[:m.value])


def union_graphs(graphs):
    """Block-diagonal union of graphs (same weight mode): (vstart[G+1], Graph)."""
    sizes = [g.vertex_num for g in graphs]
    vstart = np.zeros(len(graphs) + 1, np.int32)
    vstart[1:] = np.cumsum(sizes)
    nbr = np.concatenate([g.neighbors + vstart[i] for i, g in enumerate(graphs)]).astype(np.int32)
    ends, base = [], 0
    for g in graphs:
        ends.append(g.degrees + base)
        base += int(g.degrees[-1]) if g.vertex_num else 0
    end = np.concatenate(ends).astype(np.int64)
    mode = graphs[0].weight_mode
    w = None if graphs[0].weights is None else np.concatenate([g.weights for g in graphs])
    return vstart, Graph(end, nbr, w, mode)


def batch_sbm_params(n_graphs=4096, seed=0, n_lo=64, n_hi=20000):
    """Parameters of the batched SBM workload: vertex counts log-uniform in
    [n_lo, n_hi], 4..64 blocks, p_in uniform in [0.05, 0.3],
    p_out = 0.1 p_in / blocks, one seed per graph."""
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(np.log(n_lo), np.log(n_hi), n_graphs)).astype(np.int32)
    blocks = np.minimum(rng.integers(4, 65, n_graphs), n).astype(np.int32)
    p_in = rng.uniform(0.05, 0.3, n_graphs)
    p_out = 0.1 * p_in / blocks
    seeds = (np.uint64(seed) * np.uint64(1000003) + np.arange(n_graphs, dtype=np.uint64)).astype(np.uint64)
    return n, blocks, p_in, p_out, seeds


# ---------------------------------------------------------------------------
# engine
# ---------------------------------------------------------------------------
class Engine:
    """One device context (gala_create ... gala_destroy)."""

    def __init__(self, device=0, min_modularity=1e-6, dump_iterations=False, dump_rounds=False,
                 max_rounds=0, max_iterations=0, profile_bins=False, frontier_mode=0, learned_tau=0.5,
                 hip_graphs=True, serial_bins=False, learned_min_degree=0, learned_min_vertices=0,
                 queue_sweeps=False):
        L = lib()
        self.opts = Opts()
        L.gala_default_opts(ctypes.byref(self.opts))
        self.opts.min_modularity = min_modularity
        self.opts.dump_iterations = int(dump_iterations)
        self.opts.dump_rounds = int(dump_rounds)
        self.opts.max_rounds = max_rounds
        self.opts.max_iterations = max_iterations
        self.opts.profile_bins = int(profile_bins)
        self.opts.frontier_mode = int(frontier_mode)
        self.opts.learned_tau = float(learned_tau)
        self.opts.no_hip_graphs = 0 if hip_graphs else 1
        self.opts.serial_bins = int(serial_bins)
        self.opts.learned_min_degree = int(learned_min_degree)
        self.opts.learned_min_vertices = int(learned_min_vertices)
        self.opts.queue_sweeps = int(queue_sweeps)
        self.ctx = VP()
        _check(L.gala_create(ctypes.byref(self.ctx), device, ctypes.byref(self.opts)), "gala_create")

    def close(self):
        if self.ctx:
            lib().gala_destroy(self.ctx)
            self.ctx = VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_graph(self, g):
        w = g.weights
        _check(lib().gala_set_graph(self.ctx, g.vertex_num, _p(g.degrees, ctypes.c_int64),
                                    _p(g.neighbors, ctypes.c_int32),
                                    None if w is None else w.ctypes.data_as(VP), g.weight_mode),
               "gala_set_graph")

    def load_bin(self, path, weight_mode=W_NONE):
        _check(lib().gala_load_bin(self.ctx, path.encode(), weight_mode), "gala_load_bin")

    def load_text(self, path, weight_mode=W_NONE):
        """gala_load_text: device ingest of an edge list, set as the graph."""
        _check(lib().gala_load_text(self.ctx, path.encode(), weight_mode), "gala_load_text")

    def ingest_text(self, path, weight_mode=W_NONE):
        """gala_ingest_text_device: the edge list as a Graph, ingested on the GPU."""
        return Graph.from_text(path, weight_mode, engine=self)

    def ingest_times(self):
        """Timings (ms) of the last device ingest: read, load (file to HBM),
        device, d2h, wall."""
        v = [ctypes.c_double() for _ in range(5)]
        _check(lib().gala_ingest_times(self.ctx, *[ctypes.byref(x) for x in v]), "gala_ingest_times")
        return dict(zip(("read_ms", "load_ms", "device_ms", "d2h_ms", "wall_ms"), (x.value for x in v)))

    def run(self, trace=None, rounds=None):
        """gala_run.  trace: list receiving (round, it, Q, moved, cur|None);
        rounds: list receiving dicts with the contracted CSR when dump_rounds."""
        def icb(user, rnd, it, q, moved, cur, n):
            if trace is not None:
                vec = np.ctypeslib.as_array(cur, shape=(n,)).copy() if cur else None
                trace.append((rnd, it, q, int(moved), vec))

        def rcb(user, rnd, q, n, comm, C, end, nbr, wi, wf):
            if rounds is not None:
                d = dict(round=rnd, q=q, n=n, C=C,
                         comm=np.ctypeslib.as_array(comm, shape=(n,)).copy())
                if end:
                    d["end"] = np.ctypeslib.as_array(end, shape=(max(C, 1),))[:C].copy()
                    m = int(d["end"][-1]) if C else 0
                    d["nbr"] = np.ctypeslib.as_array(nbr, shape=(max(m, 1),))[:m].copy()
                    if wi:  # exact integer weights (the wide path's scaled sums)
                        d["w"] = np.ctypeslib.as_array(wi, shape=(max(m, 1),))[:m].copy()
                    if wf:  # the wide path: the same sums on the input weights' scale
                        d["wf"] = np.ctypeslib.as_array(wf, shape=(max(m, 1),))[:m].copy()
                rounds.append(d)

        # no callback unless the caller collects its output: a round callback
        # makes the engine copy the round's community vector to the host
        self._icb = ITER_CB(icb) if trace is not None else ITER_CB()
        self._rcb = ROUND_CB(rcb) if rounds is not None else ROUND_CB()
        res = Result()
        _check(lib().gala_run(self.ctx, self._icb, self._rcb, None, ctypes.byref(res)), "gala_run")
        return res.to_dict()

    # ---- decide kernel variants ----
    def set_variants(self, v):
        v = np.ascontiguousarray(v, np.int32)
        assert v.size == N_VARIANT_SLOTS
        _check(lib().gala_set_variants(self.ctx, _p(v, ctypes.c_int32)), "gala_set_variants")

    def get_variants(self):
        v = np.zeros(N_VARIANT_SLOTS, np.int32)
        _check(lib().gala_get_variants(self.ctx, _p(v, ctypes.c_int32)), "gala_get_variants")
        return v

    def autotune(self):
        """Times every decide variant on the loaded graph and keeps the fastest
        per bin, then tunes the degree-bin cut points (get_bin_cuts)."""
        v = np.zeros(N_VARIANT_SLOTS, np.int32)
        _check(lib().gala_autotune(self.ctx, _p(v, ctypes.c_int32)), "gala_autotune")
        return v

    def set_bin_cuts(self, lo):
        lo = np.ascontiguousarray(lo, np.int32)
        assert lo.size == 9
        _check(lib().gala_set_bin_cuts(self.ctx, _p(lo, ctypes.c_int32)), "gala_set_bin_cuts")

    def get_bin_cuts(self):
        lo = np.zeros(9, np.int32)
        _check(lib().gala_get_bin_cuts(self.ctx, _p(lo, ctypes.c_int32)), "gala_get_bin_cuts")
        return lo

    def autotune_bins(self):
        """gala_autotune_bins: the degree-bin cut points, tuned on the loaded graph."""
        lo = np.zeros(9, np.int32)
        _check(lib().gala_autotune_bins(self.ctx, _p(lo, ctypes.c_int32)), "gala_autotune_bins")
        return lo

    # ---- batched graphs ----
    def batch_set(self, vstart, g):
        """Upload a batch given as its union graph and vstart[G+1]."""
        vstart = np.ascontiguousarray(vstart, np.int32)
        w = g.weights
        _check(lib().gala_batch_set(self.ctx, len(vstart) - 1, _p(vstart, ctypes.c_int32), g.vertex_num,
                                    _p(g.degrees, ctypes.c_int64), _p(g.neighbors, ctypes.c_int32),
                                    None if w is None else w.ctypes.data_as(VP), g.weight_mode), "gala_batch_set")
        self._vstart = vstart

    def batch_run(self):
        res = BatchResult()
        _check(lib().gala_batch_run(self.ctx, ctypes.byref(res)), "gala_batch_run")
        return res.to_dict()

    def batch_results(self):
        """Per graph: modularity, communities, rounds, iterations (numpy arrays)."""
        G = len(self._vstart) - 1
        q = np.zeros(G, np.float64)
        nc, rd, it = (np.zeros(G, np.int32) for _ in range(3))
        _check(lib().gala_batch_graph_results(self.ctx, _p(q, ctypes.c_double), _p(nc, ctypes.c_int32),
                                              _p(rd, ctypes.c_int32), _p(it, ctypes.c_int32)),
               "gala_batch_graph_results")
        return dict(modularity=q, communities=nc, rounds=rd, iterations=it)

    def batch_communities(self):
        """List of per-graph community vectors (ids 0.. within each graph)."""
        vs = self._vstart
        out = np.zeros(max(int(vs[-1]), 1), np.int32)
        _check(lib().gala_batch_communities(self.ctx, _p(out, ctypes.c_int32)), "gala_batch_communities")
        return [out[vs[g]:vs[g + 1]] for g in range(len(vs) - 1)]

    # ---- learned frontier ----
    def set_logger(self, sink):
        """sink(round, iteration, features[n, 6], moved[n]) for every plain sweep after the first."""
        def cb(user, rnd, it, n, feat, moved):
            f = np.ctypeslib.as_array(feat, shape=(n, MLP_NFEAT)).copy() if n else np.zeros((0, MLP_NFEAT), np.float32)
            m = np.ctypeslib.as_array(moved, shape=(n,)).copy() if n else np.zeros(0, np.uint8)
            sink(rnd, it, f, m)
        self._logcb = LOG_CB(cb) if sink is not None else LOG_CB()
        _check(lib().gala_set_logger(self.ctx, self._logcb, None), "gala_set_logger")

    def set_mlp(self, params):
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == MLP_NPARAM
        _check(lib().gala_set_mlp(self.ctx, _p(p, ctypes.c_float)), "gala_set_mlp")

    def get_mlp(self):
        p = np.zeros(MLP_NPARAM, np.float32)
        _check(lib().gala_get_mlp(self.ctx, _p(p, ctypes.c_float)), "gala_get_mlp")
        return p

    def mlp_train(self, X, y, epochs=20, batch=4096, lr=1e-2, pos_weight=1.0, seed=0, init=None):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.uint8)
        p = np.zeros(MLP_NPARAM, np.float32) if init is None else np.ascontiguousarray(init, np.float32).copy()
        loss = np.zeros(epochs, np.float32)
        _check(lib().gala_mlp_train(self.ctx, _p(X, ctypes.c_float), _p(y, ctypes.c_uint8), len(y), epochs, batch,
                                    lr, pos_weight, seed, int(init is not None), _p(p, ctypes.c_float),
                                    _p(loss, ctypes.c_float)), "gala_mlp_train")
        return p, loss

    def mlp_forward(self, params, X):
        X = np.ascontiguousarray(X, np.float32)
        p = np.ascontiguousarray(params, np.float32)
        out = np.zeros(len(X), np.float32)
        _check(lib().gala_mlp_forward(self.ctx, _p(p, ctypes.c_float), _p(X, ctypes.c_float), len(X),
                                      _p(out, ctypes.c_float)), "gala_mlp_forward")
        return out

    def mlp_grad(self, params, X, y, pos_weight=1.0):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.uint8)
        p = np.ascontiguousarray(params, np.float32)
        g = np.zeros(MLP_NPARAM, np.float32)
        loss = np.zeros(1, np.float32)
        _check(lib().gala_mlp_grad(self.ctx, _p(p, ctypes.c_float), _p(X, ctypes.c_float), _p(y, ctypes.c_uint8),
                                   len(y), pos_weight, _p(g, ctypes.c_float), _p(loss, ctypes.c_float)),
               "gala_mlp_grad")
        return g, float(loss[0])

    def communities(self, n):
        out = np.zeros(max(n, 1), np.int32)
        _check(lib().gala_get_communities(self.ctx, _p(out, ctypes.c_int32)), "gala_get_communities")
        return out[:n]

    # ---- phase-level ----
    def phase_decide(self, g, m2, K, Self, Tot, size, prev, frontier=None):
        n = g.vertex_num
        cur = np.zeros(n, np.int32)
        a = [np.ascontiguousarray(x, np.int32) for x in (K, Self, Tot, size, prev)]
        fr = None if frontier is None else np.ascontiguousarray(frontier, np.uint8)
        w = None if g.weights is None else np.ascontiguousarray(g.weights, np.int32)
        _check(lib().gala_phase_decide(self.ctx, n, _p(g.degrees, ctypes.c_int64),
                                       _p(g.neighbors, ctypes.c_int32), _p(w, ctypes.c_int32), int(m2),
                                       *[_p(x, ctypes.c_int32) for x in a], _p(fr, ctypes.c_uint8),
                                       _p(cur, ctypes.c_int32)), "gala_phase_decide")
        return cur

    def phase_weight(self, g, m2, K, Self, Tot, cur, min_tot):
        n = g.vertex_num
        In = np.zeros(n, np.int32)
        act = np.zeros(n, np.uint8)
        a = [np.ascontiguousarray(x, np.int32) for x in (K, Self, Tot, cur)]
        w = None if g.weights is None else np.ascontiguousarray(g.weights, np.int32)
        _check(lib().gala_phase_weight(self.ctx, n, _p(g.degrees, ctypes.c_int64),
                                       _p(g.neighbors, ctypes.c_int32), _p(w, ctypes.c_int32), int(m2),
                                       *[_p(x, ctypes.c_int32) for x in a], int(min_tot),
                                       _p(In, ctypes.c_int32), _p(act, ctypes.c_uint8)),
               "gala_phase_weight")
        return In, act

    def phase_contract(self, g, comm):
        n = g.vertex_num
        comm = np.ascontiguousarray(comm, np.int32)
        w = None if g.weights is None else np.ascontiguousarray(g.weights, np.int32)
        C, m = ctypes.c_int32(), ctypes.c_int64()
        nid = np.zeros(n, np.int32)
        args = (self.ctx, n, _p(g.degrees, ctypes.c_int64), _p(g.neighbors, ctypes.c_int32),
                _p(w, ctypes.c_int32), _p(comm, ctypes.c_int32), ctypes.byref(C), ctypes.byref(m))
        _check(lib().gala_phase_contract(*args, _p(nid, ctypes.c_int32), None, None, None),
               "gala_phase_contract")
        end = np.zeros(max(C.value, 1), np.int64)
        nbr = np.zeros(max(m.value, 1), np.int32)
        ww = np.zeros(max(m.value, 1), np.int64)
        _check(lib().gala_phase_contract(*args, None, _p(end, ctypes.c_int64), _p(nbr, ctypes.c_int32),
                                         _p(ww, ctypes.c_int64)), "gala_phase_contract")
        return nid, C.value, end[:C.value], nbr[:m.value], ww[:m.value]


    def phase_modularity(self, g, comm):
        """Modularity of `comm` on g (gala_phase_modularity, host arrays)."""
        comm = np.ascontiguousarray(comm, np.int32)
        q = ctypes.c_double()
        w = None if g.weights is None else g.weights
        _check(lib().gala_phase_modularity(self.ctx, g.vertex_num, _p(g.degrees, ctypes.c_int64),
                                           _p(g.neighbors, ctypes.c_int32),
                                           None if w is None else w.ctypes.data_as(ctypes.c_void_p),
                                           g.weight_mode if w is not None else W_NONE,
                                           _p(comm, ctypes.c_int32), ctypes.byref(q)),
               "gala_phase_modularity")
        return q.value

    def modularity_noalloc(self, n, d_end_incl, d_nbr, d_w, weight_mode, d_comm):
        """gala_modularity_noalloc on device addresses (ints): returns (rc, Q)."""
        q = ctypes.c_double()
        rc = lib().gala_modularity_noalloc(self.ctx, n, d_end_incl, d_nbr, d_w, weight_mode, d_comm,
                                           ctypes.byref(q))
        return rc, q.value


def louvain_gpu(graph, min_modularity=1e-6, device=0):
    """louvain_gpu (louvain.cu:136-229): returns (final Q, composed communities)."""
    eng = Engine(device=device, min_modularity=min_modularity)
    try:
        eng.set_graph(graph)
        res = eng.run()
        return res["modularity"], eng.communities(graph.vertex_num)
    finally:
        eng.close()
