"""ctypes binding of the CPU oracle (oracle/louvain_oracle.c).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker and never as the product
path.  Every function mirrors one entry point of louvain_oracle.c; the C file's
header cites the reference lines each step restates.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

ITER_CB = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                           ctypes.POINTER(ctypes.c_int32), ctypes.c_int32, ctypes.c_int64)

_P = np.ctypeslib.ndpointer


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        i64p = ctypes.POINTER(ctypes.c_int64)
        i32p = ctypes.POINTER(ctypes.c_int32)
        L.orc_louvain.restype = ctypes.c_int
        L.orc_louvain.argtypes = [ctypes.c_int32, i64p, i32p, i32p, ctypes.c_double, ctypes.c_int,
                                  ctypes.c_int, ITER_CB, ctypes.c_void_p, i32p,
                                  ctypes.POINTER(ctypes.c_double), i32p,
                                  ctypes.POINTER(ctypes.c_double), i32p, i64p]
        L.orc_louvain_w64.restype = ctypes.c_int
        L.orc_louvain_w64.argtypes = [ctypes.c_int32, i64p, i32p, i64p, ctypes.c_double, ctypes.c_int,
                                      ctypes.c_int, ITER_CB, ctypes.c_void_p, i32p,
                                      ctypes.POINTER(ctypes.c_double), i32p,
                                      ctypes.POINTER(ctypes.c_double), i32p, i64p]
        L.orc_scale_f32.restype = ctypes.c_int
        L.orc_scale_f32.argtypes = [ctypes.c_int64, ctypes.POINTER(ctypes.c_float), i64p, i32p]
        L.orc_round.restype = ctypes.c_int
        L.orc_round.argtypes = [ctypes.c_int32, i64p, i32p, i64p, ctypes.c_int64, ctypes.c_double,
                                ctypes.c_int, ctypes.c_int, ITER_CB, ctypes.c_void_p, i32p,
                                ctypes.POINTER(ctypes.c_double), i32p]
        L.orc_contract.restype = ctypes.c_int
        L.orc_contract.argtypes = [ctypes.c_int32, i64p, i32p, i64p, i32p, i32p, i32p,
                                   ctypes.POINTER(i64p), ctypes.POINTER(i32p), ctypes.POINTER(i64p)]
        L.orc_free.argtypes = [ctypes.c_void_p]
        L.orc_decide_step.restype = ctypes.c_int
        L.orc_decide_step.argtypes = [ctypes.c_int32, i64p, i32p, i64p, ctypes.c_int64, i64p, i64p,
                                      i64p, i32p, i32p, ctypes.POINTER(ctypes.c_uint8), i32p]
        L.orc_louvain_f.restype = ctypes.c_int
        L.orc_louvain_f.argtypes = [ctypes.c_int32, i64p, i32p, ctypes.POINTER(ctypes.c_float),
                                    ctypes.c_double, ctypes.c_int, i32p,
                                    ctypes.POINTER(ctypes.c_double), i32p,
                                    ctypes.POINTER(ctypes.c_double), i32p]
        L.orc_modularity.restype = ctypes.c_double
        L.orc_modularity.argtypes = [ctypes.c_int32, i64p, i32p, i64p, i32p]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_max_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def set_threads(t):
    lib().orc_set_threads(int(t))


def max_threads():
    return lib().orc_max_threads()


class Trace:
    """Collects the per-iteration callback: (round, iteration, Q, moved[, cur copy])."""

    def __init__(self, keep_vectors=True):
        self.keep = keep_vectors
        self.rows = []

        def cb(user, rnd, it, q, cur, n, moved):
            vec = np.ctypeslib.as_array(cur, shape=(n,)).copy() if (self.keep and n > 0) else None
            self.rows.append((rnd, it, q, int(moved), vec))
        self.cfun = ITER_CB(cb)


def louvain(n, end, nbr, w=None, min_mod=1e-6, prune=1, max_rounds=0, trace=None):
    """Full multi-round run on integer weights (None = unit, int32, or int64
    e.g. scaled float32 weights).  Returns dict(final_comm, Q, nc, round_q,
    round_nc, stats)."""
    L = lib()
    end = np.ascontiguousarray(end, np.int64)
    nbr = np.ascontiguousarray(nbr, np.int32)
    wide = w is not None and np.asarray(w).dtype == np.int64
    if wide:
        w = np.ascontiguousarray(w, np.int64)
    else:
        w = np.ones(len(nbr), np.int32) if w is None else np.ascontiguousarray(w, np.int32)
    comm = np.zeros(max(n, 1), np.int32)
    q = ctypes.c_double()
    nc = ctypes.c_int32()
    rq = np.zeros(256, np.float64)
    rnc = np.zeros(256, np.int32)
    stats = np.zeros(4, np.int64)
    cb = trace.cfun if trace is not None else ITER_CB()
    fn = L.orc_louvain_w64 if wide else L.orc_louvain
    rc = fn(n, _p(end, ctypes.c_int64), _p(nbr, ctypes.c_int32),
            _p(w, ctypes.c_int64 if wide else ctypes.c_int32),
            min_mod, prune, max_rounds, cb, None, _p(comm, ctypes.c_int32),
            ctypes.byref(q), ctypes.byref(nc), _p(rq, ctypes.c_double),
            _p(rnc, ctypes.c_int32), _p(stats, ctypes.c_int64))
    if rc != 0:
        raise RuntimeError(f"orc_louvain failed: {rc}")
    r = int(stats[0])
    return dict(final_comm=comm[:n], Q=q.value, nc=nc.value, round_q=rq[:r].copy(),
                round_nc=rnc[:r].copy(), stats=stats.copy())


def scale_f32(wf):
    """Exact integer image of positive float32 weights -> (W int64[m], k) with
    wf == W / 2^k exactly and k the smallest such shift."""
    L = lib()
    wf = np.ascontiguousarray(wf, np.float32)
    out = np.zeros(max(len(wf), 1), np.int64)
    k = ctypes.c_int32()
    rc = L.orc_scale_f32(len(wf), _p(wf, ctypes.c_float), _p(out, ctypes.c_int64), ctypes.byref(k))
    if rc != 0:
        raise RuntimeError(f"orc_scale_f32 failed: {rc}")
    return out[:len(wf)], k.value


def louvain_f(n, end, nbr, wf, min_mod=1e-6, max_rounds=0):
    L = lib()
    end = np.ascontiguousarray(end, np.int64)
    nbr = np.ascontiguousarray(nbr, np.int32)
    wf = np.ascontiguousarray(wf, np.float32)
    comm = np.zeros(max(n, 1), np.int32)
    q = ctypes.c_double()
    nc = ctypes.c_int32()
    rq = np.zeros(256, np.float64)
    rnc = np.zeros(256, np.int32)
    rc = L.orc_louvain_f(n, _p(end, ctypes.c_int64), _p(nbr, ctypes.c_int32),
                         _p(wf, ctypes.c_float), min_mod, max_rounds, _p(comm, ctypes.c_int32),
                         ctypes.byref(q), ctypes.byref(nc), _p(rq, ctypes.c_double),
                         _p(rnc, ctypes.c_int32))
    if rc != 0:
        raise RuntimeError(f"orc_louvain_f failed: {rc}")
    return dict(final_comm=comm[:n], Q=q.value, nc=nc.value)


def round_(n, end, nbr, w, m2, min_mod=1e-6, prune=1, rnd=0, trace=None):
    """One phase-1 round on an int64-weighted graph -> (comm, Q, iterations)."""
    L = lib()
    end = np.ascontiguousarray(end, np.int64)
    nbr = np.ascontiguousarray(nbr, np.int32)
    w = np.ascontiguousarray(w, np.int64)
    comm = np.zeros(max(n, 1), np.int32)
    q = ctypes.c_double()
    it = ctypes.c_int32()
    cb = trace.cfun if trace is not None else ITER_CB()
    rc = L.orc_round(n, _p(end, ctypes.c_int64), _p(nbr, ctypes.c_int32), _p(w, ctypes.c_int64),
                     int(m2), min_mod, prune, rnd, cb, None, _p(comm, ctypes.c_int32),
                     ctypes.byref(q), ctypes.byref(it))
    if rc != 0:
        raise RuntimeError(f"orc_round failed: {rc}")
    return comm[:n], q.value, it.value


def contract(n, end, nbr, w, comm):
    """Contraction -> (new_id[n], C, end_incl[C], nbr, w(int64))."""
    L = lib()
    end = np.ascontiguousarray(end, np.int64)
    nbr = np.ascontiguousarray(nbr, np.int32)
    w = np.ascontiguousarray(w, np.int64)
    comm = np.ascontiguousarray(comm, np.int32)
    nid = np.zeros(max(n, 1), np.int32)
    on = ctypes.c_int32()
    pe = ctypes.POINTER(ctypes.c_int64)()
    pn = ctypes.POINTER(ctypes.c_int32)()
    pw = ctypes.POINTER(ctypes.c_int64)()
    rc = L.orc_contract(n, _p(end, ctypes.c_int64), _p(nbr, ctypes.c_int32), _p(w, ctypes.c_int64),
                        _p(comm, ctypes.c_int32), _p(nid, ctypes.c_int32), ctypes.byref(on),
                        ctypes.byref(pe), ctypes.byref(pn), ctypes.byref(pw))
    if rc != 0:
        raise RuntimeError(f"orc_contract failed: {rc}")
    C = on.value
    e2 = np.ctypeslib.as_array(pe, shape=(max(C, 1),))[:C].copy()
    m = int(e2[-1]) if C > 0 else 0
    nb = np.ctypeslib.as_array(pn, shape=(max(m, 1),))[:m].copy()
    ww = np.ctypeslib.as_array(pw, shape=(max(m, 1),))[:m].copy()
    for p in (pe, pn, pw):
        L.orc_free(ctypes.cast(p, ctypes.c_void_p))
    return nid[:n], C, e2, nb, ww


def decide_step(n, end, nbr, w, m2, K, Self, Tot, size, prev, frontier=None):
    L = lib()
    cur = np.zeros(max(n, 1), np.int32)
    args = [np.ascontiguousarray(a, t) for a, t in
            ((end, np.int64), (nbr, np.int32), (w, np.int64), (K, np.int64), (Self, np.int64),
             (Tot, np.int64), (size, np.int32), (prev, np.int32))]
    fr = None if frontier is None else np.ascontiguousarray(frontier, np.uint8)
    rc = L.orc_decide_step(n, _p(args[0], ctypes.c_int64), _p(args[1], ctypes.c_int32),
                           _p(args[2], ctypes.c_int64), int(m2), _p(args[3], ctypes.c_int64),
                           _p(args[4], ctypes.c_int64), _p(args[5], ctypes.c_int64),
                           _p(args[6], ctypes.c_int32), _p(args[7], ctypes.c_int32),
                           None if fr is None else _p(fr, ctypes.c_uint8), _p(cur, ctypes.c_int32))
    if rc != 0:
        raise RuntimeError(f"orc_decide_step failed: {rc}")
    return cur[:n]


def modularity(n, end, nbr, w, comm):
    L = lib()
    end = np.ascontiguousarray(end, np.int64)
    nbr = np.ascontiguousarray(nbr, np.int32)
    w = np.ascontiguousarray(w, np.int64)
    comm = np.ascontiguousarray(comm, np.int32)
    return L.orc_modularity(n, _p(end, ctypes.c_int64), _p(nbr, ctypes.c_int32),
                            _p(w, ctypes.c_int64), _p(comm, ctypes.c_int32))


def read_bin(path, weighted=False):
    """Reader for the reference .bin layout (graph.cpp:177-219): int32 n,
    int64 end[n] (inclusive), int32 nbr[end[n-1]], optional int32 w[...]."""
    b = open(path, "rb").read()
    n = int(np.frombuffer(b[:4], np.int32)[0])
    end = np.frombuffer(b[4:4 + 8 * n], np.int64).copy()
    m = int(end[-1]) if n > 0 else 0
    o = 4 + 8 * n
    nbr = np.frombuffer(b[o:o + 4 * m], np.int32).copy()
    w = np.frombuffer(b[o + 4 * m:o + 8 * m], np.int32).copy() if weighted else None
    return n, end, nbr, w
