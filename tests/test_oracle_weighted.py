"""Pins the oracle's weighted paths before they check the GPU (CPU only).

Float32 weights are run exactly: orc_scale_f32 maps them to integers W / 2^k
(k the smallest shift that makes every weight integral) and the exact-integer
oracle runs on those.  Decisions and Q are invariant under a common scale of
the weights, so that run is Louvain in exact arithmetic on the float values.
It is pinned here against:

* the IEEE-754 value of every weight (Fraction), and the minimality of k;
* brute force in exact rationals of every candidate community of a sweep
  (Fractions of the float weights), on graphs with self-loops and with weights
  drawn from a small set so that exact ties occur;
* networkx's weighted modularity of final partitions;
* contraction invariants (m2 preserved, Q of the singleton partition of the
  contracted graph equal to Q before);
* scale invariance (all weights doubled: identical run).

The FP64 oracle (orc_louvain_f: double sums in row order) is pinned by
agreeing with the exact run: same partition and Q within 1e-12 on random
weighted graphs.  A truncating weight cast or single-precision accumulation in
either oracle fails these tests.
"""
import itertools
from fractions import Fraction

import numpy as np
import pytest

import gala
import oracle
from test_oracle import brute_decide


def weighted_graph(rng, n, p, weights, loops=0.2):
    """Symmetric random graph with float32 weights drawn by `weights(rng)`."""
    e = [(i, j) for i, j in itertools.combinations(range(n), 2) if rng.random() < p]
    lp = [(i, i) for i in range(n) if rng.random() < loops]
    rows = [dict() for _ in range(n)]
    for i, j in e + lp:
        wt = np.float32(weights(rng))
        rows[i][j] = wt
        rows[j][i] = wt
    off, nbr, w = [0], [], []
    for i in range(n):
        for j in sorted(rows[i]):
            nbr.append(j)
            w.append(rows[i][j])
        off.append(len(nbr))
    return np.array(off, np.int64), np.array(nbr, np.int32), np.array(w, np.float32)


def uniform_w(rng):
    return rng.uniform(0.5, 2.0)


def tie_w(rng):
    return [0.5, 0.75, 1.25][int(rng.integers(0, 3))]


# ---------------------------------------------------------------------------
# the exact scaling
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("kind", ["uniform", "integers", "decimal", "tiny", "mixed"])
def test_scale_f32_is_exact_and_minimal(kind):
    rng = np.random.default_rng(3)
    m = 5000
    wf = {
        "uniform": lambda: rng.uniform(0.5, 2.0, m),
        "integers": lambda: rng.integers(1, 1000, m).astype(np.float64),
        "decimal": lambda: rng.choice([0.1, 0.2, 0.3, 1.7], m),
        "tiny": lambda: rng.uniform(1e-30, 1e-29, m),
        "mixed": lambda: rng.choice([1e-3, 1.0, 3e3], m) * rng.uniform(1.0, 2.0, m),
    }[kind]().astype(np.float32)
    W, k = oracle.scale_f32(wf)
    assert k >= 0
    for x, X in zip(wf.tolist(), W.tolist()):
        assert Fraction(x) == Fraction(X, 2 ** k)
    if k > 0:
        assert np.any(W % 2 == 1)  # one bit less would make some weight fractional
    if kind == "integers":
        assert k == 0 and np.array_equal(W, wf.astype(np.int64))


def test_scale_f32_rejects_bad_weights_and_range():
    for bad in ([1.0, 0.0], [1.0, -2.0], [np.inf], [np.nan]):
        with pytest.raises(RuntimeError):
            oracle.scale_f32(np.array(bad, np.float32))
    # exact sum beyond 2^62 (needs 2^110 of range)
    with pytest.raises(RuntimeError):
        oracle.scale_f32(np.array([1e-20, 1e13], np.float32))
    # subnormals are exact too
    W, k = oracle.scale_f32(np.array([1e-45, 3e-45], np.float32))
    assert k == 149 and list(W) == [1, 2]


# ---------------------------------------------------------------------------
# one sweep against exact rationals
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("seed", range(40))
def test_decide_step_brute_force_fp32(seed):
    rng = np.random.default_rng(1000 + seed)
    n = int(rng.integers(3, 13))
    off, nbr, wf = weighted_graph(rng, n, 0.45, tie_w if seed % 2 else uniform_w)
    if len(nbr) == 0:
        return
    W, k = oracle.scale_f32(wf)
    m2 = int(W.sum())
    K = np.array([W[off[v]:off[v + 1]].sum() for v in range(n)], np.int64)
    Self = np.array([sum(int(W[j]) for j in range(off[v], off[v + 1]) if nbr[j] == v) for v in range(n)],
                    np.int64)
    prev = rng.integers(0, n, n).astype(np.int32)
    Tot = np.zeros(n, np.int64)
    size = np.zeros(n, np.int32)
    for v in range(n):
        Tot[prev[v]] += K[v]
        size[prev[v]] += 1
    got = oracle.decide_step(n, off[1:], nbr, W, m2, K, Self, Tot, size, prev)
    # the same sweep in exact rationals of the float values themselves
    F = [Fraction(float(x)) for x in wf]
    KF = [sum(F[off[v]:off[v + 1]], Fraction(0)) for v in range(n)]
    SF = [sum((F[j] for j in range(off[v], off[v + 1]) if nbr[j] == v), Fraction(0)) for v in range(n)]
    TF = [Fraction(0)] * n
    for v in range(n):
        TF[prev[v]] += KF[v]
    exp = brute_decide(off, nbr, F, prev, KF, SF, TF, size, sum(F, Fraction(0)))
    assert np.array_equal(got, exp)


# ---------------------------------------------------------------------------
# full runs
# ---------------------------------------------------------------------------
def nx_modularity(n, off, nbr, wf, comm):
    nx = pytest.importorskip("networkx")
    G = nx.Graph()
    G.add_nodes_from(range(n))
    for u in range(n):
        for j in range(off[u], off[u + 1]):
            v = int(nbr[j])
            if u <= v:
                G.add_edge(u, v, weight=float(wf[j]))
    parts = {}
    for v, c in enumerate(comm):
        parts.setdefault(int(c), set()).add(v)
    return nx.community.modularity(G, list(parts.values()), weight="weight")


@pytest.mark.parametrize("seed", range(4))
def test_weighted_run_matches_networkx(seed):
    rng = np.random.default_rng(2000 + seed)
    n = 150
    # no self-loops: networkx counts a loop twice in the degree, the CSR once
    off, nbr, wf = weighted_graph(rng, n, 0.05, uniform_w, loops=0.0)
    W, k = oracle.scale_f32(wf)
    r = oracle.louvain(n, off[1:], nbr, W)
    assert r["Q"] == pytest.approx(nx_modularity(n, off, nbr, wf, r["final_comm"]), abs=1e-12)
    assert r["nc"] < n


@pytest.mark.parametrize("seed", range(6))
def test_weighted_contraction_invariants_exact(seed):
    rng = np.random.default_rng(3000 + seed)
    n = int(rng.integers(8, 60))
    off, nbr, wf = weighted_graph(rng, n, 0.2, uniform_w)
    if len(nbr) == 0:
        return
    W, k = oracle.scale_f32(wf)
    comm = rng.integers(0, max(1, n // 3), n).astype(np.int32)
    nid, C, end, nb, ww = oracle.contract(n, off[1:], nbr, W, comm)
    assert int(ww.sum()) == int(W.sum())  # m2 preserved exactly
    q1 = oracle.modularity(n, off[1:], nbr, W, comm)
    q2 = oracle.modularity(C, end, nb, ww, np.arange(C, dtype=np.int32))
    assert q1 == q2  # identical exact sums, identical doubles
    # the merged weights are the exact sums of the float values
    F = [Fraction(float(x)) for x in wf]
    off2 = np.concatenate([[0], end])
    for c in range(C):
        members = [v for v in range(n) if nid[comm[v]] == c]
        assert Fraction(int(ww[off2[c]:off2[c + 1]].sum()), 2 ** k) == sum(
            (F[j] for v in members for j in range(off[v], off[v + 1])), Fraction(0))


def test_scale_invariance():
    rng = np.random.default_rng(11)
    off, nbr, wf = weighted_graph(rng, 200, 0.04, uniform_w)
    W, k = oracle.scale_f32(wf)
    t1, t2 = oracle.Trace(), oracle.Trace()
    r1 = oracle.louvain(200, off[1:], nbr, W, trace=t1)
    r2 = oracle.louvain(200, off[1:], nbr, W * 8, trace=t2)
    assert r1["Q"] == r2["Q"] and np.array_equal(r1["final_comm"], r2["final_comm"])
    assert len(t1.rows) == len(t2.rows)
    for a, b in zip(t1.rows, t2.rows):
        assert a[:4] == b[:4] and np.array_equal(a[4], b[4])


def test_integer_valued_floats_take_scale_zero():
    rng = np.random.default_rng(12)
    off, nbr, wf = weighted_graph(rng, 120, 0.06, lambda r: float(r.integers(1, 6)))
    W, k = oracle.scale_f32(wf)
    assert k == 0
    a = oracle.louvain(120, off[1:], nbr, wf.astype(np.int32))
    b = oracle.louvain(120, off[1:], nbr, W)
    assert a["Q"] == b["Q"] and np.array_equal(a["final_comm"], b["final_comm"])


# ---------------------------------------------------------------------------
# the FP64 oracle against the exact run
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("seed", range(6))
def test_fp64_oracle_agrees_with_exact_run(seed):
    rng = np.random.default_rng(4000 + seed)
    n = 400
    off, nbr, wf = weighted_graph(rng, n, 0.02, uniform_w, loops=0.05)
    W, k = oracle.scale_f32(wf)
    ex = oracle.louvain(n, off[1:], nbr, W)
    fp = oracle.louvain_f(n, off[1:], nbr, wf)
    assert np.array_equal(ex["final_comm"], fp["final_comm"])
    assert abs(ex["Q"] - fp["Q"]) <= 1e-12


def test_fp64_oracle_agrees_on_rmat():
    g = gala.gen_rmat(12, 16, seed=12)
    W, k = oracle.scale_f32(g.weights)
    ex = oracle.louvain(g.vertex_num, g.degrees, g.neighbors, W)
    fp = oracle.louvain_f(g.vertex_num, g.degrees, g.neighbors, g.weights)
    assert abs(ex["Q"] - fp["Q"]) <= 1e-12
    assert ex["nc"] == fp["nc"]
