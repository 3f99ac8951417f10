"""Pins the CPU oracle (oracle/louvain_oracle.c) before it is trusted as the
checker of the HIP path: the reference README's karate trajectory, closed
forms, brute force on tiny graphs, an independent modularity routine, and
contraction invariants.  CPU only."""
import itertools
import json
import os
from fractions import Fraction

import numpy as np
import pytest

import gala
import oracle


def load(golden, name):
    return gala.Graph.load_bin_graph(os.path.join(golden, name))


# ---------------------------------------------------------------------------
# known answer: README.md:58-89
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("prune", [0, 1])
def test_karate_readme_trajectory(golden, prune):
    g = load(golden, "karate.txt.bin")
    ref = json.load(open(os.path.join(golden, "karate_readme.json")))
    t = oracle.Trace(keep_vectors=False)
    r = oracle.louvain(g.vertex_num, g.degrees, g.neighbors, prune=prune, trace=t)
    for rnd, exp in enumerate(ref["rounds"]):
        got = [q for (a, _, q, _, _) in t.rows if a == rnd]
        assert [f"{q:.6f}" for q in got] == [f"{q:.6f}" for q in exp["q"]], rnd
        assert r["round_nc"][rnd] == exp["communities"]
        assert f"{r['round_q'][rnd]:.6f}" == f"{exp['modularity']:.6f}"
    assert len(r["round_q"]) == len(ref["rounds"])
    assert r["nc"] == ref["final"]["to"] and f"{r['Q']:.6f}" == "0.408695"


def test_q0_closed_form(golden):
    # Q0 = -sum k^2 / m2^2 on a simple graph (README.md:62: -0.049803, m2 = 156)
    g = load(golden, "karate.txt.bin")
    deg = np.diff(g.offsets())
    m2 = deg.sum()
    t = oracle.Trace(keep_vectors=False)
    oracle.louvain(g.vertex_num, g.degrees, g.neighbors, trace=t, max_rounds=1)
    assert m2 == 156 and (deg ** 2).sum() == 1212
    assert t.rows[0][2] == pytest.approx(-1212 / 156 ** 2, abs=1e-15)


# ---------------------------------------------------------------------------
# independent modularity (networkx, a textbook routine) on the final partition
# ---------------------------------------------------------------------------
def test_modularity_matches_networkx(golden):
    nx = pytest.importorskip("networkx")
    for name in ("karate.txt.bin", "sbm1024_s0.bin"):
        g = load(golden, name)
        r = oracle.louvain(g.vertex_num, g.degrees, g.neighbors)
        off = g.offsets()
        G = nx.Graph()
        G.add_nodes_from(range(g.vertex_num))
        for u in range(g.vertex_num):
            for v in g.neighbors[off[u]:off[u + 1]]:
                G.add_edge(u, int(v))
        parts = {}
        for v, c in enumerate(r["final_comm"]):
            parts.setdefault(int(c), set()).add(v)
        q = nx.community.modularity(G, list(parts.values()))
        assert r["Q"] == pytest.approx(q, abs=1e-12)


def test_sbm_fixture_trace(golden):
    # regression check against the committed trace (written by the oracle);
    # the trajectory itself is pinned by test_sbm_trajectory_from_definition
    g = load(golden, "sbm1024_s0.bin")
    exp = json.load(open(os.path.join(golden, "sbm1024_s0.json")))
    t = oracle.Trace(keep_vectors=False)
    r = oracle.louvain(g.vertex_num, g.degrees, g.neighbors, trace=t)
    assert [(a, b, q, mv) for a, b, q, mv, _ in t.rows] == [tuple(x) for x in exp["iterations"]]
    assert list(r["final_comm"]) == exp["final_comm"]


# ---------------------------------------------------------------------------
# brute force: exact rational gains, every candidate community
# ---------------------------------------------------------------------------
def brute_decide(off, nbr, w, prev, K, Self, Tot, size, m2):
    n = len(off) - 1
    cur = prev.copy()
    for v in range(n):
        c = prev[v]
        if off[v + 1] == off[v]:
            continue
        wsum = {}
        for j in range(off[v], off[v + 1]):
            d = prev[nbr[j]]
            wsum[d] = wsum.get(d, 0) + w[j]
        wc = wsum.get(c, 0)
        best, bdq = None, None
        for d in sorted(wsum):
            if d == c:
                continue
            dq = 2 * Fraction(wsum[d]) - 2 * Fraction(K[v] * (Tot[d] - Tot[c] + K[v]), m2)
            if dq > 0 and (bdq is None or dq > bdq):
                best, bdq = d, dq
        if best is None:
            continue
        gain = bdq - 2 * wc + 2 * Self[v]
        if gain > 0 and not (size[best] == 1 and size[c] == 1 and best > c):
            cur[v] = best
    return cur


def random_graph(rng, n, p, wmax):
    e = [(i, j) for i, j in itertools.combinations(range(n), 2) if rng.random() < p]
    loops = [(i, i) for i in range(n) if rng.random() < 0.2]
    rows = [dict() for _ in range(n)]
    for i, j in e + loops:
        wt = int(rng.integers(1, wmax + 1))
        rows[i][j] = wt
        rows[j][i] = wt
    off = [0]
    nbr, w = [], []
    for i in range(n):
        for j in sorted(rows[i]):
            nbr.append(j)
            w.append(rows[i][j])
        off.append(len(nbr))
    return np.array(off, np.int64), np.array(nbr, np.int32), np.array(w, np.int64)


@pytest.mark.parametrize("seed", range(40))
def test_decide_step_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(3, 13))
    off, nbr, w = random_graph(rng, n, 0.4, 3 if seed % 2 else 1)
    if len(nbr) == 0:
        return
    m2 = int(w.sum())
    K = np.array([w[off[v]:off[v + 1]].sum() for v in range(n)], np.int64)
    Self = np.array([sum(w[j] for j in range(off[v], off[v + 1]) if nbr[j] == v) for v in range(n)],
                    np.int64)
    prev = rng.integers(0, n, n).astype(np.int32)
    Tot = np.zeros(n, np.int64)
    size = np.zeros(n, np.int32)
    for v in range(n):
        Tot[prev[v]] += K[v]
        size[prev[v]] += 1
    got = oracle.decide_step(n, off[1:], nbr, w, m2, K, Self, Tot, size, prev)
    exp = brute_decide(off, nbr, w, prev, K, Self, Tot, size, m2)
    assert np.array_equal(got, exp)


def test_singleton_guard():
    # two isolated-pair singletons: 0-1.  Both would move to each other; the
    # guard (kernel_functions.cu:386) lets only the higher id move down.
    off = np.array([0, 1, 2], np.int64)
    nbr = np.array([1, 0], np.int32)
    w = np.array([1, 1], np.int64)
    K = np.array([1, 1], np.int64)
    got = oracle.decide_step(2, off[1:], nbr, w, 2, K, np.zeros(2, np.int64), K.copy(),
                             np.ones(2, np.int32), np.array([0, 1], np.int32))
    assert list(got) == [0, 0]


# ---------------------------------------------------------------------------
# contraction invariants
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("seed", range(10))
def test_contraction_invariants(seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(5, 60))
    off, nbr, w = random_graph(rng, n, 0.2, 4)
    if len(nbr) == 0:
        return
    comm = rng.integers(0, max(1, n // 3), n).astype(np.int32)
    nid, C, end, nb, ww = oracle.contract(n, off[1:], nbr, w, comm)
    used = sorted(set(comm.tolist()))
    assert C == len(used)
    assert [nid[c] for c in used] == list(range(C))  # rank renumbering
    assert ww.sum() == w.sum()  # m2 preserved
    off2 = np.concatenate([[0], end])
    for c in range(C):
        row = nb[off2[c]:off2[c + 1]]
        assert np.all(np.diff(row) > 0)  # canonical: ascending, unique
    # K of the new vertex = sum of member K; self-loop = sum In of members
    K = np.array([w[off[v]:off[v + 1]].sum() for v in range(n)])
    for c in range(C):
        members = [v for v in range(n) if nid[comm[v]] == c]
        assert ww[off2[c]:off2[c + 1]].sum() == K[members].sum()
        In = sum(w[j] for v in members for j in range(off[v], off[v + 1]) if comm[nbr[j]] == comm[v])
        row = nb[off2[c]:off2[c + 1]]
        selfw = ww[off2[c]:off2[c + 1]][row == c].sum()
        assert selfw == In
    # Q of the partition == Q of the singleton partition of the contracted graph
    q1 = oracle.modularity(n, off[1:], nbr, w, comm)
    q2 = oracle.modularity(C, end, nb, ww, np.arange(C, dtype=np.int32))
    assert q1 == pytest.approx(q2, abs=1e-14)


def test_prune_is_sound_on_sbm(golden):
    g = load(golden, "sbm1024_s0.bin")
    t0, t1 = oracle.Trace(), oracle.Trace()
    r0 = oracle.louvain(g.vertex_num, g.degrees, g.neighbors, prune=0, trace=t0)
    r1 = oracle.louvain(g.vertex_num, g.degrees, g.neighbors, prune=1, trace=t1)
    assert len(t0.rows) == len(t1.rows)
    for a, b in zip(t0.rows, t1.rows):
        assert a[:4] == b[:4] and np.array_equal(a[4], b[4])
    assert np.array_equal(r0["final_comm"], r1["final_comm"])
    assert r1["stats"][2] < r0["stats"][2]  # pruning evaluated fewer vertices


def test_weighted_fp_oracle_matches_int_on_integer_weights():
    # with integer-valued float weights FP64 sums are exact: same trajectory
    rng = np.random.default_rng(7)
    off, nbr, w = random_graph(rng, 80, 0.08, 5)
    n = 80
    ri = oracle.louvain(n, off[1:], nbr, w.astype(np.int32))
    rf = oracle.louvain_f(n, off[1:], nbr, w.astype(np.float32))
    assert ri["nc"] == rf["nc"] and ri["Q"] == pytest.approx(rf["Q"], abs=1e-12)
    assert np.array_equal(ri["final_comm"], rf["final_comm"])


# ---------------------------------------------------------------------------
# C1 (the 1,024-vertex SBM) end to end from the definition: every sweep by
# exact-rational brute force, Q in exact rationals, the break/revert rule,
# contraction by a dictionary merge -- none of it calls the oracle
# ---------------------------------------------------------------------------
def q_exact(off, nbr, w, comm):
    m2 = sum(int(x) for x in w)
    n = len(off) - 1
    s_in = 0
    tot = {}
    for v in range(n):
        for j in range(off[v], off[v + 1]):
            if comm[nbr[j]] == comm[v]:
                s_in += int(w[j])
            tot[comm[v]] = tot.get(comm[v], 0) + int(w[j])
    return Fraction(s_in, m2) - Fraction(sum(t * t for t in tot.values()), m2 * m2)


def contract_dict(off, nbr, w, comm):
    """Contraction from the definition: new id = rank of the community among
    the non-empty ones, every member row merged, rows ascending."""
    n = len(off) - 1
    used = sorted(set(int(c) for c in comm))
    nid = {c: i for i, c in enumerate(used)}
    rows = [dict() for _ in used]
    for v in range(n):
        r = rows[nid[int(comm[v])]]
        for j in range(off[v], off[v + 1]):
            d = nid[int(comm[nbr[j]])]
            r[d] = r.get(d, 0) + int(w[j])
    o2, nb2, w2 = [0], [], []
    for r in rows:
        for d in sorted(r):
            nb2.append(d)
            w2.append(r[d])
        o2.append(len(nb2))
    return nid, np.array(o2, np.int64), np.array(nb2, np.int32), np.array(w2, np.int64)


def louvain_from_definition(off, nbr, w, min_mod=1e-6):
    """Rows (round, iteration, Q as Fraction, cur) and the round counts."""
    rows, nc = [], []
    final = np.arange(len(off) - 1)
    prev_mod = None
    for rnd in range(64):
        n = len(off) - 1
        m2 = int(w.sum())
        K = np.array([int(w[off[v]:off[v + 1]].sum()) for v in range(n)], np.int64)
        Self = np.array([sum(int(w[j]) for j in range(off[v], off[v + 1]) if nbr[j] == v) for v in range(n)],
                        np.int64)
        prev = np.arange(n, dtype=np.int32)
        qprev = q_exact(off, nbr, w, prev)
        rows.append((rnd, 0, qprev, prev.copy()))
        for it in range(1, 100000):
            Tot = np.zeros(n, np.int64)
            size = np.zeros(n, np.int32)
            for v in range(n):
                Tot[prev[v]] += K[v]
                size[prev[v]] += 1
            cur = brute_decide(off, nbr, w, prev, K, Self, Tot, size, m2)
            q = q_exact(off, nbr, w, cur)
            rows.append((rnd, it, q, cur.copy()))
            if float(q - qprev) < min_mod:
                break
            prev, qprev = cur, q
        nid, off, nbr, w = contract_dict(off, nbr, w, prev)
        final = np.array([nid[int(prev[c])] for c in final])
        nc.append(len(off) - 1)
        cur_mod = qprev
        if prev_mod is not None and not (float(cur_mod - prev_mod) >= min_mod):
            break
        prev_mod = cur_mod
    return rows, nc, final


def test_sbm_trajectory_from_definition(golden):
    g = load(golden, "sbm1024_s0.bin")
    off = g.offsets()
    w = np.ones(len(g.neighbors), np.int64)
    rows, nc, final = louvain_from_definition(off, g.neighbors, w)
    t = oracle.Trace()
    r = oracle.louvain(g.vertex_num, g.degrees, g.neighbors, trace=t)
    assert len(t.rows) == len(rows)
    for (a, b, q, cur), (ra, rb, rq, _, rvec) in zip(rows, t.rows):
        assert (a, b) == (ra, rb)
        assert np.array_equal(cur, rvec), (a, b)
        assert abs(float(q) - rq) <= 1e-15, (a, b)
    assert list(r["round_nc"]) == nc
    assert np.array_equal(final, r["final_comm"])
