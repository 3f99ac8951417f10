"""Batched Louvain: many independent graphs as one block-diagonal union.  Every
graph of a batch must get exactly what its own run gives (the oracle's final
communities, Q, rounds and iterations), whatever else shares the launches."""
import numpy as np
import pytest

import gala
import oracle


def say(msg):
    print(msg, flush=True)


def small_batch(n_graphs, seed, n_lo=16, n_hi=600):
    P = gala.batch_sbm_params(n_graphs, seed, n_lo=n_lo, n_hi=n_hi)
    return P, gala.gen_sbm_batch(*P)


def split(vstart, g):
    """Per-graph Graphs back out of a union."""
    off = g.offsets()
    out = []
    for i in range(len(vstart) - 1):
        a, b = int(vstart[i]), int(vstart[i + 1])
        deg = g.degrees[a:b] - off[a]
        nbr = g.neighbors[off[a]:off[b]] - a
        w = None if g.weights is None else g.weights[off[a]:off[b]]
        out.append(gala.Graph(deg, nbr, w, g.weight_mode))
    return out


# ---- CPU: generator and union layout ----
def test_batch_generator_equals_per_graph_generator():
    P, (vstart, g) = small_batch(24, 5)
    parts = split(vstart, g)
    for i, gi in enumerate(parts):
        ref = gala.gen_sbm(int(P[0][i]), int(P[1][i]), float(P[2][i]), float(P[3][i]), int(P[4][i]))
        assert np.array_equal(gi.degrees, ref.degrees) and np.array_equal(gi.neighbors, ref.neighbors), i


def test_batch_params_follow_the_workload_definition():
    n, blocks, p_in, p_out, seeds = gala.batch_sbm_params(4096, 0)
    assert n.min() >= 64 and n.max() <= 20000
    # log-uniform: the median sits near the geometric mean of the range
    assert abs(np.log(np.median(n)) - 0.5 * (np.log(64) + np.log(20000))) < 0.2
    assert blocks.min() >= 4 and blocks.max() <= 64
    assert p_in.min() >= 0.05 and p_in.max() <= 0.3
    assert np.allclose(p_out, 0.1 * p_in / blocks)
    assert len(set(seeds.tolist())) == 4096


def test_union_helper_roundtrip():
    gs = [gala.gen_sbm(50, 5, 0.3, 0.01, s) for s in range(3)]
    vstart, u = gala.union_graphs(gs)
    for a, b in zip(split(vstart, u), gs):
        assert np.array_equal(a.degrees, b.degrees) and np.array_equal(a.neighbors, b.neighbors)


# ---- GPU ----
def check_against_oracle(parts, res, comm):
    for i, gi in enumerate(parts):
        if gi.edge_num == 0:
            assert res["modularity"][i] == 0.0 and res["rounds"][i] == 0
            assert np.array_equal(comm[i], np.arange(gi.vertex_num))
            continue
        ref = oracle.louvain(gi.vertex_num, gi.degrees, gi.neighbors, gi.weights)
        assert np.array_equal(comm[i], ref["final_comm"]), i
        assert res["modularity"][i] == ref["Q"], (i, res["modularity"][i], ref["Q"])
        assert res["communities"][i] == ref["nc"], i
        assert res["rounds"][i] == ref["stats"][0], i
        assert res["iterations"][i] == ref["stats"][1], i


@pytest.mark.gpu
@pytest.mark.parametrize("graphs", [True, False])
def test_batch_matches_oracle_per_graph(graphs):
    say(f"batch of 48 SBMs vs oracle (hip graphs {graphs})")
    P, (vstart, g) = small_batch(48, 11)
    eng = gala.Engine(hip_graphs=graphs)
    eng.batch_set(vstart, g)
    r = eng.batch_run()
    res = eng.batch_results()
    comm = eng.batch_communities()
    eng.close()
    assert r["n_graphs"] == 48 and r["rounds"] == res["rounds"].max()
    check_against_oracle(split(vstart, g), res, comm)


@pytest.mark.gpu
def test_batch_edge_cases():
    say("batch edge cases")
    rng = np.random.default_rng(3)
    graphs = [
        gala.Graph(np.zeros(5, np.int64), np.zeros(0, np.int32)),          # no edge at all
        gala.Graph(np.array([1, 2], np.int64), np.array([1, 0], np.int32)),  # one edge
        gala.gen_sbm(200, 4, 0.2, 0.01, 1),
        # star: hub 0 with 300 leaves, plus isolated vertices at the end
        gala.Graph(np.concatenate([[300], 300 + np.arange(1, 301), [600, 600]]).astype(np.int64),
                   np.concatenate([np.arange(1, 301), np.zeros(300)]).astype(np.int32)),
        gala.gen_chung_lu(3000, 2.1, 8.0, 400.0, 2),                        # hubs in wide bins
        gala.gen_sbm(64, 64, 0.3, 0.1 * 0.3 / 64, 4),                       # one vertex per block
    ]
    # a triangle with a self-loop on vertex 0 (rows ascending, self entry once)
    graphs.append(gala.Graph(np.array([3, 5, 7], np.int64), np.array([0, 1, 2, 0, 2, 0, 1], np.int32)))
    for _ in range(5):
        graphs.append(gala.gen_sbm(int(rng.integers(10, 300)), int(rng.integers(2, 9)), 0.25, 0.01,
                                   int(rng.integers(1 << 30))))
    vstart, u = gala.union_graphs(graphs)
    eng = gala.Engine()
    eng.batch_set(vstart, u)
    eng.batch_run()
    res = eng.batch_results()
    comm = eng.batch_communities()
    eng.close()
    check_against_oracle(split(vstart, u), res, comm)


@pytest.mark.gpu
def test_batch_weighted_int():
    say("weighted batch")
    rng = np.random.default_rng(9)
    graphs = []
    for s in range(12):
        g = gala.gen_sbm(int(rng.integers(30, 400)), int(rng.integers(2, 10)), 0.2, 0.02, 100 + s)
        # symmetric weights: weight of (u,v) = weight of (v,u)
        off = g.offsets()
        src = np.repeat(np.arange(g.vertex_num), np.diff(off))
        key = np.minimum(src, g.neighbors).astype(np.int64) * g.vertex_num + np.maximum(src, g.neighbors)
        _, inv = np.unique(key, return_inverse=True)
        w = (1 + (inv * 7919) % 8).astype(np.int32)
        graphs.append(gala.Graph(g.degrees, g.neighbors, w, gala.W_I32))
    vstart, u = gala.union_graphs(graphs)
    eng = gala.Engine()
    eng.batch_set(vstart, u)
    eng.batch_run()
    res = eng.batch_results()
    comm = eng.batch_communities()
    eng.close()
    check_against_oracle(split(vstart, u), res, comm)


@pytest.mark.gpu
def test_batch_matches_single_graph_engine():
    # the same graphs through the single-graph path, one run each
    say("batch vs per-graph engine")
    P, (vstart, g) = small_batch(150, 21, n_lo=64, n_hi=4000)
    eng = gala.Engine()
    eng.batch_set(vstart, g)
    eng.batch_run()
    res = eng.batch_results()
    comm = eng.batch_communities()
    for i, gi in enumerate(split(vstart, g)):
        if gi.edge_num == 0:
            continue
        eng.set_graph(gi)
        r = eng.run()
        assert r["modularity"] == res["modularity"][i], i
        assert np.array_equal(eng.communities(gi.vertex_num), comm[i]), i
        assert r["rounds"] == res["rounds"][i] and r["total_iterations"] == res["iterations"][i], i
    eng.close()


@pytest.mark.gpu
def test_batch_rejects_cross_graph_edges():
    gs = [gala.gen_sbm(40, 4, 0.3, 0.01, 1), gala.gen_sbm(40, 4, 0.3, 0.01, 2)]
    vstart, u = gala.union_graphs(gs)
    bad = u.neighbors.copy()
    bad[0] = 50  # vertex 0 of graph 0 names a vertex of graph 1
    eng = gala.Engine()
    with pytest.raises(gala.GalaError):
        eng.batch_set(vstart, gala.Graph(u.degrees, bad))
    eng.close()


@pytest.mark.gpu
def test_full_batch_4096_sbms():
    # The batched workload at its full size: 4,096 SBMs, vertex counts
    # log-uniform in [64, 20,000] (13.9M vertices, 1.29B directed entries in
    # the union).  Every graph's reported Q is the modularity of its returned
    # partition (exact integer sums); the oracle's own run matches every graph
    # of 5,000+ vertices (the dense ones, degree bins 6-7) and a seeded sample
    # of 256 of the rest, partition, Q, rounds and iterations.
    from concurrent.futures import ThreadPoolExecutor
    say("full batch: generate 4,096 SBMs")
    P = gala.batch_sbm_params(4096, 0)
    vstart, g = gala.gen_sbm_batch(*P)
    say(f"  union: {g.vertex_num} vertices, {len(g.neighbors)} entries; run")
    eng = gala.Engine()
    eng.batch_set(vstart, g)
    r = eng.batch_run()
    res = eng.batch_results()
    comm = np.concatenate(eng.batch_communities())
    eng.close()
    say(f"  batch run {r['t_total_ms']:.1f} ms, {r['rounds']} rounds; checking Q of every graph")
    G = len(vstart) - 1
    off = g.offsets()
    deg = np.diff(off)
    gid = np.repeat(np.arange(G), np.diff(vstart))
    gcomm = comm + vstart[gid].astype(np.int32)  # globally unique community ids
    src = np.repeat(np.arange(g.vertex_num, dtype=np.int32), deg)
    same = gcomm[g.neighbors] == gcomm[src]
    s_in = np.bincount(gid[src[same]], minlength=G).astype(np.int64)
    m2 = np.bincount(gid, weights=deg, minlength=G).astype(np.int64)
    tot = np.bincount(gcomm, weights=deg, minlength=g.vertex_num).astype(np.int64)
    cgid = np.repeat(np.arange(G), np.diff(vstart))  # graph of every community slot
    t2 = np.zeros(G, np.int64)
    np.add.at(t2, cgid, tot * tot)
    del src, same
    for i in range(G):
        if m2[i] == 0:
            assert res["modularity"][i] == 0.0
            continue
        q = int(s_in[i]) / int(m2[i]) - float(int(t2[i])) / (int(m2[i]) * int(m2[i]))
        assert res["modularity"][i] == q, i
    n = P[0]
    rng = np.random.default_rng(0)
    small = np.where(n < 5000)[0]
    check = np.concatenate([np.where(n >= 5000)[0], rng.choice(small, 256, replace=False)])
    say(f"  oracle on {len(check)} graphs")
    oracle.set_threads(1)

    def one(i):
        a, b = int(vstart[i]), int(vstart[i + 1])
        d = g.degrees[a:b] - off[a]
        nb = g.neighbors[off[a]:off[b]] - a
        if len(nb) == 0:
            return i, None
        return i, oracle.louvain(b - a, d, nb)

    with ThreadPoolExecutor(max_workers=16) as ex:
        for i, ref in ex.map(one, check.tolist()):
            if ref is None:
                continue
            a, b = int(vstart[i]), int(vstart[i + 1])
            assert np.array_equal(comm[a:b], ref["final_comm"]), i
            assert res["modularity"][i] == ref["Q"], i
            assert res["communities"][i] == ref["nc"], i
            assert res["rounds"][i] == ref["stats"][0], i
            assert res["iterations"][i] == ref["stats"][1], i
