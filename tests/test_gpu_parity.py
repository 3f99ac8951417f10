"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bar: bit-exact for every integer quantity -- per-iteration community vectors,
moved counts, per-round renumbered contracted CSR (rows canonical when the
CSR is handed out), the composed final partition -- and identical doubles for Q, which both sides
compute as (double)sumIn/m2 - (double)sumTot2/m2^2 from exact integer sums.
"""
import os

import numpy as np
import pytest

import gala
import oracle

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def full_trace_compare(eng, g, keep_vectors=True):
    trace, rounds = [], []
    eng.set_graph(g)
    res = eng.run(trace=trace, rounds=rounds)
    t = oracle.Trace(keep_vectors=keep_vectors)
    ref = oracle.louvain(g.vertex_num, g.degrees, g.neighbors,
                         None if g.weights is None else g.weights, trace=t)
    for a, b in zip(trace, t.rows):
        assert a[0] == b[0] and a[1] == b[1], (a[:2], b[:2])
        assert a[2] == b[2], f"Q differs at round {a[0]} it {a[1]}: {a[2]!r} vs {b[2]!r}"
        assert a[3] == b[3], f"moved differs at round {a[0]} it {a[1]}"
        if keep_vectors and a[4] is not None:
            assert np.array_equal(a[4], b[4]), f"community vector differs at round {a[0]} it {a[1]}"
    assert len(trace) == len(t.rows), (len(trace), len(t.rows))
    assert res["rounds"] == len(ref["round_q"])
    assert list(res["round_nc"]) == list(ref["round_nc"])
    assert res["modularity"] == ref["Q"]
    assert np.array_equal(eng.communities(g.vertex_num), ref["final_comm"])
    return res, rounds


def test_karate_known_answer(golden):
    say("karate")
    g = gala.Graph.load_bin_graph(os.path.join(golden, "karate.txt.bin"))
    eng = gala.Engine(dump_iterations=True, dump_rounds=True)
    res, rounds = full_trace_compare(eng, g)
    assert f"{res['modularity']:.6f}" == "0.408695" and res["n_communities"] == 4
    assert [r["C"] for r in rounds] == [11, 4, 4]
    eng.close()


def test_sbm1024_per_iteration_and_contraction(golden):
    say("sbm1024")
    g = gala.Graph.load_bin_graph(os.path.join(golden, "sbm1024_s0.bin"))
    eng = gala.Engine(dump_iterations=True, dump_rounds=True)
    res, rounds = full_trace_compare(eng, g)
    # every contracted graph equals the oracle's contraction of the same partition
    cur = g
    w = np.ones(len(g.neighbors), np.int64)
    for r in rounds:
        nid, C, end, nb, ww = oracle.contract(cur.vertex_num, cur.degrees, cur.neighbors, w, r["comm"])
        assert C == r["C"]
        assert np.array_equal(end, r["end"]) and np.array_equal(nb, r["nbr"])
        assert np.array_equal(ww, r["w"])
        cur = gala.Graph(end, nb, ww.astype(np.int32), gala.W_I32)
        w = ww
    eng.close()


@pytest.mark.parametrize("seed", [0, 1])
def test_chung_lu_hub_bins(seed):
    # 200k vertices with hubs up to ~12k: every degree bin incl. the global-hash hubs
    say(f"chung_lu 200k seed {seed}")
    g = gala.gen_chung_lu(200000, 2.1, 17.35, 12000.0, seed)
    deg = np.diff(g.offsets())
    assert deg.max() >= 8192
    eng = gala.Engine(dump_iterations=True, dump_rounds=False)
    full_trace_compare(eng, g)
    eng.close()


def test_serial_bins_match_oracle():
    # every degree bin on one stream (the long-row bins otherwise run on a
    # second stream beside the others): the same trajectory
    say("serial bins")
    g = gala.gen_chung_lu(200000, 2.1, 17.35, 12000.0, 1)
    eng = gala.Engine(dump_iterations=True, serial_bins=True)
    full_trace_compare(eng, g)
    eng.close()


def test_queued_rounds_match_oracle():
    # sweeps queued ahead of the host's readback, the stopping rule applied on
    # the device: every iteration's Q and moved count, the round counts and
    # the final partition equal the oracle's; every round ran queued and each
    # round that stopped on its Q left one sweep with an emptied frontier
    say("queued rounds")
    g = gala.gen_chung_lu(200000, 2.1, 17.35, 12000.0, 0)
    eng = gala.Engine(queue_sweeps=True)
    res, _ = full_trace_compare(eng, g, keep_vectors=False)
    eng.close()
    assert res["queued_rounds"] == res["rounds"]
    assert res["inert_sweeps"] == res["rounds"]


def same_run(a, b):
    for k in ("modularity", "rounds", "round_nc", "round_iters", "round_q", "total_iterations",
              "swept_vertices", "swept_edges", "decide_bytes", "bin_swept_edges", "round_contracted_entries"):
        assert a[k] == b[k], k


@pytest.mark.parametrize("case", ["karate", "sbm1024", "star", "rmat_fp32", "cap2"])
def test_queued_rounds_equal_synchronous(golden, case):
    # the same runs with the host reading every sweep before queueing the next
    say(f"queued vs synchronous: {case}")
    kw = {}
    if case == "karate":
        g = gala.Graph.load_bin_graph(os.path.join(golden, "karate.txt.bin"))
    elif case == "sbm1024":
        g = gala.Graph.load_bin_graph(os.path.join(golden, "sbm1024_s0.bin"))
    elif case == "star":
        g = star_plus_random(30000, 40000, 3)
    elif case == "rmat_fp32":  # the wide path: 128-bit sums of Tot^2
        g = gala.gen_rmat(16, seed=5)
    else:  # at most 2 sweeps a round: rounds cut by the cap queue no sweep after it
        g = gala.gen_chung_lu(50000, 2.1, 17.35, 3000.0, 2)
        kw = dict(max_iterations=2)
    out = []
    for queue in (True, False):
        trace = []
        eng = gala.Engine(queue_sweeps=queue, **kw)
        eng.set_graph(g)
        r = eng.run(trace=trace)
        out.append((r, trace, eng.communities(g.vertex_num)))
        eng.close()
    (rq, tq, cq), (rs, ts, cs) = out
    same_run(rq, rs)
    assert [t[:4] for t in tq] == [t[:4] for t in ts]
    assert np.array_equal(cq, cs)
    assert rq["queued_rounds"] == rq["rounds"] and rs["queued_rounds"] == 0
    assert rs["inert_sweeps"] == 0
    capped = sum(1 for it in rq["round_iters"] if kw and it >= kw["max_iterations"])
    assert rq["inert_sweeps"] == rq["rounds"] - capped
    if case == "cap2":
        assert capped > 0 and max(rq["round_iters"]) == 2


def test_identity_round_contraction_skipped():
    # a last round whose first sweep is refused keeps the singleton partition;
    # its contraction (the identity) is skipped unless the round callback
    # asks for the contracted graph: both runs report the same rounds,
    # community counts, entries, Q and partition
    say("identity round")
    g = gala.gen_chung_lu(200000, 2.1, 17.35, 12000.0, 3)
    out = []
    for with_rounds in (False, True):
        eng = gala.Engine()
        eng.set_graph(g)
        r = eng.run(rounds=[] if with_rounds else None)
        out.append((r, eng.communities(g.vertex_num)))
        eng.close()
    (a, ca), (b, cb) = out
    assert a["round_iters"][-1] == 1  # the workload ends on such a round
    for k in ("modularity", "rounds", "round_nc", "round_q", "round_iters", "round_entries",
              "round_contracted_entries", "round_vertices", "n_communities"):
        assert a[k] == b[k], k
    assert np.array_equal(ca, cb)
    assert a["round_build_ms"][-1] < b["round_build_ms"][-1]


def variant_set(k):
    return np.array([min(k, c - 1) for c in gala.variant_counts()], np.int32)


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 6, 7, 8])
def test_decide_variants_match_oracle(k):
    # every decide kernel variant takes the same decisions as the oracle
    say(f"decide variants {k}")
    g = gala.gen_chung_lu(200000, 2.1, 17.35, 12000.0, 0)
    eng = gala.Engine(dump_iterations=True)
    eng.set_variants(variant_set(k))
    assert np.array_equal(eng.get_variants(), variant_set(k))
    full_trace_compare(eng, g)
    eng.close()


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 6, 7, 8])
def test_decide_variants_random_states(k):
    say(f"decide variants {k} on random states")
    rng = np.random.default_rng(40 + k)
    g = gala.gen_chung_lu(20000, 2.1, 17.35, 3000.0, 9)
    eng = gala.Engine()
    eng.set_variants(variant_set(k))
    for ncomm in (3, 50, 5000):
        K, Self, Tot, size, prev = random_state(rng, g, ncomm)
        m2 = int(K.sum())
        got = eng.phase_decide(g, m2, K, Self, Tot, size, prev)
        exp = oracle.decide_step(g.vertex_num, g.degrees, g.neighbors, np.ones(len(g.neighbors), np.int64), m2,
                                 K, Self, Tot, size, prev)
        assert np.array_equal(got, exp), ncomm
    eng.close()


def test_autotune_keeps_results():
    say("autotune")
    g = gala.gen_chung_lu(100000, 2.1, 17.35, 8000.0, 6)
    eng = gala.Engine()
    eng.set_graph(g)
    r0 = eng.run()
    c0 = eng.communities(g.vertex_num)
    v = eng.autotune()
    counts = gala.variant_counts()
    assert all(0 <= v[i] < counts[i] for i in range(len(v)))
    assert np.array_equal(eng.get_variants(), v)
    lo = eng.get_bin_cuts()
    d = gala.BIN_LO_DEFAULT
    assert all(lo[b] == d[b] for b in range(4)) and all(d[b - 1] < lo[b] <= d[b] for b in range(4, 9)), lo
    r1 = eng.run()
    assert r1["modularity"] == r0["modularity"] and np.array_equal(eng.communities(g.vertex_num), c0)
    eng.close()


def lowest_cuts():
    d = gala.BIN_LO_DEFAULT
    return [d[b] if b < 4 else d[b - 1] + 1 for b in range(9)]


def test_bin_cuts_validation(engine):
    say("bin cut validation")
    d = list(gala.BIN_LO_DEFAULT)
    assert list(engine.get_bin_cuts()) == d
    for bad in ([0] + d[1:], d[:3] + [15] + d[4:], d[:4] + [16] + d[5:], d[:8] + [8193],
                d[:5] + [32] + d[6:], d[:8] + [4096]):
        with pytest.raises(gala.GalaError):
            engine.set_bin_cuts(bad)
    engine.set_bin_cuts(lowest_cuts())
    assert list(engine.get_bin_cuts()) == lowest_cuts()
    engine.set_bin_cuts(d)


@pytest.mark.parametrize("cuts", ["lowest", "mixed"])
def test_bin_cuts_match_oracle(cuts):
    # every kernel gets shorter rows than its default range when the cuts are
    # lowered: the trajectory must not change
    say(f"bin cuts {cuts}")
    g = gala.gen_chung_lu(200000, 2.1, 17.35, 12000.0, 0)
    eng = gala.Engine(dump_iterations=True)
    lo = lowest_cuts() if cuts == "lowest" else [1, 4, 8, 16, 20, 100, 300, 3000, 6000]
    eng.set_bin_cuts(lo)
    full_trace_compare(eng, g)
    eng.close()


def test_bin_cuts_variants_split_hub():
    # the lowered hub cut sends rows of 4097+ entries through every bin-8 path
    say("bin cuts with split hub variants")
    g = star_plus_random(30000, 40000, 3)
    for variant in (0, 2, 4, 6):
        eng = gala.Engine(dump_iterations=True)
        eng.set_bin_cuts(lowest_cuts())
        v = eng.get_variants()
        v[8] = variant
        eng.set_variants(v)
        full_trace_compare(eng, g)
        eng.close()


def test_weighted_int_graph():
    say("weighted int32")
    rng = np.random.default_rng(3)
    g = gala.gen_chung_lu(30000, 2.3, 12.0, 3000.0, 5)
    off = g.offsets()
    w = np.zeros(len(g.neighbors), np.int32)
    # symmetric integer weights 1..7
    for v in range(g.vertex_num):
        for j in range(off[v], off[v + 1]):
            u = g.neighbors[j]
            if u > v:
                x = int(rng.integers(1, 8))
                w[j] = x
                k = off[u] + np.searchsorted(g.neighbors[off[u]:off[u + 1]], v)
                w[k] = x
    gw = gala.Graph(g.degrees, g.neighbors, w, gala.W_I32)
    eng = gala.Engine(dump_iterations=True)
    full_trace_compare(eng, gw)
    eng.close()


def random_state(rng, g, ncomm):
    n = g.vertex_num
    off = g.offsets()
    deg = np.diff(off)
    src = np.repeat(np.arange(n), deg)
    K = deg.astype(np.int64) if g.weights is None else np.bincount(src, weights=g.weights, minlength=n).astype(np.int64)
    Self = np.zeros(n, np.int64)
    prev = rng.integers(0, ncomm, n).astype(np.int32)
    Tot = np.bincount(prev, weights=K, minlength=n).astype(np.int64)
    size = np.bincount(prev, minlength=n).astype(np.int32)
    return K, Self, Tot, size, prev


@pytest.mark.parametrize("ncomm", [3, 50, 5000])
def test_phase_decide_random_states(engine, ncomm):
    # random assignments create many equal-weight candidates: exercises the
    # lowest-id tie-break and the singleton guard in every bin
    say(f"phase decide ncomm={ncomm}")
    rng = np.random.default_rng(ncomm)
    g = gala.gen_chung_lu(60000, 2.1, 17.35, 12000.0, 11)
    K, Self, Tot, size, prev = random_state(rng, g, ncomm)
    m2 = int(K.sum())
    w = np.ones(len(g.neighbors), np.int64)
    fr = (rng.random(g.vertex_num) < 0.7).astype(np.uint8)
    for frontier in (None, fr):
        got = engine.phase_decide(g, m2, K, Self, Tot, size, prev, frontier)
        exp = oracle.decide_step(g.vertex_num, g.degrees, g.neighbors, w, m2, K, Self, Tot, size, prev,
                                 frontier)
        assert np.array_equal(got, exp)


def test_phase_decide_singletons(engine):
    say("phase decide singletons")
    g = gala.gen_chung_lu(50000, 2.1, 17.35, 9000.0, 4)
    n = g.vertex_num
    K = np.diff(g.offsets()).astype(np.int64)
    prev = np.arange(n, dtype=np.int32)
    got = engine.phase_decide(g, int(K.sum()), K, np.zeros(n), K, np.ones(n), prev)
    exp = oracle.decide_step(n, g.degrees, g.neighbors, np.ones(len(g.neighbors), np.int64), int(K.sum()), K,
                             np.zeros(n, np.int64), K, np.ones(n, np.int32), prev)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("ncomm", [7, 900])
def test_phase_weight(engine, ncomm):
    say(f"phase weight ncomm={ncomm}")
    rng = np.random.default_rng(ncomm + 1)
    g = gala.gen_chung_lu(40000, 2.1, 17.35, 9000.0, 2)
    K, Self, Tot, size, cur = random_state(rng, g, ncomm)
    m2 = int(K.sum())
    nz = Tot[Tot > 0]
    min_tot = int(nz.min())
    In, act = engine.phase_weight(g, m2, K, Self, Tot, cur, min_tot)
    off = g.offsets()
    src = np.repeat(np.arange(g.vertex_num), np.diff(off))
    same = cur[g.neighbors] == cur[src]
    In_ref = np.bincount(src[same], minlength=g.vertex_num).astype(np.int64)
    assert np.array_equal(In, In_ref)
    c = cur.astype(np.int64)
    P = [m2 * (2 * int(In_ref[v]) - int(Self[v]) - int(K[v])) + int(K[v]) * (min_tot - (int(Tot[c[v]]) - int(K[v])))
         for v in range(g.vertex_num)]
    act_ref = np.array([0 if p > 0 else 1 for p in P], np.uint8)
    assert np.array_equal(act, act_ref)


@pytest.mark.parametrize("ncomm", [1, 2, 40, 3000, 30000])
def test_phase_contract_random(engine, ncomm):
    # ncomm small => very long rows (global-hash + bitmap path); large => LDS sort path
    say(f"phase contract ncomm={ncomm}")
    rng = np.random.default_rng(ncomm + 7)
    g = gala.gen_chung_lu(60000, 2.1, 17.35, 12000.0, 3)
    comm = rng.integers(0, ncomm, g.vertex_num).astype(np.int32)
    nid, C, end, nb, ww = engine.phase_contract(g, comm)
    rnid, rC, rend, rnb, rww = oracle.contract(g.vertex_num, g.degrees, g.neighbors,
                                               np.ones(len(g.neighbors), np.int64), comm)
    assert C == rC
    used = np.bincount(comm, minlength=g.vertex_num) > 0
    assert np.array_equal(nid[used], rnid[used]) and np.all(nid[~used] == -1)
    assert np.array_equal(end, rend) and np.array_equal(nb, rnb) and np.array_equal(ww, rww)


def test_determinism_twice(golden):
    say("determinism")
    g = gala.gen_chung_lu(100000, 2.1, 17.35, 10000.0, 9)
    eng = gala.Engine()
    eng.set_graph(g)
    r1 = eng.run()
    c1 = eng.communities(g.vertex_num)
    r2 = eng.run()  # same context, input buffers are read-only
    c2 = eng.communities(g.vertex_num)
    assert r1["modularity"] == r2["modularity"] and np.array_equal(c1, c2)
    eng.close()


def test_gapped_and_canonical_contraction_agree():
    # inside a run the contracted rows keep hash order and stay in their
    # staging segments (gapped CSR); with dump_rounds they are compacted and
    # sorted.  Every per-iteration vector, Q and the final partition agree.
    say("gapped vs canonical contraction")
    g = gala.gen_chung_lu(200000, 2.1, 17.35, 30000.0, 4)
    out = []
    for dump in (False, True):
        eng = gala.Engine(dump_iterations=True, dump_rounds=dump)
        eng.set_graph(g)
        trace = []
        r = eng.run(trace=trace)
        out.append((trace, r["modularity"], eng.communities(g.vertex_num)))
        eng.close()
    (ta, qa, ca), (tb, qb, cb) = out
    assert len(ta) == len(tb) and qa == qb and np.array_equal(ca, cb)
    for a, b in zip(ta, tb):
        assert a[:4] == b[:4] and np.array_equal(a[4], b[4]), a[:2]


def star_plus_random(nleaf, extra, seed):
    """Hub 0 joined to 1..nleaf, plus `extra` random edges among the leaves."""
    rng = np.random.default_rng(seed)
    n = nleaf + 1
    e = {(0, v) for v in range(1, n)}
    while len(e) < nleaf + extra:
        a, b = rng.integers(1, n, 2)
        if a != b:
            e.add((min(a, b), max(a, b)))
    rows = [[] for _ in range(n)]
    for a, b in e:
        rows[a].append(b)
        rows[b].append(a)
    end = np.cumsum([len(r) for r in rows]).astype(np.int64)
    nbr = np.concatenate([np.sort(np.array(r, np.int32)) for r in rows]).astype(np.int32)
    return gala.Graph(end, nbr)


def test_phase_decide_hub_overflow(engine):
    # a 50k-degree hub whose neighbours sit in ~40k distinct communities: more
    # than the 16384-slot LDS table, so part of its row goes to the global slice
    say("phase decide hub overflow")
    g = star_plus_random(50000, 60000, 1)
    rng = np.random.default_rng(5)
    K, Self, Tot, size, prev = random_state(rng, g, 40000)
    m2 = int(K.sum())
    w = np.ones(len(g.neighbors), np.int64)
    for rep in range(2):  # the second call checks the overflow slice was left zeroed
        got = engine.phase_decide(g, m2, K, Self, Tot, size, prev)
        exp = oracle.decide_step(g.vertex_num, g.degrees, g.neighbors, w, m2, K, Self, Tot, size, prev)
        assert np.array_equal(got, exp), rep


@pytest.mark.parametrize("variant", [5, 6])
def test_phase_decide_hub_full_part_table(variant):
    # a 16,383-degree hub whose neighbours all sit in distinct communities: with
    # one part per hub (variant 5) its 16,384-slot LDS table cannot take them
    # all, so the commit kernel decides the hub from its global slice; with two
    # parts (variant 6) the table holds them.  The second call checks that the
    # slice and the table were left zeroed.
    say(f"phase decide full part table, variant {variant}")
    g = star_plus_random(16383, 20000, 4)
    rng = np.random.default_rng(8)
    n = g.vertex_num
    off = g.offsets()
    K = np.diff(off).astype(np.int64)
    Self = np.zeros(n, np.int64)
    prev = rng.permutation(n).astype(np.int32)  # every vertex its own community
    Tot = np.zeros(n, np.int64)
    Tot[prev] = K
    size = np.zeros(n, np.int32)
    size[prev] = 1
    m2 = int(K.sum())
    w = np.ones(len(g.neighbors), np.int64)
    eng = gala.Engine()
    v = eng.get_variants()
    v[8] = variant
    eng.set_variants(v)
    for rep in range(2):
        got = eng.phase_decide(g, m2, K, Self, Tot, size, prev)
        exp = oracle.decide_step(g.vertex_num, g.degrees, g.neighbors, w, m2, K, Self, Tot, size, prev)
        assert np.array_equal(got, exp), rep
    eng.close()


@pytest.mark.parametrize("variant", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("ncomm", [3, 40000])
def test_phase_decide_split_hub(ncomm, variant):
    # bin 8 split over row chunks and table chunks (variant 1): a 50k-degree hub
    # whose neighbours sit in few or in ~40k distinct communities; the second
    # call checks the global slices were left zeroed
    say(f"phase decide split hub {ncomm} variant {variant}")
    g = star_plus_random(50000, 60000, 1)
    rng = np.random.default_rng(6)
    K, Self, Tot, size, prev = random_state(rng, g, ncomm)
    m2 = int(K.sum())
    w = np.ones(len(g.neighbors), np.int64)
    eng = gala.Engine()
    v = eng.get_variants()
    v[8] = variant
    eng.set_variants(v)
    for rep in range(2):
        got = eng.phase_decide(g, m2, K, Self, Tot, size, prev)
        exp = oracle.decide_step(g.vertex_num, g.degrees, g.neighbors, w, m2, K, Self, Tot, size, prev)
        assert np.array_equal(got, exp), rep
    eng.close()


@pytest.mark.parametrize("variant", [1, 2, 3, 4, 5, 6])
def test_star_graph_split_hub_full_run(variant):
    say(f"star full run, split hub variant {variant}")
    g = star_plus_random(30000, 40000, 3)
    eng = gala.Engine(dump_iterations=True)
    v = eng.get_variants()
    v[8] = variant
    eng.set_variants(v)
    full_trace_compare(eng, g)
    eng.close()


def test_phase_contract_huge_row(engine):
    # the hub's community row has 50k distinct neighbours: bitmap extraction path
    say("phase contract huge row")
    g = star_plus_random(50000, 20000, 2)
    comm = np.arange(g.vertex_num, dtype=np.int32)
    comm[1:50] = 0
    nid, C, end, nb, ww = engine.phase_contract(g, comm)
    rnid, rC, rend, rnb, rww = oracle.contract(g.vertex_num, g.degrees, g.neighbors,
                                               np.ones(len(g.neighbors), np.int64), comm)
    assert C == rC and np.array_equal(end, rend) and np.array_equal(nb, rnb) and np.array_equal(ww, rww)


def test_phase_contract_wide_keys(engine):
    # more than 2^21 communities: the dense rows' key-range digit is 10 bits wide
    say("phase contract wide keys")
    n = 2_200_000
    rng = np.random.default_rng(11)
    a = [np.arange(n)]
    b = [(np.arange(n) + 1) % n]
    for hub, k in ((0, 60000), (1_000_000, 30000)):
        nb = rng.choice(np.arange(hub + 2, hub + 900_000), k, replace=False)
        a.append(np.full(k, hub))
        b.append(nb)
    a, b = np.concatenate(a), np.concatenate(b)
    src, dst = np.concatenate([a, b]), np.concatenate([b, a])
    o = np.lexsort((dst, src))
    src, dst = src[o], dst[o]
    g = gala.Graph(np.cumsum(np.bincount(src, minlength=n)).astype(np.int64), dst.astype(np.int32))
    comm = np.arange(n, dtype=np.int32)
    comm[1:50] = 0
    comm[1_000_001:1_000_010] = 1_000_000
    nid, C, end, nb, ww = engine.phase_contract(g, comm)
    assert C > 1 << 21
    rnid, rC, rend, rnb, rww = oracle.contract(g.vertex_num, g.degrees, g.neighbors,
                                               np.ones(len(g.neighbors), np.int64), comm)
    assert C == rC and np.array_equal(end, rend) and np.array_equal(nb, rnb) and np.array_equal(ww, rww)


def test_single_engine_edge_cases():
    # degree-0 vertices (interleaved and trailing), one edge, a self-loop,
    # components that never touch, a path: every per-iteration vector, Q and
    # the final partition equal the oracle's
    say("single-graph edge cases")
    cases = [
        ("one edge", gala.Graph(np.array([1, 2], np.int64), np.array([1, 0], np.int32))),
        ("star + isolated", gala.Graph(np.concatenate([[300], 300 + np.arange(1, 301), [600, 600]]).astype(np.int64),
                                       np.concatenate([np.arange(1, 301), np.zeros(300)]).astype(np.int32))),
        ("triangle + self-loop", gala.Graph(np.array([3, 5, 7], np.int64),
                                            np.array([0, 1, 2, 0, 2, 0, 1], np.int32))),
    ]
    # isolated vertices between two 6-cliques, and a 40-vertex path
    rows = [[] for _ in range(15)]
    for base in (0, 8):
        for a in range(base, base + 6):
            rows[a] = [b for b in range(base, base + 6) if b != a]
    cases.append(("cliques + gaps", gala.Graph(np.cumsum([len(r) for r in rows]).astype(np.int64),
                                               np.concatenate([np.array(r, np.int32) for r in rows]))))
    prow = [[v - 1] * (v > 0) + [v + 1] * (v < 39) for v in range(40)]
    cases.append(("path", gala.Graph(np.cumsum([len(r) for r in prow]).astype(np.int64),
                                     np.concatenate([np.array(r, np.int32) for r in prow]))))
    for name, g in cases:
        say(f"  {name}")
        eng = gala.Engine(dump_iterations=True)
        full_trace_compare(eng, g)
        eng.close()
    # no edge at all: m2 = 0 is rejected by both
    g = gala.Graph(np.zeros(5, np.int64), np.zeros(0, np.int32))
    eng = gala.Engine()
    eng.set_graph(g)
    with pytest.raises(gala.GalaError):
        eng.run()
    eng.close()
    with pytest.raises(RuntimeError):
        oracle.louvain(g.vertex_num, g.degrees, g.neighbors)


def test_star_graph_full_run():
    say("star full run")
    g = star_plus_random(30000, 40000, 3)
    eng = gala.Engine(dump_iterations=True)
    full_trace_compare(eng, g)
    eng.close()


def test_full_size_chung_lu():
    # the 4.0M-vertex / 34.4M-edge power-law workload at full size: final
    # partition, Q and per-round community counts equal the oracle's, and Q is
    # the modularity of the returned partition (exact integer sums)
    say("full-size Chung-Lu 4.0M")
    g = gala.gen_chung_lu(4_000_000, 2.1, 17.35, 15000.0, 0)
    eng = gala.Engine()
    eng.set_graph(g)
    res = eng.run()
    comm = eng.communities(g.vertex_num)
    eng.close()
    oracle.set_threads(min(16, os.cpu_count() or 1))
    ref = oracle.louvain(g.vertex_num, g.degrees, g.neighbors)
    assert np.array_equal(comm, ref["final_comm"])
    assert res["modularity"] == ref["Q"]
    assert list(res["round_nc"]) == list(ref["round_nc"])
    off = g.offsets()
    deg = np.diff(off)
    src = np.repeat(np.arange(g.vertex_num), deg)
    m2 = int(len(g.neighbors))
    s_in = int(np.count_nonzero(comm[g.neighbors] == comm[src]))
    tot = np.bincount(comm, weights=deg, minlength=g.vertex_num).astype(np.int64)
    q = s_in / m2 - float(int((tot * tot).sum())) / (m2 * m2)
    assert q == res["modularity"]
