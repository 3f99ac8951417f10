"""The exact wide path on the GPU (int64 sums, int128 gains).

Float32-weighted graphs run as the exact integers w * 2^k, integer graphs with
m2 > 1e9 as themselves; both must be bit-exact to the exact oracle run on the
same integers: every iteration's community vector, moved count and Q (the
same three FP64 operations on identical exact sums), every contracted CSR
(integer weights), the composed partition.  Exact sums make the path
deterministic: two runs are identical bit for bit.  The north star's FP64
tolerance is checked as well: final Q within 1e-9 of the FP64 CPU oracle.
"""
import os

import numpy as np
import pytest

import gala
import oracle

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def exact_weights(g):
    """The oracle's integer image of the graph's weights (None: unit)."""
    if g.weights is None:
        return None, 0
    if g.weight_mode == gala.W_F32:
        return oracle.scale_f32(g.weights)
    return g.weights.astype(np.int64), 0


def trace_compare(eng, g, keep_vectors=True, rounds=None):
    trace = []
    eng.set_graph(g)
    res = eng.run(trace=trace, rounds=rounds)
    W, k = exact_weights(g)
    t = oracle.Trace(keep_vectors=keep_vectors)
    ref = oracle.louvain(g.vertex_num, g.degrees, g.neighbors, W, trace=t)
    assert len(trace) == len(t.rows), (len(trace), len(t.rows))
    for a, b in zip(trace, t.rows):
        assert a[:2] == b[:2]
        assert a[2] == b[2], f"Q differs at round {a[0]} it {a[1]}: {a[2]!r} vs {b[2]!r}"
        assert a[3] == b[3], f"moved differs at round {a[0]} it {a[1]}"
        if keep_vectors and a[4] is not None:
            assert np.array_equal(a[4], b[4]), f"community vector differs at round {a[0]} it {a[1]}"
    assert list(res["round_nc"]) == list(ref["round_nc"])
    assert res["modularity"] == ref["Q"]
    assert np.array_equal(eng.communities(g.vertex_num), ref["final_comm"])
    return res, ref, W, k


def weighted(g, seed, lo=0.5, hi=2.0):
    """The same graph with symmetric float32 weights U[lo, hi]."""
    rng = np.random.default_rng(seed)
    off = g.offsets()
    src = np.repeat(np.arange(g.vertex_num), np.diff(off))
    a = np.minimum(src, g.neighbors).astype(np.int64)
    b = np.maximum(src, g.neighbors).astype(np.int64)
    key = a * g.vertex_num + b
    uk, inv = np.unique(key, return_inverse=True)
    w = rng.uniform(lo, hi, len(uk)).astype(np.float32)[inv]
    return gala.Graph(g.degrees, g.neighbors, w, gala.W_F32)


def q_float_partition(g, comm):
    off = g.offsets()
    src = np.repeat(np.arange(g.vertex_num), np.diff(off))
    w = g.weights.astype(np.float64)
    m2 = w.sum()
    K = np.bincount(src, weights=w, minlength=g.vertex_num)
    tot = np.bincount(comm, weights=K, minlength=g.vertex_num)
    return w[comm[g.neighbors] == comm[src]].sum() / m2 - (tot ** 2).sum() / m2 ** 2


def test_sbm_fp32_per_iteration_and_contraction(golden):
    say("sbm1024 fp32 weights: trace and contracted graphs")
    g = weighted(gala.Graph.load_bin_graph(os.path.join(golden, "sbm1024_s0.bin")), 1)
    eng = gala.Engine(dump_iterations=True, dump_rounds=True)
    rounds = []
    res, ref, W, k = trace_compare(eng, g, rounds=rounds)
    eng.close()
    assert res["wide"] == 1 and res["weight_shift"] == k > 0
    cur = (g.vertex_num, g.degrees, g.neighbors, W)
    for r in rounds:
        nid, C, end, nb, ww = oracle.contract(*cur, r["comm"])
        assert C == r["C"]
        assert np.array_equal(end, r["end"]) and np.array_equal(nb, r["nbr"])
        assert np.array_equal(ww, r["w"])  # exact integer sums
        assert np.array_equal(r["wf"], ww / 2.0 ** k)
        cur = (C, end, nb, ww)


@pytest.mark.parametrize("scale", [12, 15, 18])
def test_weighted_fp32_rmat(scale):
    say(f"weighted fp32 rmat scale {scale}")
    g = gala.gen_rmat(scale, 16, seed=scale)
    eng = gala.Engine(dump_iterations=scale <= 15)
    res, ref, W, k = trace_compare(eng, g, keep_vectors=scale <= 15)
    comm = eng.communities(g.vertex_num)
    eng.close()
    assert res["wide"] == 1
    # the FP64 CPU oracle (double sums in row order): Q within 1e-9
    fp = oracle.louvain_f(g.vertex_num, g.degrees, g.neighbors, g.weights)
    assert abs(res["modularity"] - fp["Q"]) <= 1e-9
    # the reported Q is the modularity of the returned partition (FP64 sums)
    assert abs(q_float_partition(g, comm) - res["modularity"]) <= 1e-9


def test_wide_path_is_deterministic():
    say("fp32 rmat scale 18 twice")
    g = gala.gen_rmat(18, 16, seed=5)
    out = []
    for _ in range(2):
        eng = gala.Engine()
        eng.set_graph(g)
        r = eng.run()
        out.append((r["modularity"], list(r["round_nc"]), r["total_iterations"], eng.communities(g.vertex_num)))
        eng.close()
    a, b = out
    assert a[0] == b[0] and a[1] == b[1] and a[2] == b[2] and np.array_equal(a[3], b[3])


def test_int32_weights_beyond_1e9_take_the_wide_path(golden):
    # m2 = sum of weights ~ 8e9: past the int32 path's bound, run exactly on
    # int64 sums (the int32 path rejected this graph with GALA_E_OVERFLOW)
    say("int32 weights, m2 > 1e9")
    g0 = gala.Graph.load_bin_graph(os.path.join(golden, "sbm1024_s0.bin"))
    rng = np.random.default_rng(2)
    off = g0.offsets()
    src = np.repeat(np.arange(g0.vertex_num), np.diff(off))
    key = np.minimum(src, g0.neighbors).astype(np.int64) * 4096 + np.maximum(src, g0.neighbors)
    uk, inv = np.unique(key, return_inverse=True)
    w = rng.integers(500_000, 1_500_000, len(uk)).astype(np.int32)[inv]
    g = gala.Graph(g0.degrees, g0.neighbors, w, gala.W_I32)
    assert int(w.astype(np.int64).sum()) > 10 ** 9
    eng = gala.Engine(dump_iterations=True)
    res, ref, W, k = trace_compare(eng, g)
    eng.close()
    assert res["wide"] == 1 and k == 0


def test_integer_valued_floats_take_the_int32_path():
    say("integer-valued fp32 weights")
    g0 = gala.gen_chung_lu(20000, 2.1, 17.35, 2000.0, 3)
    rng = np.random.default_rng(4)
    off = g0.offsets()
    src = np.repeat(np.arange(g0.vertex_num), np.diff(off))
    key = np.minimum(src, g0.neighbors).astype(np.int64) * g0.vertex_num + np.maximum(src, g0.neighbors)
    uk, inv = np.unique(key, return_inverse=True)
    w = rng.integers(1, 6, len(uk)).astype(np.float32)[inv]
    g = gala.Graph(g0.degrees, g0.neighbors, w, gala.W_F32)
    eng = gala.Engine(dump_iterations=True)
    res, ref, W, k = trace_compare(eng, g)
    eng.close()
    assert res["wide"] == 0 and k == 0


def star_fp32(nleaf, extra, seed):
    from test_gpu_parity import star_plus_random
    return weighted(star_plus_random(nleaf, extra, seed), seed + 10)


@pytest.mark.parametrize("variant", [1, 4, 5, 6])
def test_fp32_star_split_hub_full_run(variant):
    # a 30k-degree hub with float32 weights: bin 8 on the wide path (split by
    # community with P >= 3 parts and the resolved rows, the global-table split)
    say(f"fp32 star full run, hub variant {variant}")
    g = star_fp32(30000, 40000, 3)
    eng = gala.Engine(dump_iterations=True)
    v = eng.get_variants()
    v[8] = variant
    eng.set_variants(v)
    trace_compare(eng, g)
    eng.close()


def test_fp32_chung_lu_all_bins():
    # every degree bin of the wide path, including 7 (global arena there)
    say("fp32 chung-lu 200k")
    g = weighted(gala.gen_chung_lu(200000, 2.1, 17.35, 12000.0, 0), 7)
    assert np.diff(g.offsets()).max() >= 8192
    eng = gala.Engine(dump_iterations=True)
    trace_compare(eng, g)
    eng.close()


def test_weighted_rmat24_full_size():
    # The weighted workload at full size: R-MAT scale 24 (16.8M vertices, 260M
    # edges, hubs past 1e5), float32 weights U[0.5, 2].  Bit-exact to the
    # exact oracle (partition, Q, per-round counts) and within 1e-9 of the FP64
    # CPU oracle's Q.
    say("rmat-24 fp32: generate")
    g = gala.gen_rmat(24, 16, seed=0)
    deg = np.diff(g.offsets())
    assert deg.max() > 100000
    say("  gpu run")
    eng = gala.Engine()
    eng.set_graph(g)
    res = eng.run()
    comm = eng.communities(g.vertex_num)
    eng.close()
    say(f"  {res['t_total_ms']:.1f} ms; exact oracle (16 threads)")
    W, k = oracle.scale_f32(g.weights)
    oracle.set_threads(min(16, os.cpu_count() or 1))
    ref = oracle.louvain(g.vertex_num, g.degrees, g.neighbors, W)
    assert np.array_equal(comm, ref["final_comm"])
    assert res["modularity"] == ref["Q"]
    assert list(res["round_nc"]) == list(ref["round_nc"])
    say("  FP64 oracle")
    fp = oracle.louvain_f(g.vertex_num, g.degrees, g.neighbors, g.weights)
    assert abs(res["modularity"] - fp["Q"]) <= 1e-9
