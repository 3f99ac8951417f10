"""gala_phase_modularity: Q of an arbitrary partition on the GPU from exact
integer sums, equal bit for bit to the oracle's modularity of the same
partition (unit, int32 and float32 weights) and to the Q gala_run reports
for the partition it returns (modularity_op, kernel_functions.cuh:77-87)."""
import os

import numpy as np
import pytest

import gala
import oracle


def say(msg):
    print(msg, flush=True)


def exact_w(g):
    if g.weights is None:
        return np.ones(len(g.neighbors), np.int64)
    if g.weight_mode == gala.W_F32:
        return oracle.scale_f32(g.weights)[0]
    return g.weights.astype(np.int64)


def graphs(golden):
    sbm = gala.Graph.load_bin_graph(os.path.join(golden, "sbm1024_s0.bin"))
    rng = np.random.default_rng(0)
    yield "sbm unit", sbm
    yield "sbm int32", gala.Graph(sbm.degrees, sbm.neighbors,
                                  rng.integers(1, 9, len(sbm.neighbors)).astype(np.int32), gala.W_I32)
    yield "rmat fp32", gala.gen_rmat(14, 16, seed=3)
    yield "chung-lu hubs", gala.gen_chung_lu(200000, 2.1, 17.35, 12000.0, 2)


@pytest.mark.gpu
def test_phase_modularity_matches_oracle(engine, golden):
    rng = np.random.default_rng(1)
    for name, g in graphs(golden):
        say(f"modularity {name}")
        n = g.vertex_num
        W = exact_w(g)
        for ncomm in (1, 7, max(1, n // 10), n):
            comm = rng.integers(0, ncomm, n).astype(np.int32)
            q = engine.phase_modularity(g, comm)
            assert q == oracle.modularity(n, g.degrees, g.neighbors, W, comm), (name, ncomm)


@pytest.mark.gpu
def test_phase_modularity_equals_run_q(golden):
    for name, g in graphs(golden):
        say(f"run Q vs phase Q {name}")
        eng = gala.Engine()
        eng.set_graph(g)
        r = eng.run()
        comm = eng.communities(g.vertex_num)
        assert eng.phase_modularity(g, comm) == r["modularity"], name
        eng.close()


@pytest.mark.gpu
def test_phase_modularity_edge_cases(engine):
    g = gala.Graph(np.zeros(4, np.int64), np.zeros(0, np.int32))
    assert engine.phase_modularity(g, np.arange(4, dtype=np.int32)) == 0.0  # no edge
    tri = gala.Graph(np.array([2, 4, 6], np.int64), np.array([1, 2, 0, 2, 0, 1], np.int32))
    with pytest.raises(gala.GalaError):  # community id out of range
        engine.phase_modularity(tri, np.array([0, 1, 3], np.int32))
    bad = gala.Graph(tri.degrees, tri.neighbors, np.array([1, 1, 1, 0, 1, 1], np.int32), gala.W_I32)
    with pytest.raises(gala.GalaError):  # non-positive weight
        engine.phase_modularity(bad, np.zeros(3, np.int32))
    # one community: Q = 0 exactly; singletons of a triangle: -1/3
    assert engine.phase_modularity(tri, np.zeros(3, np.int32)) == 0.0
    assert engine.phase_modularity(tri, np.arange(3, dtype=np.int32)) == 0.0 / 6 - 12.0 / 36
