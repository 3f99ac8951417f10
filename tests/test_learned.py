"""Learned frontier end to end on the GPU: move logs, training, learned runs
certified by the plain sweep."""
import numpy as np
import pytest

import gala
import train_frontier as TF

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def trained():
    say("collect + train")
    eng = gala.Engine()
    X, y = TF.collect_logs(eng, TF.sbm_training_graphs(range(1, 9)))
    assert len(y) > 10000 and 0 < y.sum() < len(y)
    p, loss, tau = TF.train(eng, X, y, epochs=15)
    assert loss[-1] < loss[0]
    eng.close()
    return p, tau


def run(g, mode, params=None, tau=0.5, trace=None):
    eng = gala.Engine(frontier_mode=mode, learned_tau=tau, dump_iterations=trace is not None)
    if params is not None:
        eng.set_mlp(params)
    eng.set_graph(g)
    r = eng.run(trace=trace)
    c = eng.communities(g.vertex_num)
    eng.close()
    return r, c


def test_logger_features_in_range():
    say("logger")
    eng = gala.Engine()
    X, y = TF.collect_logs(eng, [gala.gen_chung_lu(20000, 2.1, 17.35, 2000.0, 7)])
    eng.close()
    assert X.shape[1] == 6 and len(X) == len(y) > 0
    assert np.all(X >= -1.0) and np.all(X <= 1.0)
    assert np.all((X[:, 3] == 0) | (X[:, 3] == 1))
    assert 0 < y.mean() < 1


def test_predict_nobody_equals_plain():
    # a model that predicts no mover leaves every learned sweep empty; the
    # certification sweeps then replay the plain trajectory exactly
    say("predict nobody")
    g = gala.gen_chung_lu(100000, 2.1, 17.35, 8000.0, 3)
    nobody = np.zeros(gala.MLP_NPARAM, np.float32)
    nobody[-1] = -30.0
    rp, cp = run(g, 0)
    rl, cl = run(g, 1, nobody)
    assert rl["modularity"] == rp["modularity"] and np.array_equal(cl, cp)
    assert rl["learned_sweeps"] > 0 and rl["certify_sweeps"] > 0


def test_learned_close_to_plain_and_cheaper(trained):
    p, tau = trained
    say(f"learned vs plain (tau={tau:.3f})")
    g = gala.gen_chung_lu(300000, 2.1, 17.35, 15000.0, 1001)
    rp, _ = run(g, 0)
    rl, cl = run(g, 1, p, tau)
    assert abs(rl["modularity"] - rp["modularity"]) <= 0.002, (rl["modularity"], rp["modularity"])
    assert rl["learned_sweeps"] > 0
    # what the network gains: fewer vertices evaluated by decide-and-move
    assert rl["swept_vertices"] < rp["swept_vertices"], (rl["swept_vertices"], rp["swept_vertices"])
    # the reported Q is the modularity of the returned partition
    import oracle
    q = oracle.modularity(g.vertex_num, g.degrees, g.neighbors, np.ones(len(g.neighbors), np.int64), cl)
    assert abs(q - rl["modularity"]) < 1e-12


def test_learned_sbm_quality(trained):
    p, tau = trained
    say("learned sbm held out")
    g = gala.gen_sbm(1024, 16, 0.10, 0.002, 999)
    rp, _ = run(g, 0)
    rl, _ = run(g, 1, p, tau)
    assert abs(rl["modularity"] - rp["modularity"]) <= 0.002
