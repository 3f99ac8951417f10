"""Learned frontier end to end on the GPU: move logs, training on the engine's
own MFMA kernels, learned runs whose stalled sweeps are certified, and what
the trained network gains over the method without it."""
import os

import numpy as np
import pytest

import gala
import train_frontier as TF


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


def run(g, mode, params=None, tau=0.5, trace=None, **settings):
    if "learned_tau" in settings:
        tau = settings.pop("learned_tau")
    eng = gala.Engine(frontier_mode=mode, learned_tau=tau, dump_iterations=trace is not None, **settings)
    if params is not None:
        eng.set_mlp(params)
    eng.set_graph(g)
    r = eng.run(trace=trace)
    c = eng.communities(g.vertex_num)
    eng.close()
    return r, c


def auc(score, y):
    """Area under the ROC curve (rank statistic, ties averaged)."""
    order = np.argsort(score, kind="mergesort")
    s = score[order]
    ranks = np.empty(len(s))
    i = 0
    while i < len(s):
        j = i
        while j + 1 < len(s) and s[j + 1] == s[i]:
            j += 1
        ranks[order[i:j + 1]] = 0.5 * (i + j) + 1
        i = j + 1
    pos = y > 0
    npos, nneg = pos.sum(), (~pos).sum()
    return (ranks[pos].sum() - npos * (npos + 1) / 2) / (npos * nneg)


def test_auc_helper():
    assert auc(np.array([0.1, 0.2, 0.3, 0.4]), np.array([0, 0, 1, 1])) == 1.0
    assert auc(np.array([0.4, 0.3, 0.2, 0.1]), np.array([0, 0, 1, 1])) == 0.0
    assert auc(np.array([0.5, 0.5, 0.5, 0.5]), np.array([0, 1, 0, 1])) == 0.5


@pytest.mark.gpu
def test_logger_features_in_range():
    say("logger")
    eng = gala.Engine()
    X, y = TF.collect_logs(eng, [gala.gen_chung_lu(20000, 2.1, 17.35, 2000.0, 7)])
    eng.close()
    assert X.shape[1] == 6 and len(X) == len(y) > 0
    assert np.all(X >= -1.0) and np.all(X <= 1.0)
    assert np.all((X[:, 3] == 0) | (X[:, 3] == 1))
    assert 0 < y.mean() < 1


@pytest.mark.gpu
def test_predict_nobody_equals_plain():
    # a network that predicts no mover leaves every learned sweep empty; each
    # such sweep stalls and its certification sweeps the whole frontier, which
    # replays the plain trajectory exactly
    say("predict nobody")
    g = gala.gen_chung_lu(100000, 2.1, 17.35, 8000.0, 3)
    nobody = np.zeros(gala.MLP_NPARAM, np.float32)
    nobody[-1] = -30.0
    tp, tl = [], []
    rp, cp = run(g, 0, trace=tp)
    rl, cl = run(g, 1, nobody, trace=tl)
    assert rl["modularity"] == rp["modularity"] and np.array_equal(cl, cp)
    assert [(a, b, q, m) for a, b, q, m, _ in tl] == [(a, b, q, m) for a, b, q, m, _ in tp]
    assert [np.array_equal(x[4], y[4]) for x, y in zip(tl, tp)] == [True] * len(tp)
    assert rl["learned_sweeps"] > 0 and rl["certify_sweeps"] == rl["learned_sweeps"]
    assert rl["learned_kept"] == 0 and rl["learned_frontier"] > 0


@pytest.mark.gpu
def test_predict_everybody_equals_plain():
    # keeping the whole frontier makes every learned sweep the plain sweep
    say("predict everybody")
    g = gala.gen_chung_lu(100000, 2.1, 17.35, 8000.0, 4)
    everybody = np.zeros(gala.MLP_NPARAM, np.float32)
    everybody[-1] = 30.0
    rp, cp = run(g, 0)
    rl, cl = run(g, 1, everybody)
    assert rl["modularity"] == rp["modularity"] and np.array_equal(cl, cp)
    assert rl["learned_kept"] == rl["learned_frontier"] > 0
    # the stalled last sweep of a level sweeps an empty complement
    assert rl["swept_vertices"] == rp["swept_vertices"]


@pytest.mark.gpu
def test_learned_close_to_plain(trained):
    p, tau = trained
    say(f"learned vs plain (tau={tau:.3f})")
    g = gala.gen_chung_lu(300000, 2.1, 17.35, 15000.0, 1001)
    rp, _ = run(g, 0)
    rl, cl = run(g, 1, p, tau)
    say(f"  Q plain {rp['modularity']:.6f} learned {rl['modularity']:.6f} "
        f"kept {rl['learned_kept']}/{rl['learned_frontier']}")
    assert abs(rl["modularity"] - rp["modularity"]) <= 0.002, (rl["modularity"], rp["modularity"])
    assert rl["learned_sweeps"] > 0
    # the network skips part of the plain frontier
    assert 0 < rl["learned_kept"] < rl["learned_frontier"]
    # the reported Q is the modularity of the returned partition
    import oracle
    q = oracle.modularity(g.vertex_num, g.degrees, g.neighbors, np.ones(len(g.neighbors), np.int64), cl)
    assert abs(q - rl["modularity"]) < 1e-12


@pytest.mark.gpu
def test_network_gain_on_held_out_graph(trained):
    # What the trained network gains over the method without it: the plain
    # frontier (sound prune only) has to sweep every active vertex; the network
    # ranks them so that a threshold keeps most movers while dropping many
    # non-movers.  Measured on move logs of a held-out power-law graph.
    p, tau = trained
    say("network gain")
    eng = gala.Engine()
    X, y = TF.collect_logs(eng, [gala.gen_chung_lu(100000, 2.1, 17.35, 8000.0, 1002)])
    pr = eng.mlp_forward(p, X)
    eng.close()
    a = auc(pr.astype(np.float64), y)
    kept = float((pr > tau).mean())
    recall = float((pr[y > 0] > tau).mean())
    say(f"  AUC {a:.3f}; at tau {tau:.3f}: kept {kept:.3f} of the frontier, recall {recall:.3f}")
    assert a > 0.75
    # a filter without information keeps movers in proportion to what it keeps
    assert recall > kept + 0.1


@pytest.mark.gpu
def test_learned_sbm_quality(trained):
    p, tau = trained
    say("learned sbm held out")
    g = gala.gen_sbm(1024, 16, 0.10, 0.002, 999)
    rp, _ = run(g, 0)
    rl, _ = run(g, 1, p, tau)
    assert abs(rl["modularity"] - rp["modularity"]) <= 0.002


@pytest.mark.gpu
def test_degree_proportional_logs_and_filter_settings():
    # the power-law training logs are sampled in proportion to the degree (the
    # work a kept vertex costs): their mean degree exceeds the full logs'; a
    # learned run that filters only long rows of large rounds keeps every
    # short-row vertex of the frontier and still ends within 0.002 of plain
    say("degree-proportional logs")
    g = gala.gen_chung_lu(50000, 2.1, 17.35, 3000.0, 11)
    eng = gala.Engine()
    Xf, yf = TF.collect_logs(eng, [g])
    Xs, ys = TF.collect_logs(eng, [g], per_sweep=2000, seed=1)
    eng.close()
    assert 0 < len(ys) < len(yf)
    assert TF.degree_of(Xs).mean() > 2 * TF.degree_of(Xf).mean()
    nobody = np.zeros(gala.MLP_NPARAM, np.float32)
    nobody[-1] = -30.0  # predicts no mover: only the rows kept by the settings are swept
    rp, cp = run(g, 0)
    rl, cl = run(g, 1, nobody, learned_min_degree=16, learned_min_vertices=10000)
    assert rl["learned_sweeps"] > 0 and 0 < rl["learned_kept"] < rl["learned_frontier"]
    assert abs(rl["modularity"] - rp["modularity"]) <= 0.002
    # rounds below the vertex bound run plain sweeps only
    rb, cb = run(g, 1, nobody, learned_min_vertices=10 ** 9)
    assert rb["learned_sweeps"] == 0 and rb["modularity"] == rp["modularity"] and np.array_equal(cb, cp)


def q_of_partition(g, comm):
    """Modularity of `comm` on an unweighted graph from exact integer sums,
    finished in the order the engine uses: sumIn/m2 - sumTot2/m2^2."""
    off = g.offsets()
    deg = np.diff(off)
    src = np.repeat(np.arange(g.vertex_num), deg)
    m2 = int(len(g.neighbors))
    s_in = int(np.count_nonzero(comm[g.neighbors] == comm[src]))
    tot = np.bincount(comm, weights=deg, minlength=g.vertex_num).astype(np.int64)
    return s_in / m2 - float(int((tot * tot).sum())) / (m2 * m2)


@pytest.mark.gpu
def test_committed_model_on_held_out_4M():
    # the committed network (trained on move logs of other seeds of the
    # power-law family) with its stored filter settings, on a held-out seed of
    # the 4.0M-vertex workload: final Q within 0.002 of the plain run and equal
    # to the modularity of the returned partition
    if not os.path.exists(TF.DEFAULT_MODEL):
        pytest.skip("no committed model")
    params, _ = TF.load(TF.DEFAULT_MODEL)
    settings = TF.load_settings(TF.DEFAULT_MODEL)
    say("committed model, Chung-Lu 4.0M seed 1001")
    g = gala.gen_chung_lu(4_000_000, 2.1, 17.35, 15000.0, 1001)
    rp, cp = run(g, 0)
    rl, cl = run(g, 1, params, **settings)
    say(f"  Q plain {rp['modularity']:.6f} learned {rl['modularity']:.6f} "
        f"kept {rl['learned_kept']}/{rl['learned_frontier']} certified {rl['certify_sweeps']}")
    assert rl["learned_sweeps"] > 0
    assert abs(rl["modularity"] - rp["modularity"]) <= 0.002
    assert q_of_partition(g, cl) == rl["modularity"]
    assert q_of_partition(g, cp) == rp["modularity"]
