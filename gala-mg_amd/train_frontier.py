"""Trains the learned-frontier MLP on move logs of plain runs.

Move logs: for every plain sweep after the first of each round, the six
features of every swept vertex and whether the sweep moved it.  Training data,
seeds 1..32 of either family:
  chung_lu  power-law graphs of the 4.0M-vertex workload's shape (gamma 2.1,
            mean degree 17.35, degree cap 15,000) at --n vertices; each sweep's
            log is sampled with probability proportional to the vertex degree,
            so the loss weighs a vertex by the row it would sweep (the work
            the filter can save), and the threshold is set for that
            edge-weighted mover recall
  sbm       the 1,024-vertex planted-partition SBMs (16 blocks, p_in 0.10,
            p_out 0.002), every sample kept
The held-out evaluation graphs are seeds 1001-1003 of the 4.0M workload.
Training runs on the GPU (the engine's own BF16/MFMA forward and backward
kernels, Adam).

    python gala-mg_amd/train_frontier.py [--graphs chung_lu] [--seeds 32] [--out gala-mg_amd/models/frontier_mlp.json]
"""
import argparse
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import gala  # noqa: E402

DEFAULT_MODEL = os.path.join(HERE, "models", "frontier_mlp.json")


def degree_of(X):
    """Row length from feature 0 (log2(1 + degree) / 16)."""
    return np.round(np.exp2(X[:, 0].astype(np.float64) * 16.0) - 1.0)


def collect_logs(engine, graphs, per_sweep=0, seed=0):
    """Plain runs of `graphs` with the move logger on -> (X[n, 6] float32, y[n] uint8).
    per_sweep > 0: each sweep's log is sampled with probability proportional
    to the degree, about per_sweep samples per sweep."""
    xs, ys = [], []
    rng = np.random.default_rng(seed)

    def sink(rnd, it, f, m):
        if per_sweep > 0 and len(m):
            d = degree_of(f)
            keep = rng.random(len(m)) < np.minimum(1.0, d * (per_sweep / max(d.sum(), 1.0)))
            f, m = f[keep], m[keep]
        xs.append(f)
        ys.append(m)

    engine.set_logger(sink)
    try:
        for g in graphs:
            engine.set_graph(g)
            engine.run()
    finally:
        engine.set_logger(None)
    if not xs:
        return np.zeros((0, gala.MLP_NFEAT), np.float32), np.zeros(0, np.uint8)
    return np.concatenate(xs).astype(np.float32), np.concatenate(ys).astype(np.uint8)


def sbm_training_graphs(seeds):
    return [gala.gen_sbm(1024, 16, 0.10, 0.002, s) for s in seeds]


def chung_lu_training_graphs(seeds, n):
    for s in seeds:  # one at a time: each is generated, run and dropped
        yield gala.gen_chung_lu(n, 2.1, 17.35, 15000.0, s)


def train(engine, X, y, epochs=40, batch=4096, lr=5e-3, seed=0, recall=0.98):
    """Class-balanced training; returns (params, per-epoch loss, threshold).
    The threshold keeps the given fraction of the movers of the training logs."""
    pos = max(int(y.sum()), 1)
    pos_w = float(len(y) - pos) / pos
    p, loss = engine.mlp_train(X, y, epochs=epochs, batch=batch, lr=lr, pos_weight=pos_w, seed=seed)
    pr = engine.mlp_forward(p, X)
    movers = np.sort(pr[y > 0])
    tau = float(movers[int((1.0 - recall) * len(movers))]) if len(movers) else 0.5
    return p, loss, tau


def save(path, params, tau, meta):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump({"params": [float(x) for x in params], "tau": tau, **meta}, f)


def load(path=DEFAULT_MODEL):
    with open(path) as f:
        d = json.load(f)
    return np.array(d["params"], np.float32), float(d["tau"])


def load_settings(path=DEFAULT_MODEL):
    """The filter settings stored with a model: Engine keyword arguments."""
    with open(path) as f:
        d = json.load(f)
    return {"learned_tau": float(d["tau"]), "learned_min_degree": int(d.get("min_degree", 0)),
            "learned_min_vertices": int(d.get("min_vertices", 0))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--graphs", choices=["chung_lu", "sbm"], default="chung_lu")
    ap.add_argument("--n", type=int, default=1_000_000, help="vertices of the chung_lu training graphs")
    ap.add_argument("--per-sweep", type=int, default=20000, help="chung_lu: samples drawn per logged sweep")
    ap.add_argument("--seeds", type=int, default=32)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--recall", type=float, default=None,
                    help="mover recall the threshold keeps on the training samples (edge-weighted for "
                         "chung_lu); default 0.95 for chung_lu, 0.98 for sbm")
    ap.add_argument("--min-degree", type=int, default=32,
                    help="stored filter setting: rows shorter than this are always swept")
    ap.add_argument("--min-vertices", type=int, default=1_000_000,
                    help="stored filter setting: rounds of smaller graphs run plain sweeps")
    ap.add_argument("--out", default=DEFAULT_MODEL)
    args = ap.parse_args()
    eng = gala.Engine()
    seeds = list(range(1, args.seeds + 1))
    if args.graphs == "chung_lu":
        recall = 0.95 if args.recall is None else args.recall
        X, y = collect_logs(eng, chung_lu_training_graphs(seeds, args.n), per_sweep=args.per_sweep)
        what = (f"Chung-Lu n={args.n} gamma=2.1 mean_degree=17.35 cap=15000 seeds 1..{args.seeds}, "
                f"degree-proportional samples ({args.per_sweep} per sweep)")
    else:
        recall = 0.98 if args.recall is None else args.recall
        X, y = collect_logs(eng, sbm_training_graphs(seeds))
        what = f"SBM n=1024 blocks=16 p_in=0.10 p_out=0.002 seeds 1..{args.seeds}"
    p, loss, tau = train(eng, X, y, epochs=args.epochs, recall=recall)
    save(args.out, p, tau, {"train": what, "recall_target": recall, "epochs": args.epochs,
                            "samples": int(len(y)), "movers": int(y.sum()),
                            "loss_first": float(loss[0]), "loss_last": float(loss[-1]),
                            "min_degree": args.min_degree, "min_vertices": args.min_vertices})
    print(f"samples={len(y)} movers={int(y.sum())} loss {loss[0]:.4f} -> {loss[-1]:.4f} tau={tau:.4f} -> {args.out}")
    eng.close()


if __name__ == "__main__":
    main()
