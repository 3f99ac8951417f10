"""Trains the learned-frontier MLP on move logs of plain runs.

Move logs: for every plain sweep after the first of each round, the six
features of every swept vertex and whether the sweep moved it.  Training data:
the 1,024-vertex planted-partition SBMs (16 blocks, p_in 0.10, p_out 0.002)
with seeds 1..32.  Training runs on the GPU (the engine's own BF16/MFMA
forward and backward kernels, Adam).

    python gala-mg_amd/train_frontier.py [--seeds 32] [--out gala-mg_amd/models/frontier_mlp.json]
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


def collect_logs(engine, graphs):
    """Plain runs of `graphs` with the move logger on -> (X[n, 6] float32, y[n] uint8)."""
    xs, ys = [], []

    def sink(rnd, it, f, m):
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds", type=int, default=32)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--recall", type=float, default=0.98)
    ap.add_argument("--out", default=DEFAULT_MODEL)
    args = ap.parse_args()
    eng = gala.Engine()
    seeds = list(range(1, args.seeds + 1))
    X, y = collect_logs(eng, sbm_training_graphs(seeds))
    p, loss, tau = train(eng, X, y, epochs=args.epochs, recall=args.recall)
    save(args.out, p, tau, {"train": f"SBM n=1024 blocks=16 p_in=0.10 p_out=0.002 seeds 1..{args.seeds}",
                            "recall_target": args.recall, "epochs": args.epochs,
                            "samples": int(len(y)), "movers": int(y.sum()),
                            "loss_first": float(loss[0]), "loss_last": float(loss[-1])})
    print(f"samples={len(y)} movers={int(y.sum())} loss {loss[0]:.4f} -> {loss[-1]:.4f} tau={tau:.4f} -> {args.out}")
    eng.close()


if __name__ == "__main__":
    main()
