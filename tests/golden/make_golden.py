"""Regenerates the committed fixtures under tests/golden/.

  karate.txt.bin   : the reference's own preprocess (compiled from its sources by
                     oracle/Makefile into oracle/_ref/) run on karate.txt
  sbm1024_s0.bin   : the 1,024-vertex planted-partition SBM (16 blocks,
                     p_in 0.10, p_out 0.002, seed 0) from gala_gen_sbm, written
                     as an edge list and converted by the reference preprocess
  sbm1024_s0.json  : the oracle's per-iteration Q / moved counts and per-round
                     community counts on that graph, plus the final partition
  karate_readme.json : the known-answer trajectory printed in the reference
                     README (README.md:58-89), typed in as data
  snap_sparse.txt(.bin) : a SNAP-style edge list (seeded): '#' header lines,
                     blank lines, ids scattered below 2^24, leading zeros and
                     '+' signs, tab / multi-space separators, repeated and
                     reversed edges, no self-loops (the reference truncates
                     those); the .bin is the reference preprocess's output

Run from the repository root after `make -C oracle && make -C gala-mg_amd`.
"""
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "gala-mg_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import gala  # noqa: E402
import oracle  # noqa: E402

HERE = os.path.join(ROOT, "tests", "golden")
REF_PRE = os.path.join(ROOT, "oracle", "_ref", "preprocess_ref")

README_KARATE = {
    "source": "reference README.md:58-89 (karate club, 34 vertices, 78 edges)",
    "rounds": [
        {"q": [-0.049803, 0.136834, 0.191239, 0.169214], "communities": 11, "modularity": 0.191239},
        {"q": [0.191239, 0.313116, 0.401134, 0.408695, 0.408695], "communities": 4,
         "modularity": 0.408695},
        {"q": [0.408695, 0.408695], "communities": 4, "modularity": 0.408695},
    ],
    "final": {"from": 34, "to": 4, "modularity": 0.408695},
}


def ref_preprocess(txt):
    subprocess.run([REF_PRE, "-f", txt], check=True, stdout=subprocess.DEVNULL)
    return txt + ".bin"


def snap_sparse_text(seed=7, lines=1500, pool=400):
    rng = np.random.default_rng(seed)
    ids = np.unique(rng.integers(0, 1 << 24, pool * 2))[:pool]
    rng.shuffle(ids)
    pick = lambda: int(ids[int(pool * rng.random() ** 2)])  # noqa: E731  (skewed toward the first ids)
    out = ["# Directed graph (each unordered pair of nodes is saved once)",
           "# Nodes: ? Edges: ?", "# FromNodeId\tToNodeId"]
    seen = []
    for i in range(lines):
        r = rng.random()
        if r < 0.02:
            out.append("")
            continue
        if r < 0.10 and seen:
            a, b = seen[int(rng.integers(len(seen)))]
            if rng.random() < 0.5:
                a, b = b, a
        else:
            a, b = pick(), pick()
            while b == a:
                b = pick()
            seen.append((a, b))
        fa = f"{a:07d}" if rng.random() < 0.05 else (f"+{a}" if rng.random() < 0.03 else str(a))
        sep = ["\t", " ", "   ", " \t "][int(rng.integers(4))]
        lead = " " if rng.random() < 0.03 else ""
        tail = " " if rng.random() < 0.05 else ""
        out.append(f"{lead}{fa}{sep}{b}{tail}")
    return "\n".join(out) + "\n"


def main():
    with open(os.path.join(HERE, "karate_readme.json"), "w") as f:
        json.dump(README_KARATE, f, indent=1)
    if os.path.exists(REF_PRE):
        ref_preprocess(os.path.join(HERE, "karate.txt"))
        snap = os.path.join(HERE, "snap_sparse.txt")
        with open(snap, "w") as f:
            f.write(snap_sparse_text())
        ref_preprocess(snap)
    g = gala.gen_sbm(1024, 16, 0.10, 0.002, 0)
    off = g.offsets()
    assert np.all(np.diff(off) > 0), "SBM fixture has an isolated vertex; pick another seed"
    txt = os.path.join("/tmp", "sbm1024_s0")
    with open(txt, "w") as f:
        for u in range(g.vertex_num):
            for v in g.neighbors[off[u]:off[u + 1]]:
                if u < v:
                    f.write(f"{u} {v}\n")
    if os.path.exists(REF_PRE):
        b = ref_preprocess(txt)
        ref = gala.Graph.load_bin_graph(b)
        assert np.array_equal(ref.degrees, g.degrees) and np.array_equal(ref.neighbors, g.neighbors)
        os.replace(b, os.path.join(HERE, "sbm1024_s0.bin"))
    else:
        g.store_bin_graph(os.path.join(HERE, "sbm1024_s0.bin"))
    os.remove(txt)
    t = oracle.Trace(keep_vectors=False)
    r = oracle.louvain(g.vertex_num, g.degrees, g.neighbors, trace=t)
    out = {
        "graph": "SBM n=1024 blocks=16 p_in=0.10 p_out=0.002 seed=0",
        "n": int(g.vertex_num), "undirected_edges": int(g.edge_num),
        "iterations": [[int(a), int(b), float(q), int(mv)] for a, b, q, mv, _ in t.rows],
        "round_nc": [int(x) for x in r["round_nc"]],
        "round_q": [float(x) for x in r["round_q"]],
        "final_q": float(r["Q"]), "final_nc": int(r["nc"]),
        "final_comm": [int(x) for x in r["final_comm"]],
    }
    with open(os.path.join(HERE, "sbm1024_s0.json"), "w") as f:
        json.dump(out, f)
    print("golden fixtures written", out["undirected_edges"], out["final_nc"], out["final_q"])


if __name__ == "__main__":
    main()
