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


def main():
    with open(os.path.join(HERE, "karate_readme.json"), "w") as f:
        json.dump(README_KARATE, f, indent=1)
    if os.path.exists(REF_PRE):
        ref_preprocess(os.path.join(HERE, "karate.txt"))
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
