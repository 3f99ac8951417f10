"""The `gala_main` drop-in CLI and its stdout contract.

The reference prints, per run, per round and per iteration, the lines of its
README example (README.md:57-89; printers at src/gala_main.cu:8-57,
src/louvain_gpu/louvain_main_process.cu:547-548,1002-1003,1040-1044,
src/louvain_gpu/louvain.cu:180,191,218,221-224, src/graph/graph.cpp:213).
The GPU test runs the CLI on the README's own input (the karate club) and checks
every line against those formats, the Q trajectory against the README's
printed values, and the -o partition against the oracle.  It also exercises
`gala_load_bin` (file straight into the engine), which the Python binding does
not use.
"""
import json
import os
import re
import subprocess

import numpy as np
import pytest

import gala
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "gala-mg_amd", "gala_main")

FLOAT = r"(-?\d+\.\d{6})"
MS = r"(\d+\.\d{6})ms"
PATTERNS = {
    "vertex": re.compile(r"^vertex number:(\d+) edge number:(\d+)$"),
    "load": re.compile(r"^load success$"),
    "round": re.compile(r"^===============round:(\d+)===============$"),
    "iter": re.compile(r"^(gpu-\d+: )?Iteration:(\d+) Q:" + FLOAT + "$"),
    "t_round": re.compile(r"^time without data init = " + MS + "$"),
    "t_split": re.compile(r"^decideandmove time = " + MS + r" weight updating time = " + MS +
                          r" remaining time = (-?\d+\.\d{6})ms$"),
    "t_louvain0": re.compile(r"^louvain time in the first round = " + MS + "$"),
    "t_build0": re.compile(r"^build time in the first round = " + MS + "$"),
    "ncomm": re.compile(r"^number of communities:(\d+) modularity:" + FLOAT + "$"),
    "sep": re.compile(r"^=====================================$"),
    "final": re.compile(r"^final number of communities:(\d+) -> (\d+) final modularity:" + FLOAT + "$"),
    "t_exec": re.compile(r"^execution time without data transfer = " + MS + "$"),
    "t_elapsed": re.compile(r"^elapsed time = " + MS + "$"),
}


def classify(line):
    for k, p in PATTERNS.items():
        m = p.match(line)
        if m:
            return k, m
    raise AssertionError(f"line matches no reference format: {line!r}")


def parse(stdout):
    """Splits the CLI output into its parts, checking the line grammar of the
    README example: header, rounds of (round, Iteration+, timers, communities),
    then separator, final line, execution and elapsed time."""
    lines = [classify(x) for x in stdout.splitlines() if x.strip()]
    kinds = [k for k, _ in lines]
    grammar = re.compile(r"^vertex load (round (iter )+t_round t_split (t_louvain0 t_build0 )?ncomm )+"
                         r"sep final t_exec t_elapsed $")
    assert grammar.match(" ".join(kinds) + " "), kinds
    out = {"rounds": []}
    for k, m in lines:
        if k == "vertex":
            out["n"], out["edges"] = int(m.group(1)), int(m.group(2))
        elif k == "round":
            assert int(m.group(1)) == len(out["rounds"])
            out["rounds"].append({"q": [], "first": False})
        elif k == "iter":
            r = out["rounds"][-1]
            assert int(m.group(2)) == len(r["q"])
            r["q"].append(m.group(3))
        elif k == "t_louvain0":
            out["rounds"][-1]["first"] = True
        elif k == "ncomm":
            out["rounds"][-1]["communities"] = int(m.group(1))
            out["rounds"][-1]["modularity"] = m.group(2)
        elif k == "final":
            out["final"] = (int(m.group(1)), int(m.group(2)), m.group(3))
    return out


def test_parser_accepts_the_readme_example():
    # the README's own example (README.md:58-89), retyped from its Q values and
    # counts, parses to the fixture's trajectory: the grammar is not too strict
    ref = json.load(open(os.path.join(ROOT, "tests", "golden", "karate_readme.json")))
    lines = ["vertex number:34 edge number:78", "load success"]
    for r, rd in enumerate(ref["rounds"]):
        lines.append(f"===============round:{r}===============")
        lines += [f"Iteration:{i} Q:{q:.6f}" for i, q in enumerate(rd["q"])]
        lines += ["time without data init = 1.000000ms",
                  "decideandmove time = 0.100000ms weight updating time = 0.200000ms remaining time = 0.700000ms"]
        if r == 0:
            lines += ["louvain time in the first round = 2.000000ms", "build time in the first round = 1.000000ms"]
        lines.append(f"number of communities:{rd['communities']} modularity:{rd['modularity']:.6f}")
    lines += ["=====================================",
              "final number of communities:34 -> 4 final modularity:0.408695",
              "execution time without data transfer = 50.241943ms", "elapsed time = 313.437988ms"]
    out = parse("\n".join(lines) + "\n")
    assert [r["q"] for r in out["rounds"]] == [[f"{q:.6f}" for q in rd["q"]] for rd in ref["rounds"]]
    # and a line outside the grammar is rejected
    with pytest.raises(AssertionError):
        parse("\n".join(lines[:3] + ["Iteration 0 Q 0.1"] + lines[3:]))


def test_cli_usage_and_missing_file(tmp_path):
    # argument and file errors are reported before any device is touched
    if not os.path.exists(CLI):
        pytest.skip("gala_main not built")
    r = subprocess.run([CLI], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Usage" in r.stderr
    r = subprocess.run([CLI, "-f", str(tmp_path / "absent.bin")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "cannot read" in r.stderr


@pytest.mark.gpu
def test_cli_karate_readme_output(golden, tmp_path):
    print("gala_main karate", flush=True)
    assert os.path.exists(CLI), "gala_main not built"
    path = os.path.join(golden, "karate.txt.bin")
    out_file = tmp_path / "comm.txt"
    r = subprocess.run([CLI, "-f", path, "-o", str(out_file)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = parse(r.stdout)
    ref = json.load(open(os.path.join(golden, "karate_readme.json")))
    assert out["n"] == 34 and out["edges"] == 78
    assert len(out["rounds"]) == len(ref["rounds"])
    for got, exp in zip(out["rounds"], ref["rounds"]):
        assert got["q"] == [f"{q:.6f}" for q in exp["q"]]
        assert got["communities"] == exp["communities"]
        assert got["modularity"] == f"{exp['modularity']:.6f}"
    assert [r_["first"] for r_ in out["rounds"]] == [True] + [False] * (len(out["rounds"]) - 1)
    assert out["final"] == (34, 4, "0.408695")
    comm = np.array([int(x) for x in out_file.read_text().split()], np.int32)
    g = gala.Graph.load_bin_graph(path)
    ref_run = oracle.louvain(g.vertex_num, g.degrees, g.neighbors)
    assert np.array_equal(comm, ref_run["final_comm"])


@pytest.mark.gpu
def test_cli_sbm1024_matches_oracle(golden, tmp_path):
    # C1 through the CLI: every printed Q equals the oracle's Q to the printed
    # six decimals, and the -o partition equals the oracle's composed partition
    print("gala_main sbm1024", flush=True)
    path = os.path.join(golden, "sbm1024_s0.bin")
    out_file = tmp_path / "comm.txt"
    r = subprocess.run([CLI, "-f", path, "-o", str(out_file)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = parse(r.stdout)
    g = gala.Graph.load_bin_graph(path)
    t = oracle.Trace(keep_vectors=False)
    ref_run = oracle.louvain(g.vertex_num, g.degrees, g.neighbors, trace=t)
    for rnd, got in enumerate(out["rounds"]):
        exp = [f"{q:.6f}" for (a, _, q, _, _) in t.rows if a == rnd]
        assert got["q"] == exp, rnd
        assert got["communities"] == ref_run["round_nc"][rnd]
    assert out["final"][:2] == (g.vertex_num, ref_run["nc"])
    comm = np.array([int(x) for x in out_file.read_text().split()], np.int32)
    assert np.array_equal(comm, ref_run["final_comm"])
