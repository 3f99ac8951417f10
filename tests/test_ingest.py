"""Text edge-list ingest: the host gala_ingest_text and the device
gala_ingest_text_device against the pure-Python oracle (oracle/ingest.py),
which is pinned to the reference's own preprocess (graph.cpp:18-151, compiled
from its sources into oracle/_ref/) on the committed fixtures and on random
SNAP-style files, and against each other byte for byte."""
import os
import subprocess

import numpy as np
import pytest

import gala
import ingest as oing

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
REF_PRE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                       "preprocess_ref")

HARD_DECIMALS = [
    "0.1", "0.2", "0.3", "1e-3", "2.5", ".5", "5.", "+1.25", "7E2", "1e22", "1e23", "1e-22",
    "9007199254740993", "9007199254740995", "123456789012345678901234567890",
    "3.14159265358979323846264338327950288", "2.2250738585072011e-308", "2.2250738585072014e-308",
    "4.9406564584124654e-324", "2.4703282292062328e-324", "1.7976931348623157e308",
    "1.000000000000000111022302462515654042363", "1.000000000000000111022302462515654042362",
    "0.000000000000000000000000000000000000000000001",
    "8.589973e9", "1234567890123456e-30", "123456789012345e22", "0000.000125", "1.5E+3",
]


def _write(tmp_path, name, data):
    p = tmp_path / name
    p.write_bytes(data if isinstance(data, bytes) else data.encode())
    return str(p)


def crafted_cases():
    """(name, text, weight modes) of edge lists that exercise every line rule."""
    f32 = [t for t in HARD_DECIMALS if abs(float(t)) < 3e38]  # sums must fit float32
    dec = "\n".join(f"{i} {i + 1 + (i % 3)} {t}" for i, t in enumerate(f32))
    return [
        ("empty", "", (0, 1, 2)),
        ("comments_only", "# a\n#b\n\n", (0, 1, 2)),
        ("no_trailing_newline", "1 2\n2 3", (0, 1, 2)),
        ("crlf", "1 2\r\n\r\n3 4\r\n# x\r\n", (0, 1, 2)),
        ("junk_lines", "% sym unweighted\nsrc dst\n1 2\n12abc 3\n4 5x\n  6\t7  \n8\n", (0, 1, 2)),
        ("self_loops", "0 0\n0 1\n1 1\n1 1\n2 2 5\n", (0, 1, 2)),
        ("duplicates", "5 9 2\n9 5 3\n5 9 4\n9 9 1\n5 5 7\n10 5\n", (0, 1, 2)),
        ("leading_zeros_signs", "007 +3\n-0 3\n+0 007 +9\n", (0, 1, 2)),
        ("sparse_ids", "2147483647 0\n1000000 2147483647 3\n5 1000000\n", (0, 1, 2)),
        ("weights_extra_tokens", "1 2 3 junk more\n2 3 4\n", (0, 1, 2)),
        ("hard_decimals", dec + "\n", (0, 2)),
        ("int_weights_edge", "1 2 2147483647\n3 4 -2147483648\n5 6 -7\n", (0, 1)),
        # lines longer than a 4 KiB tile and its staged kilobyte before it
        ("long_lines", "1 2\n" + " " * 6000 + "3 4 5\n#" + "y" * 9000 + "\n6\t" + " " * 5000 + "7 8\n" +
         "x" * 5000 + " 9 10\n11 12", (0, 1, 2)),
    ]


ERROR_CASES = [
    ("negative_id", "1 2\n-3 4\n", (0, 1, 2)),
    ("id_too_large", "1 2147483648\n", (0, 1, 2)),
    ("id_huge", "1 99999999999999999999999\n", (0, 1, 2)),
    ("bad_float", "1 2 1.2.3\n", (2,)),
    ("bad_exponent", "1 2 1e\n", (2,)),
    ("alpha_weight", "1 2 abc\n", (1, 2)),
    ("inf_weight", "1 2 1e400\n", (2,)),
    ("nan_text", "1 2 nan\n", (1, 2)),
    ("too_many_digits", "1 2 1." + "1" * 40 + "\n", (2,)),
    ("int_weight_float", "1 2 2.5\n", (1,)),
    ("int_weight_range", "1 2 2147483648\n", (1,)),
    ("int_sum_range", "1 2 2147483647\n2 1 1\n", (1,)),
    ("float_sum_range", "1 2 3e38\n2 1 3e38\n", (2,)),
]


def snap_like(seed, lines, id_bits=20, self_loops=False, weights=None):
    rng = np.random.default_rng(seed)
    pool = np.unique(rng.integers(0, 1 << id_bits, 4 * max(lines // 8, 8)))
    out = ["# generated"]
    for i in range(lines):
        a = int(pool[int(len(pool) * rng.random() ** 2)])
        b = int(pool[int(len(pool) * rng.random() ** 3)])
        if a == b and not self_loops:
            continue
        sep = ["\t", " ", "  "][int(rng.integers(3))]
        line = f"{a}{sep}{b}"
        if weights == "int":
            line += f" {int(rng.integers(1, 1000))}"
        elif weights == "float":
            k = int(rng.integers(4))
            x = float(rng.random() * 10 ** int(rng.integers(-5, 5)))
            line += " " + [f"{x:.17g}", f"{x:.3f}", f"{x:.6e}", str(int(rng.integers(1, 50)))][k]
        out.append(line)
        if rng.random() < 0.01:
            out.append("")
    return "\n".join(out) + "\n"


def host_ingest(path, mode):
    return gala.Graph.from_text(path, mode)


def same_graph(g, n, end, nbr, w):
    assert g.vertex_num == n
    assert np.array_equal(g.degrees, end) and np.array_equal(g.neighbors, nbr)
    if w is None:
        assert g.weights is None
    else:
        assert g.weights.dtype == w.dtype and g.weights.tobytes() == w.tobytes()


# ---------------------------------------------------------------------------
# oracle pinned to the reference preprocess
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", ["karate.txt", "snap_sparse.txt"])
def test_oracle_matches_reference_fixture(name):
    data = open(os.path.join(GOLDEN, name), "rb").read()
    n, end, nbr, _ = oing.ingest(data)
    ref = gala.Graph.load_bin_graph(os.path.join(GOLDEN, name + ".bin"))
    assert n == ref.vertex_num
    assert np.array_equal(end, ref.degrees) and np.array_equal(nbr, ref.neighbors)


@pytest.mark.skipif(not os.path.exists(REF_PRE), reason="reference preprocess not built (oracle/_ref)")
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_oracle_matches_reference_preprocess_random(tmp_path, seed):
    path = _write(tmp_path, f"r{seed}.txt", snap_like(seed, 3000, id_bits=14 + 3 * seed))
    subprocess.run([REF_PRE, "-f", path], check=True, stdout=subprocess.DEVNULL)
    ref = gala.Graph.load_bin_graph(path + ".bin")
    n, end, nbr, _ = oing.ingest(open(path, "rb").read())
    assert n == ref.vertex_num
    assert np.array_equal(end, ref.degrees) and np.array_equal(nbr, ref.neighbors)


def test_oracle_decimal_rules():
    # decimal weights: Python's float() (correctly rounded) after the grammar check
    for t in HARD_DECIMALS:
        (_, _, x), = oing.parse_lines(f"1 2 {t}\n".encode(), oing.W_F32)
        assert x == float(t)
    assert oing._sig_digits("0001.2300e5") == 3 and oing._sig_digits("-0.000") == 0
    for bad in ["1.2.3", "1e", ".", "+", "e5", "1e+", "0x10", "inf", "1_0"]:
        with pytest.raises(oing.IngestError):
            oing.parse_lines(f"1 2 {bad}\n".encode(), oing.W_F32)


def test_oracle_self_loop_and_weight_invariants():
    # a self-loop line is one entry carrying its weight twice; weights symmetric
    n, end, nbr, w = oing.ingest(b"0 0 3\n0 1 2\n1 0 5\n", oing.W_I32)
    assert n == 2 and list(end) == [2, 3] and list(nbr) == [0, 1, 0] and list(w) == [6, 7, 7]
    text = snap_like(5, 2000, id_bits=10, self_loops=True, weights="float").encode()
    n, end, nbr, w = oing.ingest(text, oing.W_F32)
    off = np.concatenate([[0], end])
    W = {}
    for u in range(n):
        for j in range(off[u], off[u + 1]):
            W[(u, int(nbr[j]))] = float(w[j])
    assert all(W[(v, u)] == x for (u, v), x in W.items())
    # the stored total equals twice the line weights (self-loops pushed twice) to float32 rounding
    lines = oing.parse_lines(text, oing.W_F32)
    assert abs(sum(W.values()) - 2 * sum(x for _, _, x in lines)) <= 1e-6 * 2 * sum(x for _, _, x in lines)


# ---------------------------------------------------------------------------
# host ingest == oracle
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("case", crafted_cases(), ids=lambda c: c[0])
def test_host_matches_oracle(tmp_path, case):
    name, text, modes = case
    path = _write(tmp_path, name + ".txt", text)
    for mode in modes:
        n, end, nbr, w = oing.ingest(text.encode(), mode)
        same_graph(host_ingest(path, mode), n, end, nbr, w)


@pytest.mark.parametrize("case", ERROR_CASES, ids=lambda c: c[0])
def test_host_errors_like_oracle(tmp_path, case):
    name, text, modes = case
    path = _write(tmp_path, name + ".txt", text)
    for mode in modes:
        with pytest.raises(oing.IngestError):
            oing.ingest(text.encode(), mode)
        with pytest.raises(gala.GalaError):
            host_ingest(path, mode)


@pytest.mark.parametrize("weights,mode", [(None, 0), ("int", 1), ("float", 2)])
def test_host_matches_oracle_random(tmp_path, weights, mode):
    text = snap_like(11, 4000, id_bits=14, self_loops=True, weights=weights)
    path = _write(tmp_path, "r.txt", text)
    same_graph(host_ingest(path, mode), *oing.ingest(text.encode(), mode))


def test_host_float_ingest_of_integer_weights_equals_int_ingest(tmp_path):
    text = snap_like(12, 3000, id_bits=12, self_loops=True, weights="int")
    path = _write(tmp_path, "i.txt", text)
    gi, gf = host_ingest(path, gala.W_I32), host_ingest(path, gala.W_F32)
    assert np.array_equal(gi.neighbors, gf.neighbors)
    assert np.array_equal(gi.weights.astype(np.float32), gf.weights)


# ---------------------------------------------------------------------------
# device ingest == host ingest, byte for byte
# ---------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("name", ["karate.txt", "snap_sparse.txt"])
def test_device_matches_reference_fixture(engine, tmp_path, name):
    g = engine.ingest_text(os.path.join(GOLDEN, name))
    out = str(tmp_path / "d.bin")
    g.store_bin_graph(out)
    assert open(out, "rb").read() == open(os.path.join(GOLDEN, name + ".bin"), "rb").read()


@pytest.mark.gpu
def test_device_matches_host_crafted(engine, tmp_path):
    for name, text, modes in crafted_cases():
        path = _write(tmp_path, name + ".txt", text)
        for mode in modes:
            print(f"device ingest {name} mode {mode}", flush=True)
            h = host_ingest(path, mode)
            same_graph(engine.ingest_text(path, mode), h.vertex_num, h.degrees, h.neighbors, h.weights)


@pytest.mark.gpu
def test_device_errors_like_host(engine, tmp_path):
    for name, text, modes in ERROR_CASES:
        path = _write(tmp_path, name + ".txt", text)
        for mode in modes:
            print(f"device ingest error case {name} mode {mode}", flush=True)
            with pytest.raises(gala.GalaError):
                engine.ingest_text(path, mode)
    with pytest.raises(gala.GalaError):
        engine.ingest_text(str(tmp_path / "missing.txt"))


@pytest.mark.gpu
@pytest.mark.parametrize("weights,mode,lines,id_bits", [
    (None, 0, 400000, 22), ("int", 1, 300000, 18), ("float", 2, 300000, 20), ("float", 2, 50000, 31)])
def test_device_matches_host_random(engine, tmp_path, weights, mode, lines, id_bits):
    text = snap_like(21 + mode, lines, id_bits=id_bits, self_loops=True, weights=weights)
    path = _write(tmp_path, "big.txt", text)
    print(f"device ingest random {lines} lines mode {mode}", flush=True)
    h = host_ingest(path, mode)
    same_graph(engine.ingest_text(path, mode), h.vertex_num, h.degrees, h.neighbors, h.weights)


@pytest.mark.gpu
def test_device_hard_decimals_bulk(engine, tmp_path):
    # 200k weight tokens of every shape the decimal conversion distinguishes
    rng = np.random.default_rng(3)
    toks = []
    for i in range(200000):
        k = i % 5
        if k == 0:
            toks.append(repr(float(np.frombuffer(rng.bytes(8), np.float64)[0])))
        elif k == 1:
            toks.append(f"{rng.random():.{int(rng.integers(1, 25))}e}")
        elif k == 2:
            toks.append(f"{rng.random() * 1e-300:.17g}")
        elif k == 3:
            toks.append(f"{int(rng.integers(1, 10**15))}e{int(rng.integers(-40, 40))}")
        else:
            toks.append(HARD_DECIMALS[i % len(HARD_DECIMALS)])
    # finite, and within float32 once summed (every pair occurs once)
    toks = [t for t in toks if "inf" not in t and "nan" not in t and abs(float(t)) < 3e38]
    lines = [f"{i} {i + 1} {t}" for i, t in enumerate(toks)]
    path = _write(tmp_path, "dec.txt", "\n".join(lines) + "\n")
    try:
        h = host_ingest(path, gala.W_F32)
    except gala.GalaError:
        pytest.fail("hard-decimal file should ingest")
    same_graph(engine.ingest_text(path, gala.W_F32), h.vertex_num, h.degrees, h.neighbors, h.weights)


@pytest.mark.gpu
def test_device_two_call_protocol_and_load_text(engine, tmp_path):
    a = _write(tmp_path, "a.txt", snap_like(31, 20000, id_bits=16))
    b = _write(tmp_path, "b.txt", snap_like(32, 20000, id_bits=16, weights="int"))
    ga = engine.ingest_text(a)
    gb = engine.ingest_text(b, gala.W_I32)
    ga2 = engine.ingest_text(a)
    assert np.array_equal(ga.neighbors, ga2.neighbors) and not np.array_equal(ga.degrees[:10], gb.degrees[:10])
    t = engine.ingest_times()
    assert t["wall_ms"] > 0 and t["device_ms"] > 0
    # the graph set by gala_load_text runs like the host-ingested one
    engine.load_text(b, gala.W_I32)
    r1 = engine.run()
    engine.set_graph(host_ingest(b, gala.W_I32))
    r2 = engine.run()
    assert r1["modularity"] == r2["modularity"] and r1["n_communities"] == r2["n_communities"]


# ---------------------------------------------------------------------------
# the preprocess drop-in CLI (preprocess.cpp:7-47), on the host and with -g on the GPU
# ---------------------------------------------------------------------------
CLI_PRE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gala-mg_amd", "preprocess")


def run_preprocess(src, out_dir, gpu=None):
    cmd = [CLI_PRE, "-f", src, "-o", str(out_dir)] + ([] if gpu is None else ["-g", str(gpu)])
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[0] == os.path.join(str(out_dir), os.path.basename(src) + ".bin")
    assert lines[-1] == "store successfully. "
    return open(lines[0], "rb").read()


@pytest.mark.parametrize("name", ["karate.txt", "snap_sparse.txt"])
def test_preprocess_cli_matches_reference(tmp_path, name):
    src = os.path.join(GOLDEN, name)
    assert run_preprocess(src, tmp_path) == open(src + ".bin", "rb").read()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["karate.txt", "snap_sparse.txt"])
def test_preprocess_cli_on_gpu_matches_reference(tmp_path, name):
    src = os.path.join(GOLDEN, name)
    print(f"preprocess -g 0 {name}", flush=True)
    assert run_preprocess(src, tmp_path, gpu=0) == open(src + ".bin", "rb").read()
