"""C-ABI boundary checks that need no GPU: libgala.so loads and exports every
function include/gala.h declares; .bin IO is byte-exact against the
reference's own preprocess output; text ingest and generators behave."""
import ctypes
import os
import re

import numpy as np
import pytest

import gala

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "gala.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(gala_\w+)\s*\(", src, flags=re.M)
    return sorted(set(n for n in names if not n.endswith("_cb")))


def test_header_lists_expected_entry_points():
    names = declared_functions()
    assert set(gala.EXPORTS) <= set(names), set(gala.EXPORTS) - set(names)
    assert len(names) >= 19


def test_library_exports_every_declared_symbol():
    L = gala.lib()
    for name in declared_functions():
        assert hasattr(L, name), name
    assert gala.lib().gala_version().startswith(b"gala-mi355x")


def test_variant_counts_match_bench_tables():
    # the bench names the kernel of every (bin, variant) the library can launch
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_tables", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    counts = gala.variant_counts()
    assert len(counts) == 27 and min(counts) >= 1
    assert [len(x) for x in bench.KERNEL_GENERAL] == counts[:9]
    assert [len(x) for x in bench.KERNEL_FIRST] == counts[9:18]
    assert all(len(set(x)) == len(x) for x in bench.KERNEL_GENERAL)


def test_strerror_codes():
    L = gala.lib()
    assert L.gala_strerror(0) == b"ok"
    for code in range(-6, 0):
        assert L.gala_strerror(code) not in (b"ok", b"unknown error")


def test_bin_roundtrip_byte_exact(golden, tmp_path):
    src = os.path.join(golden, "karate.txt.bin")
    g = gala.Graph.load_bin_graph(src)
    assert g.vertex_num == 34 and g.edge_num == 78
    out = str(tmp_path / "k.bin")
    g.store_bin_graph(out)
    assert open(out, "rb").read() == open(src, "rb").read()


def test_ingest_matches_reference_preprocess(golden):
    # our ingest of karate.txt == the reference preprocess's .bin (golden)
    g = gala.Graph.from_text(os.path.join(golden, "karate.txt"))
    ref = gala.Graph.load_bin_graph(os.path.join(golden, "karate.txt.bin"))
    assert np.array_equal(g.degrees, ref.degrees) and np.array_equal(g.neighbors, ref.neighbors)


def test_ingest_renumbers_and_dedups(tmp_path):
    p = tmp_path / "e.txt"
    p.write_text("# comment\n10 20\n20 10\n10 30\n\n30 10\n")
    g = gala.Graph.from_text(str(p))
    assert g.vertex_num == 3
    assert list(g.degrees) == [2, 3, 4]
    assert list(g.neighbors) == [1, 2, 0, 0]


def test_ingest_weighted_and_selfloop(tmp_path):
    p = tmp_path / "w.txt"
    p.write_text("0 1 3\n1 1 2\n1 2 5\n")
    g = gala.Graph.from_text(str(p), gala.W_I32)
    # self-loop kept as one entry; its weight is summed twice (both directions pushed)
    assert list(g.degrees) == [1, 4, 5]
    assert list(g.neighbors) == [1, 0, 1, 2, 1]
    assert list(g.weights) == [3, 3, 4, 5, 5]


def test_truncated_and_malformed_bin(tmp_path, golden):
    L = gala.lib()
    src = open(os.path.join(golden, "karate.txt.bin"), "rb").read()
    bad = tmp_path / "t.bin"
    bad.write_bytes(src[:-10])
    with pytest.raises(gala.GalaError):
        gala.Graph.load_bin_graph(str(bad))
    with pytest.raises(gala.GalaError):
        gala.Graph.load_bin_graph(str(tmp_path / "missing.bin"))
    # neighbour id out of range
    arr = bytearray(src)
    arr[4 + 8 * 34: 4 + 8 * 34 + 4] = (999).to_bytes(4, "little")
    bad.write_bytes(bytes(arr))
    with pytest.raises(gala.GalaError):
        gala.Graph.load_bin_graph(str(bad))
    # offsets not monotone
    arr = bytearray(src)
    arr[4:12] = (50).to_bytes(8, "little")
    bad.write_bytes(bytes(arr))
    with pytest.raises(gala.GalaError):
        gala.Graph.load_bin_graph(str(bad))
    # an entry count so large that the size test would overflow if it
    # multiplied before comparing: rejected as a format error, no allocation
    for last in (2 ** 62, 2 ** 63 - 1, -5):
        arr = bytearray(src)
        arr[4 + 8 * 33: 4 + 8 * 34] = last.to_bytes(8, "little", signed=True)
        bad.write_bytes(bytes(arr))
        for wm in (gala.W_NONE, gala.W_I32):
            n, m = ctypes.c_int32(), ctypes.c_int64()
            rc = L.gala_read_bin(str(bad).encode(), wm, ctypes.byref(n), ctypes.byref(m), None, None, None)
            assert rc == -2, (last, wm, rc)
    assert L is not None


def test_generators_deterministic_and_canonical():
    a = gala.gen_sbm(1024, 16, 0.1, 0.002, 0)
    b = gala.gen_sbm(1024, 16, 0.1, 0.002, 0)
    c = gala.gen_sbm(1024, 16, 0.1, 0.002, 1)
    assert np.array_equal(a.neighbors, b.neighbors) and not np.array_equal(a.neighbors, c.neighbors)
    off = a.offsets()
    for v in range(a.vertex_num):
        row = a.neighbors[off[v]:off[v + 1]]
        assert np.all(np.diff(row) > 0) and v not in row
    # symmetric
    e = set()
    for v in range(a.vertex_num):
        for u in a.neighbors[off[v]:off[v + 1]]:
            e.add((v, int(u)))
    assert all((u, v) in e for v, u in e)


def test_sbm_expected_density():
    g = gala.gen_sbm(1024, 16, 0.10, 0.002, 0)
    # E ~ 16*C(64,2)*0.1 + (C(1024,2) - 16*C(64,2))*0.002 ~ 4.2k
    assert 3900 < g.edge_num < 4500


def test_chung_lu_small_shape():
    g = gala.gen_chung_lu(20000, 2.1, 17.35, 1500.0, 0)
    deg = np.diff(g.offsets())
    assert 14 < deg.mean() < 19
    assert deg.max() <= 2500 and deg.max() > 300


def test_rmat_small():
    g = gala.gen_rmat(10, 8, seed=0)
    assert g.vertex_num == 1024 and g.weights is not None
    assert np.all(g.weights > 0)
    off = g.offsets()
    for v in range(0, 1024, 37):
        row = g.neighbors[off[v]:off[v + 1]]
        assert np.all(np.diff(row) > 0)
