"""Benchmark: end-to-end Louvain on the 4.0M-vertex / 34.4M-edge Chung-Lu
power-law graph (gamma 2.1, mean degree 17.35, degree cap 15,000, seed 0,
unweighted) on one MI355X.

One step = one full multi-round Louvain run (all rounds, contraction included)
on the graph already resident in HBM (the input CSR stays read-only on the
device between steps).  value = edge-TEPS = undirected input edges / T_e2e,
with T_e2e the host wall clock of the timed steps (each run ends with a
device synchronisation); the engine's own HIP-event time is reported beside it.

Also reported:
  roofline      the dominant decide-and-move kernel (largest total time in a run
                with per-bin HIP events on the engine's stream) against HBM
                peak: its algorithmic bytes (16*d + 40 per swept vertex of
                degree d, 12*d + 40 with implicit unit weights; first sweeps
                12*d + 40 / 8*d + 40) over its launch time; the whole decide
                phase is reported beside it
  contraction   the contraction phase against HBM peak: per round
                12*2E + 8*2E' + 16*V + 8*C bytes (read the round CSR and the
                community of every entry, write the contracted CSR) over its time
  cpu_baseline  the oracle (exact-integer C port of the same algorithm, OpenMP)
                on the same graph on the host's CPU share, 1 run; and the same
                on one thread
  parity        final partition of the last GPU step == oracle's, bit for bit
  learned       the learned-frontier form (network trained on SBM move logs,
                gala-mg_amd/models/frontier_mlp.json) on held-out seeds
                1001-1003 of the same workload: time per run, speedup over the
                plain path on the same graph and final-Q difference
  batched       4,096 independent SBMs run as one union against the same graphs
                run one at a time
  weighted_rmat R-MAT scale 24 with fp32 weights (run exactly: the weights
                scaled to integers, int64 sums, int128 gains)
  text_ingest   the same graph as a 34.4M-line text edge list, ingested on the
                GPU (gala_ingest_text_device) against the host ingest and the
                reference's own preprocess on a 2M-line sample
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gala-mg_amd"))

HBM_PEAK_GBS = 8000.0
BIN_DEG = ["1-3", "4-7", "8-15", "16-31", "32-127", "128-511", "512-4095", "4096-8191", ">=8192"]
# decide-and-move kernel of every (degree bin, variant) on the integer path,
# as the engine launches them: general sweeps, and the first sweep of a round
KERNEL_GENERAL = [
    ["k_decide_ring<4, IntW>", "k_decide_wave<8, 4, IntW>", "k_decide_pipe<4, false, IntW>",
     "k_decide_tpv<3, false, IntW>"],
    ["k_decide_ring<8, IntW>", "k_decide_wave<16, 8, IntW>", "k_decide_pipe<8, false, IntW>",
     "k_decide_tpv<7, false, IntW>"],
    ["k_decide_ring<16, IntW>", "k_decide_wave<32, 16, IntW>", "k_decide_pipe<16, false, IntW>",
     "k_decide_tpv<15, false, IntW>", "k_decide_hlead<32, 8, 2, false, IntW>"],
    ["k_decide_ring<32, IntW>", "k_decide_wave<64, 32, IntW>", "k_decide_pipe<32, false, IntW>",
     "k_decide_tpv<31, false, IntW>", "k_decide_hlead<64, 16, 2, false, IntW>"],
    ["k_decide_wave<256, 64, IntW>", "k_decide_wave<256, 32, IntW>", "k_decide_wave<256, 16, IntW>",
     "k_decide_hpipe<256, 32, 4, IntW>", "k_decide_hpipe<256, 16, 8, IntW>",
     "k_decide_hlead<256, 32, 4, false, IntW>", "k_decide_hlead<256, 16, 8, false, IntW>",
     "k_decide_hlead<256, 8, 16, false, IntW>", "k_decide_glead<256, 16, 8, false, IntW>"],
    ["k_decide_wave<1024, 64, IntW>", "k_decide_wave<1024, 32, IntW>", "k_decide_hpipe<1024, 64, 8, IntW>",
     "k_decide_hlead<1024, 64, 8, false, IntW>", "k_decide_hlead<1024, 32, 16, false, IntW>",
     "k_decide_glead<1024, 64, 8, false, IntW>"],
    ["k_decide_block<256, 8192, IntW>", "k_decide_blead<256, 8192, 16, 1, IntW>"],
    ["k_decide_block<1024, 16384, IntW>", "k_decide_blead<1024, 16384, 8, 1, IntW>"],
    ["k_decide_hub<16384, IntW>", "k_hub_plan/agg/scan/commit<IntW>",
     "k_decide_hub<16384, IntW> + k_hub_plan/agg/scan/commit<IntW>",
     "k_decide_blead<1024, 16384, 16, 2, IntW> + k_hub_plan/agg/scan/commit<IntW>",
     "k_decide_hub<16384, IntW> + k_hub_part<16384, IntW>",
     "k_decide_hub<16384, IntW> + k_hub_part<16384, IntW> (one part per hub)",
     "bin-8 split sequence: k_hub_plan, k_hub_part<16384, IntW>, k_hub_commit (k_hub_resolve first when a "
     "row has 3+ parts; k_hub_plan, k_hub_agg, k_hub_scan, k_hub_commit after it for rows too long to split)"]]
KERNEL_FIRST = [
    ["k_decide_first<4, IntW>", "k_decide_pipe<4, true, IntW>", "k_decide_tpv<3, true, IntW>"],
    ["k_decide_first<8, IntW>", "k_decide_pipe<8, true, IntW>", "k_decide_tpv<7, true, IntW>"],
    ["k_decide_first<16, IntW>", "k_decide_pipe<16, true, IntW>", "k_decide_tpv<15, true, IntW>"],
    ["k_decide_first<32, IntW>", "k_decide_pipe<32, true, IntW>", "k_decide_tpv<31, true, IntW>",
     "k_decide_hlead<1, 16, 2, true, IntW>"],
    ["k_decide_first<64, IntW>", "k_decide_first<32, IntW>", "k_decide_hlead<1, 16, 8, true, IntW>",
     "k_decide_glead<1, 16, 8, true, IntW>"],
    ["k_decide_first<64, IntW>", "k_decide_hlead<1, 64, 8, true, IntW>", "k_decide_glead<1, 64, 8, true, IntW>"],
    ["k_decide_first_block<256, IntW>"],
    ["k_decide_first_block<1024, IntW>"],
    ["k_decide_first_block<1024, IntW>"]]


# HBM bytes per launch of the decide kernels on this workload, from two
# rocprofv3 --pmc runs of this script (FETCH_SIZE, then WRITE_SIZE, KiB per
# dispatch, averaged over the kernel's dispatches), each scaled by what the
# same counter reads for a plain 256 MiB device copy: HBM bytes = 2.00 x
# FETCH_SIZE + 1.00 x WRITE_SIZE.  Commands and the full table:
# profiles/decide_traffic_chung_lu_4M.md.  A counter run cannot be made from
# inside the timed process, so the figures are recorded here and reported
# beside the live timing of the same kernel.
PMC_FETCH_FACTOR, PMC_WRITE_FACTOR = 2.00, 1.00
KERNEL_HUB_PART = "k_hub_part<16384, IntW>"  # the part kernel of the bin-8 split sequence
PMC_KIB = {  # kernel: (FETCH_SIZE KiB, WRITE_SIZE KiB) per dispatch
    KERNEL_HUB_PART: (196282.3, 26.4),
    "k_decide_glead<256, 16, 8, false, IntW>": (193498.7, 7295.9),
    "k_decide_blead<256, 8192, 16, 1, IntW>": (124281.4, 580.8),
    "k_decide_glead<1024, 64, 8, false, IntW>": (107480.0, 1433.8),
    "k_decide_glead<1, 16, 8, true, IntW>": (105940.6, 5874.6),
    "k_decide_tpv<31, false, IntW>": (103958.6, 6662.0),
    "k_decide_tpv<15, false, IntW>": (85845.0, 9329.2),
    "k_decide_first_block<256, IntW>": (77679.1, 671.7),
    "k_decide_tpv<7, false, IntW>": (64397.8, 11764.5),
    "k_decide_glead<1, 64, 8, true, IntW>": (58218.3, 1062.2),
    "k_decide_tpv<15, true, IntW>": (56296.8, 8347.0),
    "k_decide_tpv<31, true, IntW>": (55553.6, 5671.7),
    "k_decide_blead<1024, 16384, 8, 1, IntW>": (50872.7, 84.3),
    "k_decide_first_block<1024, IntW>": (49557.4, 104.9),
    "k_decide_tpv<7, true, IntW>": (41649.4, 9658.7),
    "k_decide_tpv<3, false, IntW>": (25767.5, 8027.9),
    "k_decide_pipe<4, true, IntW>": (16396.7, 5978.3),
}


def pmc_traffic(kernel):
    """HBM bytes per launch of a decide kernel (None if not measured)."""
    if kernel not in PMC_KIB:
        return None
    f, w = PMC_KIB[kernel]
    return 1024.0 * (PMC_FETCH_FACTOR * f + PMC_WRITE_FACTOR * w)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_threads(args):
    """Host threads for the CPU legs: the CPU share this process may use (its
    affinity, and OMP_NUM_THREADS where the host sets it), unless given."""
    if args.cpu_threads:
        return args.cpu_threads
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return min(n, int(omp)) if omp.isdigit() and int(omp) > 0 else n


def contraction_roofline(r):
    """Contraction against HBM peak: per round 12*2E + 8*2E' + 16*V + 8*C bytes
    (read the round CSR with each entry's community, write the contracted CSR).
    A round whose first sweep was refused keeps the singleton partition; the
    engine skips its contraction (the identity), so it counts no bytes."""
    rounds = []
    tot_b = 0
    for i in range(r["rounds"]):
        V, m_in, m_out, C = (r["round_vertices"][i], r["round_entries"][i], r["round_contracted_entries"][i],
                             r["round_nc"][i])
        skipped = r["round_iters"][i] == 1 and C == V and m_out == m_in
        b = 0 if skipped else 12 * m_in + 8 * m_out + 16 * V + 8 * C
        ms = r["round_build_ms"][i]
        tot_b += b
        rounds.append({"round": i, "vertices": V, "entries": m_in, "contracted_entries": m_out,
                       "communities": C, "bytes": b, "ms": ms, "skipped": skipped,
                       "GBps": b / (ms * 1e-3) / 1e9 if ms > 0 else 0.0})
    ms = r["t_contract_ms"]
    gbs = tot_b / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    return {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
            "algorithmic_bytes": tot_b, "time_ms": ms, "rounds": rounds}


def learned_leg(gala, args):
    """Learned frontier on held-out seeds 1001-1003 of the workload (the network
    never saw them: it is trained on move logs of seeds 1-32 at 1M vertices).  Per seed:
    the plain path and the learned form on the same graph, K timed runs each
    after one warmup."""
    import numpy as np
    import train_frontier as TF
    if not os.path.exists(TF.DEFAULT_MODEL):
        log("learned leg skipped: no trained model")
        return None
    params, tau = TF.load(TF.DEFAULT_MODEL)
    settings = TF.load_settings(TF.DEFAULT_MODEL)
    per_seed = []
    for seed in (1001, 1002, 1003):
        g = gala.gen_chung_lu(args.n, 2.1, 17.35, 15000.0, seed)
        res = {}
        for mode in (0, 1):
            eng = gala.Engine(device=0, frontier_mode=mode, **settings)
            if mode:
                eng.set_mlp(params)
            eng.set_graph(g)
            eng.run()
            runs = [eng.run() for _ in range(args.steps)]
            eng.close()
            res[mode] = (float(np.mean([x["t_total_ms"] for x in runs])), runs[-1])
        (pms, rp), (lms, rl) = res[0], res[1]
        per_seed.append({"seed": seed, "plain_ms": pms, "learned_ms": lms, "speedup_vs_plain": pms / lms,
                         "plain_q": rp["modularity"], "learned_q": rl["modularity"],
                         "dQ": rl["modularity"] - rp["modularity"], "learned_sweeps": rl["learned_sweeps"],
                         "certify_sweeps": rl["certify_sweeps"], "certify_empty": rl["certify_empty"],
                         "kept_fraction": (rl["learned_kept"] / rl["learned_frontier"]) if rl["learned_frontier"] else None,
                         "swept_edges": rl["swept_edges"], "plain_swept_edges": rp["swept_edges"],
                         "phase_ms": {"decide": rl["t_decide_ms"], "update": rl["t_update_ms"],
                                      "contract": rl["t_contract_ms"]},
                         "plain_phase_ms": {"decide": rp["t_decide_ms"], "update": rp["t_update_ms"],
                                            "contract": rp["t_contract_ms"]}})
        log(f"learned seed {seed}: {lms:.2f} ms vs plain {pms:.2f} ms ({pms / lms:.2f}x) "
            f"dQ={rl['modularity'] - rp['modularity']:+.6f}")
    sp = [x["speedup_vs_plain"] for x in per_seed]
    return {"tau": tau, "settings": settings, "seeds": per_seed, "speedup_vs_plain": float(np.mean(sp)),
            "ms_per_step": float(np.mean([x["learned_ms"] for x in per_seed])),
            "plain_ms": float(np.mean([x["plain_ms"] for x in per_seed])),
            "dQ": max((x["dQ"] for x in per_seed), key=abs)}


def batch_leg(gala, args):
    """Batched workload: 4,096 SBMs (vertex counts log-uniform in [64, 20,000],
    4..64 blocks, p_in in [0.05, 0.3], p_out = 0.1 p_in / blocks, unweighted)
    run as one union (hipGraph-replayed iterations, then kernel-by-kernel),
    against the same graphs run one at a time (per-graph launches).  Every
    graph's result is checked against the single-graph path for a sample."""
    import numpy as np
    t0 = time.time()
    P = gala.batch_sbm_params(args.batch_graphs, seed=0)
    vstart, u = gala.gen_sbm_batch(*P)
    log(f"batch: {args.batch_graphs} graphs, n={u.vertex_num} undirected_edges={u.edge_num} "
        f"gen {time.time() - t0:.1f}s")
    out = {"graphs": args.batch_graphs, "vertices": int(u.vertex_num), "undirected_edges": int(u.edge_num)}
    for name, hg in (("batched_hipgraph", True), ("batched_launches", False)):
        eng = gala.Engine(device=0, hip_graphs=hg)
        eng.batch_set(vstart, u)
        eng.batch_run()
        ts = []
        for _ in range(args.steps):
            t1 = time.perf_counter()
            r = eng.batch_run()
            ts.append(time.perf_counter() - t1)
        res = eng.batch_results()
        comm = eng.batch_communities() if hg else None
        eng.close()
        out[name] = {"wall_ms": 1e3 * float(np.mean(ts)), "device_ms": r["t_total_ms"],
                     "phase1_ms": r["t_phase1_ms"], "contract_ms": r["t_contract_ms"], "rounds": r["rounds"],
                     "iterations": r["iterations"], "round_open": r["round_open"]}
        log(f"{name}: {out[name]['wall_ms']:.1f} ms wall, {r['t_total_ms']:.1f} ms device, "
            f"rounds {r['rounds']} iterations {r['iterations']}")
        if hg:
            bres, bcomm = res, comm
    # per-graph launches: the single-graph path, graph by graph (upload excluded)
    off = u.offsets()
    eng = gala.Engine(device=0)
    tot = 0.0
    dev = 0.0
    rng = np.random.default_rng(1)
    sample = set(rng.choice(args.batch_graphs, size=min(64, args.batch_graphs), replace=False).tolist())
    same = True
    for i in range(args.batch_graphs):
        a, b = int(vstart[i]), int(vstart[i + 1])
        gi = gala.Graph(u.degrees[a:b] - off[a], u.neighbors[off[a]:off[b]] - a)
        if gi.edge_num == 0:
            continue
        eng.set_graph(gi)
        t1 = time.perf_counter()
        r = eng.run()
        tot += time.perf_counter() - t1
        dev += r["t_total_ms"]
        if i in sample:
            same &= r["modularity"] == bres["modularity"][i] and bool(
                np.array_equal(eng.communities(gi.vertex_num), bcomm[i]))
    eng.close()
    out["per_graph_launches"] = {"wall_ms": 1e3 * tot, "device_ms": dev}
    out["speedup_vs_per_graph"] = out["per_graph_launches"]["wall_ms"] / out["batched_hipgraph"]["wall_ms"]
    out["hipgraph_gain"] = out["batched_launches"]["wall_ms"] / out["batched_hipgraph"]["wall_ms"]
    out["parity_sample_vs_single_graph"] = bool(same)
    out["mean_q"] = float(np.mean(bres["modularity"]))
    log(f"per-graph launches: {1e3 * tot:.1f} ms wall ({dev:.1f} ms device); batched speedup "
        f"{out['speedup_vs_per_graph']:.2f}x; parity sample {same}")
    return out


def rmat_leg(gala, args):
    """Weighted R-MAT: scale S (2^S vertices), edge factor 16, a=0.57 b=c=0.19,
    fp32 weights uniform [0.5, 2.0], self-loops/duplicates merged; run exactly
    (weights scaled to integers, int64 sums, int128 gains).  Optionally checked
    against the exact oracle (identical partition and Q) and the FP64 oracle
    (Q within 1e-9)."""
    import numpy as np
    t0 = time.time()
    g = gala.gen_rmat(args.rmat_scale, 16, 0.57, 0.19, 0.19, 0, 0.5, 2.0)
    deg = np.diff(g.offsets())
    log(f"rmat: scale {args.rmat_scale} n={g.vertex_num} undirected_edges={g.edge_num} max_degree={deg.max()} "
        f"gen {time.time() - t0:.1f}s")
    eng = gala.Engine(device=0)
    eng.set_graph(g)
    eng.run()
    runs = [eng.run() for _ in range(args.steps)]
    comm = eng.communities(g.vertex_num)
    eng.close()
    r = runs[-1]
    ms = float(np.mean([x["t_total_ms"] for x in runs]))
    dm_gbs = r["decide_bytes"] / (r["t_decide_ms"] * 1e-3) / 1e9 if r["t_decide_ms"] > 0 else 0.0
    out = {"scale": args.rmat_scale, "vertices": int(g.vertex_num), "undirected_edges": int(g.edge_num),
           "max_degree": int(deg.max()), "hubs_over_1e5": int((deg > 100000).sum()), "ms_per_run": ms,
           "edge_teps": g.edge_num / (ms * 1e-3), "modularity": r["modularity"], "rounds": r["rounds"],
           "iterations": r["total_iterations"],
           "phase_ms": {"decide": r["t_decide_ms"], "update": r["t_update_ms"], "contract": r["t_contract_ms"]},
           "decide_GBps": dm_gbs, "wide": r["wide"], "weight_shift": r["weight_shift"],
           "dtype": "int64 sums / int128 gains (fp32 weights scaled exactly)", "oracle_dq": None,
           "contraction": contraction_roofline(r)}
    log(f"rmat: {ms:.1f} ms Q={r['modularity']:.9f} rounds {r['rounds']} decide {dm_gbs:.0f} GB/s")
    if args.rmat_oracle:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        oracle.set_threads(cpu_threads(args))
        t2 = time.perf_counter()
        W, _ = oracle.scale_f32(g.weights)
        ex = oracle.louvain(g.vertex_num, g.degrees, g.neighbors, W)
        out["exact_oracle_s"] = time.perf_counter() - t2
        out["exact_oracle_identical"] = bool(np.array_equal(ex["final_comm"], comm)) and ex["Q"] == r["modularity"]
        t2 = time.perf_counter()
        ref = oracle.louvain_f(g.vertex_num, g.degrees, g.neighbors, g.weights)
        out["oracle_dq"] = r["modularity"] - ref["Q"]
        out["oracle_s"] = time.perf_counter() - t2
        out["oracle_same_partition"] = bool(np.array_equal(ref["final_comm"], comm))
        log(f"rmat oracles: exact identical {out['exact_oracle_identical']} ({out['exact_oracle_s']:.1f} s); "
            f"FP64 dQ={out['oracle_dq']:.3e} ({out['oracle_s']:.1f} s)")
    return out


def write_edge_list(path, g, chunk=4_000_000):
    """The graph's undirected edges (u < v) as a SNAP-style text edge list:
    a '#' header line, then "u<TAB>v" per line."""
    import numpy as np
    off = np.concatenate([[0], g.degrees])
    src = np.repeat(np.arange(g.vertex_num, dtype=np.int64), np.diff(off))
    dst = g.neighbors.astype(np.int64)
    sel = src < dst
    u, v = src[sel], dst[sel]
    width = len(str(max(g.vertex_num - 1, 0)))
    ids = np.arange(g.vertex_num, dtype=np.int64)
    digits = ((ids[:, None] // (10 ** np.arange(width - 1, -1, -1, dtype=np.int64))[None, :]) % 10 + 48).astype(
        np.uint8)
    nd = 1 + sum((ids >= 10 ** k).astype(np.int64) for k in range(1, width))  # digits of each id
    keep_id = np.arange(width)[None, :] >= (width - nd)[:, None]
    with open(path, "wb") as f:
        f.write(b"# FromNodeId\tToNodeId\n")
        for s0 in range(0, len(u), chunk):
            uu, vv = u[s0:s0 + chunk], v[s0:s0 + chunk]
            n = len(uu)
            row = np.concatenate([digits[uu], np.full((n, 1), 9, np.uint8), digits[vv],
                                  np.full((n, 1), 10, np.uint8)], axis=1)
            keep = np.concatenate([keep_id[uu], np.ones((n, 1), bool), keep_id[vv], np.ones((n, 1), bool)], axis=1)
            f.write(row[keep].tobytes())
    return len(u)


def ingest_leg(gala, g, args):
    """Text edge-list ingest of the same graph (one line per undirected edge):
    gala_ingest_text_device (file read, copy to HBM, parse, renumber, sort,
    dedup, copy-out) against the host gala_ingest_text (same output, checked
    byte for byte) and the reference's own preprocess (graph.cpp:18-151, built
    from its sources into oracle/_ref/) on the first 2M lines."""
    import subprocess
    import tempfile
    import numpy as np
    tmp = tempfile.mkdtemp(prefix="gala_ingest_")
    path = os.path.join(tmp, "edges.txt")
    t0 = time.time()
    lines = write_edge_list(path, g)
    size = os.path.getsize(path)
    log(f"ingest: {lines} lines, {size / 1e6:.1f} MB text written in {time.time() - t0:.1f}s")
    eng = gala.Engine(device=0)
    runs = []
    for i in range(1 + max(args.steps, 1)):
        d = eng.ingest_text(path)
        runs.append(eng.ingest_times())
    t = runs[-1]
    dev_ms = float(np.mean([r["device_ms"] for r in runs[1:]]))
    wall_ms = float(np.mean([r["wall_ms"] for r in runs[1:]]))
    t2 = time.perf_counter()
    h = gala.Graph.from_text(path)
    host_s = time.perf_counter() - t2
    same = bool(np.array_equal(h.degrees, d.degrees) and np.array_equal(h.neighbors, d.neighbors))
    out = {"lines": int(lines), "text_bytes": int(size), "vertices": int(d.vertex_num),
           "entries": int(len(d.neighbors)), "wall_ms": wall_ms, "device_ms": dev_ms,
           "read_ms": t["read_ms"], "load_ms": t["load_ms"], "d2h_ms": t["d2h_ms"],
           "lines_per_s": lines / (wall_ms * 1e-3), "text_GBps_wall": size / (wall_ms * 1e-3) / 1e9,
           "text_GBps_device": size / (dev_ms * 1e-3) / 1e9, "host_ingest_s": host_s,
           "device_equals_host": same, "reference": None}
    ref = os.path.join(ROOT, "oracle", "_ref", "preprocess_ref")
    if os.path.exists(ref):
        sample = os.path.join(tmp, "sample.txt")
        with open(path, "rb") as fi, open(sample, "wb") as fo:
            for k, line in enumerate(fi):
                if k > 2_000_000:
                    break
                fo.write(line)
        t2 = time.perf_counter()
        subprocess.run([ref, "-f", sample], check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        ref_s = time.perf_counter() - t2
        ds = eng.ingest_text(sample)
        rb = gala.Graph.load_bin_graph(sample + ".bin")
        out["reference"] = {"kind": "reference", "sample": "first 2,000,000 edge lines", "seconds": ref_s,
                            "lines_per_s": 2_000_000 / ref_s, "cores": 1,
                            "device_equals_reference": bool(np.array_equal(rb.degrees, ds.degrees) and
                                                            np.array_equal(rb.neighbors, ds.neighbors))}
    eng.close()
    for f in os.listdir(tmp):
        os.remove(os.path.join(tmp, f))
    os.rmdir(tmp)
    log(f"ingest: device {wall_ms:.1f} ms wall ({dev_ms:.1f} ms device passes), host {host_s:.2f} s, "
        f"identical {same}; reference {out['reference']['seconds'] if out['reference'] else None}")
    return out


def copy_probe(mib):
    """A plain device-to-device copy of known size: FETCH_SIZE and WRITE_SIZE of
    its dispatches against mib MiB read and written give the factor that turns
    those counters into HBM bytes."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.int32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"copy probe: 5 copies of {mib} MiB")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=4_000_000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--serial-bins", action="store_true",
                    help="timed steps with every degree bin on one stream (profiling: each kernel's duration "
                         "is then its own; the default overlaps the long-row bins with the others)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--batch-graphs", type=int, default=4096, help="graphs of the batched leg (0: skip it)")
    ap.add_argument("--rmat-scale", type=int, default=24, help="scale of the weighted R-MAT leg (0: skip it)")
    ap.add_argument("--rmat-oracle", action="store_true", help="check the R-MAT leg against the FP64 oracle")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-frontier leg")
    ap.add_argument("--no-ingest", action="store_true", help="skip the text-ingest leg")
    ap.add_argument("--ingest-only", action="store_true",
                    help="only the text-ingest leg (prints its JSON object; for profiling the ingest kernels)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-MiB device buffer 5 times and exit (scales the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on exactly one GPU (asked for {args.gpus}); "
              "the multi-GPU design is described but not built", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
        return

    import numpy as np
    import gala

    t0 = time.time()
    g = gala.gen_chung_lu(args.n, 2.1, 17.35, 15000.0, args.seed)
    log(f"graph: n={g.vertex_num} undirected_edges={g.edge_num} gen {time.time() - t0:.1f}s")
    if args.ingest_only:
        print(json.dumps(ingest_leg(gala, g, args)))
        return

    eng = gala.Engine(device=0, serial_bins=args.serial_bins)
    eng.set_graph(g)
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync = lambda: None  # noqa: E731

    for i in range(args.warmup):
        r = eng.run()
        log(f"warmup {i}: {r['t_total_ms']:.2f} ms Q={r['modularity']:.6f}")
    sync()
    t1 = time.perf_counter()
    runs = []
    for i in range(args.steps):
        r = eng.run()
        runs.append(r)
        log(f"step {i}: {r['t_total_ms']:.2f} ms rounds={r['rounds']} iters={r['total_iterations']} "
            f"Q={r['modularity']:.6f}")
    sync()
    wall = time.perf_counter() - t1
    # the final partition stays in HBM after a run (as the input CSR does
    # before it); its readback is timed on its own, outside the steps
    t3 = time.perf_counter()
    comm = eng.communities(g.vertex_num)
    t_d2h_ms = 1e3 * (time.perf_counter() - t3)
    eng.close()
    learned = None if args.no_learned else learned_leg(gala, args)
    batched = batch_leg(gala, args) if args.batch_graphs > 0 else None
    rmat = rmat_leg(gala, args) if args.rmat_scale > 0 else None
    ingest = None if args.no_ingest else ingest_leg(gala, g, args)
    # one more run with per-bin HIP events (outside the timed steps)
    peng = gala.Engine(device=0, profile_bins=True)
    peng.set_graph(g)
    prof = peng.run()
    var = peng.get_variants()
    peng.close()
    bins = []
    for b in range(9):
        for kind, name, ms_b, by, ent, nl in (
                ("general", KERNEL_GENERAL[b][var[b]], prof["bin_decide_ms"][b], prof["bin_decide_bytes"][b],
                 prof["bin_swept_edges"][b], prof["bin_launches"][b]),
                ("first", KERNEL_FIRST[b][var[9 + b]], prof["bin_first_ms"][b], prof["bin_first_bytes"][b],
                 prof["bin_first_edges"][b], prof["bin_first_launches"][b])):
            bins.append({"bin": b, "degrees": BIN_DEG[b], "sweep": kind, "kernel": name, "bytes": int(by),
                         "entries": int(ent), "launches": int(nl), "ms": ms_b,
                         "avg_launch_us": 1e3 * ms_b / nl if nl else 0.0,
                         "GBps": (by / ms_b / 1e6) if ms_b > 0 else 0.0})
    # bin 8's general sweeps run the split sequence (plan, part, commit); its
    # part kernel does all of the bin's row work and is timed by its own pair
    # of events, so the bin's entry is that kernel's time (the sequence's,
    # with the launch gaps, is kept beside it)
    if prof.get("hub_part_launches", 0) > 0:
        for x in bins:
            if x["bin"] == 8 and x["sweep"] == "general":
                x["sequence_ms"] = x["ms"]
                x["sequence_kernels"] = x["kernel"]
                x["kernel"] = KERNEL_HUB_PART
                x["ms"] = prof["hub_part_ms"]
                x["launches"] = prof["hub_part_launches"]
                x["avg_launch_us"] = 1e3 * x["ms"] / x["launches"]
                x["GBps"] = x["bytes"] / x["ms"] / 1e6 if x["ms"] > 0 else 0.0
    # the dominant decide kernel: largest total time over one run
    dom = max(bins, key=lambda x: x["ms"])

    dev_ms = float(np.mean([r["t_total_ms"] for r in runs]))
    wall_ms = wall * 1e3 / args.steps
    ms = wall_ms  # the headline: host wall clock per step
    edges = g.edge_num
    last = runs[-1]
    # decide-and-move algorithmic bytes (per swept vertex of degree d): 16d + 40,
    # 12d + 40 with implicit unit weights; the first sweep of a round needs no
    # prev[] gathers: 12d + 40 (8d + 40 unit).  Summed by the engine per sweep.
    dm_bytes = last["decide_bytes"]
    dm_s = last["t_decide_ms"] * 1e-3
    achieved = dm_bytes / dm_s / 1e9 if dm_s > 0 else 0.0

    cpu = None
    parity = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        th = cpu_threads(args)
        oracle.set_threads(th)
        t2 = time.perf_counter()
        ref = oracle.louvain(g.vertex_num, g.degrees, g.neighbors)
        cpu_s = time.perf_counter() - t2
        parity = bool(np.array_equal(ref["final_comm"], comm)) and ref["Q"] == last["modularity"]
        cpu = {"value": edges / cpu_s, "unit": "edge-TEPS", "cores": th, "nproc": os.cpu_count(), "kind": "port",
               "sample": f"full workload, one end-to-end run ({cpu_s:.1f} s, {th} OpenMP threads: the CPU share "
                         f"of this process, nproc {os.cpu_count()})"}
        log(f"cpu oracle: {cpu_s:.2f} s on {th} threads (nproc {os.cpu_count()}) Q={ref['Q']:.6f} parity={parity}")
        # one thread on the same full workload (the oracle's run is about as
        # long on one thread: most of its time is in sequential steps)
        oracle.set_threads(1)
        t2 = time.perf_counter()
        ref1 = oracle.louvain(g.vertex_num, g.degrees, g.neighbors)
        s1 = time.perf_counter() - t2
        cpu["single_thread"] = {"value": edges / s1, "unit": "edge-TEPS", "cores": 1, "seconds": s1,
                                "sample": "full workload, one end-to-end run on one thread",
                                "same_result": bool(np.array_equal(ref1["final_comm"], ref["final_comm"]))}
        log(f"cpu oracle: 1 thread {s1:.2f} s")

    out = {
        "metric": "end-to-end Louvain s + edge-TEPS, 4M/35M power-law; dQ learned vs plain",
        "value": edges / (ms * 1e-3),
        "unit": "edge-TEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": "plain GPU Louvain, Chung-Lu power-law graph",
                   "vertices": int(g.vertex_num), "undirected_edges": int(edges),
                   "gamma": 2.1, "mean_degree": 17.35, "degree_cap": 15000, "seed": args.seed,
                   "weights": "unit (int32 CSR)", "mode": "plain"},
        "roofline": {"bound": "hbm", "achieved": dom["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom["GBps"] / HBM_PEAK_GBS, "traffic": pmc_traffic(dom["kernel"]),
                     "traffic_source": "profiles/decide_traffic_chung_lu_4M.md (rocprofv3 --pmc FETCH_SIZE, "
                                       "WRITE_SIZE; copy-scaled)" if pmc_traffic(dom["kernel"]) else None,
                     "kernel": dom["kernel"], "bin": dom["bin"], "sweep": dom["sweep"],
                     "launches": dom["launches"], "bytes_per_launch": dom["bytes"] / max(dom["launches"], 1),
                     "avg_launch_us": dom["avg_launch_us"],
                     "decide_phase": {"achieved": achieved, "frac": achieved / HBM_PEAK_GBS,
                                      "algorithmic_bytes": int(dm_bytes), "time_ms": last["t_decide_ms"]}},
        "cpu_baseline": cpu,
        "seconds_e2e": ms * 1e-3,
        "wall_ms_per_step": wall_ms,
        "device_ms_per_step": dev_ms,
        "t_h2d_ms": last["t_h2d_ms"],
        "t_partition_d2h_ms": t_d2h_ms,
        "edge_teps_device": edges / (dev_ms * 1e-3),
        "contraction": contraction_roofline(last),
        "modularity": last["modularity"],
        "communities": last["n_communities"],
        "rounds": last["rounds"],
        "iterations": last["total_iterations"],
        "phase_ms": {"decide": last["t_decide_ms"], "update": last["t_update_ms"],
                     "contract": last["t_contract_ms"]},
        "decide_bins": bins,
        "rounds_detail": {"communities": last["round_nc"], "iterations": last["round_iters"],
                          "phase1_ms": last["round_ms"], "decide_ms": last["round_decide_ms"],
                          "update_ms": last["round_update_ms"], "contract_ms": last["round_build_ms"]},
        "parity_vs_oracle": parity,
        "learned": learned,
        "batched": batched,
        "weighted_rmat": rmat,
        "text_ingest": ingest,
        "dQ_learned_vs_plain": learned["dQ"] if learned else None,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
