"""Benchmark of the MI355X CSR analytics path on a seeded R-MAT graph.

One "step" is one pass of the metric's algorithms over the resident graph:
BFS and SSSP from the max-out-degree vertex and PageRank (20 iterations,
damping 0.85).  The graph is built and uploaded before the timed region;
inputs are resident in HBM when it starts and results stay on the device.
BFS and SSSP use the trained schedule networks (graphhet_amd/sched_net.txt:
BFS mode per level, SSSP bucket width per phase; trained on R-MAT scales
18-23, so this graph is held out) when they load, else the heuristic
(`--step-schedule heuristic` forces it).  The heuristic forms are timed
beside the step on the same source and reported under `plain_heuristic`;
the `learned` section compares the two over nine sources.

Prints ONE JSON line.  `value` is BFS GTEPS (m_r / t, m_r = out-edges of the
reached vertices), the headline of the metric; SSSP GTEPS and PageRank HBM
GB/s ride along.  `roofline` covers the PageRank iteration kernel, the path's
bandwidth-bound kernel and the largest share of a step, timed with HIP events
on its own stream.  Also reported, outside the timed step: WCC, CDLP and TC
on the same graph; the batched forms against the plain form on the same GPU
(256 BFS and 256 SSSP sources on a scale-22 R-MAT; many small graphs for WCC
and TC); and `cpu_baseline`, GraphHET's CPU BFS (oracle/_ref, or the oracle
restatement when _ref is absent) on the same graph on this host.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--scale 24]
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
INT_INF = 2**31 - 1


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pr_iter_bytes(n: int, m: int) -> int:
    """Algorithmic HBM bytes of one PageRank iteration in this design:
    in-offsets 4(n+1) + in-col 4m + contribution read 4n (compulsory) +
    contribution write 4n + out-degree read 4n.  (The FP64 rank is written
    only on the last iteration.)"""
    return 4 * (n + 1) + 4 * m + 4 * n + 4 * n + 4 * n


PR_SOURCES = ("pr.hip", "pr_bin.hip", "pr_common.hpp")


def pr_source_hash() -> str:
    """First 16 hex digits of the SHA-256 of the PageRank kernels' sources
    (both forms and their shared kernels, concatenated in a fixed order)."""
    import hashlib
    h = hashlib.sha256()
    for name in PR_SOURCES:
        with open(os.path.join(ROOT, "graphhet_amd", "csrc", name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def pr_stream_bytes(out_rowptr, n: int, m: int) -> int:
    """Bytes the binned form (pr_bin.hip) moves per iteration: 2 B source +
    4 B value written + 4 B value read + 2 B destination per edge, a 128 KB
    window of contribution words per gather item (each source window of 2^15
    ids cut into items of at most 2^17 edges), and per vertex the out-degree
    read and the next contribution written."""
    import numpy as np
    starts = np.minimum(np.arange(0, n + (1 << 15), 1 << 15), n)
    per_win = np.diff(np.asarray(out_rowptr, np.int64)[starts])
    items = int(((per_win + (1 << 17) - 1) >> 17).sum())
    return 12 * m + items * 4 * (1 << 15) + 8 * n


def measured_traffic(workload: str, seed: int):
    """L2-to-fabric bytes per PR iteration (both kernels of the default form)
    from the rocprofv3 PMC runs recorded under profiles/ for this workload,
    and a note.  None when not measured, or when the PR sources have changed
    since the runs (the file records the hash of the sources it was measured
    with)."""
    path = os.path.join(ROOT, "profiles", f"pr_iteration__{workload}.traffic.txt")
    if not os.path.exists(path):
        return None, "not measured for this workload"
    kv = {}
    with open(path) as f:
        for line in f:
            if line.startswith("#"):
                continue
            for tok in line.split():
                if "=" in tok:
                    k, v = tok.split("=", 1)
                    kv[k] = v
    if kv.get("workload") != workload or kv.get("seed") != str(seed):
        return None, "not measured for this workload"
    if kv.get("pr_src_sha") != pr_source_hash():
        return None, "stale: the PR sources changed since the PMC runs"
    return int(kv["traffic_bytes"]), ("L2-to-fabric bytes (2.0 x FETCH_SIZE + 1.0 x WRITE_SIZE, MALL hits included), "
                                      f"from {os.path.relpath(path, ROOT)}, measured with these PR sources")


def live_pr_traffic(args, kernels):
    """L2-to-fabric bytes per PR iteration, measured in this run: two
    rocprofv3 PMC passes (FETCH_SIZE, then WRITE_SIZE, one counter each)
    over a child run of this script that runs the PR step alone on the same
    graph (the step's iterations, whose last also writes the FP64 ranks, and
    one 1 GiB device copy).  Each counter is
    scaled by what it reads for that copy (FETCH_SIZE about half the true
    bytes, WRITE_SIZE all of them), and the per-dispatch averages of the
    iteration's kernels are summed.  Returns (bytes, details) or (None,
    reason).  A pass that exceeds its time limit is killed with its process
    group.  (rocprofv3 may exit non-zero after the program has finished and
    its files are whole: the counters are used if they are there.)"""
    import csv
    import re
    import shutil
    import signal
    import subprocess
    import tempfile
    exe = shutil.which("rocprofv3")
    if exe is None:
        return None, "rocprofv3 not found"
    child = [sys.executable, os.path.join(ROOT, "bench.py"), "--algos", "pr", "--steps", "1", "--warmup", "0",
             "--pr-iters", str(args.pr_iters), "--scale", str(args.scale), "--edge-factor", str(args.edge_factor),
             "--seed", str(args.seed), "--gen", args.gen, "--no-extras", "--no-batched", "--no-dynamic",
             "--no-learned", "--no-cpu-baseline", "--no-pmc", "--copy-probe-mb", "1024"]
    if args.graph_cache:
        child += ["--graph-cache", args.graph_cache]
    copy_kib = 1024 * 1024  # 1 GiB
    per = {}
    with tempfile.TemporaryDirectory(dir="/tmp") as td:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            d = os.path.join(td, counter)
            cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "run", "--"] + child
            t0 = time.time()
            log(f"PR traffic: {counter} pass (rocprofv3 over a child run of the PR step)")
            try:
                proc = subprocess.Popen(cmd, cwd="/tmp", env={**os.environ, "TMPDIR": "/tmp"},
                                        stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, start_new_session=True)
            except OSError as e:
                return None, f"rocprofv3 did not start: {e}"
            try:
                proc.wait(timeout=150)
            except subprocess.TimeoutExpired:
                os.killpg(proc.pid, signal.SIGKILL)
                proc.wait()
                return None, f"the {counter} pass exceeded 150 s"
            path = os.path.join(d, "run_counter_collection.csv")
            if not os.path.exists(path):
                return None, f"the {counter} pass wrote no counters (exit {proc.returncode})"
            vals, names = {}, {}
            with open(path) as f:
                for r in csv.DictReader(f):
                    dsp = int(r["Dispatch_Id"])
                    vals[dsp] = vals.get(dsp, 0.0) + float(r["Counter_Value"])
                    mt = re.search(r"(k_\w+)", r["Kernel_Name"])
                    names[dsp] = mt.group(1) if mt else r["Kernel_Name"]
            copies = [v for dsp, v in vals.items() if "copyBuffer" in names[dsp]]
            if not copies or max(copies) <= 0:
                return None, f"no calibration copy in the {counter} pass"
            factor = copy_kib / max(copies)
            avg = {}
            for k in kernels:
                xs = [v for dsp, v in vals.items() if names[dsp] == k]
                if not xs:
                    return None, f"no {k} dispatch in the {counter} pass"
                avg[k] = statistics.mean(xs)
            per[counter] = {"factor": round(factor, 3), "kib_per_dispatch": {k: round(v, 1) for k, v in avg.items()},
                            "pass_s": round(time.time() - t0, 1)}
            log(f"PR traffic: {counter} pass done in {time.time() - t0:.0f}s")
    total = sum(per[c]["factor"] * v * 1024 for c in per for v in per[c]["kib_per_dispatch"].values())
    return int(total), per


def algorithmic_bandwidth(n, m, runs, per_alg_ms, other, pr_roofline):
    """Each algorithm's algorithmic bytes (the per-unit figures the roofline
    counts), its device time in this run, and the rate against the 8 TB/s
    HBM peak.  BFS: 4 m_r + 8 n_r + 4 n + n/8; SSSP: 12 m_r + 8 n_r + 8 n;
    PR: per iteration, as the roofline counts it; WCC: one edge pass,
    4m + 4(n+1) + 8n, over the whole run; CDLP: per iteration, 4(2m) + 4(2m)
    + 4(n+1) + 4n over the mean iteration; TC: the compulsory 4 m_dag +
    4(n+1) (the intersection traffic is reported beside it).  A direction-
    optimising BFS reads fewer edges than 4 m_r counts, so its figure is an
    effective rate."""
    out = {}

    def put(name, nbytes, ms, unit):
        if ms:
            gbs = nbytes / (ms * 1e-3) / 1e9
            out[name] = {"bytes": int(nbytes), "ms": round(ms, 4), "gbs": round(gbs, 1),
                         "frac": round(gbs / PEAK_HBM_GBS, 4), "unit": unit}

    for alg, per_edge in (("bfs", 4), ("sssp", 12)):
        st = runs[-1].get(alg)
        if st is not None:
            m_r, n_r = st.edges_traversed, st.vertices_reached
            extra = n / 8 if alg == "bfs" else 4 * n
            put(alg, per_edge * m_r + 8 * n_r + 4 * n + extra, per_alg_ms.get(alg), "one call")
    if pr_roofline is not None:
        put("pr", pr_roofline["bytes_per_launch"], pr_roofline["kernel_avg_us"] * 1e-3, "one iteration")
    if "wcc_ms" in other:
        put("wcc", 4 * m + 4 * (n + 1) + 8 * n, other["wcc_ms"], "one edge pass over the whole call")
    if "cdlp_iter_ms" in other:
        put("cdlp", 16 * m + 4 * (n + 1) + 4 * n, other["cdlp_iter_ms"], "one iteration (mean of the call's)")
    if "tc_ms" in other:
        put("tc", 4 * other["tc_dag_edges"] + 4 * (n + 1), other["tc_ms"], "compulsory bytes of one call")
    return out


def gteps(edges, ms):
    return edges / (ms * 1e-3) / 1e9 if ms > 0 else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--edge-factor", type=float, default=16)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--pr-iters", type=int, default=20)
    ap.add_argument("--cdlp-iters", type=int, default=10)
    ap.add_argument("--algos", default="bfs,sssp,pr", help="algorithms of the timed step")
    ap.add_argument("--no-extras", action="store_true", help="skip WCC/CDLP/TC")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-vs-plain section")
    ap.add_argument("--batched-scale", type=int, default=22)
    ap.add_argument("--no-dynamic", action="store_true", help="skip the dynamic-graph update section")
    ap.add_argument("--batch-sources", type=int, default=256)
    ap.add_argument("--many-graphs", type=int, default=1024)
    ap.add_argument("--many-max-scale", type=int, default=18,
                    help="largest small graph is 2^this vertices (vertex counts log-uniform in [2^10, 2^this])")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-vs-heuristic section")
    ap.add_argument("--no-learned-section", action="store_true",
                    help="keep the learned schedule in the step but skip the learned-vs-heuristic section "
                         "(profiles of the step alone)")
    ap.add_argument("--step-schedule", choices=("heuristic", "learned"), default="learned",
                    help="schedule of the timed step's BFS and SSSP (learned falls back to the heuristic "
                         "when no network loads)")
    ap.add_argument("--sched-file", default=os.path.join(ROOT, "graphhet_amd", "sched_net.txt"),
                    help="trained schedule networks (python -m graphhet_amd.learn)")
    ap.add_argument("--gen", choices=("device", "host"), default="device",
                    help="R-MAT generation on the GPU (default) or on the host")
    ap.add_argument("--graph-cache", default=os.environ.get("GHET_GRAPH_CACHE", ""),
                    help="directory of binary CSR caches for the R-MAT graphs (default: $GHET_GRAPH_CACHE, "
                         "empty = always generate)")
    ap.add_argument("--no-pmc", action="store_true",
                    help="do not measure PR's L2-to-fabric traffic in this run (two rocprofv3 PMC passes over a "
                         "child run of the PR step); use the committed profiles/ figure instead")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="also run one plain device-to-device copy of this size (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        log("bench.py measures one GPU: the library has no multi-GPU path "
            "(graphs are not sharded across devices); run with --gpus 1.")
        sys.exit(2)
    if args.copy_probe_mb:
        copy_probe(args.copy_probe_mb)

    import graphhet_amd as G

    if G.device_count() < 1:
        log("no GPU visible")
        sys.exit(1)
    algos = set(args.algos.split(","))
    t0 = time.time()
    extras = not args.no_extras
    dg, o, hg = make_graph(G, args, args.scale, args.edge_factor, with_all=extras)
    n, m = o.n, o.m
    log(f"R-MAT scale {args.scale}: n={n} m={m} on the GPU in {time.time() - t0:.1f}s "
        f"({args.gen} generation, {dg.device_bytes() / 1e9:.2f} GB)")
    src = int(np.argmax(o.degrees()))
    # The step's BFS and SSSP run the learned schedule when the networks
    # load (trained on R-MAT scales 18-23; this graph is held out), and the
    # heuristic otherwise or with --step-schedule heuristic.  Outputs are
    # identical whichever schedule runs; the heuristic forms are timed beside
    # the step (plain_heuristic) and over nine sources (learned section).
    have_net = not args.no_learned and G.sched_load(args.sched_file)
    sched = G.SCHED_LEARNED if (args.step_schedule == "learned" and have_net) else G.SCHED_HEURISTIC
    sched_name = "learned BFS modes and SSSP widths" if sched == G.SCHED_LEARNED else "heuristic"

    def run(alg):
        if alg == "bfs":
            return dg.bfs(src, schedule=sched, out=False)[1]
        if alg == "sssp":
            return dg.sssp(src, schedule=sched, out=False)[1]
        return dg.pr(args.pr_iters, 0.85, out=False)[1]

    def step():
        return {alg: run(alg) for alg in ("bfs", "sssp", "pr") if alg in algos}

    # First call of each algorithm on the fresh graph, wall clock: includes
    # the one-time per-graph layouts (SSSP weight-ordered rows, PR tiles, BFS
    # first in-neighbours) that later calls reuse.
    first_call = {}
    for alg in ("bfs", "sssp", "pr"):
        if alg in algos:
            t1 = time.perf_counter()
            run(alg)
            first_call[alg] = round(1e3 * (time.perf_counter() - t1), 1)
    log(f"first calls (ms, layouts included): {first_call}")

    for w in range(args.warmup):
        step()
        log(f"warmup {w + 1}/{args.warmup} done")
    runs = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        runs.append(step())
        log(f"step {k + 1}/{args.steps} done")
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / max(args.steps, 1)

    def med(alg, f):
        return statistics.median(f(r[alg]) for r in runs) if alg in algos else None

    per_alg_ms = {alg: med(alg, lambda s: s.device_ms) for alg in sorted(algos)}
    bfs_gteps = med("bfs", lambda s: gteps(s.edges_traversed, s.device_ms))
    sssp_gteps = med("sssp", lambda s: gteps(s.edges_traversed, s.device_ms))
    roofline = None
    pr_gbs = None
    if "pr" in algos:
        k_ms = med("pr", lambda s: s.hot_kernel_ms / max(s.hot_kernel_launches, 1))
        bytes_it = pr_iter_bytes(n, m)
        pr_gbs = bytes_it / (k_ms * 1e-3) / 1e9
        traffic, traffic_note = measured_traffic(f"rmat-s{args.scale}-ef{int(args.edge_factor)}", args.seed)
        binned = runs[0]["pr"].extra[3] == 2
        roofline = {"bound": "hbm", "achieved": round(pr_gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": round(pr_gbs / PEAK_HBM_GBS, 4), "traffic": traffic, "traffic_note": traffic_note,
                    "kernel": "k_prb_gather+k_prb_accum" if binned else "k_pr_iter+k_pr_long",
                    "kernel_avg_us": round(k_ms * 1e3, 2), "bytes_per_launch": bytes_it}
        if binned:
            # the two launches of an iteration, and the bytes the form streams
            g_ms = med("pr", lambda s: s.extra[2] / max(s.hot_kernel_launches, 1))
            sb = pr_stream_bytes(o.rowptr, n, m)
            roofline["phases_us"] = {"gather": round(g_ms * 1e3, 2), "accumulate": round((k_ms - g_ms) * 1e3, 2)}
            roofline["stream_bytes"] = sb
            roofline["stream_gbs"] = round(sb / (k_ms * 1e-3) / 1e9, 1)
        # after the timed region: the pull-gather form on the same graph, its
        # phases timed apart, and the per-edge gather pattern's own rate over
        # its layout
        gf = dg.pr(args.pr_iters, 0.85, out=False, form="gather", phases=True)[1]
        hot_ms = gf.extra[2] / max(gf.hot_kernel_launches, 1)
        split_ms = gf.hot_kernel_ms / max(gf.hot_kernel_launches, 1)
        ceil = dg.pr_gather_ceiling(5)
        roofline["gather_form"] = {
            "iteration_us": round(split_ms * 1e3, 2),
            "phases_us": {"hot": round(hot_ms * 1e3, 2), "cold_and_finish": round((split_ms - hot_ms) * 1e3, 2),
                          "note": "timed as two launches per iteration"},
            "gather_ceiling_us": {"hot": round(ceil["hot_ms"] * 1e3, 2), "cold": round(ceil["cold_ms"] * 1e3, 2),
                                  "hot_edges": ceil["hot_edges"], "cold_edges": ceil["cold_edges"]},
            "frac_of_gather_ceiling": round((ceil["hot_ms"] + ceil["cold_ms"]) / split_ms, 4) if split_ms > 0 else None}

    plain = {}
    if sched != G.SCHED_HEURISTIC:
        for alg, fn in (("bfs", dg.bfs), ("sssp", dg.sssp)):
            if alg in algos:
                sts = [fn(src, schedule=G.SCHED_HEURISTIC, out=False)[1] for _ in range(3)]
                ms = statistics.median(x.device_ms for x in sts)
                plain[alg] = {"ms": round(ms, 3), "gteps": round(gteps(sts[0].edges_traversed, ms), 3)}
    learned = None
    if have_net and not args.no_learned_section:
        learned = learned_section(G, dg, o, src)

    other = {}
    want_cpu = not args.no_cpu_baseline and "bfs" in algos
    inn = None
    if want_cpu:
        inn = hg.view(G.VIEW_IN, copy=False) if hg is not None else dg.download(G.VIEW_IN)
    sym = None
    if extras:
        t1 = time.perf_counter()
        dg.wcc(out=False)
        first_call["wcc"] = round(1e3 * (time.perf_counter() - t1), 1)
        _, st = dg.wcc(out=False)
        other["wcc_ms"] = round(st.device_ms, 3)
        t1 = time.perf_counter()
        dg.cdlp(args.cdlp_iters, out=False)
        first_call["cdlp"] = round(1e3 * (time.perf_counter() - t1), 1)
        _, st = dg.cdlp(args.cdlp_iters, out=False)
        other["cdlp_ms"] = round(st.device_ms, 3)
        other["cdlp_iter_ms"] = round(st.hot_kernel_ms / max(args.cdlp_iters, 1), 3)
        del dg
        if hg is None:
            hg = G.HostGraph.from_csr(o)
        sym = hg.view(G.VIEW_SYM, copy=False)
        sg = G.DeviceGraph(sym, undirected=True, weights=False)
        t1 = time.perf_counter()
        sg.tc(out=False)
        first_call["tc"] = round(1e3 * (time.perf_counter() - t1), 1)
        _, st = sg.tc(out=False)
        other["tc_ms"] = round(st.device_ms, 3)
        other["tc_dag_edges"] = int(st.extra[1])
        other["tc_intersection_bytes"] = int(st.extra[2])
        other["tc_intersection_gbs"] = round(st.extra[2] / (st.device_ms * 1e-3) / 1e9, 1)
        del sg
        log("extras done")
    else:
        del dg

    bfs_stats = runs[-1].get("bfs")
    cpu = None
    if want_cpu:
        gpu_ms = dict(per_alg_ms)
        gpu_ms.update({k[:-3]: v for k, v in other.items() if k in ("wcc_ms", "cdlp_ms", "tc_ms")})
        cpu = cpu_baseline(o, inn, sym, src, bfs_stats, args, gpu_ms)
    del hg, o, inn, sym

    batched = None
    if not args.no_batched:
        batched = batched_section(G, args)
    dynamic = None
    if not args.no_dynamic:
        dynamic = dynamic_section(G, args)

    if roofline is not None and not args.no_pmc:
        # after every timed and reported section: PR's traffic measured now
        kernels = roofline["kernel"].split("+")
        live, detail = live_pr_traffic(args, kernels)
        if live is not None:
            roofline["traffic_committed"] = roofline["traffic"]
            roofline["traffic"] = live
            roofline["traffic_note"] = ("L2-to-fabric bytes per iteration measured in this run (MALL hits "
                                        "included): rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, over a child run "
                                        "of the PR step, each scaled by what it reads for a 1 GiB device copy in "
                                        "the same pass; the committed profiles/ figure is traffic_committed")
            roofline["traffic_pmc"] = detail
        else:
            roofline["traffic_note"] = f"not measured in this run ({detail}); " + roofline["traffic_note"]
        log("PR traffic measured")

    out = {
        "metric": "GTEPS (BFS/SSSP) + PR HBM GB/s, scale-24 R-MAT; learned/batched vs plain",
        "value": round(bfs_gteps, 3) if bfs_gteps else None,
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"rmat-s{args.scale}-ef{int(args.edge_factor)} BFS+SSSP+PR pass",
                   "scale": args.scale, "edge_factor": args.edge_factor, "seed": args.seed, "n": n, "m": m,
                   "source": src, "pr_iters": args.pr_iters, "step_algos": sorted(algos),
                   "schedule": sched_name},
        "plain_heuristic": plain or None,
        "sssp_gteps": round(sssp_gteps, 3) if sssp_gteps else None,
        "pr_hbm_gbs": round(pr_gbs, 1) if pr_gbs else None,
        "per_algorithm_ms": {k: (round(v, 3) if v is not None else None) for k, v in per_alg_ms.items()},
        "bfs_levels": bfs_stats.levels if bfs_stats else None,
        "edges_traversed": bfs_stats.edges_traversed if bfs_stats else None,
        "other_algorithms": other,
        "first_call_ms": first_call,
        "learned": learned,
        "batched": batched,
        "dynamic": dynamic,
        "roofline": roofline,
        "algorithmic_bandwidth": algorithmic_bandwidth(n, m, runs, per_alg_ms, other, roofline),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def dynamic_section(G, args):
    """Dynamic store on the batched-section graph: batches of new random edges
    inserted, then the same edges deleted (as GraphHET's GPMA benchmark does
    with 1000-edge batches), and a snapshot of the whole graph.  Update rates
    count distinct edge updates per second of device time (upload of the
    batch included)."""
    res = {"graph": f"rmat-s{args.batched_scale}-ef16"}
    dg = G.DeviceGraph.rmat(args.batched_scale, 16, seed=args.seed, with_all=False)
    dyn = G.DynGraph.from_graph(dg)
    del dg
    n = 1 << args.batched_scale
    rng = np.random.default_rng(args.seed + 7)
    for size in (1000, 65536, 1 << 20):
        s = rng.integers(0, n, size).astype(np.int32)
        d = rng.integers(0, n, size).astype(np.int32)
        w = rng.random(size) + 1e-3
        ins = dyn.update(s, d, w)
        dele = dyn.delete(s, d)
        res[f"batch_{size}"] = {"insert_ms": round(ins.device_ms, 3), "delete_ms": round(dele.device_ms, 3),
                                "inserted": int(ins.extra[0]), "deleted": int(dele.extra[1]),
                                "insert_mups": round(ins.edges_traversed / (ins.device_ms * 1e3), 2),
                                "delete_mups": round(dele.edges_traversed / (dele.device_ms * 1e3), 2)}
        log(f"dynamic batch {size}: {res[f'batch_{size}']}")
    t0 = time.perf_counter()
    g = dyn.snapshot(with_all=False)
    res["snapshot_ms"] = round(1e3 * (time.perf_counter() - t0), 1)
    res["edges"] = dyn.m
    del g, dyn
    return res


def make_graph(G, args, scale, edge_factor, with_all):
    """The R-MAT input as (device graph, host out-CSR, host graph or None).
    Device generation builds it in GPU memory (bit-identical to the host
    generator) and downloads the out-CSR for the host-side work; host
    generation, or a graph cache, builds and uploads it."""
    if args.gen == "device" and not args.graph_cache:
        dg = G.DeviceGraph.rmat(scale, edge_factor, seed=args.seed, with_all=with_all)
        return dg, dg.download(G.VIEW_OUT), None
    hg = G.HostGraph.rmat_cached(args.graph_cache, scale, edge_factor, seed=args.seed)
    o = hg.view(G.VIEW_OUT, copy=False)
    i = hg.view(G.VIEW_IN, copy=False)
    a = hg.view(G.VIEW_ALL, copy=False) if with_all else None
    return G.DeviceGraph(o, i, a, with_all=with_all), o, hg


def learned_section(G, dg, o, src):
    """Learned schedule vs the heuristic (and GraphHET's n/100 rule) on this
    graph, which the networks never saw (they are trained on R-MAT graphs of
    scales 18-23).  Outputs are compared bit for bit."""
    from graphhet_amd import learn
    srcs = np.concatenate([[src], learn.pick_sources(o, 8, 24)]).astype(np.int32)
    learn.compare_bfs(dg, srcs[:2])  # warm-up
    b = learn.compare_bfs(dg, srcs)
    s = learn.compare_sssp(dg, srcs[:4])
    log("learned section done")
    return {
        "sources": len(srcs),
        "bfs_ms": {k: round(b[k], 3) for k in ("heuristic", "learned", "reference")},
        "bfs_gain_vs_heuristic": round(b["gain_vs_heuristic"], 3),
        "bfs_gain_vs_reference_rule": round(b["gain_vs_reference"], 3),
        "sssp_ms": {k: round(s[k], 3) for k in ("heuristic", "learned")},
        "sssp_gain_vs_heuristic": round(s["gain_vs_heuristic"], 3),
        "identical": bool(b["identical"] and s["identical"]),
    }


def batched_section(G, args):
    """Batched forms vs the plain form on the same GPU (same inputs, results
    compared on a sample of instances)."""
    res = {}
    dg, o, hg = make_graph(G, args, args.batched_scale, 16, with_all=False)
    indeg = np.bincount(o.col, minlength=o.n)
    cand = np.nonzero((o.degrees() + indeg) > 0)[0]
    rng = np.random.default_rng(args.seed)
    srcs = rng.choice(cand, size=args.batch_sources, replace=False).astype(np.int32)
    # warm both forms (layouts, workspaces)
    dg.bfs_multi(srcs[:8], out=False)
    dg.bfs(int(srcs[0]), out=False)
    _, mst = dg.bfs_multi(srcs, out=False)
    plain_ms, plain_edges = 0.0, 0
    for s in srcs:
        _, st = dg.bfs(int(s), out=False)
        plain_ms += st.device_ms
        plain_edges += st.edges_traversed
    res["bfs"] = {"sources": len(srcs), "batched_ms": round(mst.device_ms, 3), "plain_ms": round(plain_ms, 3),
                  "batched_gteps": round(gteps(mst.edges_traversed, mst.device_ms), 3),
                  "plain_gteps": round(gteps(plain_edges, plain_ms), 3),
                  "speedup": round(plain_ms / mst.device_ms, 3), "edges_match": mst.edges_traversed == plain_edges}
    log(f"batched BFS: {res['bfs']}")
    dg.sssp_multi(srcs[:4], out=False)
    _, mst = dg.sssp_multi(srcs, out=False)
    plain_ms, plain_edges = 0.0, 0
    for s in srcs:
        _, st = dg.sssp(int(s), out=False)
        plain_ms += st.device_ms
        plain_edges += st.edges_traversed
    res["sssp"] = {"sources": len(srcs), "batched_ms": round(mst.device_ms, 3), "plain_ms": round(plain_ms, 3),
                   "batched_gteps": round(gteps(mst.edges_traversed, mst.device_ms), 3),
                   "plain_gteps": round(gteps(plain_edges, plain_ms), 3),
                   "speedup": round(plain_ms / mst.device_ms, 3), "edges_match": mst.edges_traversed == plain_edges}
    log(f"batched SSSP: {res['sssp']}")
    del dg, hg
    res["graph"] = f"rmat-s{args.batched_scale}-ef16"

    # many small graphs: n = 2^scale with scale uniform over the integers in
    # [10, max] (log-uniform over powers of two; R-MAT needs them), edge
    # factor uniform in [4, 32]; generated on the GPU and downloaded
    count = args.many_graphs
    t0 = time.time()
    graphs_dir, graphs_sym = [], []
    for k in range(count):
        scale = int(rng.integers(10, args.many_max_scale + 1))
        ef = float(rng.uniform(4, 32))
        dgk = G.DeviceGraph.rmat(scale, ef, seed=args.seed * 100003 + k, weights=False, with_in=False,
                                 with_all=False)
        od = dgk.download(G.VIEW_OUT, weights=False)
        dgk.close()
        graphs_dir.append(od)
        graphs_sym.append(G.HostGraph.from_csr(od).view(G.VIEW_SYM))
    log(f"{count} small graphs built in {time.time() - t0:.1f}s")
    G.wcc_many(graphs_dir[:4])
    labs, wst = G.wcc_many(graphs_dir)
    cnts, tst = G.tc_many(graphs_sym)
    plain_w = plain_t = 0.0
    ok = True
    for k, (gd, gs) in enumerate(zip(graphs_dir, graphs_sym)):
        d = G.DeviceGraph(gd, with_in=False, weights=False)
        lab, st = d.wcc()
        plain_w += st.device_ms
        ok &= bool(np.array_equal(lab, labs[k]))
        del d
        u = G.DeviceGraph(gs, undirected=True, weights=False)
        c, st = u.tc()
        plain_t += st.device_ms
        ok &= bool(np.array_equal(c, cnts[k]))
        del u
    res["many_graphs"] = {
        "count": count, "vertices": int(sum(g.n for g in graphs_dir)), "edges": int(sum(g.m for g in graphs_dir)),
        "scale_range": [10, args.many_max_scale],
        "wcc_batched_ms": round(wst.device_ms, 3), "wcc_plain_ms": round(plain_w, 3),
        "wcc_speedup": round(plain_w / wst.device_ms, 3),
        "tc_batched_ms": round(tst.device_ms, 3), "tc_plain_ms": round(plain_t, 3),
        "tc_speedup": round(plain_t / tst.device_ms, 3), "results_match": ok}
    log(f"many graphs: {res['many_graphs']}")
    return res


def copy_probe(mb: int) -> None:
    """One hipMemcpy device-to-device of `mb` MiB (the runtime's copy kernel):
    a plain copy of known size, to scale FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    nbytes = mb << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    for _ in range(2):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy probe: 2 x {mb} MiB device-to-device")


def cpu_baseline(o, inn, sym, src, gpu_stats, args, gpu_ms):
    """GraphHET's own CPU algorithms (oracle/_ref, compiled from the
    reference headers; the oracle restatement when _ref is absent) on this
    host, on the same graph, each on a bounded sample of the workload:
    BFS and SSSP from the step's source over the whole graph (1 thread, as
    in the reference), PageRank for 2 iterations and CDLP for 1 (the
    reference's 100-thread pool), WCC to completion (hardware_concurrency()
    threads), and TC on a seeded vertex sample of the symmetrised graph
    (1 thread; the full time is extrapolated by merge steps, whose total
    has the closed form 2 * sum(deg^2) as an upper bound).  The graph
    construction is outside every timing, as in the reference
    (LDBC_GPU_csr.cu:86-90).  The headline entry is BFS; the others are
    under per_algorithm, each with the GPU's speed-up on the same work."""
    import oracle as O

    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    per = {}
    try:
        kind = "reference" if O.ref_available() else "port"
        t0 = time.time()
        if kind == "reference":
            rg = O.RefGraph(o.rowptr, o.col, o.w, inn.rowptr, inn.col, inn.w)
            log(f"cpu baseline: reference graph_structure built in {time.time() - t0:.1f}s")
        deg = np.diff(o.rowptr)

        def timed(fn_ref, fn_port):
            if kind == "reference":
                return fn_ref()
            t1 = time.perf_counter()
            r = fn_port()
            return time.perf_counter() - t1, r

        secs, depth = timed(lambda: rg.bfs(src), lambda: O.bfs(o.rowptr, o.col, src))
        reached = depth != INT_INF
        m_r = int(deg[reached].sum())
        per["bfs"] = {"value": round(m_r / secs / 1e9, 4), "unit": "GTEPS", "cores": 1, "kind": kind,
                      "sample": f"one full BFS from vertex {src} ({m_r} edges)", "seconds": round(secs, 3),
                      "matches_gpu_reached": (gpu_stats is None or int(reached.sum()) == gpu_stats.vertices_reached)}
        log(f"cpu baseline bfs: {secs:.2f}s")
        secs, dist = timed(lambda: rg.sssp(src), lambda: O.sssp(o.rowptr, o.col, o.w, src))
        m_r = int(deg[dist != np.finfo(np.float64).max].sum())
        per["sssp"] = {"value": round(m_r / secs / 1e9, 4), "unit": "GTEPS", "cores": 1, "kind": kind,
                       "sample": f"one full SSSP (binary-heap Dijkstra) from vertex {src} ({m_r} edges)",
                       "seconds": round(secs, 3)}
        log(f"cpu baseline sssp: {secs:.2f}s")
        del depth, dist
        it = 2
        secs, _ = timed(lambda: rg.pr(0.85, it), lambda: O.pr(inn.rowptr, inn.col, o.rowptr, 0.85, it, 100))
        per["pr"] = {"value": round(pr_iter_bytes(o.n, o.m) * it / secs / 1e9, 2), "unit": "GB/s",
                     "iteration_ms": round(1e3 * secs / it, 1), "cores": cores, "threads": 100, "kind": kind,
                     "sample": f"{it} iterations over the whole graph (bytes as the GPU's roofline counts them)",
                     "seconds": round(secs, 3)}
        log(f"cpu baseline pr: {secs:.2f}s")
        secs, _ = timed(lambda: rg.wcc(), lambda: O.wcc(o.rowptr, o.col, 0))
        per["wcc"] = {"value": round(1e3 * secs, 1), "unit": "ms", "cores": cores, "threads": os.cpu_count(),
                      "kind": kind, "sample": "the whole graph to the fixpoint", "seconds": round(secs, 3)}
        log(f"cpu baseline wcc: {secs:.2f}s")
        secs, _ = timed(lambda: rg.cdlp(1), lambda: O.cdlp(inn.rowptr, inn.col, o.rowptr, o.col, 1, 100))
        per["cdlp"] = {"value": round(1e3 * secs, 1), "unit": "ms per iteration", "cores": cores, "threads": 100,
                       "kind": kind, "sample": "1 iteration over the whole graph", "seconds": round(secs, 3)}
        log(f"cpu baseline cdlp: {secs:.2f}s")
        if kind == "reference":
            rg.close()
        if sym is not None:
            sdeg = np.diff(sym.rowptr).astype(np.int64)
            bound = np.zeros(sym.n, np.int64)  # per vertex: sum over v in N(u) of deg u + deg v
            np.add.at(bound, np.repeat(np.arange(sym.n), sdeg), sdeg[sym.col])
            bound += sdeg * sdeg
            rng = np.random.default_rng(args.seed)
            cand = np.nonzero(sdeg >= 2)[0]
            order = rng.permutation(cand)
            secs = 0.0
            taken = 0
            k = 1024
            while secs < 10.0 and taken < len(order):
                batch = np.sort(order[taken:taken + k]).astype(np.int32)
                t1 = time.perf_counter()
                O.tc_vertices(sym.rowptr, sym.col, batch)
                secs += time.perf_counter() - t1
                taken += len(batch)
                k *= 2
            sample_bound = int(bound[order[:taken]].sum())
            est = secs * float(bound.sum()) / max(sample_bound, 1)
            per["tc"] = {"value": round(est, 1), "unit": "s (estimated, whole graph)", "cores": 1, "kind": "port",
                         "sample": f"{taken} seeded random vertices of degree >= 2 ({secs:.1f}s), extrapolated by "
                                   "the merge-step bound 2*sum(deg^2)", "seconds": round(secs, 3)}
            log(f"cpu baseline tc: sample of {taken} vertices in {secs:.2f}s, estimate {est:.0f}s")
    except Exception as e:  # the baseline is reported, never required
        log(f"cpu baseline failed: {e}")
        if "bfs" not in per:
            return None
    for alg, entry in per.items():
        g = gpu_ms.get(alg)
        if not g:
            continue
        if alg in ("bfs", "sssp"):
            entry["gpu_speedup"] = round(entry["seconds"] * 1e3 / g, 1)
        elif alg == "pr":
            entry["gpu_speedup"] = round(entry["iteration_ms"] / (g / args.pr_iters), 1)
        elif alg == "cdlp":
            entry["gpu_speedup"] = round(entry["value"] / (g / args.cdlp_iters), 1)
        elif alg == "wcc":
            entry["gpu_speedup"] = round(entry["value"] / g, 1)
        elif alg == "tc":
            entry["gpu_speedup"] = round(entry["value"] * 1e3 / g, 1)
    head = dict(per["bfs"])
    head["per_algorithm"] = per
    return head


if __name__ == "__main__":
    main()
