"""Benchmark of the MI355X CSR analytics path on a seeded R-MAT graph.

One "step" is one pass of the six algorithms over the resident graph: BFS and
SSSP from the max-out-degree vertex, PageRank (20 iterations, damping 0.85),
WCC, CDLP (10 iterations) and TC on the symmetrised graph.  The graph is built
and uploaded before the timed region; inputs are resident in HBM when it
starts, and results stay on the device (no D2H inside the timed region).

Prints ONE JSON line.  `value` is BFS GTEPS (m_r / t, m_r = out-edges of the
reached vertices), the headline of the metric; SSSP GTEPS and PageRank HBM
GB/s ride along.  `roofline` covers the PageRank iteration kernel, the path's
bandwidth-bound kernel, timed with HIP events on its own stream.
`cpu_baseline` times GraphHET's CPU BFS (oracle/_ref, or the oracle
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pr_iter_bytes(n: int, m: int) -> int:
    """Algorithmic HBM bytes of one PageRank iteration in this design:
    in-offsets 4(n+1) + in-col 4m + contribution read 4n (compulsory) +
    contribution write 4n + out-degree read 4n.  (The FP64 rank is written
    only on the last iteration.)"""
    return 4 * (n + 1) + 4 * m + 4 * n + 4 * n + 4 * n


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
    ap.add_argument("--algos", default="bfs,sssp,pr,wcc,cdlp,tc")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-s", type=float, default=20.0, help="budget for the CPU baseline")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="also run one plain device-to-device copy of this size (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe_mb:
        copy_probe(args.copy_probe_mb)

    if args.gpus != 1:
        log("bench.py measures one GPU: this build has no multi-GPU path "
            "(the sharding design is described in DESIGN.md); run with --gpus 1.")
        sys.exit(2)

    import graphhet_amd as G

    if G.device_count() < 1:
        log("no GPU visible")
        sys.exit(1)
    algos = set(args.algos.split(","))
    t0 = time.time()
    hg = G.HostGraph.rmat(args.scale, args.edge_factor, seed=args.seed)
    o = hg.view(G.VIEW_OUT, copy=False)
    i = hg.view(G.VIEW_IN, copy=False)
    a = hg.view(G.VIEW_ALL, copy=False) if "cdlp" in algos else None
    n, m = o.n, o.m
    log(f"R-MAT scale {args.scale}: n={n} m={m} built in {time.time() - t0:.1f}s")
    t0 = time.time()
    dg = G.DeviceGraph(o, i, a, with_all="cdlp" in algos)
    sg = None
    if "tc" in algos:
        s = hg.view(G.VIEW_SYM, copy=False)
        sg = G.DeviceGraph(s, undirected=True, weights=False)
    log(f"uploaded in {time.time() - t0:.1f}s ({dg.device_bytes() / 1e9:.2f} GB)")
    src = int(np.argmax(o.degrees()))

    def step():
        r = {}
        if "bfs" in algos:
            _, r["bfs"] = dg.bfs(src, out=False)
        if "sssp" in algos:
            _, r["sssp"] = dg.sssp(src, out=False)
        if "pr" in algos:
            _, r["pr"] = dg.pr(args.pr_iters, 0.85, out=False)
        if "wcc" in algos:
            _, r["wcc"] = dg.wcc(out=False)
        if "cdlp" in algos:
            _, r["cdlp"] = dg.cdlp(args.cdlp_iters, out=False)
        if "tc" in algos:
            _, r["tc"] = sg.tc(out=False)
        return r

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
    bfs_gteps = med("bfs", lambda s: s.edges_traversed / (s.device_ms * 1e-3) / 1e9)
    sssp_gteps = med("sssp", lambda s: s.edges_traversed / (s.device_ms * 1e-3) / 1e9)
    roofline = None
    pr_gbs = None
    if "pr" in algos:
        k_ms = med("pr", lambda s: s.hot_kernel_ms / max(s.hot_kernel_launches, 1))
        bytes_it = pr_iter_bytes(n, m)
        pr_gbs = bytes_it / (k_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(pr_gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": round(pr_gbs / PEAK_HBM_GBS, 4), "traffic": None,
                    "kernel": "k_pr_iter", "kernel_avg_us": round(k_ms * 1e3, 2), "bytes_per_launch": bytes_it}

    bfs_stats = runs[-1].get("bfs")
    cpu = None
    if not args.no_cpu_baseline and "bfs" in algos:
        cpu = cpu_baseline(o, i, src, args, bfs_stats)

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
        "config": {"workload": f"rmat-s{args.scale}-ef{int(args.edge_factor)} six-algorithm pass",
                   "scale": args.scale, "edge_factor": args.edge_factor, "seed": args.seed, "n": n, "m": m,
                   "source": src, "pr_iters": args.pr_iters, "cdlp_iters": args.cdlp_iters,
                   "algos": sorted(algos)},
        "sssp_gteps": round(sssp_gteps, 3) if sssp_gteps else None,
        "pr_hbm_gbs": round(pr_gbs, 1) if pr_gbs else None,
        "per_algorithm_ms": {k: (round(v, 3) if v is not None else None) for k, v in per_alg_ms.items()},
        "bfs_levels": bfs_stats.levels if bfs_stats else None,
        "edges_traversed": bfs_stats.edges_traversed if bfs_stats else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


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


def cpu_baseline(o, i, src, args, gpu_stats):
    """GraphHET's CPU BFS (CPU_BFS.hpp:9-41) on this host, same graph and
    source; single-threaded as in the reference."""
    import oracle as O

    cores = 1
    try:
        if O.ref_available():
            t0 = time.time()
            rg = O.RefGraph(o.rowptr, o.col, o.w)
            log(f"cpu baseline: reference graph_structure built in {time.time() - t0:.1f}s")
            secs, depth = rg.bfs(src)
            rg.close()
            kind = "reference"
        else:
            t0 = time.perf_counter()
            depth = O.bfs(o.rowptr, o.col, src)
            secs = time.perf_counter() - t0
            kind = "port"
    except Exception as e:  # the baseline is reported, never required
        log(f"cpu baseline failed: {e}")
        return None
    reached = depth != 2**31 - 1
    m_r = int(np.diff(o.rowptr)[reached].sum())
    return {"value": round(m_r / secs / 1e9, 4), "unit": "GTEPS", "cores": cores, "kind": kind,
            "sample": f"one full BFS from vertex {src} over the whole graph ({m_r} edges), {secs:.2f}s",
            "seconds": round(secs, 3),
            "matches_gpu_reached": (gpu_stats is None or int(reached.sum()) == gpu_stats.vertices_reached)}


if __name__ == "__main__":
    main()
