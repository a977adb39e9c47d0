"""Learned schedule: per-level trace collection and training of the schedule
networks on the GPU (the library's own forward/backward/Adam kernels).

BFS.  A BFS visits the same vertex sets level by level whichever direction
computes each level, so one source run under each forced schedule (push,
pull, bitmap) yields, for every level, the device time of every mode at
identical frontier features.  The labels are the cheapest mode path through
the levels (dynamic programming over those times plus the cost of changing
frontier representation).  The "previous mode" features of a sample are
those of the label path (the mode the network would have picked one level
earlier), level 0 starting from push as the library does.

SSSP.  The network picks the bucket width class (delta = base *
2^(class-1)) at every threshold advance.  Runs under different widths are not
phase-aligned, but they are threshold-aligned: after a bucket with threshold
t every vertex closer than t is final, whatever widths came before.  Each
source is run at the four classes; every decision of the base-width run is
labelled with the class whose run crossed the distance window up to the
next decision fastest.

Outputs never depend on the networks: every schedule computes the same
depths and distances, which the tests check bit for bit.

Command line (runs on a GPU):
    python -m graphhet_amd.learn --out graphhet_amd/sched_net.txt
trains both networks on R-MAT graphs of scale 18-23 with a in [0.45, 0.65]
(b = c = (1-a)*0.19/0.43, d = the rest) and edge factor in [8, 32].  The
range reaches past 2^22.4 vertices, where a bottom-up sweep of the tail
starts to cost more than a top-down level on MI355X, so that both sides of
that crossover are in the data; scale 24 stays held out.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import (DeviceGraph, HostGraph, MLP_IN, MLP_OUT, NET_BFS, NET_SSSP, SCHED_BITMAP, SCHED_HEURISTIC,
               SCHED_LEARNED, SCHED_PULL, SCHED_PUSH, mlp_forward, mlp_init, mlp_train_step, sched_save,
               sched_set_params)

FORCED = ((0, SCHED_PUSH), (1, SCHED_PULL), (2, SCHED_BITMAP))
PREV_MODE = slice(9, 12)  # one-hot of the previous level's mode in the feature vector


def rmat_params(a: float) -> tuple[float, float, float]:
    """(a, b, c) with b = c = (1-a)*0.19/0.43 (d keeps the rest)."""
    bc = (1.0 - a) * 0.19 / 0.43
    return a, bc, bc


def pick_sources(hg, k: int, seed: int) -> np.ndarray:
    """k distinct sources among vertices with at least one out-edge (hg: a
    HostGraph or an out-CSR)."""
    o = hg.view(0, copy=False) if isinstance(hg, HostGraph) else hg
    cand = np.nonzero(o.degrees() > 0)[0]
    rng = np.random.default_rng(seed)
    return rng.choice(cand, size=min(k, len(cand)), replace=False).astype(np.int32)


HEUR = 13  # BFS feature slot holding the heuristic rule's pick (push 0, bitmap 0.5, pull 1)


def heuristic_pick(f: np.ndarray, prev: int) -> int:
    """The heuristic rule (sched.cpp heuristic_bfs) evaluated on a level's
    features with `prev` as the previous mode: frontier/n = f[4], frontier
    out-edges/m = f[5], unvisited in-edges/m = f[7], growth sign = sign of
    f[12], level > 0 when f[8] > 0 (graphs with an in-CSR, m_in = m)."""
    alpha, beta = 14.0, 24.0
    small = f[4] * 64.0 < 1.0
    if prev == 1 and f[8] > 0:
        if f[4] < 1.0 / beta and f[12] < 0:
            return 0 if small else 2
        return 1
    if f[5] > f[7] / alpha:
        return 1
    return 0 if small else 2


def _trace_arrays(trace):
    f = np.array([list(t.features) for t in trace], np.float32).reshape(-1, MLP_IN)
    ms = np.array([t.ms for t in trace], np.float64)
    return f, ms


def switch_ms(n: int) -> dict:
    """Extra device time (ms) of changing frontier representation between
    levels on an n-vertex graph: a push level after a bottom-up or bitmap
    level first turns the bitmap into a queue (`k_bitmap_to_queue`, a pass
    over n bits: 9.9 us at 2^24 vertices on MI355X); a bottom-up or bitmap
    level after a push level needs the frontier as a bitmap, which a large
    push level writes as it goes and a small one scatters (a few us)."""
    s = n / float(1 << 20)
    b2q = 0.0025 + 0.00046 * s
    q2b = 0.004
    return {(0, 1): q2b, (0, 2): q2b, (1, 0): b2q, (2, 0): b2q}


def label_path(ms: np.ndarray, n: int) -> np.ndarray:
    """Cheapest mode sequence over the levels (dynamic programming over the
    per-level mode times [L,3] plus representation-switch costs); level 0
    starts from a queue holding the source."""
    sw = switch_ms(n)
    L = ms.shape[0]
    cost = ms[0] + np.array([sw.get((0, m), 0.0) for m in range(3)])
    back = np.zeros((L, 3), np.int64)
    for lv in range(1, L):
        new = np.empty(3)
        for m in range(3):
            c = [cost[p] + sw.get((p, m), 0.0) for p in range(3)]
            back[lv, m] = int(np.argmin(c))
            new[m] = c[back[lv, m]] + ms[lv, m]
        cost = new
    path = np.empty(L, np.int64)
    path[-1] = int(np.argmin(cost))
    for lv in range(L - 1, 0, -1):
        path[lv - 1] = back[lv, path[lv]]
    return path


def regret_weights(T: np.ndarray) -> np.ndarray:
    """Sampling weight of a level: log of its worst-to-best mode time ratio
    (a wrong pick at level 0 can cost 1000x), plus a floor."""
    return np.log(T.max(axis=1) / np.maximum(T.min(axis=1), 1e-6)) + 0.05


def bfs_samples(dg: DeviceGraph, sources, reps: int = 3):
    """Features [N,16] and fastest-mode labels [N] from forced-mode runs;
    also the per-level mode times [N,3]."""
    X, Y, T = [], [], []
    for src in sources:
        src = int(src)
        times, feats = [], None
        for mode, sched in FORCED:
            best = None
            for _ in range(reps):
                _, st, tr = dg.bfs(src, schedule=sched, out=False, trace_cap=256)
                f, ms = _trace_arrays(tr)
                best = ms if best is None else np.minimum(best, ms)
                if mode == 0:
                    feats = f
            times.append(best)
        L = min(len(t) for t in times)
        ms = np.stack([t[:L] for t in times], axis=1)
        lab = label_path(ms, dg.n)
        f = feats[:L].copy()
        prev = np.concatenate([[0], lab[:-1]])
        f[:, PREV_MODE] = 0.0
        f[np.arange(L), 9 + prev] = 1.0
        for lv in range(L):  # the heuristic's pick after the label path's previous mode
            h = heuristic_pick(f[lv], int(prev[lv]))
            f[lv, HEUR] = (0.0, 1.0, 0.5)[h]
        X.append(f)
        Y.append(lab)
        T.append(ms)
    return np.concatenate(X), np.concatenate(Y).astype(np.int32), np.concatenate(T)


def _curve_at(th: np.ndarray, ms: np.ndarray, total: float, x: float) -> float:
    """Device time at which a run's threshold reached x: linear between its
    decision records (threshold, ms so far); the run's total past the last
    (x = inf: the end of the run)."""
    if len(th) == 0 or x > th[-1]:
        return float(total)
    return float(np.interp(x, th, ms))


def sssp_samples(dg: DeviceGraph, sources, tie: float = 1.03, trace_cap: int = 4096):
    """Per-decision samples: the features of every bucket-width decision of
    the base-width run, labelled with the width class that crosses the next
    distance window fastest.

    After a bucket with threshold t, every vertex closer than t is final
    whatever the widths used before, so the runs at the four fixed widths can
    be aligned on the threshold axis: each run's (threshold, device time so
    far) records give its time to cross [t_k, t_k + 4 base), the widest
    class's bucket from the base run's decision k.  (A narrower window would
    credit a wide class with a share of one large bucket it must process
    whole.)  The base class wins ties within `tie` (timing noise).  The
    previous-class features follow the label path, as the network's own
    earlier picks would.  A sample's weight is the device time at stake in
    its window (largest minus smallest class cost), not a ratio: picking the
    width of a window that holds most of the run's work matters more than a
    short one."""
    X, Y, W = [], [], []
    for src in sources:
        src = int(src)
        _, st0 = dg.sssp(src, out=False)
        base = st0.extra[0]
        curves, feats = [], None
        for cls in range(MLP_OUT):
            best = None
            for _ in range(2):
                _, st, tr = dg.sssp(src, out=False, delta=base * 2.0 ** (cls - 1), trace_cap=trace_cap)
                th = np.array([t.threshold for t in tr], np.float64)
                ms = np.array([t.ms for t in tr], np.float64)
                if best is None or st.device_ms < best[2]:
                    best = (th, ms, float(st.device_ms), tr)
            curves.append(best[:3])
            if cls == 1:
                feats, _ = _trace_arrays(best[3])
        th_b = curves[1][0]
        span = base * 2.0 ** (MLP_OUT - 2)  # the widest class's bucket
        prev = 1
        for k in range(len(th_b)):
            a = th_b[k]
            b = a + span
            cost = np.array([_curve_at(*c, b) - _curve_at(*c, a) for c in curves])
            cost = np.maximum(cost, 1e-4)
            lab = int(np.argmin(cost))
            if cost[1] <= tie * cost[lab]:
                lab = 1
            f = feats[k].copy()
            f[9:13] = 0.0
            f[9 + prev] = 1.0
            X.append(f)
            Y.append(lab)
            # the device time at stake in this window: the few windows that
            # hold most of a run's time decide its length
            W.append(cost.max() - cost.min() + 0.01 * cost.mean())
            prev = lab
    return np.array(X, np.float32).reshape(-1, MLP_IN), np.array(Y, np.int32), np.array(W, np.float64)


def train(kind: int, X: np.ndarray, y: np.ndarray, steps: int = 1500, lr: float = 2e-3, batch: int = 256,
          seed: int = 0, weights: np.ndarray | None = None) -> list[float]:
    """Fresh network, Adam on the GPU; minibatches drawn with probability
    proportional to `weights` (uniform if None).  Returns the loss history."""
    sched_set_params(kind, mlp_init(seed))
    rng = np.random.default_rng(seed)
    prob = None if weights is None else weights / weights.sum()
    hist = []
    for _ in range(steps):
        idx = rng.choice(len(X), size=min(batch, len(X)), p=prob)
        loss, _ = mlp_train_step(kind, X[idx], y[idx], lr)
        hist.append(loss)
    return hist


def accuracy(kind: int, X: np.ndarray, y: np.ndarray, classes: int) -> float:
    logits = mlp_forward(kind, X)[:, :classes]
    return float(np.mean(np.argmax(logits, axis=1) == y))


def compare_bfs(dg: DeviceGraph, sources) -> dict:
    """Learned vs heuristic vs reference-rule device time over the sources;
    depths must be identical."""
    from . import SCHED_REFERENCE
    tot = {"heuristic": 0.0, "learned": 0.0, "reference": 0.0}
    same = True
    for src in sources:
        src = int(src)
        d0, s0 = dg.bfs(src, schedule=SCHED_HEURISTIC)
        d1, s1 = dg.bfs(src, schedule=SCHED_LEARNED)
        d2, s2 = dg.bfs(src, schedule=SCHED_REFERENCE)
        same &= bool(np.array_equal(d0, d1) and np.array_equal(d0, d2))
        tot["heuristic"] += s0.device_ms
        tot["learned"] += s1.device_ms
        tot["reference"] += s2.device_ms
    tot["identical"] = same
    tot["gain_vs_heuristic"] = tot["heuristic"] / max(tot["learned"], 1e-9)
    tot["gain_vs_reference"] = tot["reference"] / max(tot["learned"], 1e-9)
    return tot


def compare_sssp(dg: DeviceGraph, sources) -> dict:
    tot = {"heuristic": 0.0, "learned": 0.0}
    same = True
    for src in sources:
        src = int(src)
        d0, s0 = dg.sssp(src, schedule=SCHED_HEURISTIC)
        d1, s1 = dg.sssp(src, schedule=SCHED_LEARNED)
        same &= bool(np.array_equal(d0, d1))
        tot["heuristic"] += s0.device_ms
        tot["learned"] += s1.device_ms
    tot["identical"] = same
    tot["gain_vs_heuristic"] = tot["heuristic"] / max(tot["learned"], 1e-9)
    return tot


def trace_set(scales, graphs_per_scale: int, sources: int, seed: int, log=print):
    """Traces from R-MAT graphs with a in [0.45, 0.65], ef in [8, 32]:
    BFS (features, labels, weights) and SSSP (features, labels, weights)."""
    rng = np.random.default_rng(seed)
    bx, by, bw, sx, sy, sw = [], [], [], [], [], []
    for sc in scales:
        for gi in range(graphs_per_scale):
            a = float(rng.uniform(0.45, 0.65))
            ef = float(rng.uniform(8, 32))
            gseed = int(rng.integers(1, 2**31))
            t0 = time.time()
            hg = HostGraph.rmat(sc, ef, *rmat_params(a), seed=gseed)
            dg = DeviceGraph.from_host(hg, with_all=False)
            srcs = pick_sources(hg, sources, gseed)
            X, y, T = bfs_samples(dg, srcs)
            bx.append(X)
            by.append(y)
            bw.append(regret_weights(T))
            X, y, w = sssp_samples(dg, srcs)
            sx.append(X)
            sy.append(y)
            sw.append(w)
            log(f"traces: scale {sc} a={a:.3f} ef={ef:.1f}: {len(by[-1])} BFS levels "
                f"(labels {np.bincount(by[-1], minlength=3).tolist()}), SSSP decisions by class "
                f"{np.bincount(sy[-1], minlength=MLP_OUT).tolist()} in {time.time() - t0:.1f}s")
            dg.close()
            hg.close()
    cat = np.concatenate
    return (cat(bx), cat(by), cat(bw)), (cat(sx), cat(sy), cat(sw))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", default=None, help="schedule file to write")
    ap.add_argument("--scales", default="18,19,20,21,22,23")
    ap.add_argument("--graphs", type=int, default=2, help="graphs per scale")
    ap.add_argument("--sources", type=int, default=8, help="sources per graph")
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--seed", type=int, default=7)
    args = ap.parse_args(argv)
    scales = [int(s) for s in args.scales.split(",")]
    (bx, by, bw), (sx, sy, sw) = trace_set(scales, args.graphs, args.sources, args.seed,
                                           log=lambda s: print(s, flush=True))
    hb = train(NET_BFS, bx, by, steps=args.steps, seed=args.seed, weights=bw)
    hs = train(NET_SSSP, sx, sy, steps=args.steps, seed=args.seed + 1, weights=sw)
    report = {
        "bfs_samples": int(len(by)), "bfs_label_counts": np.bincount(by, minlength=3).tolist(),
        "bfs_loss": [hb[0], hb[-1]], "bfs_train_accuracy": accuracy(NET_BFS, bx, by, 3),
        "sssp_samples": int(len(sy)), "sssp_label_counts": np.bincount(sy, minlength=4).tolist(),
        "sssp_loss": [hs[0], hs[-1]], "sssp_train_accuracy": accuracy(NET_SSSP, sx, sy, 4),
    }
    print(json.dumps(report), flush=True)
    if args.out:
        sched_save(args.out)
        print(f"wrote {args.out}", flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
