"""Writes tests/golden/*.npz: inputs and GraphHET's own CPU outputs.

The outputs come from oracle/_ref/libghet_ref.so, i.e. the reference's CPU
algorithms (include/CPU_adj_list/algorithm/*.hpp) compiled unchanged from its
sources by oracle/Makefile.  Run in the build container (where the reference
sources exist):  python tests/golden/make_golden.py

Cases: the two sample graphs of the reference's examples, small R-MAT and ER
graphs, and shaped graphs (path, star, K_n, self-loops + isolated vertices).
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import graphhet_amd as G  # noqa: E402
import oracle as O  # noqa: E402


def edges_to_host(n, edges):
    src = [e[0] for e in edges]
    dst = [e[1] for e in edges]
    w = [e[2] if len(e) > 2 else 1.0 for e in edges]
    return G.HostGraph.from_edges(n, src, dst, w)


def cases():
    out = {}
    # src/GPU_csr/GPU_csr_example.cu:25-50 after its removals: one..five = 0..4,
    # "six" reuses the tombstoned id 5 of "R".
    out["gpu_example"] = (edges_to_host(6, [(0, 1, 0.8), (1, 2, 1.0), (0, 2, 1.0), (0, 3, 1.0), (3, 2, 1.0),
                                            (3, 4, 1.0), (0, 5, 1.0)]), 0)
    # src/CPU_adj_list/CPU_example.cpp:21-40 (Andy..Leo = 0..5, a self-loop on
    # Bob); its BFS source "one" is absent, operator[] maps it to id 0.
    out["cpu_example"] = (edges_to_host(6, [(0, 1), (1, 2), (1, 3), (0, 3), (1, 1), (0, 5), (5, 3), (3, 5)]), 0)
    for scale, seed in ((8, 1), (10, 2)):
        hg = G.HostGraph.rmat(scale, 16, seed=seed)
        o = hg.view(G.VIEW_OUT)
        out[f"rmat_s{scale}"] = (hg, int(np.argmax(o.degrees())))
    hg = G.HostGraph.er(1000, 8000, seed=3)
    out["er_1000"] = (hg, int(np.argmax(hg.view(G.VIEW_OUT).degrees())))
    n = 12
    out["path_12"] = (edges_to_host(n, [(i, i + 1, 0.5) for i in range(n - 1)]), 0)
    out["star_9"] = (edges_to_host(9, [(0, i, 1.0 / i) for i in range(1, 9)] + [(i, 0, 1.0) for i in range(1, 9)]), 3)
    out["k6"] = (edges_to_host(6, [(i, j, 1.0 + i + j) for i in range(6) for j in range(6) if i != j]), 0)
    rng = np.random.default_rng(7)
    e = [(int(a), int(b), float(rng.random()) + 0.01) for a, b in rng.integers(0, 40, size=(160, 2))]
    e += [(5, 5, 0.3), (7, 7, 0.2), (5, 5, 0.9)]  # self-loops and a duplicate (last weight wins)
    out["loops_isolated_50"] = (edges_to_host(50, e), 5)
    return out


def main():
    if not O.ref_available():
        O.build()
    assert O.ref_available(), "oracle/_ref not built: the reference sources are needed"
    for name, (hg, src) in cases().items():
        o, i, s = hg.view(G.VIEW_OUT), hg.view(G.VIEW_IN), hg.view(G.VIEW_SYM)
        res = dict(
            n=np.int64(o.n), src=np.int64(src),
            out_rowptr=o.rowptr, out_col=o.col, out_w=o.w,
            in_rowptr=i.rowptr, in_col=i.col, in_w=i.w,
            sym_rowptr=s.rowptr, sym_col=s.col,
            bfs=O.ref_bfs(o.rowptr, o.col, src),
            sssp=O.ref_sssp(o.rowptr, o.col, o.w, src),
            wcc=O.ref_wcc(o.rowptr, o.col, i.rowptr, i.col),
            pr=O.ref_pr(o.rowptr, o.col, i.rowptr, i.col, 0.85, 10),
            cdlp=O.ref_cdlp(o.rowptr, o.col, i.rowptr, i.col, 10),
            tc_sym=O.ref_tc(s.rowptr, s.col),
            tc_dir=O.ref_tc(o.rowptr, o.col),
        )
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **res)
        print(f"{name}: n={o.n} m={o.m}")


if __name__ == "__main__":
    main()
