"""LDBC host layer: the Graphalytics-style loader, the checkers and the
LDBC-style GPU driver (include/LDBC, graphhet_amd/drivers/LDBC_GPU_csr.cpp).

A small dataset is written in the Graphalytics format (.properties, .v, .e,
and <name>-<ALGO> baselines).  Vertex names are sparse strings, so the
name <-> id mapping is exercised.  The baselines come from the oracle:
    BFS   depth, 9223372036854775807 when unreached
    SSSP  distance with 17 significant digits, "infinity" when unreached
    WCC   component label = name of the smallest member
    PR    rank; CDLP label name
The CPU test runs ldbc_selftest: the loader's parsed configuration, and
every checker accepting the baseline's own values and rejecting a copy with
one vertex changed.  The GPU test runs the driver end to end and expects
every algorithm's cell to be a time (the checkers accepted the GPU results).
"""
import os
import subprocess

import numpy as np
import pytest

import graphhet_amd as G
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "graphhet_amd", "bin")
TEST_BIN = os.path.join(ROOT, "tests", "bin")
LLONG_MAX = 9223372036854775807
DBL_MAX = np.finfo(np.float64).max
INT_MAX = 2**31 - 1


def _build_bins():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "graphhet_amd", "drivers")], check=True,
                   stdout=subprocess.DEVNULL)


def write_dataset(d, name, directed=True, scale=9, seed=3):
    """Writes the dataset; returns (names, source index)."""
    hg = G.HostGraph.rmat(scale, 8, seed=seed)
    o = hg.view(G.VIEW_OUT)
    n = o.n
    names = [f"v{7 * i + 3}" for i in range(n)]
    src = int(np.argmax(o.degrees()))
    rows = np.repeat(np.arange(n), o.degrees())
    if directed:
        eu, ev, ew = rows, o.col, o.w
    else:  # each undirected edge once; the loader adds the reverse
        keep = rows < o.col
        eu, ev, ew = rows[keep], o.col[keep], o.w[keep]
    with open(os.path.join(d, f"{name}.v"), "w") as f:
        f.write("".join(f"{s}\n" for s in names))
    with open(os.path.join(d, f"{name}.e"), "w") as f:
        f.write("".join(f"{names[u]} {names[v]} {float(w)!r}\n" for u, v, w in zip(eu, ev, ew)))
    p = f"graph.{name}"
    with open(os.path.join(d, f"{name}.properties"), "w") as f:
        f.write(f"# test dataset\n{p}.vertex-file = {name}.v\n{p}.edge-file = {name}.e\n"
                f"{p}.meta.vertices = {n}\n{p}.meta.edges = {len(eu)}\n"
                f"{p}.directed = {'true' if directed else 'false'}\n{p}.edge-properties.names = weight\n"
                f"{p}.algorithms = bfs, cdlp, pr, sssp, wcc\n{p}.bfs.source-vertex = {names[src]}\n"
                f"{p}.cdlp.max-iterations = 10\n{p}.pr.damping-factor = 0.85\n{p}.pr.num-iterations = 10\n"
                f"{p}.sssp.weight-property = weight\n{p}.sssp.source-vertex = {names[src]}\n")
    # the graph as the loader builds it (undirected: both directions)
    if directed:
        gs = hg
    else:
        gs = G.HostGraph.from_edges(n, np.concatenate([eu, ev]), np.concatenate([ev, eu]), np.concatenate([ew, ew]))
    oo, ii = gs.view(G.VIEW_OUT), gs.view(G.VIEW_IN)
    depth = O.bfs(oo.rowptr, oo.col, src)
    dist = O.sssp(oo.rowptr, oo.col, oo.w, src)
    comp = O.wcc(oo.rowptr, oo.col)
    rank = O.pr(ii.rowptr, ii.col, oo.rowptr, 0.85, 10)
    lab = O.cdlp(ii.rowptr, ii.col, oo.rowptr, oo.col, 10)

    def put(algo, vals):
        with open(os.path.join(d, f"{name}-{algo}"), "w") as f:
            f.write("".join(f"{names[i]} {vals[i]}\n" for i in range(n)))

    put("BFS", [LLONG_MAX if x == INT_MAX else int(x) for x in depth])
    put("SSSP", ["infinity" if x == DBL_MAX else f"{x:.17g}" for x in dist])
    put("WCC", [names[c] for c in comp])
    put("PR", [f"{x:.17g}" for x in rank])
    put("CDLP", [names[c] for c in lab])
    return names, src, n, int(oo.m), bool(np.any(depth == INT_MAX))


@pytest.mark.parametrize("directed", [True, False])
def test_loader_and_checkers(tmp_path, directed):
    _build_bins()
    names, src, n, m, has_unreached = write_dataset(str(tmp_path), "g", directed)
    out = subprocess.run([os.path.join(TEST_BIN, "ldbc_selftest"), str(tmp_path), "g"], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    cfg = dict(kv.split("=") for kv in
               next(l for l in out.stdout.splitlines() if l.startswith("config ")).split()[1:])
    assert cfg["directed"] == ("1" if directed else "0")
    assert cfg["weight"] == "1"
    assert all(cfg[a] == "1" for a in ("bfs", "cdlp", "pr", "sssp", "wcc"))
    assert int(cfg["bfs_src"]) == src and int(cfg["sssp_src"]) == src
    assert cfg["cdlp_its"] == "10" and cfg["pr_its"] == "10" and float(cfg["damping"]) == 0.85
    assert int(cfg["V"]) == n and int(cfg["E"]) == m
    checks = {l.split()[1]: l for l in out.stdout.splitlines() if l.startswith("check ")}
    assert set(checks) == {"BFS", "SSSP", "PR", "CDLP", "WCC"}
    for algo, line in checks.items():
        assert "accept=1" in line and "reject=1" in line, line
    assert has_unreached or not directed  # the directed case exercises the unreached markers


@pytest.mark.gpu
@pytest.mark.parametrize("directed", [True, False])
def test_ldbc_driver_end_to_end(tmp_path, directed):
    exe = os.path.join(BIN, "LDBC_GPU_csr")
    assert os.path.exists(exe), "graphhet_amd/bin/LDBC_GPU_csr not built"
    write_dataset(str(tmp_path), "g", directed)
    out = subprocess.run([exe], input=f"{tmp_path}\ng\n", capture_output=True, text=True, timeout=120,
                         cwd=str(tmp_path))
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    row = lines[lines.index("Result: ") + 1].split(",")
    assert len(row) == 5, row
    for cell in row:  # BFS, SSSP, WCC, PR, CDLP: a time means the checker accepted
        float(cell)
    for msg in ("BFS results are correct!", "SSSP results are correct!", "WCC results are correct!",
                "PageRank results are correct!", "CDLP results are correct!"):
        assert msg in out.stdout
    with open(os.path.join(tmp_path, "result-gpu-csr.csv")) as f:
        assert f.read().strip().split(",") == row
