"""Pins the CPU oracle (oracle/ghet_oracle.cpp) before anything is checked
against it: golden vectors from GraphHET's own CPU code, independent library
routines (networkx / scipy), and closed forms chosen so that a plausible slip
(off-by-one depth, a dropped dangling term, a wrong tie rule, in- vs
out-lists swapped) fails at least one of them."""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as csg

import graphhet_amd as G
import oracle as O

INT_INF = 2**31 - 1
DBL_MAX = np.finfo(np.float64).max


def test_oracle_matches_reference_golden(golden):
    g = golden
    orp, ocol, ow = g["out_rowptr"], g["out_col"], g["out_w"]
    irp, icol = g["in_rowptr"], g["in_col"]
    src = int(g["src"])
    np.testing.assert_array_equal(O.bfs(orp, ocol, src), g["bfs"])
    np.testing.assert_array_equal(O.sssp(orp, ocol, ow, src), g["sssp"])
    np.testing.assert_array_equal(O.wcc(orp, ocol), g["wcc"])
    # same operation order as CPU_PageRank.hpp -> bit-identical
    np.testing.assert_array_equal(O.pr(irp, icol, orp, 0.85, 10), g["pr"])
    np.testing.assert_array_equal(O.cdlp(irp, icol, orp, ocol, 10), g["cdlp"])
    np.testing.assert_array_equal(O.tc(g["sym_rowptr"], g["sym_col"]), g["tc_sym"])
    np.testing.assert_array_equal(O.tc(orp, ocol), g["tc_dir"])


def _csr_matrix(rowptr, col, w=None):
    n = len(rowptr) - 1
    data = np.ones(len(col)) if w is None else w
    return sp.csr_matrix((data, col, rowptr), shape=(n, n))


def test_bfs_against_networkx_and_invariants(golden):
    nx = pytest.importorskip("networkx")
    g = golden
    orp, ocol, src = g["out_rowptr"], g["out_col"], int(g["src"])
    d = O.bfs(orp, ocol, src)
    G_ = nx.DiGraph()
    n = len(orp) - 1
    G_.add_nodes_from(range(n))
    for u in range(n):
        for e in range(orp[u], orp[u + 1]):
            G_.add_edge(u, int(ocol[e]))
    lens = nx.single_source_shortest_path_length(G_, src)
    ref = np.full(n, INT_INF, np.int32)
    for v, l in lens.items():
        ref[v] = l
    np.testing.assert_array_equal(d, ref)


def test_bfs_closed_forms():
    n = 31  # complete binary tree: depth(v) = floor(log2(v+1))
    src = [v for v in range(1, n)]
    par = [(v - 1) // 2 for v in range(1, n)]
    hg = G.HostGraph.from_edges(n, par, src)
    o = hg.view(G.VIEW_OUT)
    d = O.bfs(o.rowptr, o.col, 0)
    np.testing.assert_array_equal(d, [int(math.log2(v + 1)) for v in range(n)])
    # 5x7 grid, 4-neighbour, undirected: depth = manhattan distance from (0,0)
    R, Cc = 5, 7
    e = []
    for r in range(R):
        for c in range(Cc):
            v = r * Cc + c
            if c + 1 < Cc:
                e += [(v, v + 1), (v + 1, v)]
            if r + 1 < R:
                e += [(v, v + Cc), (v + Cc, v)]
    hg = G.HostGraph.from_edges(R * Cc, [a for a, _ in e], [b for _, b in e])
    o = hg.view(G.VIEW_OUT)
    d = O.bfs(o.rowptr, o.col, 0)
    np.testing.assert_array_equal(d, [r + c for r in range(R) for c in range(Cc)])


def test_sssp_against_scipy_and_invariants(golden):
    g = golden
    orp, ocol, ow, src = g["out_rowptr"], g["out_col"], g["out_w"], int(g["src"])
    d = O.sssp(orp, ocol, ow, src)
    ref = csg.dijkstra(_csr_matrix(orp, ocol, ow), directed=True, indices=src)
    ref = np.where(np.isinf(ref), DBL_MAX, ref)
    np.testing.assert_allclose(d, ref, rtol=1e-12, atol=0)
    # d(v) <= d(u) + w for every edge; tight edge for every reached v != src
    n = len(orp) - 1
    u = np.repeat(np.arange(n), np.diff(orp))
    reach = d[u] != DBL_MAX
    assert np.all(d[ocol[reach]] <= d[u[reach]] + ow[reach])


def test_wcc_against_scipy(golden):
    g = golden
    orp, ocol = g["out_rowptr"], g["out_col"]
    n = len(orp) - 1
    ncomp, lab = csg.connected_components(_csr_matrix(orp, ocol), directed=True, connection="weak")
    # closed form: label = smallest vertex id of the component
    mins = np.full(ncomp, n, np.int64)
    np.minimum.at(mins, lab, np.arange(n))
    np.testing.assert_array_equal(O.wcc(orp, ocol), mins[lab])


def test_pr_closed_forms():
    n = 17  # directed cycle: every rank stays 1/n
    hg = G.HostGraph.from_edges(n, list(range(n)), [(v + 1) % n for v in range(n)])
    o, i = hg.view(G.VIEW_OUT), hg.view(G.VIEW_IN)
    r = O.pr(i.rowptr, i.col, o.rowptr, 0.85, 20)
    np.testing.assert_allclose(r, 1.0 / n, rtol=1e-14)
    # two vertices, 0 -> 1 only (1 is a sink): closed form per iteration
    hg = G.HostGraph.from_edges(2, [0], [1])
    o, i = hg.view(G.VIEW_OUT), hg.view(G.VIEW_IN)
    d = 0.85
    r0 = np.array([0.5, 0.5])
    for it in range(1, 6):
        s = r0[1]
        base = s * d / 2 + (1 - d) / 2
        r0 = np.array([base, base + d * r0[0]])
        np.testing.assert_allclose(O.pr(i.rowptr, i.col, o.rowptr, d, it), r0, rtol=1e-15)


def test_pr_mass_is_conserved(golden):
    g = golden
    r = O.pr(g["in_rowptr"], g["in_col"], g["out_rowptr"], 0.85, 10)
    assert abs(r.sum() - 1.0) < 1e-12


def test_pr_step_is_one_iteration_of_pr(golden):
    """pr_step (used to check every iteration of a GPU run) composes to pr."""
    g = golden
    n = int(g["n"])
    r = np.full(n, 1.0 / n)
    for k in range(1, 11):
        r = O.pr_step(g["in_rowptr"], g["in_col"], g["out_rowptr"], r, 0.85)
        if k in (1, 10):
            np.testing.assert_array_equal(r, O.pr(g["in_rowptr"], g["in_col"], g["out_rowptr"], 0.85, k))
    np.testing.assert_array_equal(r, g["pr"])  # the reference's own 10-iteration output


def _cdlp_brute(n, orp, ocol, irp, icol, iters):
    lab = list(range(n))
    for _ in range(iters):
        new = lab[:]
        for v in range(n):
            nb = [lab[int(u)] for u in icol[irp[v]:irp[v + 1]]] + [lab[int(u)] for u in ocol[orp[v]:orp[v + 1]]]
            if not nb:
                continue
            cnt = {}
            for x in nb:
                cnt[x] = cnt.get(x, 0) + 1
            best = max(cnt.values())
            new[v] = min(x for x, c in cnt.items() if c == best)
        lab = new
    return np.array(lab, np.int32)


def test_cdlp_brute_force(golden):
    g = golden
    if int(g["n"]) > 300:
        pytest.skip("brute force only on small cases")
    n = int(g["n"])
    for iters in (1, 2, 10):
        np.testing.assert_array_equal(
            O.cdlp(g["in_rowptr"], g["in_col"], g["out_rowptr"], g["out_col"], iters),
            _cdlp_brute(n, g["out_rowptr"], g["out_col"], g["in_rowptr"], g["in_col"], iters))


def test_cdlp_special_cases():
    # K_5 (both directions): after one iteration every v != 0 takes 0, 0 takes 1
    n = 5
    e = [(i, j) for i in range(n) for j in range(n) if i != j]
    hg = G.HostGraph.from_edges(n, [a for a, _ in e], [b for _, b in e])
    o, i = hg.view(G.VIEW_OUT), hg.view(G.VIEW_IN)
    np.testing.assert_array_equal(O.cdlp(i.rowptr, i.col, o.rowptr, o.col, 1), [1, 0, 0, 0, 0])
    # isolated vertices keep their label
    hg = G.HostGraph.from_edges(4, [0], [1])
    o, i = hg.view(G.VIEW_OUT), hg.view(G.VIEW_IN)
    lab = O.cdlp(i.rowptr, i.col, o.rowptr, o.col, 3)
    assert lab[2] == 2 and lab[3] == 3


def test_tc_against_networkx(golden):
    nx = pytest.importorskip("networkx")
    g = golden
    rp, col = g["sym_rowptr"], g["sym_col"]
    n = len(rp) - 1
    U = nx.Graph()
    U.add_nodes_from(range(n))
    for u in range(n):
        for e in range(rp[u], rp[u + 1]):
            U.add_edge(u, int(col[e]))
    tri = nx.triangles(U)
    np.testing.assert_array_equal(O.tc(rp, col), [6 * tri[v] for v in range(n)])


def test_tc_complete_graph():
    for n in (3, 4, 7):
        e = [(i, j) for i in range(n) for j in range(n) if i != j]
        hg = G.HostGraph.from_edges(n, [a for a, _ in e], [b for _, b in e])
        s = hg.view(G.VIEW_SYM)
        np.testing.assert_array_equal(O.tc(s.rowptr, s.col), [6 * math.comb(n - 1, 2)] * n)


def test_planted_mistakes_are_caught(golden):
    """A perturbed restatement must fail the golden check: shows the fixtures
    discriminate (damping, iteration count, BFS source)."""
    g = golden
    irp, icol, orp = g["in_rowptr"], g["in_col"], g["out_rowptr"]
    if int(g["n"]) < 10:
        pytest.skip("too small to discriminate")
    assert not np.array_equal(O.pr(irp, icol, orp, 0.86, 10), g["pr"])
    assert not np.array_equal(O.pr(irp, icol, orp, 0.85, 9), g["pr"])
    # in- and out-lists swapped in CDLP's counting changes nothing (it is symmetric),
    # but using only the out-list does:
    empty_rp = np.zeros_like(irp)
    if not np.array_equal(O.cdlp(empty_rp, np.zeros(0, np.int32), orp, g["out_col"], 10), g["cdlp"]):
        return
    pytest.skip("graph symmetric enough that dropping in-lists is invisible")
