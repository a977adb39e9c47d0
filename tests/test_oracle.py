"""Pins the CPU oracle (oracle/ghet_oracle.cpp) before anything is checked
against it: golden vectors from GraphHET's own CPU code, independent library
routines (networkx / scipy), and closed forms chosen so that a plausible slip
(off-by-one depth, a dropped dangling term, a wrong tie rule, in- vs
out-lists swapped) fails at least one of them."""
import math
import os
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as csg

import graphhet_amd as G
import oracle as O

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from conftest import golden_names, load_golden  # noqa: E402

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


# ---------------------------------------------------------------- parents --
# The parent rules are pinned from the out-CSR by a route that shares nothing
# with the oracle's in-list scan: every out-edge (u, v) that satisfies the
# level (or distance) condition proposes u to v, and a scatter-min keeps the
# smallest proposal.  Graphs with many equal-length paths (unit or quantised
# weights) make the "smallest" rule visible.

def _edges(rowptr, col):
    n = len(rowptr) - 1
    return np.repeat(np.arange(n, dtype=np.int64), np.diff(rowptr)), np.asarray(col, np.int64)


def _bfs_parent_by_out_edges(out_rowptr, out_col, depth):
    n = len(out_rowptr) - 1
    u, v = _edges(out_rowptr, out_col)
    d = np.asarray(depth, np.int64)
    ok = (d[u] != INT_INF) & (d[v] == d[u] + 1)
    best = np.full(n, n, np.int64)
    np.minimum.at(best, v[ok], u[ok])
    return np.where(best == n, -1, best).astype(np.int32)


def _sssp_parent_by_out_edges(out_rowptr, out_col, out_w, dist, src):
    n = len(out_rowptr) - 1
    u, v = _edges(out_rowptr, out_col)
    d = np.asarray(dist, np.float64)
    ok = (u != v) & (d[u] != DBL_MAX) & (d[u] + np.asarray(out_w, np.float64) == d[v]) & (v != src)
    best = np.full(n, n, np.int64)
    np.minimum.at(best, v[ok], u[ok])
    return np.where(best == n, -1, best).astype(np.int32)


def _check_bfs_parents(out_rowptr, out_col, depth, src, pre):
    """(i) pre[v] -> v is an out-edge, (ii) depth[pre[v]] == depth[v] - 1,
    (iii) no smaller in-neighbour qualifies; -1 exactly for src and unreached."""
    n = len(out_rowptr) - 1
    pre = np.asarray(pre, np.int64)
    has = pre >= 0
    reached = (np.asarray(depth) != INT_INF) & (np.arange(n) != src)
    assert np.array_equal(has, reached), "parent present exactly for reached non-source vertices"
    vs = np.nonzero(has)[0]
    key_edges = set(zip(*(_edges(out_rowptr, out_col))))
    assert all((int(pre[x]), int(x)) in key_edges for x in vs), "pre[v] -> v is not an out-edge"
    assert np.all(np.asarray(depth)[pre[vs]] == np.asarray(depth)[vs] - 1), "parent not one level up"
    np.testing.assert_array_equal(pre, _bfs_parent_by_out_edges(out_rowptr, out_col, depth))


def _check_sssp_parents(out_rowptr, out_col, out_w, dist, src, pre):
    n = len(out_rowptr) - 1
    pre = np.asarray(pre, np.int64)
    has = pre >= 0
    reached = (np.asarray(dist) != DBL_MAX) & (np.arange(n) != src)
    assert np.array_equal(has, reached), "parent present exactly for reached non-source vertices"
    u, v = _edges(out_rowptr, out_col)
    w_of = {(int(a), int(b)): float(x) for a, b, x in zip(u, v, out_w)}
    for x in np.nonzero(has)[0]:
        e = (int(pre[x]), int(x))
        assert e in w_of, "pre[v] -> v is not an out-edge"
        assert dist[e[0]] + w_of[e] == dist[x], "d[pre] + w != d[v] in FP64"
    np.testing.assert_array_equal(pre, _sssp_parent_by_out_edges(out_rowptr, out_col, out_w, dist, src))


def _parent_graphs():
    """Golden graphs plus graphs with many ties: R-MAT with quantised
    weights, ER with unit weights, a grid and K_n."""
    out = []
    for name in golden_names():
        g = load_golden(name)
        out.append((name, g["out_rowptr"], g["out_col"], g["out_w"], g["in_rowptr"], g["in_col"], g["in_w"],
                    int(g["src"])))
    for name, hg in (("rmat_s11_q", G.HostGraph.rmat(11, 8, seed=7)), ("er_2000", G.HostGraph.er(2000, 12000, seed=3))):
        o = hg.view(G.VIEW_OUT)
        q = 1.0 if name.startswith("er") else None
        w = np.ones(o.m) if q else np.ceil(o.w * 3) / 4  # {0.25, 0.5, 0.75}
        hq = G.HostGraph.from_csr(G.Csr(o.n, o.rowptr, o.col, w))
        o2, i2 = hq.view(G.VIEW_OUT), hq.view(G.VIEW_IN)
        out.append((name, o2.rowptr, o2.col, o2.w, i2.rowptr, i2.col, i2.w, int(np.argmax(o2.degrees()))))
    R, Cc = 9, 13
    e = []
    for r in range(R):
        for c in range(Cc):
            x = r * Cc + c
            if c + 1 < Cc:
                e += [(x, x + 1), (x + 1, x)]
            if r + 1 < R:
                e += [(x, x + Cc), (x + Cc, x)]
    hg = G.HostGraph.from_edges(R * Cc, [a for a, _ in e], [b for _, b in e])
    o, i = hg.view(G.VIEW_OUT), hg.view(G.VIEW_IN)
    out.append(("grid", o.rowptr, o.col, o.w, i.rowptr, i.col, i.w, (R // 2) * Cc + Cc // 2))
    return out


PARENT_GRAPHS = _parent_graphs()


@pytest.mark.parametrize("case", PARENT_GRAPHS, ids=[c[0] for c in PARENT_GRAPHS])
def test_bfs_parent_oracle_is_valid_and_minimal(case):
    _, orp, ocol, ow, irp, icol, iw, src = case
    d = O.bfs(orp, ocol, src)
    _check_bfs_parents(orp, ocol, d, src, O.bfs_parent(irp, icol, d))


@pytest.mark.parametrize("case", PARENT_GRAPHS, ids=[c[0] for c in PARENT_GRAPHS])
def test_sssp_parent_oracle_is_valid_and_minimal(case):
    _, orp, ocol, ow, irp, icol, iw, src = case
    d = O.sssp(orp, ocol, ow, src)
    _check_sssp_parents(orp, ocol, ow, d, src, O.sssp_parent(irp, icol, iw, d, src))


def _reverse_rows(rowptr, col, w=None):
    c = np.array(col, copy=True)
    ww = None if w is None else np.array(w, copy=True)
    for v in range(len(rowptr) - 1):
        a, b = rowptr[v], rowptr[v + 1]
        c[a:b] = c[a:b][::-1]
        if ww is not None:
            ww[a:b] = ww[a:b][::-1]
    return c, ww


def test_parent_checks_reject_planted_mistakes():
    """Mistakes planted in the oracle's inputs on the CPU: the largest qualifying
    in-neighbour (rows scanned in descending order) and in/out lists swapped.
    Each must fail the checks on at least one graph with ties."""
    caught = {"bfs_largest": 0, "bfs_swapped": 0, "sssp_largest": 0, "sssp_swapped": 0}
    for _, orp, ocol, ow, irp, icol, iw, src in PARENT_GRAPHS:
        d = O.bfs(orp, ocol, src)
        dist = O.sssp(orp, ocol, ow, src)
        rcol, rw = _reverse_rows(irp, icol, iw)
        for key, pre in (("bfs_largest", O.bfs_parent(irp, rcol, d)), ("bfs_swapped", O.bfs_parent(orp, ocol, d))):
            try:
                _check_bfs_parents(orp, ocol, d, src, pre)
            except AssertionError:
                caught[key] += 1
        for key, pre in (("sssp_largest", O.sssp_parent(irp, rcol, rw, dist, src)),
                         ("sssp_swapped", O.sssp_parent(orp, ocol, ow, dist, src))):
            try:
                _check_sssp_parents(orp, ocol, ow, dist, src, pre)
            except AssertionError:
                caught[key] += 1
    assert all(v > 0 for v in caught.values()), caught
    print(caught)


def test_tc_forward_against_networkx_and_restatement(golden):
    """The forward-algorithm counter that checks TC at full size: networkx
    per-vertex triangles on every golden graph's symmetrisation, and 6 t(v)
    equal to the reference restatement's per-vertex counts."""
    nx = pytest.importorskip("networkx")
    rp, col = golden["sym_rowptr"], golden["sym_col"]
    n = len(rp) - 1
    U = nx.Graph()
    U.add_nodes_from(range(n))
    for u in range(n):
        for e in range(rp[u], rp[u + 1]):
            U.add_edge(u, int(col[e]))
    tri = nx.triangles(U)
    t = O.tc_forward(rp, col)
    np.testing.assert_array_equal(t, [tri[v] for v in range(n)])
    np.testing.assert_array_equal(6 * t, golden["tc_sym"])


@pytest.mark.parametrize("make", [lambda: G.HostGraph.rmat(13, 16, seed=5), lambda: G.HostGraph.er(5000, 60000, seed=2)],
                         ids=["rmat_s13", "er_5000"])
def test_tc_forward_equals_restatement(make):
    s = make().view(G.VIEW_SYM)
    for threads in (1, 3):
        np.testing.assert_array_equal(6 * O.tc_forward(s.rowptr, s.col, threads), O.tc(s.rowptr, s.col))


def test_tc_forward_closed_forms():
    for n in (3, 5, 9):  # K_n: C(n-1, 2) triangles through every vertex
        e = [(i, j) for i in range(n) for j in range(n) if i != j]
        s = G.HostGraph.from_edges(n, [a for a, _ in e], [b for _, b in e]).view(G.VIEW_SYM)
        np.testing.assert_array_equal(O.tc_forward(s.rowptr, s.col), [math.comb(n - 1, 2)] * n)
    # complete bipartite K_{4,5}: no triangles
    e = [(i, 4 + j) for i in range(4) for j in range(5)]
    s = G.HostGraph.from_edges(9, [a for a, _ in e], [b for _, b in e]).view(G.VIEW_SYM)
    assert not O.tc_forward(s.rowptr, s.col).any()


def test_tc_vertex_sample_is_the_restatement_over_those_vertices(golden):
    """The CPU baseline's TC sample: oracle_tc's outer loop over a vertex list.
    All vertices reproduce the full count; a split into two lists sums to it;
    the merge steps equal the closed form sum over u, v in N(u) of the merge
    length (at most deg u + deg v, exactly the steps until one list ends)."""
    rp, col = golden["sym_rowptr"], golden["sym_col"]
    n = len(rp) - 1
    full = O.tc(rp, col)
    allv, steps = O.tc_vertices(rp, col, np.arange(n))
    np.testing.assert_array_equal(allv, full)
    a, sa = O.tc_vertices(rp, col, np.arange(0, n, 2))
    b, sb = O.tc_vertices(rp, col, np.arange(1, n, 2))
    np.testing.assert_array_equal(a + b, full)
    assert sa + sb == steps
    deg = np.diff(rp)
    assert steps <= int(sum(deg[u] * deg[u] + deg[col[rp[u]:rp[u + 1]]].sum() for u in range(n)))
