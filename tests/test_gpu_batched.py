"""Batched forms: many BFS/SSSP sources in one launch and many small graphs
in one captured launch.  Each must equal the per-instance results bit for
bit (and the per-instance results equal the oracle, see test_gpu_parity)."""
import numpy as np
import pytest

import graphhet_amd as G
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rmat14():
    hg = G.HostGraph.rmat(14, 16, seed=11)
    o, i = hg.view(G.VIEW_OUT), hg.view(G.VIEW_IN)
    dg = G.DeviceGraph(o, i)
    return hg, o, i, dg


def sources(o, k, seed):
    deg = o.degrees() + np.diff(np.concatenate([[0], np.cumsum(np.bincount(o.col, minlength=o.n))]))
    cand = np.nonzero(deg > 0)[0]
    rng = np.random.default_rng(seed)
    return rng.choice(cand, size=k, replace=True).astype(np.int32)


def test_bfs_multi_matches_single(rmat14):
    hg, o, i, dg = rmat14
    srcs = sources(o, 300, 1)  # two passes (256 + 44), repeated sources included
    srcs[5] = srcs[6]
    srcs[-1] = int(np.argmin(o.degrees()))  # possibly a sink / isolated vertex
    depth, st = dg.bfs_multi(srcs)
    for j in [0, 5, 6, 100, 255, 256, 299]:
        ref, _ = dg.bfs(int(srcs[j]))
        np.testing.assert_array_equal(depth[j], ref, err_msg=f"source #{j}")
    for j in [0, 257]:
        np.testing.assert_array_equal(depth[j], O.bfs(o.rowptr, o.col, int(srcs[j])))
    # every row, cheaply: reached counts equal the single runs'
    tot = 0
    for j in range(len(srcs)):
        ref, rs = dg.bfs(int(srcs[j]))
        assert np.array_equal(depth[j], ref), j
        tot += rs.edges_traversed
    assert st.edges_traversed == tot


def test_sssp_multi_matches_single(rmat14):
    hg, o, i, dg = rmat14
    srcs = sources(o, 70, 2)  # three passes of up to 32 sources (32 + 32 + 6)
    srcs[5] = srcs[6]  # one vertex carrying two sources of a pass
    srcs[-1] = int(np.argmin(o.degrees()))  # possibly a sink / isolated vertex
    dist, st = dg.sssp_multi(srcs)
    tot = 0
    for j in range(len(srcs)):
        ref, rs = dg.sssp(int(srcs[j]))
        np.testing.assert_array_equal(dist[j], ref, err_msg=f"source #{j}")
        tot += rs.edges_traversed
    assert st.edges_traversed == tot
    np.testing.assert_array_equal(dist[3], O.sssp(o.rowptr, o.col, o.w, int(srcs[3])))
    dist7, _ = dg.sssp_multi(srcs, per_pass=7)  # pass width changes nothing
    np.testing.assert_array_equal(dist7, dist)


def small_graphs(count, seed, sym):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        scale = int(rng.integers(3, 11))
        ef = float(rng.uniform(2, 12))
        hg = G.HostGraph.rmat(scale, ef, seed=seed * 1000 + k)
        out.append(hg.view(G.VIEW_SYM if sym else G.VIEW_OUT))
    out.append(G.Csr(1, np.array([0, 0], np.int32), np.zeros(0, np.int32)))  # a one-vertex graph
    return out


def test_wcc_many_matches_per_graph():
    graphs = small_graphs(60, 3, sym=False)
    labs, st = G.wcc_many(graphs)
    for g, lab in zip(graphs, labs):
        np.testing.assert_array_equal(lab, O.wcc(g.rowptr, g.col))
    labs2, _ = G.wcc_many(graphs, tune=False)
    for a, b in zip(labs, labs2):
        np.testing.assert_array_equal(a, b)


def test_tc_many_matches_per_graph():
    for sym in (True, False):
        graphs = small_graphs(40, 4 + sym, sym=sym)
        cnts, st = G.tc_many(graphs)
        assert st.extra[2] == (1.0 if sym else 0.0)
        for g, c in zip(graphs, cnts):
            np.testing.assert_array_equal(c, O.tc(g.rowptr, g.col))


def test_autotune_keeps_results_and_caches(rmat14, tmp_path):
    hg, o, i, dg = rmat14
    s = hg.view(G.VIEW_SYM)
    sg = G.DeviceGraph(s, undirected=True, weights=False)
    before, _ = sg.tc()
    cache = str(tmp_path / "tune.txt")
    ept = sg.autotune(cache)
    assert ept in (2, 4, 8, 16, 32)
    after, _ = sg.tc()
    np.testing.assert_array_equal(before, after)
    txt = open(cache).read()
    assert "tc_ept=" in txt and "sssp_class=" in txt
    sg2 = G.DeviceGraph(s, undirected=True, weights=False)
    assert sg2.autotune(cache) == ept  # read back from the cache
    d0, _ = dg.sssp(0)
    dg.autotune(None)
    d1, _ = dg.sssp(0)
    np.testing.assert_array_equal(d0, d1)
