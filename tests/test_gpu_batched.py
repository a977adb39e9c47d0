"""Batched forms: many BFS/SSSP sources in persistent cooperative launches
(BFS: 256 sources per launch; SSSP: 32 per launch, the whole bucket loop on
the device) and many small graphs in one captured launch.  Each must equal the per-instance results bit for
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


@pytest.mark.parametrize("delta", [1e-4, 0.05, 10.0])
def test_sssp_multi_bucket_widths(rmat14, delta):
    """Many narrow buckets (far-list jumps), a few, and one bucket for all:
    the device-side bucket loop gives the per-source distances every time."""
    hg, o, i, dg = rmat14
    srcs = sources(o, 40, 5)
    dist, _ = dg.sssp_multi(srcs, delta=delta)
    for j in range(0, len(srcs), 7):
        ref, _ = dg.sssp(int(srcs[j]))
        np.testing.assert_array_equal(dist[j], ref, err_msg=f"delta {delta} source #{j}")


def test_sssp_multi_long_paths_and_ties():
    """A 3000-vertex path (one vertex per round: thousands of grid-barrier
    phases in one launch) and a grid with unit weights (many equal sums)."""
    n = 3000
    hg = G.HostGraph.from_edges(n, list(range(n - 1)), list(range(1, n)), np.full(n - 1, 0.5))
    dg = G.DeviceGraph.from_host(hg)
    o = hg.view(G.VIEW_OUT)
    srcs = np.array([0, 10, 1500, 2999, 0], np.int32)
    dist, st = dg.sssp_multi(srcs, delta=0.25)
    for j, s_ in enumerate(srcs):
        np.testing.assert_array_equal(dist[j], O.sssp(o.rowptr, o.col, o.w, int(s_)))
    R, Cc = 40, 50
    e = []
    for r in range(R):
        for c in range(Cc):
            x = r * Cc + c
            if c + 1 < Cc:
                e += [(x, x + 1), (x + 1, x)]
            if r + 1 < R:
                e += [(x, x + Cc), (x + Cc, x)]
    hg = G.HostGraph.from_edges(R * Cc, [a for a, _ in e], [b for _, b in e], np.ones(len(e)))
    dg = G.DeviceGraph.from_host(hg)
    o = hg.view(G.VIEW_OUT)
    srcs = np.arange(0, R * Cc, 61, dtype=np.int32)
    dist, _ = dg.sssp_multi(srcs)
    for j, s_ in enumerate(srcs):
        np.testing.assert_array_equal(dist[j], O.sssp(o.rowptr, o.col, o.w, int(s_)))


def test_sssp_multi_phase_budget_stops_cleanly(rmat14):
    """A pass that spends its phase budget (forced tiny from the host) stops
    on the device for every block at once and the call reports it; the next
    call runs normally."""
    hg, o, i, dg = rmat14
    srcs = sources(o, 8, 9)
    with pytest.raises(G.GhetError) as e:
        dg.sssp_multi(srcs, phase_budget=3)
    assert e.value.code == G.ERR_HIP and "budget" in str(e.value)
    dist, _ = dg.sssp_multi(srcs)
    ref, _ = dg.sssp(int(srcs[0]))
    np.testing.assert_array_equal(dist[0], ref)


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
