"""Parity of the HIP path (through the C ABI) with the CPU oracle and with
GraphHET's own CPU outputs (golden fixtures).

Bars: BFS depths, WCC labels, CDLP labels and TC counts bit-exact; SSSP
within 1e-12 relative (the fixpoint argument makes it exact, and the tests
also check bit equality); PageRank L1 <= 1e-6 and max-abs <= 1e-7 against the
FP64 oracle."""
import numpy as np
import pytest

import graphhet_amd as G
import oracle as O

pytestmark = pytest.mark.gpu

INT_INF = 2**31 - 1
DBL_MAX = np.finfo(np.float64).max
PR_L1, PR_MAXABS = 1e-6, 1e-7
SCHEDULES = [G.SCHED_HEURISTIC, G.SCHED_PUSH, G.SCHED_PULL, G.SCHED_BITMAP, G.SCHED_REFERENCE]


def dev_from_golden(g):
    out = G.Csr(int(g["n"]), g["out_rowptr"], g["out_col"], g["out_w"])
    inn = G.Csr(int(g["n"]), g["in_rowptr"], g["in_col"], g["in_w"])
    return G.DeviceGraph(out, inn)


def sym_from_golden(g):
    return G.DeviceGraph(G.Csr(int(g["n"]), g["sym_rowptr"], g["sym_col"]), undirected=True, weights=False)


def test_golden_bfs_all_schedules(golden):
    dg = dev_from_golden(golden)
    for sch in SCHEDULES:
        d, st = dg.bfs(int(golden["src"]), schedule=sch)
        np.testing.assert_array_equal(d, golden["bfs"], err_msg=f"schedule {sch}")


def test_golden_sssp(golden):
    dg = dev_from_golden(golden)
    for delta in (0.0, 0.01, 0.3, 5.0):
        for heavy in (None, "push", "pull"):
            d, st = dg.sssp(int(golden["src"]), delta=delta, heavy=heavy)
            np.testing.assert_allclose(d, golden["sssp"], rtol=1e-12, atol=0)
            np.testing.assert_array_equal(d, golden["sssp"])  # fixpoint: bit-identical
            if heavy == "push":
                assert st.extra[3] == 0


def test_golden_wcc(golden):
    dg = dev_from_golden(golden)
    lab, st = dg.wcc()
    np.testing.assert_array_equal(lab, golden["wcc"])


def test_golden_pr(golden):
    dg = dev_from_golden(golden)
    r, st = dg.pr(10, 0.85)
    err = np.abs(r - golden["pr"])
    assert err.sum() <= PR_L1 and err.max() <= PR_MAXABS, (err.sum(), err.max())
    r2, _ = dg.pr(10, 0.85, normalise=False)
    err = np.abs(r2 - golden["pr"])
    assert err.sum() <= PR_L1 and err.max() <= PR_MAXABS, (err.sum(), err.max())


def test_golden_cdlp(golden):
    dg = dev_from_golden(golden)
    lab, st = dg.cdlp(10)
    np.testing.assert_array_equal(lab, golden["cdlp"])


def test_golden_tc(golden):
    sg = sym_from_golden(golden)
    c, st = sg.tc()
    np.testing.assert_array_equal(c, golden["tc_sym"])
    dg = dev_from_golden(golden)  # directed input: the reference loop itself
    c, st = dg.tc()
    np.testing.assert_array_equal(c, golden["tc_dir"])


def test_golden_parents(golden):
    g = golden
    dg = dev_from_golden(g)
    src = int(g["src"])
    d, pre, _ = dg.bfs_pre(src)
    np.testing.assert_array_equal(d, g["bfs"])
    np.testing.assert_array_equal(pre, O.bfs_parent(g["in_rowptr"], g["in_col"], g["bfs"]))
    for heavy in (None, "pull"):
        dist, pre, _ = dg.sssp_pre(src, heavy=heavy)
        np.testing.assert_array_equal(dist, g["sssp"])
        np.testing.assert_array_equal(pre, O.sssp_parent(g["in_rowptr"], g["in_col"], g["in_w"], g["sssp"], src))


@pytest.fixture(scope="module", params=[("rmat", 16, 1), ("rmat", 14, 9), ("er", 14, 4)])
def medium(request):
    kind, scale, seed = request.param
    if kind == "rmat":
        hg = G.HostGraph.rmat(scale, 16, seed=seed)
    else:
        hg = G.HostGraph.er(1 << scale, 16 << scale, seed=seed)
    o, i, a = hg.view(G.VIEW_OUT), hg.view(G.VIEW_IN), hg.view(G.VIEW_ALL)
    dg = G.DeviceGraph(o, i, a)
    src = int(np.argmax(o.degrees()))
    return dict(hg=hg, o=o, i=i, dg=dg, src=src, name=f"{kind}{scale}")


def test_medium_bfs(medium):
    o, dg, src = medium["o"], medium["dg"], medium["src"]
    ref = O.bfs(o.rowptr, o.col, src)
    for sch in SCHEDULES:
        d, st = dg.bfs(src, schedule=sch)
        np.testing.assert_array_equal(d, ref, err_msg=f"schedule {sch}")
        reached = ref != INT_INF
        assert st.vertices_reached == reached.sum()
        assert st.edges_traversed == np.diff(o.rowptr)[reached].sum()
    # a few other sources, including an isolated / sink vertex
    deg = o.degrees()
    for s in [0, int(np.argmin(deg)), o.n - 1]:
        d, _ = dg.bfs(s)
        np.testing.assert_array_equal(d, O.bfs(o.rowptr, o.col, s))


def test_medium_sssp(medium):
    o, dg, src = medium["o"], medium["dg"], medium["src"]
    ref = O.sssp(o.rowptr, o.col, o.w, src)
    for delta in (0.0, 0.05, 1.0):
        for heavy in (None, "push", "pull"):
            d, st = dg.sssp(src, delta=delta, heavy=heavy)
            np.testing.assert_array_equal(d, ref, err_msg=f"delta {delta} heavy {heavy}")
            if heavy == "pull":
                assert st.extra[3] >= 1
    # a finite `inf` like the reference's default leaves far vertices at inf
    ref_cap = np.where(ref < 0.5, ref, 0.5)
    for heavy in (None, "pull"):
        d, _ = dg.sssp(src, inf=0.5, heavy=heavy)
        np.testing.assert_array_equal(d, ref_cap)


def test_sssp_pull_forms():
    """The pull form of the heavy pass over the in-CSR: taken by the
    automatic form on an R-MAT graph whose big bucket settles most of it,
    forced for every heavy pass, on an undirected graph (the in-CSR is the
    out-CSR) and with zero-weight edges and multi-edges; every form is
    bit-identical to the oracle."""
    hg = G.HostGraph.rmat(17, 16, seed=5)
    o, i, a = hg.view(G.VIEW_OUT), hg.view(G.VIEW_IN), hg.view(G.VIEW_ALL)
    dg = G.DeviceGraph(o, i, a)
    src = int(np.argmax(o.degrees()))
    ref = O.sssp(o.rowptr, o.col, o.w, src)
    d, st = dg.sssp(src)
    np.testing.assert_array_equal(d, ref)
    assert st.extra[3] >= 1, "the automatic form did not pull on R-MAT s17"
    for heavy in ("push", "pull"):
        d, st = dg.sssp(src, heavy=heavy)
        np.testing.assert_array_equal(d, ref, err_msg=heavy)
    # no in-CSR: always the push form
    dn = G.DeviceGraph(o, with_in=False)
    d, st = dn.sssp(src, heavy="pull")
    np.testing.assert_array_equal(d, ref)
    assert st.extra[3] == 0
    # undirected: both directions in the out-CSR, which is also the in-CSR
    rng = np.random.default_rng(3)
    n = 3000
    u = rng.integers(0, n, 20000)
    v = rng.integers(0, n, 20000)
    w = np.round(rng.random(20000), 2)  # ties and zero weights
    w[:500] = 0.0
    uu, vv, ww = np.concatenate([u, v, u[:50]]), np.concatenate([v, u, v[:50]]), np.concatenate([w, w, w[:50] + 0.5])
    order = np.lexsort((vv, uu))
    rowptr = np.zeros(n + 1, np.int32)
    np.add.at(rowptr, uu + 1, 1)
    rowptr = np.cumsum(rowptr).astype(np.int32)
    col, wt = vv[order].astype(np.int32), ww[order].astype(np.float64)
    ug = G.DeviceGraph(G.Csr(n, rowptr, col, wt), undirected=True)
    for s0 in (0, 17, n - 1):
        ref = O.sssp(rowptr, col, wt, s0)
        for heavy in (None, "push", "pull"):
            for delta in (0.0, 0.02, 0.5):
                d, _ = ug.sssp(s0, delta=delta, heavy=heavy)
                np.testing.assert_array_equal(d, ref, err_msg=f"src {s0} heavy {heavy} delta {delta}")


def test_medium_wcc_pr_cdlp(medium):
    o, i, dg = medium["o"], medium["i"], medium["dg"]
    np.testing.assert_array_equal(dg.wcc()[0], O.wcc(o.rowptr, o.col))
    ref = O.pr(i.rowptr, i.col, o.rowptr, 0.85, 20)
    r, st = dg.pr(20, 0.85)
    err = np.abs(r - ref)
    assert err.sum() <= PR_L1 and err.max() <= PR_MAXABS, (err.sum(), err.max())
    np.testing.assert_array_equal(dg.cdlp(10)[0], O.cdlp(i.rowptr, i.col, o.rowptr, o.col, 10))
    # 0 and 1 iterations stay by vertex id (identity, closed form); from 2 on
    # the counting runs in rank space and is mapped back
    for it in (0, 1, 2, 3):
        np.testing.assert_array_equal(dg.cdlp(it)[0], O.cdlp(i.rowptr, i.col, o.rowptr, o.col, it),
                                      err_msg=f"{it} iterations")


def test_cdlp_undirected_rank_space(medium):
    """Undirected input (the all-list is the out-list of a symmetric CSR):
    the closed-form first iteration takes the smallest neighbour, then the
    counting runs in rank space.  The oracle with in = out counts every
    neighbour twice, which leaves the mode and its ties unchanged."""
    s = medium["hg"].view(G.VIEW_SYM)
    sg = G.DeviceGraph(s, undirected=True, weights=False)
    for it in (1, 2, 5):
        np.testing.assert_array_equal(sg.cdlp(it)[0], O.cdlp(s.rowptr, s.col, s.rowptr, s.col, it),
                                      err_msg=f"{it} iterations")


def test_medium_tc(medium):
    s = medium["hg"].view(G.VIEW_SYM)
    sg = G.DeviceGraph(s, undirected=True, weights=False)
    c, st = sg.tc()
    assert st.extra[0] == 1.0  # symmetric -> DAG form
    np.testing.assert_array_equal(c, O.tc(s.rowptr, s.col))


def test_medium_tc_directed(medium):
    """The reference loop on the directed out-CSR (no DAG): rows above the
    per-wave table go to the block kernel."""
    o = medium["o"]
    c, st = medium["dg"].tc()
    assert st.extra[0] == 0.0
    np.testing.assert_array_equal(c, O.tc(o.rowptr, o.col))


def test_tc_clique_block_rows():
    """K_600: every DAG row of rank r has 599 - r entries, so rows above 256
    entries take the block kernel; each vertex is in C(599, 2) triangles."""
    k = 600
    src = [u for u in range(k) for v in range(k) if u != v]
    dst = [v for u in range(k) for v in range(k) if u != v]
    hg = G.HostGraph.from_edges(k, src, dst)
    s = hg.view(G.VIEW_SYM)
    c, st = G.DeviceGraph(s, undirected=True, weights=False).tc()
    assert st.extra[0] == 1.0
    np.testing.assert_array_equal(c, np.full(k, 6 * (k - 1) * (k - 2) // 2, np.int64))


def test_tc_directed_hub_in_chunks():
    """A directed hub with 20,000 out-neighbours (more than one block table
    of 8,192 entries: three chunks), a path among them, self-loops on some
    vertices and a second hub: the reference loop, against the oracle."""
    n = 20003
    rng = np.random.default_rng(5)
    src = [0] * (n - 1) + list(range(1, n - 1)) + [7, 9, 11, 20001] + [1] * 9000
    dst = list(range(1, n)) + list(range(2, n)) + [7, 9, 11, 20001] + rng.choice(np.arange(2, n), 9000, replace=False).tolist()
    hg = G.HostGraph.from_edges(n, src, dst)
    o = hg.view(G.VIEW_OUT)
    c, st = G.DeviceGraph(o, weights=False).tc()
    assert st.extra[0] == 0.0
    np.testing.assert_array_equal(c, O.tc(o.rowptr, o.col))


def test_caller_stream(medium):
    """BFS and SSSP on a stream the caller owns (opts.stream): level and pass
    totals still reach the host through the mailbox, results through the
    staging copy, and they equal the graph-stream results."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")  # the runtime libghet.so is linked against
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        o, dg, src = medium["o"], medium["dg"], medium["src"]
        d0, s0 = dg.bfs(src)
        d1, s1 = dg.bfs(src, stream=stream.value)
        np.testing.assert_array_equal(d0, d1)
        assert s0.levels == s1.levels and s0.edges_traversed == s1.edges_traversed
        x0, _ = dg.sssp(src)
        x1, _ = dg.sssp(src, stream=stream.value)
        np.testing.assert_array_equal(x0, x1)
        np.testing.assert_array_equal(x1, O.sssp(o.rowptr, o.col, o.w, src))
    finally:
        assert hip.hipStreamSynchronize(stream) == 0
        hip.hipStreamDestroy(stream)


def test_results_are_deterministic(medium):
    dg = medium["dg"]
    a, _ = dg.pr(10)
    b, _ = dg.pr(10)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(dg.cdlp(5)[0], dg.cdlp(5)[0])


def test_edge_cases():
    # single vertex, no edges
    dg = G.DeviceGraph(G.Csr(1, np.array([0, 0], np.int32), np.zeros(0, np.int32), np.zeros(0)))
    assert dg.bfs(0)[0].tolist() == [0]
    assert dg.sssp(0)[0].tolist() == [0.0]
    assert dg.wcc()[0].tolist() == [0]
    np.testing.assert_allclose(dg.pr(5)[0], [1.0])
    assert dg.cdlp(3)[0].tolist() == [0]
    assert dg.tc()[0].tolist() == [0]
    # all isolated
    n = 100
    dg = G.DeviceGraph(G.Csr(n, np.zeros(n + 1, np.int32), np.zeros(0, np.int32), np.zeros(0)))
    d, _ = dg.bfs(7)
    assert d[7] == 0 and (d == INT_INF).sum() == n - 1
    np.testing.assert_array_equal(dg.wcc()[0], np.arange(n))
    np.testing.assert_array_equal(dg.cdlp(2)[0], np.arange(n))
    # invalid source
    with pytest.raises(G.GhetError) as e:
        dg.bfs(n)
    assert e.value.code == G.ERR_INVALID
    with pytest.raises(G.GhetError):
        dg.sssp(-1)


def test_hub_heavy_graph():
    """Stars with hubs above every tile/LDS threshold (PR long rows, CDLP
    global hash, BFS hub path)."""
    n = 20000
    src = list(range(1, n)) + [0] * (n - 1) + [1] * 3000 + list(range(2, 3002))
    dst = [0] * (n - 1) + list(range(1, n)) + list(range(2, 3002)) + [1] * 3000
    w = np.linspace(0.01, 1.0, len(src))
    hg = G.HostGraph.from_edges(n, src, dst, w)
    o, i, a = hg.view(G.VIEW_OUT), hg.view(G.VIEW_IN), hg.view(G.VIEW_ALL)
    dg = G.DeviceGraph(o, i, a)
    for sch in SCHEDULES:
        np.testing.assert_array_equal(dg.bfs(5, schedule=sch)[0], O.bfs(o.rowptr, o.col, 5))
    np.testing.assert_array_equal(dg.sssp(5)[0], O.sssp(o.rowptr, o.col, o.w, 5))
    ref = O.pr(i.rowptr, i.col, o.rowptr, 0.85, 10)
    err = np.abs(dg.pr(10)[0] - ref)
    assert err.sum() <= PR_L1 and err.max() <= PR_MAXABS
    np.testing.assert_array_equal(dg.cdlp(4)[0], O.cdlp(i.rowptr, i.col, o.rowptr, o.col, 4))
    np.testing.assert_array_equal(dg.wcc()[0], O.wcc(o.rowptr, o.col))
    s = hg.view(G.VIEW_SYM)
    np.testing.assert_array_equal(G.DeviceGraph(s, undirected=True, weights=False).tc()[0], O.tc(s.rowptr, s.col))


@pytest.mark.parametrize("split", [(0, 1), (4, 6), (9, 10), (12, 12), (0, 1, True), (4, 6, True), (9, 10, True),
                                   (12, 12, True), (15, 13, True), (20, 20, True)])
def test_pr_source_windows_match_oracle(medium, split):
    """Forced hot/window splits of the relabeled sources (log2 hot sources,
    log2 window): hot parts and cold segments become partial slots summed by
    k_pr_long; every split must meet the same tolerance as the default.
    With blocked=True the hot parts go through the source-blocked phase
    (k_pr_hot): several 16K-source blocks, no cold sources at all (20, 20),
    hot blocks cut short by the split (12, 12)."""
    o, i, dg = medium["o"], medium["i"], medium["dg"]
    ref = O.pr(i.rowptr, i.col, o.rowptr, 0.85, 20)
    r, st = dg.pr(20, 0.85, split=split)
    err = np.abs(r - ref)
    assert err.sum() <= PR_L1 and err.max() <= PR_MAXABS, (split, err.sum(), err.max())
    r2, _ = dg.pr(20, 0.85, split=split)
    np.testing.assert_array_equal(r, r2)  # deterministic under a split
    r0, _ = dg.pr(20, 0.85)  # back to the automatic layout
    err = np.abs(r0 - ref)
    assert err.sum() <= PR_L1 and err.max() <= PR_MAXABS


def test_pr_source_windows_with_long_segments():
    """A hub whose in-list spans windows of more than 2048 sources: cold
    segments longer than a tile are chunked into several slots."""
    n = 20000
    src = list(range(1, n)) + [0] * (n - 1) + [1] * 3000 + list(range(2, 3002))
    dst = [0] * (n - 1) + list(range(1, n)) + list(range(2, 3002)) + [1] * 3000
    hg = G.HostGraph.from_edges(n, src, dst)
    o, i = hg.view(G.VIEW_OUT), hg.view(G.VIEW_IN)
    dg = G.DeviceGraph(o, i)
    ref = O.pr(i.rowptr, i.col, o.rowptr, 0.85, 10)
    for split in ((2, 13), (6, 12), (0, 14), (2, 13, True), (14, 14, True), (20, 20, True)):
        r, _ = dg.pr(10, 0.85, split=split)
        err = np.abs(r - ref)
        assert err.sum() <= PR_L1 and err.max() <= PR_MAXABS, (split, err.sum(), err.max())


@pytest.mark.parametrize("split", [None, (9, 10), (12, 12, True)])
def test_pr_phase_timing_keeps_results(medium, split):
    """Hot and cold gathers as two launches (phase timing): same tolerance
    against the oracle, deterministic, and the hot phase is timed."""
    o, i, dg = medium["o"], medium["i"], medium["dg"]
    ref = O.pr(i.rowptr, i.col, o.rowptr, 0.85, 20)
    r, st = dg.pr(20, 0.85, split=split, phases=True)
    err = np.abs(r - ref)
    assert err.sum() <= PR_L1 and err.max() <= PR_MAXABS, (split, err.sum(), err.max())
    r2, _ = dg.pr(20, 0.85, split=split, phases=True)
    np.testing.assert_array_equal(r, r2)
    assert 0 < st.extra[2] <= st.hot_kernel_ms


def test_pr_gather_ceiling_covers_every_edge(medium):
    """The gather probe walks the hot and the cold edges of the layout:
    together every in-edge once; an unsplit layout has no cold edges."""
    i, dg = medium["i"], medium["dg"]
    c = dg.pr_gather_ceiling(3, split=(9, 10))
    assert c["hot_edges"] + c["cold_edges"] == i.m and c["cold_edges"] > 0
    assert c["hot_ms"] > 0 and c["cold_ms"] > 0
    c = dg.pr_gather_ceiling(3)
    assert c["hot_edges"] == i.m and c["cold_edges"] == 0 and c["cold_ms"] == 0
    with pytest.raises(G.GhetError):
        dg.pr_gather_ceiling(0)


def test_golden_pr_binned(golden):
    """The destination-bin form against GraphHET's own CPU outputs, in the
    reference form and normalised."""
    dg = dev_from_golden(golden)
    for norm in (True, False):
        r, st = dg.pr(10, 0.85, normalise=norm, form="binned")
        err = np.abs(r - golden["pr"])
        assert err.sum() <= PR_L1 and err.max() <= PR_MAXABS, (norm, err.sum(), err.max())
        assert st.extra[3] == 2


def test_pr_binned_medium(medium):
    """Binned and gather forms each within the bar of the FP64 oracle, and
    the binned sums exact integers: two runs are bit-identical."""
    o, i, dg = medium["o"], medium["i"], medium["dg"]
    ref = O.pr(i.rowptr, i.col, o.rowptr, 0.85, 20)
    r, st = dg.pr(20, 0.85, form="binned")
    err = np.abs(r - ref)
    assert err.sum() <= PR_L1 and err.max() <= PR_MAXABS, (err.sum(), err.max())
    np.testing.assert_array_equal(r, dg.pr(20, 0.85, form="binned")[0])
    assert st.extra[3] == 2 and 0 < st.extra[2] < st.hot_kernel_ms
    g, stg = dg.pr(20, 0.85, form="gather")
    err = np.abs(g - ref)
    assert err.sum() <= PR_L1 and err.max() <= PR_MAXABS and stg.extra[3] != 2


def test_pr_binned_windows_bins_and_items():
    """Several source windows (2^15 ids) and destination bins; a vertex whose
    in-edges exceed a bin's edge target (a bin of its own); a window with
    more edges than one gather item (vertex 0's out-edges); R-MAT with its
    hubs at low ids."""
    n = 150000
    src = list(range(1, n)) + [0] * (n - 1) + [1] * 3000 + list(range(2, 3002))
    dst = [0] * (n - 1) + list(range(1, n)) + list(range(2, 3002)) + [1] * 3000
    hg = G.HostGraph.from_edges(n, src, dst)
    for h in (hg, G.HostGraph.rmat(17, 16, seed=3)):
        o, i = h.view(G.VIEW_OUT), h.view(G.VIEW_IN)
        dg = G.DeviceGraph(o, i)
        ref = O.pr(i.rowptr, i.col, o.rowptr, 0.85, 10)
        for norm in (True, False):
            r, _ = dg.pr(10, 0.85, normalise=norm, form="binned")
            want = ref / ref.sum() if norm else ref
            err = np.abs(r - want)
            assert err.sum() <= PR_L1 and err.max() <= PR_MAXABS, (o.n, norm, err.sum(), err.max())


def test_pr_binned_edge_cases():
    dg = G.DeviceGraph(G.Csr(1, np.array([0, 0], np.int32), np.zeros(0, np.int32), np.zeros(0)))
    np.testing.assert_allclose(dg.pr(5, form="binned")[0], [1.0])
    n = 100
    dg = G.DeviceGraph(G.Csr(n, np.zeros(n + 1, np.int32), np.zeros(0, np.int32), np.zeros(0)))
    np.testing.assert_allclose(dg.pr(3, form="binned")[0], np.full(n, 1.0 / n), rtol=1e-15)
    r0, _ = dg.pr(0, form="binned")
    np.testing.assert_allclose(r0, np.full(n, 1.0 / n), rtol=1e-15)
    with pytest.raises(ValueError):
        dg.pr(3, form="binned", split=(4, 6))


def test_cdlp_hub_merge_partitions():
    """A hub whose neighbours keep many distinct labels: hub 0 points to N
    leaves, each leaf is tied both ways to two private partners (so a leaf's
    label is its smaller partner), and the hub sees N distinct labels -- its
    pairs need several merge partitions.  Also with the merge tables capped
    at 64 slots (the sub-pass path), a second hub of 9000 neighbours (two
    chunks), and a vertex of 6000 neighbours with distinct labels (more than
    its 4096-slot table holds: the block's sub-passes)."""
    N = 20000
    src, dst = [], []
    for i in range(1, N + 1):
        a, b = N + 2 * i - 1, N + 2 * i
        src += [0, i, a, i, b, a, b]
        dst += [i, a, i, b, i, b, a]
    n = 3 * N + 1 + 9001 + 1
    h2 = 3 * N + 1
    for j in range(1, 9001):
        src += [h2 + j, h2]
        dst += [h2, h2 + j]
    h3 = n - 1
    src += [h3] * 6000
    dst += list(range(1, 6001))
    hg = G.HostGraph.from_edges(n, src, dst)
    o, i, a = hg.view(G.VIEW_OUT), hg.view(G.VIEW_IN), hg.view(G.VIEW_ALL)
    dg = G.DeviceGraph(o, i, a)
    for iters in (2, 3, 5):
        ref = O.cdlp(i.rowptr, i.col, o.rowptr, o.col, iters)
        np.testing.assert_array_equal(dg.cdlp(iters)[0], ref, err_msg=f"{iters} iterations")
        np.testing.assert_array_equal(dg.cdlp(iters, small_merge_table=True)[0], ref, err_msg=f"capped, {iters}")
