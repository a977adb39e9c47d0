"""CPU-side checks of the C ABI: the library loads and exports every symbol
include/ghet.h declares, the host graph builders are correct and
deterministic, and argument validation fails loudly without a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import graphhet_amd as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "ghet.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ghet_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = G.lib()
    names = declared_symbols()
    assert len(names) >= 30
    missing = [s for s in names if not hasattr(L, s)]
    assert not missing, missing
    assert set(names) == set(G.EXPORTS)
    assert L.ghet_abi_version() == 1


def test_rmat_is_deterministic_and_thread_independent():
    a = G.HostGraph.rmat(11, 16, seed=5, threads=1).view()
    b = G.HostGraph.rmat(11, 16, seed=5, threads=8).view()
    c = G.HostGraph.rmat(11, 16, seed=6, threads=8).view()
    np.testing.assert_array_equal(a.rowptr, b.rowptr)
    np.testing.assert_array_equal(a.col, b.col)
    np.testing.assert_array_equal(a.w, b.w)
    assert not np.array_equal(a.col, c.col)


def test_rmat_shape():
    o = G.HostGraph.rmat(12, 16, seed=1).view()
    n = o.n
    assert n == 4096 and o.rowptr[0] == 0 and np.all(np.diff(o.rowptr) >= 0)
    u = np.repeat(np.arange(n), np.diff(o.rowptr))
    assert not np.any(u == o.col), "self-loops must be dropped"
    key = u.astype(np.int64) * n + o.col
    assert np.all(np.diff(key) > 0), "rows sorted and deduplicated"
    assert np.all(o.w > 0) and np.all(o.w <= 1.0)
    # (k+1)*2^-53 weights are exact multiples of 2^-53
    assert np.all(np.floor(o.w * 2.0**53) == o.w * 2.0**53)
    # skew: R-MAT a=.57 puts the heaviest rows at low ids
    deg = np.diff(o.rowptr)
    assert deg[: n // 16].sum() > deg[-n // 16:].sum() * 4
    assert 0.9 * 16 * n > o.m > 0.8 * 16 * n


def test_er_exact_edge_count():
    o = G.HostGraph.er(5000, 40000, seed=2).view()
    assert o.m == 40000
    u = np.repeat(np.arange(o.n), np.diff(o.rowptr))
    assert not np.any(u == o.col)
    key = u.astype(np.int64) * o.n + o.col
    assert np.all(np.diff(key) > 0)


def test_views_in_all_sym():
    hg = G.HostGraph.rmat(10, 8, seed=3)
    o, i, a, s = (hg.view(k) for k in (G.VIEW_OUT, G.VIEW_IN, G.VIEW_ALL, G.VIEW_SYM))
    n = o.n
    M = {(int(u), int(v)): w for u in range(n) for v, w in zip(o.col[o.rowptr[u]:o.rowptr[u + 1]],
                                                               o.w[o.rowptr[u]:o.rowptr[u + 1]])}
    for v in range(n):
        ins = i.col[i.rowptr[v]:i.rowptr[v + 1]]
        assert np.all(np.diff(ins) > 0)
        for u, w in zip(ins, i.w[i.rowptr[v]:i.rowptr[v + 1]]):
            assert M[(int(u), v)] == w
        allv = a.col[a.rowptr[v]:a.rowptr[v + 1]]
        np.testing.assert_array_equal(allv, np.concatenate([ins, o.col[o.rowptr[v]:o.rowptr[v + 1]]]))
        sym = s.col[s.rowptr[v]:s.rowptr[v + 1]]
        expect = sorted(set(ins.tolist()) | set(o.col[o.rowptr[v]:o.rowptr[v + 1]].tolist()) - {v})
        np.testing.assert_array_equal(sym, expect)
    assert i.m == o.m and a.m == 2 * o.m


def test_from_edges_last_weight_wins_and_keeps_loops():
    hg = G.HostGraph.from_edges(4, [0, 0, 2, 0, 3], [1, 1, 2, 3, 0], [0.5, 0.25, 0.75, 1.0, 2.0])
    o = hg.view()
    np.testing.assert_array_equal(o.rowptr, [0, 2, 2, 3, 4])
    np.testing.assert_array_equal(o.col, [1, 3, 2, 0])
    np.testing.assert_array_equal(o.w, [0.25, 1.0, 0.75, 2.0])


def test_from_edges_rejects_out_of_range():
    with pytest.raises(G.GhetError) as e:
        G.HostGraph.from_edges(3, [0, 5], [1, 1])
    assert e.value.code == G.ERR_INVALID


def test_graph_create_validates_before_touching_gpu():
    L = G.lib()
    h = C.c_void_p()
    rp = np.array([0, 2, 1], np.int32)  # non-monotone offsets
    col = np.array([1, 0], np.int32)
    st = L.ghet_graph_create(2, rp.ctypes.data_as(G.i32p), col.ctypes.data_as(G.i32p), None, None, None, None,
                             None, None, 0, C.byref(h))
    assert st == G.ERR_INVALID
    assert b"monotone" in L.ghet_last_error()
    rp = np.array([0, 1, 2], np.int32)
    col = np.array([1, 7], np.int32)  # column out of range
    st = L.ghet_graph_create(2, rp.ctypes.data_as(G.i32p), col.ctypes.data_as(G.i32p), None, None, None, None,
                             None, None, 0, C.byref(h))
    assert st == G.ERR_INVALID


@pytest.mark.skipif(G.device_count() > 0, reason="checks the no-GPU error path")
def test_graph_create_without_gpu_reports_nodevice():
    o = G.HostGraph.rmat(6, 4, seed=1).view()
    with pytest.raises(G.GhetError) as e:
        G.DeviceGraph(o)
    assert e.value.code == G.ERR_NODEVICE


def test_null_handle_is_rejected():
    L = G.lib()
    st = L.ghet_bfs(None, 0, None, None, None)
    assert st == G.ERR_INVALID
    assert L.ghet_sched_load(b"/nonexistent/weights.txt") == G.WARN_FALLBACK
