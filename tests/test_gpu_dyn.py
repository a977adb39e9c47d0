"""Dynamic graph store (ghet_dyn_*): after every batch of inserts, weight
changes and deletes, the snapshot's out-, in- and all-CSR equal the oracle's
edge set bit for bit (oracle/dyn.py, GPMA+ semantics), and the analytics on
the snapshot equal the CPU oracle on that edge set."""
import numpy as np
import pytest

import graphhet_amd as G
import oracle as O
from oracle.dyn import EdgeDict

pytestmark = pytest.mark.gpu


def _check_snapshot(dyn, ref, what):
    g = dyn.snapshot()
    for kind, rev in ((G.VIEW_OUT, False), (G.VIEW_IN, True)):
        rp, col, w = ref.csr(reverse=rev)
        got = g.download(kind)
        np.testing.assert_array_equal(got.rowptr, rp, err_msg=f"{what}: offsets ({kind})")
        np.testing.assert_array_equal(got.col, col, err_msg=f"{what}: targets ({kind})")
        np.testing.assert_array_equal(got.w, w, err_msg=f"{what}: weights ({kind})")
    a = g.download(G.VIEW_ALL)
    orp, ocol, _ = ref.csr()
    irp, icol, _ = ref.csr(reverse=True)
    exp = np.concatenate([np.concatenate([icol[irp[v]:irp[v + 1]], ocol[orp[v]:orp[v + 1]]])
                          for v in range(ref.n)]) if ref.e else np.zeros(0, np.int32)
    np.testing.assert_array_equal(a.col, exp, err_msg=f"{what}: all-CSR")
    assert dyn.m == len(ref.e)
    return g


def _random_batch(rng, ref, n, size):
    """Inserts of new pairs, weight changes and deletes of present edges,
    deletes of absent pairs, repeated keys and self-loops."""
    present = list(ref.e.keys())
    src, dst, w = [], [], []
    for _ in range(size):
        t = rng.random()
        if t < 0.45 or not present:
            u, v = int(rng.integers(0, n)), int(rng.integers(0, n))
            src.append(u); dst.append(v); w.append(float(rng.integers(1, 1 << 20)) / (1 << 20))
        elif t < 0.65:
            u, v = present[int(rng.integers(0, len(present)))]
            src.append(u); dst.append(v); w.append(float(rng.integers(1, 1 << 20)) / (1 << 20))
        elif t < 0.9:
            u, v = present[int(rng.integers(0, len(present)))]
            src.append(u); dst.append(v); w.append(0.0)
        else:
            src.append(int(rng.integers(0, n))); dst.append(int(rng.integers(0, n))); w.append(0.0)
    # a repeated key with different outcomes inside the batch
    if src:
        src += [src[0], src[0]]; dst += [dst[0], dst[0]]; w += [0.0, 0.625]
    return np.array(src, np.int32), np.array(dst, np.int32), np.array(w)


def test_batches_match_oracle_after_every_batch():
    hg = G.HostGraph.rmat(11, 8, seed=21)
    o = hg.view(G.VIEW_OUT)
    ref = EdgeDict(o.n, o.rowptr, o.col, o.w)
    dyn = G.DynGraph.create(o)
    _check_snapshot(dyn, ref, "initial")
    rng = np.random.default_rng(5)
    for b, size in enumerate([1, 17, 1000, 5000, 300, 20000, 64]):
        print(f"\ndyn batch {b} size {size}", flush=True)
        s, d, w = _random_batch(rng, ref, o.n, size)
        c = ref.apply(s, d, w)
        st = dyn.update(s, d, w)
        assert (st.extra[0], st.extra[1], st.extra[2]) == (c["inserted"], c["deleted"], c["modified"]), b
        assert st.edges_traversed == c["distinct"]
        _check_snapshot(dyn, ref, f"batch {b}")
    dyn.close()


def test_row_growth_relocation_and_pool_compaction():
    """One row grows far past its slack (relocations), then the whole graph
    grows past the pool (compaction); deleting everything empties it."""
    hg = G.HostGraph.rmat(8, 4, seed=3)
    o = hg.view(G.VIEW_OUT)
    ref = EdgeDict(o.n, o.rowptr, o.col, o.w)
    dyn = G.DynGraph.create(o)
    n = o.n
    compactions = 0
    for k in range(6):  # vertex 7 gains edges to every vertex, in growing chunks
        vs = np.arange(k * 40, min(n, (k + 1) * 40), dtype=np.int32)
        us = np.full(len(vs), 7, np.int32)
        ref.apply(us, vs, np.full(len(vs), 0.5))
        st = dyn.update(us, vs, np.full(len(vs), 0.5))
        compactions += st.levels
        _check_snapshot(dyn, ref, f"growth {k}")
    rng = np.random.default_rng(9)
    for k in range(3):  # several times the edge count: forces pool compaction
        s = rng.integers(0, n, 4 * len(ref.e) + 100).astype(np.int32)
        d = rng.integers(0, n, len(s)).astype(np.int32)
        w = np.full(len(s), 0.25)
        ref.apply(s, d, w)
        st = dyn.update(s, d, w)
        compactions += st.levels
        _check_snapshot(dyn, ref, f"bulk {k}")
    assert compactions > 0
    keys = list(ref.e.keys())
    s = np.array([k[0] for k in keys], np.int32)
    d = np.array([k[1] for k in keys], np.int32)
    ref.apply(s, d, np.zeros(len(s)))
    dyn.delete(s, d)
    assert dyn.m == 0
    _check_snapshot(dyn, ref, "emptied")


def test_analytics_on_snapshots_match_oracle():
    hg = G.HostGraph.rmat(10, 8, seed=4)
    o = hg.view(G.VIEW_OUT)
    ref = EdgeDict(o.n, o.rowptr, o.col, o.w)
    dyn = G.DynGraph.create(o)
    rng = np.random.default_rng(11)
    for b in range(3):
        s, d, w = _random_batch(rng, ref, o.n, 2000)
        ref.apply(s, d, w)
        dyn.update(s, d, w)
        g = dyn.snapshot()
        orp, ocol, ow = ref.csr()
        irp, icol, _ = ref.csr(reverse=True)
        src = int(np.argmax(np.diff(orp)))
        np.testing.assert_array_equal(g.bfs(src)[0], O.bfs(orp, ocol, src))
        np.testing.assert_array_equal(g.sssp(src)[0], O.sssp(orp, ocol, ow, src))
        np.testing.assert_array_equal(g.wcc()[0], O.wcc(orp, ocol))
        err = np.abs(g.pr(10)[0] - O.pr(irp, icol, orp, 0.85, 10))
        assert err.sum() <= 1e-6 and err.max() <= 1e-9
        np.testing.assert_array_equal(g.cdlp(5)[0], O.cdlp(irp, icol, orp, ocol, 5))
        np.testing.assert_array_equal(g.tc()[0], O.tc(orp, ocol))


def test_from_device_graph_and_no_op_batches():
    dg = G.DeviceGraph.rmat(12, 8, seed=2)
    o = dg.download(G.VIEW_OUT)
    ref = EdgeDict(o.n, o.rowptr, o.col, o.w)
    dyn = G.DynGraph.from_graph(dg)
    _check_snapshot(dyn, ref, "from graph")
    st = dyn.update(np.zeros(0, np.int32), np.zeros(0, np.int32))
    assert st.edges_traversed == 0
    absent = [(u, v) for u in range(5) for v in range(5) if (u, v) not in ref.e][:5]
    s = np.array([a for a, _ in absent], np.int32)
    d = np.array([b for _, b in absent], np.int32)
    st = dyn.delete(s, d)  # deletes of absent edges change nothing
    assert st.extra[1] == 0 and dyn.m == len(ref.e)
    _check_snapshot(dyn, ref, "no-op deletes")
    with pytest.raises(G.GhetError):
        dyn.update(np.array([0], np.int32), np.array([o.n], np.int32))
    with pytest.raises(G.GhetError):
        dyn.update(np.array([0], np.int32), np.array([1], np.int32), np.array([-1.0]))


def test_gpma_drop_in_driver():
    """graphhet_amd/drivers/GPU_gpma_example.cpp: the reference's GPMA example
    on the drop-in headers (include/GPU_gpma), then batches of update_gpma
    checked against the CSR drop-in on a graph_structure with the same
    changes, for all six analytics."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(G.LIB_PATH), "..", "bin", "GPU_gpma_example")
    if not os.path.exists(exe):
        pytest.skip("drivers not built")
    out = subprocess.run([exe, "--check"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "gpma selftest ok" in out.stdout
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and "Community Detection result:" in out.stdout, out.stdout + out.stderr
