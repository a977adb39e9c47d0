"""Parity at the bandwidth-scale configuration: R-MAT scale 22 (edge factor
16, a .57 b .19 c .19, seed 1; 4.19M vertices, about 65M directed edges),
generated on the GPU and downloaded for the checks.  All six algorithms run
on the one graph, as the reference's LDBC driver runs them on one dataset
(src/LDBC/LDBC_GPU_csr.cu:82-95), and each is compared with the CPU oracle
on the same input:

  BFS, SSSP, WCC, CDLP (10 iterations)   bit-exact;
  BFS / SSSP parents                     bit-exact against the parent rules;
  PR (damping 0.85, 10 iterations)       L1 <= 1e-6 against the FP64 oracle,
                                         max-abs <= 1e-9 (both forms);
  TC on the symmetrisation               6 t(v) for every vertex, t from the
                                         forward-algorithm counter.
BFS and SSSP start at the max-out-degree vertex (ties to the smallest id).
"""
import numpy as np
import pytest

import graphhet_amd as G
import oracle as O

pytestmark = pytest.mark.gpu

INT_MAX = 2**31 - 1
DBL_MAX = np.finfo(np.float64).max


@pytest.fixture(scope="module")
def s22():
    print("\n[s22] generating R-MAT scale 22 on the GPU", flush=True)
    dg = G.DeviceGraph.rmat(22, 16, seed=1)
    o, i = dg.download(G.VIEW_OUT), dg.download(G.VIEW_IN)
    src = int(np.argmax(o.degrees()))
    yield dg, o, i, src
    dg.close()


def test_s22_bfs_and_parents(s22):
    dg, o, i, src = s22
    ref = O.bfs(o.rowptr, o.col, src)
    for sched in (G.SCHED_HEURISTIC, G.SCHED_REFERENCE, G.SCHED_PUSH, G.SCHED_PULL):
        d, st = dg.bfs(src, schedule=sched)
        np.testing.assert_array_equal(d, ref, err_msg=f"schedule {sched}")
    assert st.edges_traversed == int(o.degrees()[ref != INT_MAX].sum())
    d, pre, _ = dg.bfs_pre(src)
    np.testing.assert_array_equal(d, ref)
    np.testing.assert_array_equal(pre, O.bfs_parent(i.rowptr, i.col, ref))


def test_s22_sssp_and_parents(s22):
    dg, o, i, src = s22
    ref = O.sssp(o.rowptr, o.col, o.w, src)
    for heavy in (None, "push", "pull"):
        d, _ = dg.sssp(src, heavy=heavy)
        np.testing.assert_array_equal(d, ref, err_msg=str(heavy))
    d, pre, _ = dg.sssp_pre(src)
    np.testing.assert_array_equal(d, ref)
    np.testing.assert_array_equal(pre, O.sssp_parent(i.rowptr, i.col, i.w, ref, src))


def test_s22_wcc(s22):
    dg, o, i, src = s22
    lab, _ = dg.wcc()
    np.testing.assert_array_equal(lab, O.wcc(o.rowptr, o.col, 16))


def test_s22_pr_10_iterations(s22):
    dg, o, i, src = s22
    ref = O.pr(i.rowptr, i.col, o.rowptr, 0.85, 10, threads=16)
    raw, _ = dg.pr(10, 0.85, normalise=False)
    nrm, _ = dg.pr(10, 0.85)
    err = np.abs(raw - ref)
    errn = np.abs(nrm - ref / ref.sum())
    print(f"\nPR s22 10 iterations: L1 {err.sum():.3e} max-abs {err.max():.3e}; "
          f"normalised L1 {errn.sum():.3e} max-abs {errn.max():.3e}")
    assert err.sum() <= 1e-6 and err.max() <= 1e-9
    assert errn.sum() <= 1e-6 and errn.max() <= 1e-9


def test_s22_cdlp_10_iterations(s22):
    dg, o, i, src = s22
    lab, _ = dg.cdlp(10)
    np.testing.assert_array_equal(lab, O.cdlp(i.rowptr, i.col, o.rowptr, o.col, 10, threads=16))


def test_s22_tc_every_vertex(s22):
    dg, o, i, src = s22
    s = G.HostGraph.from_csr(o).view(G.VIEW_SYM, copy=False)
    sg = G.DeviceGraph(s, undirected=True, weights=False)
    cnt, _ = sg.tc()
    sg.close()
    t = O.tc_forward(s.rowptr, s.col)
    np.testing.assert_array_equal(cnt, 6 * t)
    assert int(cnt.sum()) == 18 * (int(t.sum()) // 3)
    print(f"\nTC s22: {int(t.sum()) // 3} triangles, max per vertex {int(t.max())}")
