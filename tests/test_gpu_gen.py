"""On-GPU R-MAT generation (ghet_graph_gen_rmat): every CSR of the generated
graph equals, bit for bit, what the host generator plus ghet_graph_create
give for the same parameters; results of the algorithms then match too."""
import numpy as np
import pytest

import graphhet_amd as G
import oracle as O

pytestmark = pytest.mark.gpu

CASES = [
    (1, 16, 0.57, 0.19, 0.19, 1),     # two vertices: only loops and one pair
    (3, 4, 0.57, 0.19, 0.19, 2),
    (10, 16, 0.57, 0.19, 0.19, 3),
    (12, 5.5, 0.45, 0.25, 0.15, 4),   # fractional edge factor, other quadrants
    (14, 32, 0.6, 0.15, 0.15, 5),
    (18, 16, 0.57, 0.19, 0.19, 1),
]


def _eq(a, b, what):
    np.testing.assert_array_equal(a.rowptr, b.rowptr, err_msg=what + " offsets")
    np.testing.assert_array_equal(a.col, b.col, err_msg=what + " columns")
    if b.w is not None:
        np.testing.assert_array_equal(a.w, b.w, err_msg=what + " weights")


@pytest.mark.parametrize("case", CASES)
def test_device_rmat_equals_host_rmat(case):
    scale, ef, a, b, c, seed = case
    print(f"\ngen case scale={scale} ef={ef} a={a} b={b} c={c} seed={seed}", flush=True)
    hg = G.HostGraph.rmat(scale, ef, a, b, c, seed=seed)
    dg = G.DeviceGraph.rmat(scale, ef, a, b, c, seed=seed)
    _eq(dg.download(G.VIEW_OUT), hg.view(G.VIEW_OUT), "out")
    _eq(dg.download(G.VIEW_IN), hg.view(G.VIEW_IN), "in")
    _eq(dg.download(G.VIEW_ALL), hg.view(G.VIEW_ALL), "all")
    dg.close()


def test_flags_and_algorithms_on_a_generated_graph():
    hg = G.HostGraph.rmat(13, 16, seed=7)
    o, i = hg.view(G.VIEW_OUT), hg.view(G.VIEW_IN)
    dg = G.DeviceGraph.rmat(13, 16, seed=7)
    src = int(np.argmax(o.degrees()))
    np.testing.assert_array_equal(dg.bfs(src)[0], O.bfs(o.rowptr, o.col, src))
    np.testing.assert_array_equal(dg.sssp(src)[0], O.sssp(o.rowptr, o.col, o.w, src))
    np.testing.assert_array_equal(dg.cdlp(3)[0], O.cdlp(i.rowptr, i.col, o.rowptr, o.col, 3))
    ref = O.pr(i.rowptr, i.col, o.rowptr, 0.85, 10)
    assert np.abs(dg.pr(10)[0] - ref).sum() <= 1e-6
    lean = G.DeviceGraph.rmat(13, 16, seed=7, weights=False, with_all=False)
    p = lean.device_ptrs()
    assert p.m_all == 0 and not p.out_w and p.m_in == p.m_out == o.m
    _eq(lean.download(G.VIEW_IN, weights=False), G.Csr(i.n, i.rowptr, i.col, None), "in (no weights)")
    with pytest.raises(G.GhetError):
        lean.download(G.VIEW_ALL)
    np.testing.assert_array_equal(lean.bfs(src)[0], O.bfs(o.rowptr, o.col, src))
