"""Binary CSR cache (ghet_host_graph_save / ghet_host_graph_load) in the
vector-of-vectors layout of binary_save_read_vector_of_vectors.hpp:10-76:
an int32 row count, then per row an int32 length and its elements; targets
in `path`, weights in `path`.w.  The layout is checked byte for byte against
files written here with numpy, so writer and reader cannot agree on a wrong
layout between themselves."""
import os
import subprocess

import numpy as np
import pytest

import graphhet_amd as G

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "bin")


def _vv_bytes(rows, dtype):
    """The layout from its definition: int32 count, then (int32 len, elements)."""
    parts = [np.array([len(rows)], "<i4").tobytes()]
    for r in rows:
        parts.append(np.array([len(r)], "<i4").tobytes())
        parts.append(np.asarray(r, dtype).tobytes())
    return b"".join(parts)


def _rows(csr, arr):
    return [arr[csr.rowptr[u]:csr.rowptr[u + 1]] for u in range(csr.n)]


def _same(a, b):
    np.testing.assert_array_equal(a.rowptr, b.rowptr)
    np.testing.assert_array_equal(a.col, b.col)
    np.testing.assert_array_equal(a.w, b.w)


def test_save_writes_the_vector_of_vectors_layout(tmp_path):
    hg = G.HostGraph.rmat(9, 6, seed=4)
    o = hg.view(G.VIEW_OUT)
    p = str(tmp_path / "g.csr")
    hg.save(p)
    assert open(p, "rb").read() == _vv_bytes(_rows(o, o.col), "<i4")
    assert open(p + ".w", "rb").read() == _vv_bytes(_rows(o, o.w), "<f8")


def test_round_trip_is_bit_exact(tmp_path):
    for scale, ef, seed in ((1, 1, 1), (8, 4, 2), (12, 16, 3)):
        hg = G.HostGraph.rmat(scale, ef, seed=seed)
        p = str(tmp_path / f"g{scale}.csr")
        hg.save(p)
        _same(G.HostGraph.load(p).view(G.VIEW_OUT), hg.view(G.VIEW_OUT))
        # derived views come out the same as well
        _same(G.HostGraph.load(p, threads=3).view(G.VIEW_IN), hg.view(G.VIEW_IN))


def test_empty_graph_and_isolated_rows(tmp_path):
    p = str(tmp_path / "e.csr")
    open(p, "wb").write(_vv_bytes([[], [], []], "<i4"))
    g = G.HostGraph.load(p).view(G.VIEW_OUT)
    assert g.n == 3 and g.m == 0
    np.testing.assert_array_equal(g.rowptr, [0, 0, 0, 0])
    p0 = str(tmp_path / "z.csr")
    open(p0, "wb").write(_vv_bytes([], "<i4"))
    assert G.HostGraph.load(p0).view(G.VIEW_OUT).n == 0


def test_unsorted_rows_are_canonicalised_like_an_edge_list(tmp_path):
    rng = np.random.default_rng(5)
    n = 50
    rows = [list(rng.integers(0, n, rng.integers(0, 9))) for _ in range(n)]
    wts = [list(rng.random(len(r))) for r in rows]
    p = str(tmp_path / "u.csr")
    open(p, "wb").write(_vv_bytes(rows, "<i4"))
    open(p + ".w", "wb").write(_vv_bytes(wts, "<f8"))
    src = np.concatenate([np.full(len(r), u) for u, r in enumerate(rows)]).astype(np.int32)
    dst = np.concatenate([np.asarray(r, np.int32) for r in rows])
    w = np.concatenate([np.asarray(x, np.float64) for x in wts])
    _same(G.HostGraph.load(p).view(G.VIEW_OUT), G.HostGraph.from_edges(n, src, dst, w).view(G.VIEW_OUT))


def test_missing_weight_file_gives_unit_weights(tmp_path):
    p = str(tmp_path / "nw.csr")
    open(p, "wb").write(_vv_bytes([[1], [0, 2], []], "<i4"))
    g = G.HostGraph.load(p).view(G.VIEW_OUT)
    np.testing.assert_array_equal(g.w, [1.0, 1.0, 1.0])


@pytest.mark.parametrize("case", ["missing", "truncated", "trailing", "negative_len", "bad_target",
                                  "weights_mismatch", "negative_count"])
def test_malformed_files_raise_io_error(tmp_path, case):
    p = str(tmp_path / "bad.csr")
    good = _vv_bytes([[1, 2], [0], []], "<i4")
    if case == "truncated":
        open(p, "wb").write(good[:-3])
    elif case == "trailing":
        open(p, "wb").write(good + b"\0")
    elif case == "negative_len":
        open(p, "wb").write(np.array([1, -2], "<i4").tobytes())
    elif case == "bad_target":
        open(p, "wb").write(_vv_bytes([[1, 3], [0], []], "<i4"))
    elif case == "weights_mismatch":
        open(p, "wb").write(good)
        open(p + ".w", "wb").write(_vv_bytes([[1.0], [0.5, 0.5], []], "<f8"))
    elif case == "negative_count":
        open(p, "wb").write(np.array([-1], "<i4").tobytes())
    with pytest.raises(G.GhetError) as e:
        G.HostGraph.load(p)
    assert e.value.code == G.ERR_IO
    assert "bad.csr" in str(e.value)


def test_rmat_cached_generates_once_then_loads(tmp_path):
    d = str(tmp_path / "cache")
    a = G.HostGraph.rmat_cached(d, 10, 8, seed=9)
    files = sorted(os.listdir(d))
    assert len(files) == 2 and files[0].endswith(".csr") and files[1].endswith(".csr.w")
    b = G.HostGraph.rmat_cached(d, 10, 8, seed=9)
    _same(a.view(G.VIEW_OUT), b.view(G.VIEW_OUT))
    _same(b.view(G.VIEW_OUT), G.HostGraph.rmat(10, 8, seed=9).view(G.VIEW_OUT))


def test_header_and_library_interoperate(tmp_path):
    """The C++ header's writer/reader and the library's cache agree
    (tests/drivers/cache_selftest.cpp)."""
    exe = os.path.join(BIN, "cache_selftest")
    if not os.path.exists(exe):
        pytest.skip("drivers not built")
    out = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "cache selftest ok" in out.stdout


def test_device_generator_rejects_bad_arguments_on_the_host():
    """Argument checks run before any device work (so they hold without a GPU)."""
    for args in ((0, 16, 0.57, 0.19, 0.19, 1, 0), (10, -1, 0.57, 0.19, 0.19, 1, 0),
                 (10, 16, 0.7, 0.2, 0.2, 1, 0), (10, 16, 0.57, 0.19, 0.19, 1, G.GRAPH_UNDIRECTED)):
        h = G.C.c_void_p()
        assert G.lib().ghet_graph_gen_rmat(*args, G.C.byref(h)) == G.ERR_INVALID


def test_from_csr_adopts_canonical_rows_and_canonicalises_others():
    hg = G.HostGraph.rmat(11, 8, seed=6)
    o = hg.view(G.VIEW_OUT)
    _same(G.HostGraph.from_csr(o).view(G.VIEW_OUT), o)
    _same(G.HostGraph.from_csr(o, threads=2).view(G.VIEW_SYM), hg.view(G.VIEW_SYM))
    # unsorted row with a repeat: row 0 = [2, 1, 2] -> [1, 2] with the later weight of 2
    c = G.Csr(3, np.array([0, 3, 3, 4], np.int32), np.array([2, 1, 2, 0], np.int32), np.array([0.1, 0.2, 0.3, 0.4]))
    g = G.HostGraph.from_csr(c).view(G.VIEW_OUT)
    np.testing.assert_array_equal(g.rowptr, [0, 2, 2, 3])
    np.testing.assert_array_equal(g.col, [1, 2, 0])
    np.testing.assert_array_equal(g.w, [0.2, 0.3, 0.4])
    with pytest.raises(G.GhetError):
        G.HostGraph.from_csr(G.Csr(2, np.array([0, 1, 1], np.int32), np.array([5], np.int32), None))
