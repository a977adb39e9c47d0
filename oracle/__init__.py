"""ORACLE -- test infrastructure only.

ctypes access to the CPU restatement (libghet_oracle.so) and, where it was
built, to GraphHET's own CPU algorithms (_ref/libghet_ref.so).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module;
the product (graphhet_amd) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libghet_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libghet_ref.so")

P = C.POINTER
i32p, i64p, f64p = P(C.c_int32), P(C.c_int64), P(C.c_double)
I32, F64 = C.c_int32, C.c_double


def build(ref_dir: str | None = None) -> None:
    """Builds the restatement (and _ref when the reference sources exist)."""
    args = ["make", "-B", "-C", HERE]
    if ref_dir:
        args.append(f"REF={ref_dir}")
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)


def _load(path: str, sigs: dict) -> C.CDLL:
    L = C.CDLL(path)
    for name, args in sigs.items():
        fn = getattr(L, name)
        fn.restype = None
        fn.argtypes = args
    return L


_oracle = None
_ref = None


def lib() -> C.CDLL:
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            build()
        _oracle = _load(ORACLE_SO, {
            "oracle_bfs": [I32, i32p, i32p, I32, i32p],
            "oracle_bfs_parent": [I32, i32p, i32p, i32p, i32p],
            "oracle_sssp": [I32, i32p, i32p, f64p, I32, f64p],
            "oracle_sssp_parent": [I32, i32p, i32p, f64p, f64p, F64, I32, i32p],
            "oracle_wcc": [I32, i32p, i32p, I32, i32p],
            "oracle_pr": [I32, i32p, i32p, i32p, F64, I32, I32, f64p],
            "oracle_pr_step": [I32, i32p, i32p, i32p, F64, I32, f64p, f64p],
            "oracle_cdlp": [I32, i32p, i32p, i32p, i32p, I32, I32, i32p],
            "oracle_tc": [I32, i32p, i32p, I32, I32, i64p],
            "oracle_tc_forward": [I32, i32p, i32p, I32, i64p],
            "oracle_tc_vertices": [I32, i32p, i32p, i32p, I32, i64p],
        })
        _oracle.oracle_tc_vertices.restype = C.c_int64
    return _oracle


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref() -> C.CDLL:
    global _ref
    if _ref is None:
        _ref = _load(REF_SO, {
            "ref_bfs": [I32, i32p, i32p, I32, i32p],
            "ref_sssp": [I32, i32p, i32p, f64p, I32, f64p],
            "ref_wcc": [I32, i32p, i32p, i32p, i32p, i32p],
            "ref_pr": [I32, i32p, i32p, i32p, i32p, F64, I32, f64p],
            "ref_cdlp": [I32, i32p, i32p, i32p, i32p, I32, i32p],
            "ref_tc": [I32, i32p, i32p, i64p],
        })
        _ref.ref_graph_create.restype = C.c_void_p
        _ref.ref_graph_create.argtypes = [I32, i32p, i32p, f64p, i32p, i32p, f64p]
        _ref.ref_graph_free.restype = None
        _ref.ref_graph_free.argtypes = [C.c_void_p]
        for name, args in (("ref_graph_bfs", [C.c_void_p, I32, i32p]), ("ref_graph_sssp", [C.c_void_p, I32, f64p]),
                           ("ref_graph_pr", [C.c_void_p, F64, I32, f64p]), ("ref_graph_wcc", [C.c_void_p, i32p]),
                           ("ref_graph_cdlp", [C.c_void_p, I32, i32p])):
            fn = getattr(_ref, name)
            fn.restype = C.c_double
            fn.argtypes = args
    return _ref


def _a(x, dt):
    return np.ascontiguousarray(x, dtype=dt)


def _p(a, ct):
    return a.ctypes.data_as(P(ct))


def hw_threads() -> int:
    return os.cpu_count() or 1


# ------------------------------------------------------------ restatement --
def bfs(rowptr, col, src: int) -> np.ndarray:
    rp, c = _a(rowptr, np.int32), _a(col, np.int32)
    n = len(rp) - 1
    out = np.empty(n, np.int32)
    lib().oracle_bfs(n, _p(rp, I32), _p(c, I32), src, _p(out, I32))
    return out


def bfs_parent(in_rowptr, in_col, depth) -> np.ndarray:
    rp, c, d = _a(in_rowptr, np.int32), _a(in_col, np.int32), _a(depth, np.int32)
    n = len(rp) - 1
    out = np.empty(n, np.int32)
    lib().oracle_bfs_parent(n, _p(rp, I32), _p(c, I32), _p(d, I32), _p(out, I32))
    return out


def sssp(rowptr, col, w, src: int) -> np.ndarray:
    rp, c, ww = _a(rowptr, np.int32), _a(col, np.int32), _a(w, np.float64)
    n = len(rp) - 1
    out = np.empty(n, np.float64)
    lib().oracle_sssp(n, _p(rp, I32), _p(c, I32), _p(ww, F64), src, _p(out, F64))
    return out


def sssp_parent(in_rowptr, in_col, in_w, dist, src: int, inf=np.finfo(np.float64).max) -> np.ndarray:
    rp, c, w, d = _a(in_rowptr, np.int32), _a(in_col, np.int32), _a(in_w, np.float64), _a(dist, np.float64)
    n = len(rp) - 1
    out = np.empty(n, np.int32)
    lib().oracle_sssp_parent(n, _p(rp, I32), _p(c, I32), _p(w, F64), _p(d, F64), inf, src, _p(out, I32))
    return out


def wcc(rowptr, col, threads: int = 0) -> np.ndarray:
    rp, c = _a(rowptr, np.int32), _a(col, np.int32)
    n = len(rp) - 1
    out = np.empty(n, np.int32)
    lib().oracle_wcc(n, _p(rp, I32), _p(c, I32), threads or hw_threads(), _p(out, I32))
    return out


def pr(in_rowptr, in_col, out_rowptr, damping=0.85, iters=10, threads: int = 100) -> np.ndarray:
    irp, ic, orp = _a(in_rowptr, np.int32), _a(in_col, np.int32), _a(out_rowptr, np.int32)
    n = len(irp) - 1
    out = np.empty(n, np.float64)
    lib().oracle_pr(n, _p(irp, I32), _p(ic, I32), _p(orp, I32), damping, iters, threads, _p(out, F64))
    return out


def pr_step(in_rowptr, in_col, out_rowptr, rank, damping=0.85, threads: int = 100) -> np.ndarray:
    """One PageRank iteration from `rank` (the oracle's operation order)."""
    irp, ic, orp = _a(in_rowptr, np.int32), _a(in_col, np.int32), _a(out_rowptr, np.int32)
    r = _a(rank, np.float64)
    n = len(irp) - 1
    out = np.empty(n, np.float64)
    lib().oracle_pr_step(n, _p(irp, I32), _p(ic, I32), _p(orp, I32), damping, threads, _p(r, F64), _p(out, F64))
    return out


def cdlp(in_rowptr, in_col, out_rowptr, out_col, iters=10, threads: int = 100) -> np.ndarray:
    irp, ic, orp, oc = (_a(in_rowptr, np.int32), _a(in_col, np.int32), _a(out_rowptr, np.int32),
                        _a(out_col, np.int32))
    n = len(irp) - 1
    out = np.empty(n, np.int32)
    lib().oracle_cdlp(n, _p(irp, I32), _p(ic, I32), _p(orp, I32), _p(oc, I32), iters, threads, _p(out, I32))
    return out


def tc(rowptr, col, v_begin: int = 0, v_end: int | None = None) -> np.ndarray:
    rp, c = _a(rowptr, np.int32), _a(col, np.int32)
    n = len(rp) - 1
    out = np.empty(n, np.int64)
    lib().oracle_tc(n, _p(rp, I32), _p(c, I32), v_begin, n if v_end is None else v_end, _p(out, C.c_int64))
    return out


def tc_vertices(rowptr, col, verts):
    """oracle_tc with its outer loop over `verts` only: (counts, merge steps)."""
    rp, c, vs = _a(rowptr, np.int32), _a(col, np.int32), _a(verts, np.int32)
    n = len(rp) - 1
    out = np.empty(n, np.int64)
    steps = lib().oracle_tc_vertices(n, _p(rp, I32), _p(c, I32), _p(vs, I32), len(vs), _p(out, C.c_int64))
    return out, int(steps)


def tc_forward(rowptr, col, threads: int = 0) -> np.ndarray:
    """t(v), triangles through v, of a symmetric loop-free graph (forward
    algorithm; 6 * t(v) is the reference's per-vertex convention)."""
    rp, c = _a(rowptr, np.int32), _a(col, np.int32)
    n = len(rp) - 1
    out = np.empty(n, np.int64)
    lib().oracle_tc_forward(n, _p(rp, I32), _p(c, I32), threads or min(16, hw_threads()), _p(out, C.c_int64))
    return out


# ------------------------------------------------- the reference itself --
def ref_bfs(rowptr, col, src):
    rp, c = _a(rowptr, np.int32), _a(col, np.int32)
    out = np.empty(len(rp) - 1, np.int32)
    ref().ref_bfs(len(rp) - 1, _p(rp, I32), _p(c, I32), src, _p(out, I32))
    return out


def ref_sssp(rowptr, col, w, src):
    rp, c, ww = _a(rowptr, np.int32), _a(col, np.int32), _a(w, np.float64)
    out = np.empty(len(rp) - 1, np.float64)
    ref().ref_sssp(len(rp) - 1, _p(rp, I32), _p(c, I32), _p(ww, F64), src, _p(out, F64))
    return out


def ref_wcc(out_rowptr, out_col, in_rowptr, in_col):
    a = [_a(x, np.int32) for x in (out_rowptr, out_col, in_rowptr, in_col)]
    out = np.empty(len(a[0]) - 1, np.int32)
    ref().ref_wcc(len(a[0]) - 1, *[_p(x, I32) for x in a], _p(out, I32))
    return out


def ref_pr(out_rowptr, out_col, in_rowptr, in_col, damping=0.85, iters=10):
    a = [_a(x, np.int32) for x in (out_rowptr, out_col, in_rowptr, in_col)]
    out = np.empty(len(a[0]) - 1, np.float64)
    ref().ref_pr(len(a[0]) - 1, *[_p(x, I32) for x in a], damping, iters, _p(out, F64))
    return out


def ref_cdlp(out_rowptr, out_col, in_rowptr, in_col, iters=10):
    a = [_a(x, np.int32) for x in (out_rowptr, out_col, in_rowptr, in_col)]
    out = np.empty(len(a[0]) - 1, np.int32)
    ref().ref_cdlp(len(a[0]) - 1, *[_p(x, I32) for x in a], iters, _p(out, I32))
    return out


def ref_tc(rowptr, col):
    rp, c = _a(rowptr, np.int32), _a(col, np.int32)
    out = np.empty(len(rp) - 1, np.int64)
    ref().ref_tc(len(rp) - 1, _p(rp, I32), _p(c, I32), _p(out, C.c_int64))
    return out


class RefGraph:
    """GraphHET's graph_structure<double>, built once, for timing the
    reference CPU algorithms without their construction cost."""

    def __init__(self, out_rowptr, out_col, out_w, in_rowptr=None, in_col=None, in_w=None):
        self._keep = [_a(out_rowptr, np.int32), _a(out_col, np.int32), _a(out_w, np.float64)]
        ins = [None, None, None]
        if in_rowptr is not None:
            ins = [_a(in_rowptr, np.int32), _a(in_col, np.int32), _a(in_w, np.float64)]
            self._keep += ins
        self.n = len(self._keep[0]) - 1
        cp = lambda a, ct: None if a is None else _p(a, ct)  # noqa: E731
        self._h = ref().ref_graph_create(self.n, cp(self._keep[0], I32), cp(self._keep[1], I32),
                                         cp(self._keep[2], F64), cp(ins[0], I32), cp(ins[1], I32), cp(ins[2], F64))

    def bfs(self, src):
        out = np.empty(self.n, np.int32)
        return ref().ref_graph_bfs(self._h, src, _p(out, I32)), out

    def sssp(self, src):
        out = np.empty(self.n, np.float64)
        return ref().ref_graph_sssp(self._h, src, _p(out, F64)), out

    def pr(self, damping, iters):
        out = np.empty(self.n, np.float64)
        return ref().ref_graph_pr(self._h, damping, iters, _p(out, F64)), out

    def wcc(self):
        out = np.empty(self.n, np.int32)
        return ref().ref_graph_wcc(self._h, _p(out, I32)), out

    def cdlp(self, iters):
        out = np.empty(self.n, np.int32)
        return ref().ref_graph_cdlp(self._h, iters, _p(out, I32)), out

    def close(self):
        if self._h:
            ref().ref_graph_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
