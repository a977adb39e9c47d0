"""Python binding of libghet.so, the MI355X CSR graph-analytics library.

Thin ctypes layer over the C ABI declared in include/ghet.h.  Used by the
tests and bench.py; the product itself is the shared library and the C++
drop-in headers under include/.

The library is loaded from graphhet_amd/lib/libghet.so.  Loading fails with
an ImportError-like GhetError when it is missing: there is no fallback.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libghet.so")

OK, ERR_INVALID, ERR_NOMEM, ERR_HIP, ERR_NODEVICE, ERR_UNSUPPORTED, ERR_IO = 0, 1, 2, 3, 4, 5, 6
WARN_FALLBACK = 100

GRAPH_UNDIRECTED, GRAPH_NO_WEIGHTS, GRAPH_NO_IN, GRAPH_NO_ALL = 1, 2, 4, 8
VIEW_OUT, VIEW_IN, VIEW_ALL, VIEW_SYM = 0, 1, 2, 3
SCHED_HEURISTIC, SCHED_PUSH, SCHED_PULL, SCHED_BITMAP, SCHED_REFERENCE, SCHED_LEARNED = range(6)
MODE_PUSH, MODE_PULL, MODE_BITMAP = 0, 1, 2
NET_BFS, NET_SSSP = 0, 1
TRACE_FEATURES = 16
MLP_IN, MLP_H, MLP_OUT, MLP_PARAMS = 16, 64, 4, 5508
INT_INF = 2**31 - 1
DBL_MAX = np.finfo(np.float64).max


class GhetError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ghet status {code}: {msg}")
        self.code = code


class CsrView(C.Structure):
    _fields_ = [("n", C.c_int32), ("m", C.c_int64), ("rowptr", C.POINTER(C.c_int32)),
                ("col", C.POINTER(C.c_int32)), ("w", C.POINTER(C.c_double))]


class LevelTrace(C.Structure):
    _fields_ = [("features", C.c_float * TRACE_FEATURES), ("mode", C.c_int32), ("ms", C.c_float),
                ("threshold", C.c_double)]


class Opts(C.Structure):
    _fields_ = [("size", C.c_uint32), ("schedule", C.c_int32), ("sssp_delta", C.c_double),
                ("stream", C.c_void_p), ("trace", C.POINTER(LevelTrace)), ("trace_cap", C.c_int32),
                ("block_threads", C.c_int32), ("reserved", C.c_int32 * 6)]


class RunStats(C.Structure):
    _fields_ = [("device_ms", C.c_double), ("hot_kernel_ms", C.c_double), ("hot_kernel_launches", C.c_int32),
                ("levels", C.c_int32), ("edges_traversed", C.c_int64), ("vertices_reached", C.c_int64),
                ("trace_len", C.c_int32), ("status_flags", C.c_int32), ("extra", C.c_double * 4)]

    def as_dict(self) -> dict:
        return {"device_ms": self.device_ms, "hot_kernel_ms": self.hot_kernel_ms,
                "hot_kernel_launches": self.hot_kernel_launches, "levels": self.levels,
                "edges_traversed": self.edges_traversed, "vertices_reached": self.vertices_reached,
                "trace_len": self.trace_len, "status_flags": self.status_flags, "extra": list(self.extra)}


class DevicePtrs(C.Structure):
    _fields_ = [("n", C.c_int32), ("m_out", C.c_int64), ("m_in", C.c_int64), ("m_all", C.c_int64),
                ("out_rowptr", C.c_void_p), ("out_col", C.c_void_p), ("in_rowptr", C.c_void_p),
                ("in_col", C.c_void_p), ("all_rowptr", C.c_void_p), ("all_col", C.c_void_p),
                ("out_w", C.c_void_p), ("in_w", C.c_void_p)]


# Every symbol include/ghet.h declares (checked by the CPU tests).
EXPORTS = [
    "ghet_last_error", "ghet_abi_version", "ghet_device_count", "ghet_device_mem_info",
    "ghet_gen_rmat", "ghet_gen_er", "ghet_host_graph_from_edges", "ghet_host_graph_view", "ghet_host_graph_free",
    "ghet_host_graph_save", "ghet_host_graph_load", "ghet_host_graph_from_csr",
    "ghet_graph_create", "ghet_graph_wrap_device", "ghet_graph_destroy", "ghet_graph_device_ptrs",
    "ghet_graph_device_bytes", "ghet_graph_gen_rmat", "ghet_graph_download", "ghet_shutdown",
    "ghet_graph_set_workspace_cap",
    "ghet_dyn_create", "ghet_dyn_from_graph", "ghet_dyn_update", "ghet_dyn_edges", "ghet_dyn_snapshot",
    "ghet_dyn_destroy",
    "ghet_bfs", "ghet_bfs_pre", "ghet_bfs_noalloc", "ghet_sssp", "ghet_sssp_pre", "ghet_wcc", "ghet_pr", "ghet_pr_gather_ceiling",
    "ghet_cdlp", "ghet_tc",
    "ghet_bfs_multi", "ghet_sssp_multi", "ghet_wcc_many", "ghet_tc_many",
    "ghet_sched_load", "ghet_sched_save", "ghet_sched_clear", "ghet_sched_set_params", "ghet_sched_get_params",
    "ghet_mlp_forward", "ghet_mlp_forward_host", "ghet_mlp_train_step", "ghet_autotune",
]

_lib = None

P = C.POINTER
i32p, i64p, f64p, f32p = P(C.c_int32), P(C.c_int64), P(C.c_double), P(C.c_float)


def lib() -> C.CDLL:
    """Loads libghet.so (raises GhetError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GhetError(ERR_UNSUPPORTED, f"{LIB_PATH} not built: run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "ghet_last_error": (C.c_char_p, []),
        "ghet_abi_version": (C.c_int, []),
        "ghet_device_count": (C.c_int, []),
        "ghet_device_mem_info": (C.c_int, [P(C.c_size_t), P(C.c_size_t)]),
        "ghet_gen_rmat": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_int,
                                    P(vp)]),
        "ghet_gen_er": (C.c_int, [C.c_int32, C.c_int64, C.c_uint64, C.c_int, P(vp)]),
        "ghet_host_graph_from_edges": (C.c_int, [C.c_int32, C.c_int64, i32p, i32p, f64p, C.c_int, P(vp)]),
        "ghet_host_graph_save": (C.c_int, [vp, C.c_char_p]),
        "ghet_host_graph_from_csr": (C.c_int, [C.c_int32, i32p, i32p, f64p, C.c_int, P(vp)]),
        "ghet_graph_gen_rmat": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                          C.c_uint32, P(vp)]),
        "ghet_graph_download": (C.c_int, [vp, C.c_int, i32p, i32p, f64p]),
        "ghet_dyn_create": (C.c_int, [C.c_int32, i32p, i32p, f64p, C.c_int, P(vp)]),
        "ghet_dyn_from_graph": (C.c_int, [vp, P(vp)]),
        "ghet_dyn_update": (C.c_int, [vp, C.c_int64, i32p, i32p, f64p, vp]),
        "ghet_dyn_edges": (C.c_int64, [vp]),
        "ghet_dyn_snapshot": (C.c_int, [vp, C.c_uint32, P(vp)]),
        "ghet_dyn_destroy": (None, [vp]),
        "ghet_host_graph_load": (C.c_int, [C.c_char_p, C.c_int, P(vp)]),
        "ghet_host_graph_view": (C.c_int, [vp, C.c_int, P(CsrView)]),
        "ghet_host_graph_free": (None, [vp]),
        "ghet_graph_create": (C.c_int, [C.c_int32, i32p, i32p, f64p, i32p, i32p, f64p, i32p, i32p, C.c_uint32,
                                        P(vp)]),
        "ghet_graph_wrap_device": (C.c_int, [C.c_int32, C.c_int64, vp, vp, vp, C.c_int64, vp, vp, C.c_int64, vp, vp,
                                             C.c_uint32, P(vp)]),
        "ghet_graph_destroy": (None, [vp]),
        "ghet_shutdown": (C.c_int, []),
        "ghet_graph_device_ptrs": (C.c_int, [vp, P(DevicePtrs)]),
        "ghet_graph_device_bytes": (C.c_int64, [vp]),
        "ghet_graph_set_workspace_cap": (C.c_int, [vp, C.c_int64]),
        "ghet_bfs": (C.c_int, [vp, C.c_int32, i32p, P(Opts), P(RunStats)]),
        "ghet_bfs_pre": (C.c_int, [vp, C.c_int32, i32p, i32p, P(Opts), P(RunStats)]),
        "ghet_bfs_noalloc": (C.c_int, [C.c_int32, C.c_int64, vp, vp, C.c_int32, vp, vp]),
        "ghet_sssp": (C.c_int, [vp, C.c_int32, C.c_double, f64p, P(Opts), P(RunStats)]),
        "ghet_sssp_pre": (C.c_int, [vp, C.c_int32, C.c_double, f64p, i32p, P(Opts), P(RunStats)]),
        "ghet_wcc": (C.c_int, [vp, i32p, P(Opts), P(RunStats)]),
        "ghet_pr": (C.c_int, [vp, C.c_int32, C.c_double, f64p, P(Opts), P(RunStats)]),
        "ghet_pr_gather_ceiling": (C.c_int, [vp, C.c_int32, P(Opts), f64p]),
        "ghet_cdlp": (C.c_int, [vp, C.c_int32, i32p, P(Opts), P(RunStats)]),
        "ghet_tc": (C.c_int, [vp, i64p, P(Opts), P(RunStats)]),
        "ghet_bfs_multi": (C.c_int, [vp, C.c_int32, i32p, i32p, P(Opts), P(RunStats)]),
        "ghet_sssp_multi": (C.c_int, [vp, C.c_int32, i32p, C.c_double, f64p, P(Opts), P(RunStats)]),
        "ghet_wcc_many": (C.c_int, [C.c_int32, i32p, i32p, i32p, i32p, P(Opts), P(RunStats)]),
        "ghet_tc_many": (C.c_int, [C.c_int32, i32p, i32p, i32p, i64p, P(Opts), P(RunStats)]),
        "ghet_sched_load": (C.c_int, [C.c_char_p]),
        "ghet_sched_save": (C.c_int, [C.c_char_p]),
        "ghet_sched_clear": (C.c_int, []),
        "ghet_sched_set_params": (C.c_int, [C.c_int, f32p]),
        "ghet_sched_get_params": (C.c_int, [C.c_int, f32p]),
        "ghet_mlp_forward": (C.c_int, [C.c_int, C.c_int32, f32p, f32p]),
        "ghet_mlp_forward_host": (C.c_int, [C.c_int, C.c_int32, f32p, f32p]),
        "ghet_mlp_train_step": (C.c_int, [C.c_int, C.c_int32, f32p, i32p, C.c_float, f32p, f32p]),
        "ghet_autotune": (C.c_int, [vp, C.c_char_p, i32p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


# Handles record the generation they were made in; shutdown() starts a new
# one, after which older handles are never passed to the library again.
_GEN = [0]


def shutdown() -> None:
    """Releases every graph, dynamic store and schedule network the library
    holds on the GPU (ghet_shutdown).  Runs at interpreter exit, before the
    HIP runtime and any attached profiler are torn down, so nothing is freed
    after them.  Objects made before the call are closed; the library stays
    usable."""
    if _lib is not None:
        _GEN[0] += 1
        _lib.ghet_shutdown()


atexit.register(shutdown)


def _check(code: int, allow_fallback: bool = True) -> int:
    if code == OK or (allow_fallback and code == WARN_FALLBACK):
        return code
    msg = lib().ghet_last_error()
    raise GhetError(code, msg.decode() if msg else "")


def _ptr(a: np.ndarray | None, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(P(ctype))


def device_count() -> int:
    return lib().ghet_device_count()


@dataclass
class Csr:
    """CSR arrays (numpy, int32 offsets/indices, optional float64 weights)."""
    n: int
    rowptr: np.ndarray
    col: np.ndarray
    w: np.ndarray | None = None

    @property
    def m(self) -> int:
        return int(self.rowptr[-1])

    def degrees(self) -> np.ndarray:
        return np.diff(self.rowptr)


class HostGraph:
    """A host graph built by the library (R-MAT, ER or an edge list)."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)

    @classmethod
    def rmat(cls, scale: int, edge_factor: float = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
             seed: int = 1, threads: int = 0) -> "HostGraph":
        h = C.c_void_p()
        _check(lib().ghet_gen_rmat(scale, edge_factor, a, b, c, seed, threads, C.byref(h)))
        return cls(h.value)

    @classmethod
    def er(cls, n: int, m: int, seed: int = 1, threads: int = 0) -> "HostGraph":
        h = C.c_void_p()
        _check(lib().ghet_gen_er(n, m, seed, threads, C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_edges(cls, n: int, src, dst, w=None, threads: int = 0) -> "HostGraph":
        src = np.ascontiguousarray(src, dtype=np.int32)
        dst = np.ascontiguousarray(dst, dtype=np.int32)
        wa = None if w is None else np.ascontiguousarray(w, dtype=np.float64)
        h = C.c_void_p()
        _check(lib().ghet_host_graph_from_edges(n, len(src), _ptr(src, C.c_int32), _ptr(dst, C.c_int32),
                                                _ptr(wa, C.c_double), threads, C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_csr(cls, csr: Csr, threads: int = 0) -> "HostGraph":
        """From CSR arrays (e.g. DeviceGraph.download()); canonical rows are adopted."""
        rp = np.ascontiguousarray(csr.rowptr, dtype=np.int32)
        col = np.ascontiguousarray(csr.col, dtype=np.int32)
        w = None if csr.w is None else np.ascontiguousarray(csr.w, dtype=np.float64)
        h = C.c_void_p()
        _check(lib().ghet_host_graph_from_csr(csr.n, _ptr(rp, C.c_int32), _ptr(col, C.c_int32), _ptr(w, C.c_double),
                                              threads, C.byref(h)))
        return cls(h.value)

    @classmethod
    def load(cls, path: str, threads: int = 0) -> "HostGraph":
        """Reads a binary CSR cache (`path` targets, `path`.w weights)."""
        h = C.c_void_p()
        _check(lib().ghet_host_graph_load(os.fsencode(path), threads, C.byref(h)))
        return cls(h.value)

    def save(self, path: str) -> None:
        """Writes the out-adjacency as a binary CSR cache (`path`, `path`.w)."""
        _check(lib().ghet_host_graph_save(self._h, os.fsencode(path)))

    @classmethod
    def rmat_cached(cls, cache_dir: str | None, scale: int, edge_factor: float = 16, a: float = 0.57,
                    b: float = 0.19, c: float = 0.19, seed: int = 1, threads: int = 0) -> "HostGraph":
        """R-MAT through a binary cache in `cache_dir` (generated and saved on
        a miss; no cache when cache_dir is None or empty)."""
        if not cache_dir:
            return cls.rmat(scale, edge_factor, a, b, c, seed, threads)
        name = f"rmat-s{scale}-ef{edge_factor:g}-a{a:g}-b{b:g}-c{c:g}-seed{seed}.csr"
        path = os.path.join(cache_dir, name)
        if os.path.exists(path) and os.path.exists(path + ".w"):
            return cls.load(path, threads)
        g = cls.rmat(scale, edge_factor, a, b, c, seed, threads)
        os.makedirs(cache_dir, exist_ok=True)
        g.save(path)
        return g

    def _borrow(self, ptr, count: int, ctype, dtype) -> np.ndarray:
        """numpy array over library-owned memory that keeps this graph alive
        for as long as the array (or any view of it) exists."""
        buf = (ctype * count).from_address(C.cast(ptr, C.c_void_p).value)
        buf._owner = self
        return np.frombuffer(buf, dtype=dtype)

    def view(self, kind: int = VIEW_OUT, copy: bool = True) -> Csr:
        v = CsrView()
        _check(lib().ghet_host_graph_view(self._h, kind, C.byref(v)))
        n, m = v.n, v.m
        rp = self._borrow(v.rowptr, n + 1, C.c_int32, np.int32) if n >= 0 else np.zeros(1, np.int32)
        col = self._borrow(v.col, m, C.c_int32, np.int32) if m > 0 else np.zeros(0, np.int32)
        w = None
        if v.w and m > 0:
            w = self._borrow(v.w, m, C.c_double, np.float64)
        elif v.w and m == 0:
            w = np.zeros(0, np.float64)
        if copy:
            rp, col = rp.copy(), col.copy()
            w = None if w is None else w.copy()
        return Csr(n, rp, col, w)

    @property
    def n(self) -> int:
        return self.view(VIEW_OUT, copy=False).n

    def close(self):
        if self._h and self._h.value:
            lib().ghet_host_graph_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _opts(schedule=SCHED_HEURISTIC, sssp_delta=0.0, trace=None, flags=0, stream=None) -> Opts:
    """stream: a hipStream_t as an integer (e.g. torch.cuda.Stream().cuda_stream)
    or None for the graph's own stream."""
    o = Opts()
    o.size = C.sizeof(Opts)
    o.schedule = schedule
    o.sssp_delta = sssp_delta
    o.stream = stream
    if trace is not None:
        o.trace = C.cast(trace, P(LevelTrace))
        o.trace_cap = len(trace)
    o.reserved[0] = flags
    return o


def _heavy_flags(heavy) -> int:
    """opts reserved[0] bits of the SSSP heavy-pass form."""
    if heavy not in (None, "push", "pull"):
        raise ValueError(f"heavy {heavy!r}")
    return {None: 0, "push": 8, "pull": 16}[heavy]


class DeviceGraph:
    """A graph resident on the GPU (ghet_graph handle)."""

    def __init__(self, out: Csr, inn: Csr | None = None, all_: Csr | None = None, undirected: bool = False,
                 weights: bool = True, with_in: bool = True, with_all: bool = True):
        self.n = out.n
        flags = 0
        if undirected:
            flags |= GRAPH_UNDIRECTED
        if not weights:
            flags |= GRAPH_NO_WEIGHTS
        if not with_in:
            flags |= GRAPH_NO_IN
        if not with_all:
            flags |= GRAPH_NO_ALL
        keep = []

        def arr(a, dt):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a

        orp, ocol, ow = arr(out.rowptr, np.int32), arr(out.col, np.int32), arr(out.w if weights else None,
                                                                              np.float64)
        irp = icol = iw = arp = acol = None
        if inn is not None:
            irp, icol, iw = arr(inn.rowptr, np.int32), arr(inn.col, np.int32), arr(inn.w if weights else None,
                                                                                   np.float64)
        if all_ is not None:
            arp, acol = arr(all_.rowptr, np.int32), arr(all_.col, np.int32)
        h = C.c_void_p()
        _check(lib().ghet_graph_create(out.n, _ptr(orp, C.c_int32), _ptr(ocol, C.c_int32), _ptr(ow, C.c_double),
                                       _ptr(irp, C.c_int32), _ptr(icol, C.c_int32), _ptr(iw, C.c_double),
                                       _ptr(arp, C.c_int32), _ptr(acol, C.c_int32), flags, C.byref(h)))
        self._h = h
        self._gen = _GEN[0]

    @classmethod
    def from_host(cls, hg: HostGraph, sym: bool = False, weights: bool = True, with_all: bool = True):
        if sym:
            return cls(hg.view(VIEW_SYM, copy=False), undirected=True, weights=False)
        out = hg.view(VIEW_OUT, copy=False)
        inn = hg.view(VIEW_IN, copy=False)
        all_ = hg.view(VIEW_ALL, copy=False) if with_all else None
        return cls(out, inn, all_, weights=weights, with_all=with_all)

    @classmethod
    def rmat(cls, scale: int, edge_factor: float = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
             seed: int = 1, weights: bool = True, with_in: bool = True, with_all: bool = True) -> "DeviceGraph":
        """R-MAT generated on the GPU: the graph of HostGraph.rmat with the same
        arguments, bit for bit, without a host copy."""
        flags = (0 if weights else GRAPH_NO_WEIGHTS) | (0 if with_in else GRAPH_NO_IN) | \
                (0 if with_all else GRAPH_NO_ALL)
        h = C.c_void_p()
        _check(lib().ghet_graph_gen_rmat(scale, edge_factor, a, b, c, seed, flags, C.byref(h)))
        self = cls.__new__(cls)
        self._h = h
        self._gen = _GEN[0]
        self.n = 1 << scale
        return self

    def download(self, kind: int = VIEW_OUT, weights: bool = True) -> Csr:
        """Copies a device CSR (VIEW_OUT, VIEW_IN or VIEW_ALL) to host arrays."""
        p = self.device_ptrs()
        m = {VIEW_OUT: p.m_out, VIEW_IN: p.m_in, VIEW_ALL: p.m_all}[kind]
        rp = np.empty(p.n + 1, np.int32)
        col = np.empty(m, np.int32)
        has_w = weights and kind != VIEW_ALL and bool(p.out_w if kind == VIEW_OUT else p.in_w)
        w = np.empty(m, np.float64) if has_w else None
        _check(lib().ghet_graph_download(self._h, kind, _ptr(rp, C.c_int32), _ptr(col, C.c_int32),
                                         _ptr(w, C.c_double)))
        return Csr(p.n, rp, col, w)

    @property
    def handle(self):
        return self._h

    def device_ptrs(self) -> DevicePtrs:
        p = DevicePtrs()
        _check(lib().ghet_graph_device_ptrs(self._h, C.byref(p)))
        return p

    def device_bytes(self) -> int:
        return lib().ghet_graph_device_bytes(self._h)

    def set_workspace_cap(self, max_bytes: int) -> None:
        """Caps the graph's workspace (0 removes the cap); calls that would
        need more raise GhetError(ERR_NOMEM)."""
        _check(lib().ghet_graph_set_workspace_cap(self._h, max_bytes))

    def close(self):
        if self._h and self._h.value:
            if getattr(self, "_gen", -1) == _GEN[0]:
                lib().ghet_graph_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- runs --
    def bfs(self, src: int, schedule=SCHED_HEURISTIC, out=True, trace_cap=0, stream=None):
        depth = np.empty(self.n, np.int32) if out else None
        st = RunStats()
        tr = (LevelTrace * trace_cap)() if trace_cap else None
        _check(lib().ghet_bfs(self._h, src, _ptr(depth, C.c_int32), C.byref(_opts(schedule, trace=tr, stream=stream)),
                              C.byref(st)))
        if trace_cap:
            return depth, st, [tr[i] for i in range(st.trace_len)]
        return depth, st

    def bfs_pre(self, src: int, schedule=SCHED_HEURISTIC):
        depth = np.empty(self.n, np.int32)
        pre = np.empty(self.n, np.int32)
        st = RunStats()
        _check(lib().ghet_bfs_pre(self._h, src, _ptr(depth, C.c_int32), _ptr(pre, C.c_int32),
                                  C.byref(_opts(schedule)), C.byref(st)))
        return depth, pre, st

    def sssp(self, src: int, inf: float = DBL_MAX, delta: float = 0.0, schedule=SCHED_HEURISTIC, out=True,
             trace_cap=0, stream=None, heavy=None):
        """heavy: None (automatic), "push" or "pull" -- how a bucket's heavy
        pass runs: over the bucket's out-edges, or over the in-edges of the
        unsettled vertices (needs the in-CSR; "pull" takes it for every
        heavy pass).  stats.extra[3] = heavy passes run as pulls."""
        dist = np.empty(self.n, np.float64) if out else None
        st = RunStats()
        tr = (LevelTrace * trace_cap)() if trace_cap else None
        _check(lib().ghet_sssp(self._h, src, inf, _ptr(dist, C.c_double),
                               C.byref(_opts(schedule, sssp_delta=delta, trace=tr, stream=stream,
                                             flags=_heavy_flags(heavy))), C.byref(st)))
        if trace_cap:
            return dist, st, [tr[i] for i in range(st.trace_len)]
        return dist, st

    def sssp_pre(self, src: int, inf: float = DBL_MAX, delta: float = 0.0, heavy=None):
        dist = np.empty(self.n, np.float64)
        pre = np.empty(self.n, np.int32)
        st = RunStats()
        _check(lib().ghet_sssp_pre(self._h, src, inf, _ptr(dist, C.c_double), _ptr(pre, C.c_int32),
                                   C.byref(_opts(sssp_delta=delta, flags=_heavy_flags(heavy))), C.byref(st)))
        return dist, pre, st

    def wcc(self, out=True):
        lab = np.empty(self.n, np.int32) if out else None
        st = RunStats()
        _check(lib().ghet_wcc(self._h, _ptr(lab, C.c_int32), C.byref(_opts()), C.byref(st)))
        return lab, st

    def pr(self, iters: int = 10, damping: float = 0.85, normalise: bool = True, out=True, split=None,
           phases=False, form=None):
        """form: None (automatic), "gather" (pull gathers over the
        degree-relabeled in-CSR) or "binned" (destination bins over source
        windows; stats.extra[2] = its gather launches' ms, extra[3] = 2).
        split = (log2 hot sources, log2 window[, blocked]) forces the gather
        form's source-window layout, with the source-blocked hot phase when
        blocked is true and per-edge gathers otherwise.  phases: run the
        gather form's hot and cold gathers as separate launches and report
        the hot phase's ms in stats.extra[2]."""
        if form not in (None, "gather", "binned"):
            raise ValueError(f"form {form!r}")
        if form == "binned" and split is not None:
            raise ValueError("split applies to the gather form")
        r = np.empty(self.n, np.float64) if out else None
        st = RunStats()
        o = _opts(flags=(0 if normalise else 1) | (2 if phases else 0))
        if split is not None:
            blocked = len(split) > 2 and bool(split[2])
            o.reserved[4] = int(split[0]) | (int(split[1]) << 8) | ((2 if blocked else 1) << 16)
        elif form == "binned":
            o.reserved[4] = 3 << 16
        elif form == "gather":
            o.reserved[4] = 1 << 18
        _check(lib().ghet_pr(self._h, iters, damping, _ptr(r, C.c_double), C.byref(o), C.byref(st)))
        return r, st

    def pr_gather_ceiling(self, reps: int = 5, split=None):
        """Median ms of a pure-gather kernel (sum of contrib[col[e]], no rows)
        over the PR layout's hot edges and its cold edges, and the two edge
        counts: {"hot_ms", "cold_ms", "hot_edges", "cold_edges"}."""
        o = _opts()
        if split is not None:
            blocked = len(split) > 2 and bool(split[2])
            o.reserved[4] = int(split[0]) | (int(split[1]) << 8) | ((2 if blocked else 1) << 16)
        r = np.zeros(4, np.float64)
        _check(lib().ghet_pr_gather_ceiling(self._h, reps, C.byref(o), _ptr(r, C.c_double)))
        return {"hot_ms": float(r[0]), "cold_ms": float(r[1]), "hot_edges": int(r[2]), "cold_edges": int(r[3])}

    def cdlp(self, iters: int = 10, out=True, small_merge_table=False):
        """small_merge_table: cap the hub merge tables at 64 slots (a test
        setting that sends the merge through its sub-pass path)."""
        lab = np.empty(self.n, np.int32) if out else None
        st = RunStats()
        o = _opts(flags=4 if small_merge_table else 0)
        _check(lib().ghet_cdlp(self._h, iters, _ptr(lab, C.c_int32), C.byref(o), C.byref(st)))
        return lab, st

    def tc(self, out=True):
        cnt = np.empty(self.n, np.int64) if out else None
        st = RunStats()
        _check(lib().ghet_tc(self._h, _ptr(cnt, C.c_int64), C.byref(_opts()), C.byref(st)))
        return cnt, st

    def bfs_multi(self, srcs, out=True, schedule=SCHED_HEURISTIC, push_div=0, trace_cap=0):
        """Depths [k, n] of k sources in one batched launch; with trace_cap,
        also the first pass's per-level records (mode, features[0] = frontier
        out-edges / m, device ms)."""
        srcs = np.ascontiguousarray(srcs, dtype=np.int32)
        depth = np.empty((len(srcs), self.n), np.int32) if out else None
        st = RunStats()
        tr = (LevelTrace * trace_cap)() if trace_cap else None
        o = _opts(schedule, trace=tr)
        o.reserved[2] = push_div
        _check(lib().ghet_bfs_multi(self._h, len(srcs), _ptr(srcs, C.c_int32), _ptr(depth, C.c_int32),
                                    C.byref(o), C.byref(st)))
        if trace_cap:
            return depth, st, [tr[i] for i in range(st.trace_len)]
        return depth, st

    def autotune(self, cache_path: str | None = None) -> int:
        best = C.c_int32()
        _check(lib().ghet_autotune(self._h, cache_path.encode() if cache_path else None, C.byref(best)))
        return best.value

    def sssp_multi(self, srcs, inf: float = DBL_MAX, delta: float = 0.0, out=True, per_pass: int = 0,
                   phase_budget: int = 0):
        srcs = np.ascontiguousarray(srcs, dtype=np.int32)
        dist = np.empty((len(srcs), self.n), np.float64) if out else None
        st = RunStats()
        o = _opts(sssp_delta=delta)
        o.reserved[3] = per_pass
        o.reserved[5] = phase_budget
        _check(lib().ghet_sssp_multi(self._h, len(srcs), _ptr(srcs, C.c_int32), inf, _ptr(dist, C.c_double),
                                     C.byref(o), C.byref(st)))
        return dist, st


class DynGraph:
    """A dynamic graph on the GPU (ghet_dyn handle): batched inserts, weight
    changes and deletes (weight 0), and CSR snapshots for the analytics."""

    def __init__(self, handle, n: int):
        self._h = handle
        self._gen = _GEN[0]
        self.n = n

    @classmethod
    def create(cls, csr: Csr, threads: int = 0) -> "DynGraph":
        rp = np.ascontiguousarray(csr.rowptr, dtype=np.int32)
        col = np.ascontiguousarray(csr.col, dtype=np.int32)
        w = None if csr.w is None else np.ascontiguousarray(csr.w, dtype=np.float64)
        h = C.c_void_p()
        _check(lib().ghet_dyn_create(csr.n, _ptr(rp, C.c_int32), _ptr(col, C.c_int32), _ptr(w, C.c_double), threads,
                                     C.byref(h)))
        return cls(h, csr.n)

    @classmethod
    def from_graph(cls, dg: "DeviceGraph") -> "DynGraph":
        h = C.c_void_p()
        _check(lib().ghet_dyn_from_graph(dg.handle, C.byref(h)))
        return cls(h, dg.n)

    def update(self, src, dst, w=None) -> RunStats:
        src = np.ascontiguousarray(src, dtype=np.int32)
        dst = np.ascontiguousarray(dst, dtype=np.int32)
        wa = None if w is None else np.ascontiguousarray(w, dtype=np.float64)
        st = RunStats()
        _check(lib().ghet_dyn_update(self._h, len(src), _ptr(src, C.c_int32), _ptr(dst, C.c_int32),
                                     _ptr(wa, C.c_double), C.byref(st)))
        return st

    def delete(self, src, dst) -> RunStats:
        return self.update(src, dst, np.zeros(len(src)))

    @property
    def m(self) -> int:
        return lib().ghet_dyn_edges(self._h)

    def snapshot(self, weights: bool = True, with_in: bool = True, with_all: bool = True) -> "DeviceGraph":
        flags = (0 if weights else GRAPH_NO_WEIGHTS) | (0 if with_in else GRAPH_NO_IN) | \
                (0 if with_all else GRAPH_NO_ALL)
        h = C.c_void_p()
        _check(lib().ghet_dyn_snapshot(self._h, flags, C.byref(h)))
        g = DeviceGraph.__new__(DeviceGraph)
        g._h = h
        g._gen = _GEN[0]
        g.n = self.n
        return g

    def close(self):
        if self._h and self._h.value:
            if getattr(self, "_gen", -1) == _GEN[0]:
                lib().ghet_dyn_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pack_graphs(csrs: list[Csr]):
    """Concatenates small out-CSRs (local column ids) for the many-graph calls."""
    voff = np.zeros(len(csrs) + 1, np.int32)
    for i, g in enumerate(csrs):
        voff[i + 1] = voff[i] + g.n
    rowptr = np.zeros(int(voff[-1]) + 1, np.int64)
    cols = []
    e = 0
    for i, g in enumerate(csrs):
        rowptr[voff[i]:voff[i + 1] + 1] = g.rowptr.astype(np.int64) + e
        e += g.m
        cols.append(g.col.astype(np.int32))
    col = np.concatenate(cols) if cols else np.zeros(0, np.int32)
    return voff, rowptr.astype(np.int32), col


def _many_opts(tune: bool) -> Opts:
    o = _opts()
    o.reserved[1] = 0 if tune else 1
    return o


def wcc_many(csrs: list[Csr], tune: bool = True):
    voff, rowptr, col = pack_graphs(csrs)
    lab = np.empty(int(voff[-1]), np.int32)
    st = RunStats()
    _check(lib().ghet_wcc_many(len(csrs), _ptr(voff, C.c_int32), _ptr(rowptr, C.c_int32), _ptr(col, C.c_int32),
                               _ptr(lab, C.c_int32), C.byref(_many_opts(tune)), C.byref(st)))
    return [lab[voff[i]:voff[i + 1]] for i in range(len(csrs))], st


def tc_many(csrs: list[Csr], tune: bool = True):
    voff, rowptr, col = pack_graphs(csrs)
    cnt = np.empty(int(voff[-1]), np.int64)
    st = RunStats()
    _check(lib().ghet_tc_many(len(csrs), _ptr(voff, C.c_int32), _ptr(rowptr, C.c_int32), _ptr(col, C.c_int32),
                              _ptr(cnt, C.c_int64), C.byref(_many_opts(tune)), C.byref(st)))
    return [cnt[voff[i]:voff[i + 1]] for i in range(len(csrs))], st


# --- learned schedule networks (16 -> 64 -> 64 -> 4, FP16 weights) -----------

NET_BFS, NET_SSSP = 0, 1
MLP_IN, MLP_H, MLP_OUT = 16, 64, 4
MLP_PARAMS = MLP_H * MLP_IN + MLP_H + MLP_H * MLP_H + MLP_H + MLP_OUT * MLP_H + MLP_OUT


def mlp_init(seed: int = 0) -> np.ndarray:
    """He-initialised FP32 parameters in the library's layout
    (W1[64][16], b1, W2[64][64], b2, W3[4][64], b3)."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((MLP_IN, MLP_H), (MLP_H, MLP_H), (MLP_H, MLP_OUT)):
        parts.append((rng.standard_normal((fan_out, fan_in)) * np.sqrt(2.0 / fan_in)).ravel())
        parts.append(np.zeros(fan_out))
    return np.concatenate(parts).astype(np.float32)


def sched_set_params(kind: int, params: np.ndarray) -> None:
    p = np.ascontiguousarray(params, dtype=np.float32)
    if p.size != MLP_PARAMS:
        raise ValueError(f"expected {MLP_PARAMS} parameters, got {p.size}")
    _check(lib().ghet_sched_set_params(kind, _ptr(p, C.c_float)), allow_fallback=False)


def sched_get_params(kind: int) -> np.ndarray:
    p = np.empty(MLP_PARAMS, np.float32)
    _check(lib().ghet_sched_get_params(kind, _ptr(p, C.c_float)), allow_fallback=False)
    return p


def sched_save(path: str) -> None:
    _check(lib().ghet_sched_save(path.encode()), allow_fallback=False)


def sched_load(path: str) -> bool:
    """Installs the networks in `path`; False (heuristic stays) if unusable."""
    return _check(lib().ghet_sched_load(path.encode())) == OK


def sched_clear() -> None:
    _check(lib().ghet_sched_clear())


def mlp_forward(kind: int, x: np.ndarray, host: bool = False) -> np.ndarray:
    """Logits [B,4] of the installed network (GPU kernel, or the host mirror)."""
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, MLP_IN)
    y = np.empty((x.shape[0], MLP_OUT), np.float32)
    fn = lib().ghet_mlp_forward_host if host else lib().ghet_mlp_forward
    _check(fn(kind, x.shape[0], _ptr(x, C.c_float), _ptr(y, C.c_float)), allow_fallback=False)
    return y


def mlp_train_step(kind: int, x: np.ndarray, labels: np.ndarray, lr: float, want_grad: bool = False):
    """One Adam step on the GPU; returns (mean loss, gradient or None)."""
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, MLP_IN)
    lab = np.ascontiguousarray(labels, dtype=np.int32)
    loss = C.c_float()
    g = np.empty(MLP_PARAMS, np.float32) if want_grad else None
    _check(lib().ghet_mlp_train_step(kind, x.shape[0], _ptr(x, C.c_float), _ptr(lab, C.c_int32), lr,
                                     C.byref(loss), _ptr(g, C.c_float)), allow_fallback=False)
    return float(loss.value), g
