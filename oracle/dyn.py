"""ORACLE -- test infrastructure only.

The update semantics of GraphHET's GPMA+ store, restated over a Python dict
of edges: update_gpma (GPU_gpma.hpp:756-824) with handle_del_mod
(GPU_gpma.hpp:168-181) -- a batch entry with value VALUE_NONE (0) deletes the
edge, an entry whose key is present overwrites the value, any other entry is
inserted -- and the GPMA algorithms' view of the store (GPU_BFS_gpma.cuh:59:
slots with value 0 are skipped), i.e. the current edge set in (src, dst)
order.  Where GPMA leaves it undefined (the same key twice in one batch), the
last occurrence wins, which is the dynamic store's stated rule.

Pinned by tests/test_dyn_oracle.py (hand-computed sequences).
"""
from __future__ import annotations

import numpy as np


class EdgeDict:
    """Current edge set {(u, v): w} of an n-vertex directed graph."""

    def __init__(self, n: int, rowptr=None, col=None, w=None):
        self.n = n
        self.e: dict[tuple[int, int], float] = {}
        if rowptr is not None:
            for u in range(n):
                for k in range(int(rowptr[u]), int(rowptr[u + 1])):
                    wt = 1.0 if w is None else float(w[k])
                    if wt != 0.0:  # weight 0 is "no edge"
                        self.e[(u, int(col[k]))] = wt

    def apply(self, src, dst, w=None) -> dict:
        """One batch.  Returns the counts of inserts, deletes and weight changes."""
        last: dict[tuple[int, int], float] = {}
        for i in range(len(src)):
            last[(int(src[i]), int(dst[i]))] = 1.0 if w is None else float(w[i])
        ins = dele = mod = 0
        for key, wt in last.items():
            if wt == 0.0:
                if key in self.e:
                    del self.e[key]
                    dele += 1
            elif key in self.e:
                self.e[key] = wt
                mod += 1
            else:
                self.e[key] = wt
                ins += 1
        return {"inserted": ins, "deleted": dele, "modified": mod, "distinct": len(last)}

    def csr(self, reverse: bool = False):
        """(rowptr, col, w): rows by source (reverse: by target), ascending."""
        items = sorted(((v, u, wt) if reverse else (u, v, wt)) for (u, v), wt in self.e.items())
        rowptr = np.zeros(self.n + 1, np.int32)
        col = np.array([x[1] for x in items], np.int32)
        w = np.array([x[2] for x in items], np.float64)
        for r, _, _ in items:
            rowptr[r + 1] += 1
        np.cumsum(rowptr, out=rowptr)
        return rowptr, col, w
