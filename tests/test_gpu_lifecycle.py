"""Resource limits and teardown through the C ABI.

Workspace cap: a graph whose workspace is capped below what a call needs
returns GHET_ERR_NOMEM with the sizes in the message, runs nothing wrong,
and works again once the cap allows it (the reference instead prints a
failed cudaMalloc and carries on, GPU_csr.hpp:21-25).

Shutdown: ghet_shutdown releases every live graph, dynamic store and
schedule network; handles made before it are never used again, and the
library stays usable.  It runs in a child process so that it cannot touch
the objects of the other tests.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import graphhet_amd as G
import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def small():
    hg = G.HostGraph.rmat(12, 16, seed=3)
    o, i = hg.view(G.VIEW_OUT), hg.view(G.VIEW_IN)
    return hg, o, i, int(np.argmax(o.degrees()))


def _runs(dg, src):
    return {
        "bfs": lambda: dg.bfs(src)[0],
        "sssp": lambda: dg.sssp(src)[0],
        "wcc": lambda: dg.wcc()[0],
        "pr": lambda: dg.pr(5, 0.85)[0],
        "cdlp": lambda: dg.cdlp(3)[0],
    }


def test_workspace_cap_returns_nomem_cleanly(small):
    hg, o, i, src = small
    ref = {"bfs": O.bfs(o.rowptr, o.col, src), "sssp": O.sssp(o.rowptr, o.col, o.w, src),
           "wcc": O.wcc(o.rowptr, o.col), "cdlp": O.cdlp(i.rowptr, i.col, o.rowptr, o.col, 3)}
    pr_ref = O.pr(i.rowptr, i.col, o.rowptr, 0.85, 5)
    for name in ("bfs", "sssp", "wcc", "pr", "cdlp"):
        # what the call needs, from an uncapped run on a fresh graph
        probe = G.DeviceGraph.from_host(hg)
        arena = probe.device_bytes()
        _runs(probe, src)[name]()
        need = probe.device_bytes() - arena
        probe.close()
        dg = G.DeviceGraph.from_host(hg)
        for cap in (4096, max(4096, need - 4096)):
            dg.set_workspace_cap(cap)
            with pytest.raises(G.GhetError) as e:
                _runs(dg, src)[name]()
            assert e.value.code == G.ERR_NOMEM, (name, cap)
            assert "workspace cap" in str(e.value), str(e.value)
        dg.set_workspace_cap(need)
        got = _runs(dg, src)[name]()
        if name == "pr":
            assert np.abs(got - pr_ref / pr_ref.sum()).sum() <= 1e-6
        else:
            np.testing.assert_array_equal(got, ref[name], err_msg=name)
        dg.set_workspace_cap(0)
        dg.close()


CHILD = r"""
import numpy as np, sys
sys.path.insert(0, sys.argv[1])
import graphhet_amd as G, oracle as O
hg = G.HostGraph.rmat(10, 8, seed=2)
o = hg.view(G.VIEW_OUT)
a = G.DeviceGraph.from_host(hg)
b = G.DeviceGraph.from_host(hg)
d = G.DynGraph.create(o)
snap = d.snapshot()
G.sched_set_params(G.NET_BFS, G.mlp_init(1))
a.bfs(0); snap.wcc()
G.shutdown()
b.close(); a.close(); snap.close(); d.close()   # handles from before: never used again
c = G.DeviceGraph.from_host(hg)
dep, _ = c.bfs(5)
assert np.array_equal(dep, O.bfs(o.rowptr, o.col, 5))
G.sched_set_params(G.NET_BFS, G.mlp_init(2))
print("child ok", flush=True)
"""


def test_shutdown_releases_everything_and_library_stays_usable():
    out = subprocess.run([sys.executable, "-c", CHILD, ROOT], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "child ok" in out.stdout
