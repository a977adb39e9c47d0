"""The C++ drop-in headers (include/GPU_csr/...) end to end: a C++ program
built with a plain host compiler calls every GraphHET entry point (Cuda_*,
csr_bfs, bfs_pre, csr_sssp, csr_sssp_pre, WCC_GPU, GPU_PR, CDLP_GPU,
Cuda_TC_PerNode) through CSR_graph / toCSR; its output is compared with the
golden outputs of GraphHET's own CPU code."""
import os
import subprocess
from collections import defaultdict

import numpy as np
import pytest

from conftest import ROOT, load_golden
import oracle as O

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "tests", "bin")
EXAMPLE_BIN = os.path.join(ROOT, "graphhet_amd", "bin")
INT_INF = 2**31 - 1


def run_shim_check(g, tmp_path):
    n = int(g["n"])
    rp, col, w = g["out_rowptr"], g["out_col"], g["out_w"]
    path = tmp_path / "edges.txt"
    with open(path, "w") as f:
        f.write(f"{n}\n")
        for u in range(n):
            for e in range(rp[u], rp[u + 1]):
                f.write(f"{u} {col[e]} {float(w[e])!r}\n")
    exe = os.path.join(BIN, "shim_check")
    assert os.path.exists(exe), "tests/bin/shim_check not built"
    out = subprocess.run([exe, str(path), str(int(g["src"]))], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    # The drop-ins' own checks (the parent-chain walks of Cuda_BFS_pre and
    # Cuda_SSSP_pre) print to stdout; any such line fails the test.
    assert "Not found!" not in out.stdout, "a parent chain of Cuda_BFS_pre / Cuda_SSSP_pre is broken"
    assert "Error: pre_v is wrong" not in out.stdout, "a Cuda_SSSP_pre parent chain does not sum to its distance"
    res = defaultdict(dict)
    for line in out.stdout.splitlines():
        tag, v, val = line.split(" ", 2)
        res[tag][int(v)] = val
    return res, out.stderr


def as_array(d, n, conv):
    return np.array([conv(d[i]) for i in range(n)])


@pytest.mark.parametrize("name", ["gpu_example", "cpu_example", "rmat_s8", "loops_isolated_50", "star_9"])
def test_every_entry_point_matches_reference(name, tmp_path):
    g = load_golden(name)
    n, src = int(g["n"]), int(g["src"])
    res, err = run_shim_check(g, tmp_path)
    for tag in ("Cuda_BFS", "Cuda_BFS_pre", "csr_bfs", "bfs_pre"):
        np.testing.assert_array_equal(as_array(res[tag], n, int), g["bfs"], err_msg=tag)
    for tag in ("Cuda_SSSP", "Cuda_SSSP_pre", "csr_sssp", "csr_sssp_pre"):
        np.testing.assert_array_equal(as_array(res[tag], n, float), g["sssp"], err_msg=tag)
    exp_pre = O.bfs_parent(g["in_rowptr"], g["in_col"], g["bfs"])
    np.testing.assert_array_equal(as_array(res["Cuda_BFS_pre.pre"], n, int), exp_pre)
    exp_spre = O.sssp_parent(g["in_rowptr"], g["in_col"], g["in_w"], g["sssp"], src)
    np.testing.assert_array_equal(as_array(res["Cuda_SSSP_pre.pre"], n, int), exp_spre)
    for tag in ("Cuda_WCC", "WCC_GPU"):
        np.testing.assert_array_equal(as_array(res[tag], n, int), g["wcc"], err_msg=tag)
    for tag in ("Cuda_PR", "GPU_PR"):
        err = np.abs(as_array(res[tag], n, float) - g["pr"])
        assert err.sum() <= 1e-6 and err.max() <= 1e-7, tag
    for tag in ("Cuda_CDLP", "CDLP_GPU"):
        np.testing.assert_array_equal(as_array(res[tag], n, int), g["cdlp"], err_msg=tag)
    for tag in ("Cuda_TC", "Cuda_TC_PerNode"):
        np.testing.assert_array_equal(as_array(res[tag], n, int), g["tc_sym"], err_msg=tag)


def test_example_program_prints_reference_results():
    exe = os.path.join(EXAMPLE_BIN, "GPU_csr_example")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    bfs = lines[lines.index("BFS result:") + 1: lines.index("Connected Components result:")]
    # one->two/three/four/six at depth 1, five at depth 2 (GPU_csr_example.cu:25-50)
    assert dict(l.split() for l in bfs) == {"one": "0", "two": "1", "three": "1", "four": "1", "five": "2",
                                            "six": "1"}
    wcc = lines[lines.index("Connected Components result:") + 1: lines.index("SSSP result:")]
    assert {l.split()[1] for l in wcc} == {"one"}
    g = load_golden("gpu_example")
    sssp = lines[lines.index("SSSP result:") + 1: lines.index("PageRank result:")]
    names = ["one", "two", "three", "four", "five", "six"]
    got = {a: float(b) for a, b in (l.split() for l in sssp)}
    np.testing.assert_allclose([got[x] for x in names], g["sssp"], rtol=1e-6)
