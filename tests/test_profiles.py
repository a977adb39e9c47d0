"""The committed PageRank traffic figures match the kernels they describe:
each profiles/pr_iteration__*.traffic.txt records the hash of the PR sources
it was measured with (bench.py reports its figure as stale otherwise), and
its byte total is the calibrated sum of its own counter values."""
import glob
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_for_tests", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _fields(path):
    kv = {}
    for line in open(path):
        if line.startswith("#"):
            continue
        for tok in line.split():
            if "=" in tok:
                k, v = tok.split("=", 1)
                kv[k] = v
    return kv


FILES = sorted(glob.glob(os.path.join(ROOT, "profiles", "pr_iteration__*.traffic.txt")))


def test_traffic_files_exist():
    assert FILES, "no committed PageRank traffic figure"


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(p) for p in FILES])
def test_traffic_file_is_current_and_consistent(path):
    kv = _fields(path)
    assert kv["pr_src_sha"] == _bench().pr_source_hash(), \
        "PR sources changed since this figure was measured: re-run its rocprofv3 passes"
    ff, wf = float(kv["fetch_factor"]), float(kv["write_factor"])
    kib = sum(ff * int(kv[f"{k}_fetch_kib"]) + wf * int(kv[f"{k}_write_kib"]) for k in ("k_prb_gather", "k_prb_accum"))
    assert int(kv["traffic_bytes"]) == round(kib * 1024)
    # the workload named in the file is the one in its name
    assert os.path.basename(path) == f"pr_iteration__{kv['workload']}.traffic.txt"
