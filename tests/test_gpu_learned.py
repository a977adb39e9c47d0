"""Learned schedule on the GPU: forward/backward/Adam kernels against the FP64
oracle, the host mirror against the GPU forward (bit-exact), schedule files,
bit-exact BFS/SSSP results under any network, and the gain of a trained
network over the heuristic on a held-out graph."""
import os
import tempfile

import numpy as np
import pytest

import graphhet_amd as G
from graphhet_amd import learn
from oracle import mlp as M

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _clean_nets():
    G.sched_clear()
    yield
    G.sched_clear()


def _rand_params(seed, scale=0.3):
    return (np.random.default_rng(seed).standard_normal(G.MLP_PARAMS) * scale).astype(np.float32)


def test_forward_gpu_equals_host_mirror_and_oracle():
    p = _rand_params(0)
    G.sched_set_params(G.NET_BFS, p)
    x = np.random.default_rng(1).uniform(0, 1, size=(300, 16)).astype(np.float32)
    y_gpu = G.mlp_forward(G.NET_BFS, x)
    y_host = G.mlp_forward(G.NET_BFS, x, host=True)
    assert np.array_equal(y_gpu.view(np.uint32), y_host.view(np.uint32))
    z, _ = M.forward(M.fp16(p), x)
    np.testing.assert_allclose(y_gpu, z, rtol=1e-4, atol=1e-5)


def test_train_step_gradient_matches_oracle():
    rng = np.random.default_rng(2)
    p = _rand_params(2)
    G.sched_set_params(G.NET_SSSP, p)
    x = rng.uniform(0, 1, size=(200, 16)).astype(np.float32)  # 3 full chunks and a partial one
    y = rng.integers(0, 4, size=200).astype(np.int32)
    loss, g = G.mlp_train_step(G.NET_SSSP, x, y, lr=0.0, want_grad=True)
    g_ref = M.grad(M.fp16(p), x.astype(np.float64), y)
    assert abs(loss - M.loss(M.fp16(p), x, y)) < 1e-5
    rel = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    assert rel < 1e-3, rel
    o = 0
    for s in M.SHAPES:  # every layer on its own
        k = int(np.prod(s))
        gl, rl = g[o:o + k], g_ref[o:o + k]
        assert np.linalg.norm(gl - rl) <= 1e-3 * max(np.linalg.norm(rl), 1e-6), s
        o += k
    np.testing.assert_array_equal(G.sched_get_params(G.NET_SSSP), p)  # lr = 0 leaves the master


def test_adam_steps_match_oracle():
    rng = np.random.default_rng(3)
    p = _rand_params(3)
    G.sched_set_params(G.NET_BFS, p)
    x = rng.uniform(0, 1, size=(64, 16)).astype(np.float32)
    y = rng.integers(0, 3, size=64).astype(np.int32)
    w = p.astype(np.float64)
    m1 = np.zeros_like(w)
    m2 = np.zeros_like(w)
    for step in (1, 2, 3):
        _, g = G.mlp_train_step(G.NET_BFS, x, y, lr=1e-3, want_grad=True)
        w, m1, m2 = M.adam_step(w, m1, m2, g.astype(np.float64), step, 1e-3)
        np.testing.assert_allclose(G.sched_get_params(G.NET_BFS), w, rtol=0, atol=2e-6)


def test_training_learns_a_separable_rule():
    rng = np.random.default_rng(4)
    x = rng.uniform(0, 1, size=(4096, 16)).astype(np.float32)
    y = np.argmax(x[:, :3] * np.array([1.0, 1.2, 0.8]), axis=1).astype(np.int32)
    hist = learn.train(G.NET_BFS, x, y, steps=400, lr=3e-3, batch=256, seed=0)
    assert hist[-1] < 0.5 * hist[0]
    assert learn.accuracy(G.NET_BFS, x, y, 3) > 0.9


def test_schedule_file_round_trip_and_fallback():
    p0, p1 = _rand_params(5), _rand_params(6)
    G.sched_set_params(G.NET_BFS, p0)
    G.sched_set_params(G.NET_SSSP, p1)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "sched.txt")
        G.sched_save(path)
        G.sched_clear()
        assert G.sched_load(path)
        np.testing.assert_array_equal(G.sched_get_params(G.NET_BFS), p0)
        np.testing.assert_array_equal(G.sched_get_params(G.NET_SSSP), p1)
        bad = os.path.join(d, "bad.txt")
        with open(bad, "w") as f:
            f.write("ghet-sched 1\nnet 0 5508\n1.0\n2.0\n")
        G.sched_clear()
        assert not G.sched_load(bad)  # truncated: heuristic stays
        assert not G.sched_load(os.path.join(d, "missing.txt"))
        with pytest.raises(G.GhetError):
            G.sched_get_params(G.NET_BFS)


def _graph(scale, seed, a=0.57, ef=16):
    hg = G.HostGraph.rmat(scale, ef, *learn.rmat_params(a), seed=seed)
    return hg, G.DeviceGraph.from_host(hg, with_all=False)


def test_learned_schedule_results_are_bit_exact_under_any_network():
    hg, dg = _graph(14, 11)
    srcs = learn.pick_sources(hg, 6, 3)
    base_bfs = [dg.bfs(int(s))[0] for s in srcs]
    base_sssp = [dg.sssp(int(s))[0] for s in srcs[:3]]
    # no network: LEARNED falls back to the heuristic and says so
    d, st = dg.bfs(int(srcs[0]), schedule=G.SCHED_LEARNED)
    assert st.status_flags & 1 and np.array_equal(d, base_bfs[0])
    for seed in range(4):  # arbitrary networks pick arbitrary modes / widths
        G.sched_set_params(G.NET_BFS, _rand_params(100 + seed, scale=2.0))
        G.sched_set_params(G.NET_SSSP, _rand_params(200 + seed, scale=2.0))
        for s, ref in zip(srcs, base_bfs):
            d, st = dg.bfs(int(s), schedule=G.SCHED_LEARNED)
            assert st.status_flags & 1 == 0
            assert np.array_equal(d, ref)
        for s, ref in zip(srcs[:3], base_sssp):
            d, _ = dg.sssp(int(s), schedule=G.SCHED_LEARNED)
            assert np.array_equal(d, ref)


def test_trained_network_gain_on_held_out_graph():
    """Trains on scale 16-18 traces, evaluates on a held-out scale-20 graph
    (larger than any training graph; other seed, a and edge factor).  Prints
    the gains; results must be identical and the learned BFS must beat the
    heuristic (median of three timings of the eight sources)."""
    (bx, by, bw), (sx, sy, sw) = learn.trace_set([16, 17, 18], 2, 6, seed=21, log=lambda s: None)
    learn.train(G.NET_BFS, bx, by, steps=800, seed=1, weights=bw)
    learn.train(G.NET_SSSP, sx, sy, steps=400, seed=2, weights=sw)
    hg, dg = _graph(20, 12345, a=0.6, ef=20)
    srcs = learn.pick_sources(hg, 8, 9)
    learn.compare_bfs(dg, srcs[:2])  # warm-up
    runs = [learn.compare_bfs(dg, srcs) for _ in range(3)]
    gain = float(np.median([r["gain_vs_heuristic"] for r in runs]))
    r = runs[0]
    print(f"\nheld-out s20 BFS: heuristic {r['heuristic']:.3f} ms, learned {r['learned']:.3f} ms, "
          f"reference rule {r['reference']:.3f} ms, median gain vs heuristic {gain:.3f}, "
          f"vs reference rule {r['gain_vs_reference']:.3f}")
    assert all(x["identical"] for x in runs)
    assert gain > 1.0
    s = learn.compare_sssp(dg, srcs[:4])
    print(f"held-out s20 SSSP: heuristic {s['heuristic']:.3f} ms, learned {s['learned']:.3f} ms, "
          f"gain {s['gain_vs_heuristic']:.3f}")
    assert s["identical"]


def test_sssp_width_decision_records():
    """One trace record per bucket-width decision: before the first bucket and
    at every threshold advance, with the threshold and the device time so far
    (both nondecreasing).  Distances are unchanged when a network picks the
    width at every decision."""
    hg, dg = _graph(14, 5)
    src = int(learn.pick_sources(hg, 1, 2)[0])
    d0, st0, tr = dg.sssp(src, trace_cap=4096)
    assert st0.trace_len == len(tr) == int(st0.extra[1]) + 1
    th = np.array([t.threshold for t in tr])
    ms = np.array([t.ms for t in tr])
    assert th[0] == 0.0 and np.all(np.diff(th) > 0) and np.all(np.diff(ms) >= 0)
    assert all(t.mode == 1 for t in tr)  # the heuristic keeps the base width
    G.sched_set_params(G.NET_SSSP, _rand_params(7, scale=2.0))
    d1, st1, tr1 = dg.sssp(src, schedule=G.SCHED_LEARNED, trace_cap=4096)
    assert np.array_equal(d0, d1)
    assert st1.trace_len == int(st1.extra[1]) + 1
    assert all(0 <= t.mode < G.MLP_OUT for t in tr1)
