"""Host logic that needs no GPU: the seeded workload generators, the ego cut of
the served workload, the benchmark's command-line contract and its CPU legs."""
import os
import subprocess
import sys

import numpy as np

from gmbe_amd import BiGraph, workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def csr(g):
    v_off, v_adj = g.csr()[:2]
    return np.asarray(v_off).copy(), np.asarray(v_adj).copy()


def test_config_graphs_have_the_stated_shape():
    g1 = workloads.raw_graph(1)
    s = g1.stats()
    assert (s["U_size"], s["V_size"], s["edge_size"]) == (2000, 1500, 12000)
    g2 = workloads.raw_graph(2)
    s = g2.stats()
    assert (s["U_size"], s["V_size"], s["edge_size"]) == (100000, 200000, 1500000)
    assert max(s["U_degree"], s["V_degree"]) <= 20000


def test_generators_are_deterministic():
    a, b = workloads.raw_graph(2, 0.05), workloads.raw_graph(2, 0.05)
    for x, y in zip(csr(a), csr(b)):
        assert np.array_equal(x, y)
    s1, s2 = workloads.small_graph(workloads.batch_seed(7)), workloads.small_graph(workloads.batch_seed(7))
    for x, y in zip(csr(s1), csr(s2)):
        assert np.array_equal(x, y)


def test_small_graphs_follow_the_batched_distribution():
    sizes = []
    for i in range(64):
        st = workloads.small_graph(workloads.batch_seed(i)).stats()
        assert 50 <= st["U_size"] <= 5000 and 50 <= st["V_size"] <= 5000
        # every vertex has degree >= 1 (the covering edges come first); beyond
        # them the edge count follows the mean degree (at most 20 per U vertex)
        assert st["edge_size"] >= max(st["U_size"], st["V_size"])
        assert st["edge_size"] <= max(20 * st["U_size"] + 1, st["U_size"] + st["V_size"])
        sizes.append(st["U_size"])
    # log-uniform sizes: both ends of [50, 5000] are reached in 64 draws
    assert min(sizes) < 200 and max(sizes) > 1500


def test_ego_cut_respects_the_edge_cap():
    g = workloads.raw_graph(2, 0.1)
    nU = g.stats()["U_size"]
    for u in (0, 1, nU // 2, nU - 1):
        e = g.ego(u, 20000)
        st = e.stats()
        assert st["edge_size"] <= 20000
        # every kept row is a neighbour of u: all of them share U vertex u
        v_off, v_adj = csr(e)
        assert st["V_size"] == 0 or all(len(set(v_adj[v_off[i]:v_off[i + 1]])) > 0
                                        for i in range(st["V_size"]))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_bench_cpu_baseline_legs():
    """The CPU-baseline legs of the benchmark line (MbeaFinder, ImbeaAdvFinder
    and PmbeAdvFinder restatements) on a small graph: every field the line
    carries, and the whole sample's count when the time is ample."""
    sys.path.insert(0, ROOT)
    import bench
    import oracle
    g = workloads.run_graph(1)
    full = oracle.mbea(g.oracle_graph(), threads=4)
    total = full.maximal
    first, stride = bench.CPU_FIRST, bench.CPU_STRIDE
    try:
        # the whole graph as the sample: rank 0, stride 1
        bench.CPU_FIRST, bench.CPU_STRIDE = 0, 1
        for algo in ("mbea", "imbea", "pmbe"):
            b = bench.cpu_baseline(g, 30.0, algo=algo)
            assert set(b) == {"value", "unit", "cores", "kind", "cpu", "sample", "complete",
                              "digest"}
            assert b["unit"] == "bicliques/s" and b["kind"] == "port" and b["cores"] >= 1
            assert b["value"] > 0 and b["complete"]
            assert f"{total} maximal bicliques" in b["sample"], b["sample"]
            assert b["digest"] == [total, full.hash_sum, full.hash_xor]
            if algo == "pmbe":
                assert "containment order" in b["sample"]
    finally:
        bench.CPU_FIRST, bench.CPU_STRIDE = first, stride


def test_finder_knob_errors_are_codes():
    from gmbe_amd import GmbeError, _lib
    import ctypes as C
    g = BiGraph.from_rows(2, [[0, 1], [1]], uvtrans=1, order=1, fast_mode=1)
    out = C.c_void_p()
    rc = _lib.lib().gmbe_finder_create(g.handle, 2, 20, 1500, 1, C.byref(out))
    if rc != 0:
        assert rc == -5  # no GPU here: creation itself fails loudly
        return
    try:
        for key, val in [(b"deadline_ms", -1), (b"root_end", -2), (b"ctx_slots", 0),
                         (b"lds_kb", 5), (b"waves_per_block", 3), (b"nonsense", 1)]:
            assert _lib.lib().gmbe_finder_set(out, key, val) == -1
    finally:
        _lib.lib().gmbe_finder_destroy(out)
