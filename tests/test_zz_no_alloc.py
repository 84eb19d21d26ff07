"""gmbe_finder_execute_static (the entry point that allocates no device
memory) against gmbe_finder_create + gmbe_finder_execute on the same graphs:
the same count and (sum, xor) digest, on calls in a row through one static
array (the second graph smaller than the first, so stale contents of a larger
run are left behind for the next one), and the same as the oracle."""
import numpy as np
import pytest

import graphs
import oracle
from gmbe_amd import BiGraph, GmbeError, enumerate_graph, workloads
from gmbe_amd.finder import execute_static

pytestmark = pytest.mark.gpu


def _cases():
    # larger first, then smaller: config 1 of the benchmark family, a random
    # graph, a crown graph
    return [
        ("config1", lambda: workloads.run_graph(1)),
        ("random_60x80", lambda: graphs.random_graph(60, 80, 0.2, 7, uvtrans=1, order=1,
                                                     fast_mode=1)),
        ("crown9", lambda: BiGraph.from_rows(*graphs.crown(9)[:2], uvtrans=1, order=1,
                                             fast_mode=1)),
    ]


def test_static_matches_execute_twice(gpu):
    made = [(name, make()) for name, make in _cases()]
    want = {}
    for name, g in made:
        r = enumerate_graph(g)
        assert r.overflow == 0
        want[name] = r.digest()
        assert want[name] == oracle.gmbe(g.oracle_graph()).digest(), name
    # two rounds of calls in a row through the one array, each larger -> smaller
    for rnd in range(2):
        for name, g in made:
            print(f"static call {rnd} {name}", flush=True)
            s = execute_static(g)
            assert s.overflow == 0
            assert s.digest() == want[name], (rnd, name)


def test_static_bounds_match_execute(gpu):
    """Split thresholds change the task tree, never the result."""
    g = workloads.run_graph(1)
    want = enumerate_graph(g).digest()
    for bh, bs in ((20, 1500), (2, 16), (0, 0)):
        print(f"static bounds {bh} {bs}", flush=True)
        assert execute_static(g, bh, bs).digest() == want


def test_static_rejects_oversize_graph(gpu):
    """A graph whose copy and slabs do not fit the static array is refused
    before anything runs, with a capacity error that names the sizes."""
    n = 6_000_000  # a perfect matching: 6 M vertices a side
    off = np.arange(n + 1, dtype=np.int64)
    adj = np.arange(n, dtype=np.uint32)
    g = BiGraph.from_csr(n, off, adj, uvtrans=1, order=1, fast_mode=1)
    with pytest.raises(GmbeError) as e:
        execute_static(g)
    assert e.value.code == -4 and "static scratch" in str(e.value)
    # the array still serves the next call
    small = BiGraph.from_rows(*graphs.crown(6)[:2], uvtrans=1, order=1, fast_mode=1)
    assert execute_static(small).maximal == 2 ** 6 - 2


def test_static_argument_errors(gpu):
    g = BiGraph.from_rows(*graphs.crown(4)[:2], uvtrans=1, order=1, fast_mode=1)
    with pytest.raises(GmbeError) as e:
        execute_static(g, -1, 10)
    assert e.value.code == -1
