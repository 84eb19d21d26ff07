"""Serving loop: every request's count and hash equal the plain finder's on
the same request graph, whether it ran alone or in a mix of overlapping
launches, and the oracle's on a sample; admission and error paths.  The
resident graph is the full-size power-law graph of the served workload."""
import numpy as np
import pytest

import graphs
import oracle
from gmbe_amd import BiGraph, GmbeError, Server, enumerate_graph, workloads

pytestmark = pytest.mark.gpu
RUN = workloads.RUN_OPTION


@pytest.fixture(scope="module")
def resident():
    return workloads.raw_graph(2)


def alone(g_raw):
    r = enumerate_graph(g_raw.derive(**RUN))
    assert r.overflow == 0
    return r.digest()


def drain(s, n):
    got = {}
    while len(got) < n:
        for d in s.poll(wait=True):
            got[d.ticket] = d
    return got


def test_mixed_requests_equal_alone(gpu, resident):
    rng = np.random.default_rng(5)
    nU = resident.stats()["U_size"]
    s = Server(resident, slots=64, mem_budget=1 << 28)
    want = {}
    for j in range(40):
        if j % 2 == 0:
            u = int(rng.integers(0, nU))
            t = s.submit_ego(u, 20000)
            want[t] = alone(resident.ego(u, 20000))
        else:
            g = workloads.small_graph(workloads.serve_small_seed(j))
            t = s.submit_graph(g)
            want[t] = alone(g)
    got = drain(s, len(want))
    s.close()
    for t, d in got.items():
        assert d.status == 0
        assert (d.maximal, d.hash_sum, d.hash_xor) == want[t]
        assert d.t_done_s >= d.t_submit_s


def test_served_requests_equal_the_oracle(gpu, resident):
    """Ego subgraphs of the resident graph and small graphs, served in one
    mix, against the CPU oracle bit for bit."""
    rng = np.random.default_rng(11)
    nU = resident.stats()["U_size"]
    s = Server(resident, slots=16, mem_budget=1 << 28)
    want = {}
    for j in range(8):
        if j % 2 == 0:
            u = int(rng.integers(0, nU))
            t = s.submit_ego(u, 20000)
            g = resident.ego(u, 20000)
        else:
            g = workloads.small_graph(workloads.serve_small_seed(200 + j))
            t = s.submit_graph(g)
        want[t] = oracle.gmbe(g.derive(**RUN).oracle_graph(), threads=8).digest()
    got = drain(s, len(want))
    s.close()
    for t, d in got.items():
        assert d.status == 0
        assert (d.maximal, d.hash_sum, d.hash_xor) == want[t]


def test_tight_budget_admits_one_at_a_time(gpu, resident):
    gs = [workloads.small_graph(workloads.serve_small_seed(100 + j)) for j in range(6)]
    biggest = max(g.stats()["edge_size"] for g in gs)
    # room for about one request per lane
    s = Server(None, slots=2, mem_budget=3 * (24 * biggest + 400_000))
    want = {}
    for g in gs:
        want[s.submit_graph(g)] = alone(g)
    got = drain(s, len(want))
    s.close()
    for t, d in got.items():
        assert d.status == 0
        assert (d.maximal, d.hash_sum, d.hash_xor) == want[t]


def test_zero_budget_rejects(gpu):
    s = Server(None, slots=4, mem_budget=0)
    with pytest.raises(GmbeError) as e:
        s.submit_graph(workloads.small_graph(workloads.serve_small_seed(1)))
    assert e.value.code == -7
    s.close()


def test_list_memory_admission(gpu):
    """A request whose task roots surely need more list memory per wave than
    stack_limit_kb allows is refused at submit; one that fits is served
    exactly.  Crown graph S_60: every root has at least 58 candidates of 24
    bytes (1,392 B > 1 KB); a matching has none."""
    s = Server(None, slots=4, mem_budget=1 << 24)
    s.set("stack_limit_kb", 1)
    crown = BiGraph.from_rows(*graphs.crown(60)[:2])
    with pytest.raises(GmbeError) as e:
        s.submit_graph(crown)
    assert e.value.code == -7 and "list memory" in str(e.value)
    with pytest.raises(GmbeError):
        s.set("stack_limit_kb", 1 << 30)  # above the lanes' arena
    match = BiGraph.from_rows(*graphs.matching(7)[:2])
    t = s.submit_graph(match)
    done = {}
    while t not in done:
        done.update({d.ticket: d for d in s.poll(wait=True)})
    assert done[t].status == 0 and done[t].maximal == 7
    s.close()


def test_ego_needs_a_resident_graph(gpu):
    s = Server(None, slots=4, mem_budget=1 << 24)
    with pytest.raises(GmbeError) as e:
        s.submit_ego(0, 100)
    assert e.value.code == -9
    assert s.poll() == []
    s.close()


def test_request_deadline_returns_exact_or_cut(gpu, resident):
    """With a per-launch deadline every request either finishes exactly or is
    reported as cut short (GMBE_ERR_DEADLINE); none is silently wrong."""
    gs = [workloads.small_graph(workloads.batch_seed(i)) for i in range(24)]
    s = Server(None, slots=8, mem_budget=1 << 28)
    s.set("request_ms", 5)
    want = {}
    for g in gs:
        want[s.submit_graph(g)] = alone(g)
    got = drain(s, len(want))
    s.close()
    statuses = {d.status for d in got.values()}
    assert statuses <= {0, -10}
    for t, d in got.items():
        if d.status == 0:
            assert (d.maximal, d.hash_sum, d.hash_xor) == want[t]
        else:
            assert d.maximal <= want[t][0]
