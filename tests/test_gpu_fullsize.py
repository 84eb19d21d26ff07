"""The full-size power-law graphs on the GPU: the hub core of the 100,000 x
200,000 graph runs to completion without exceeding any capacity, root ranges
partition the result, and the result is the same under every scheduler,
list-arena and reduction setting; the 50,000 x 1,000,000 graph with planted
blocks agrees with the oracle on root samples and partitions by root range.

The hub roots hold billions of maximal bicliques each, beyond what the CPU
oracle enumerates in a test, so they are checked through these properties;
the oracle checks the lighter roots of the same graphs."""
import pytest

import oracle
from gmbe_amd import IterFinderGpu6, enumerate_graph, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c2():
    return workloads.run_graph(2)


@pytest.fixture(scope="module")
def c3():
    return workloads.run_graph(3)


def run(g, **kw):
    r = enumerate_graph(g, workloads.BOUND_HEIGHT, workloads.BOUND_SIZE, **kw)
    assert r.overflow == 0
    return r


def combine(parts):
    hx = 0
    for d in parts:
        hx ^= d[2]
    return (sum(d[0] for d in parts), sum(d[1] for d in parts) % (1 << 64), hx)


def test_hub_core_ranks_complete_and_partition(gpu, c2):
    """Ranks [0, 32): the heaviest top-level vertices (degree up to 14,046)
    finish with no capacity exceeded, and the two halves sum to the whole."""
    whole = run(c2, root_begin=0, root_end=32)
    halves = [run(c2, root_begin=a, root_end=b).digest() for a, b in ((0, 16), (16, 32))]
    assert whole.maximal > 10 ** 7
    assert combine(halves) == whole.digest()


# share=0 and dom_k=0 are left out here: without sharing one wave keeps a hub
# subtree to itself, and without the list reduction the hub subtrees hold ten
# times the nodes; either runs for minutes on these roots (both are covered on
# smaller graphs by the parity tests)
VARIANTS = [dict(), dict(group=0), dict(lds_kb=0), dict(lds_kb=16), dict(waves_per_block=8),
            dict(waves_per_block=2), dict(dom_skip=0, late_share=0), dict(late_share=0),
            dict(dom_skip=0)]


def test_hub_core_sample_same_under_every_setting(gpu, c2):
    """A stride-257 sample of ranks 0..1023 (ranks 0, 257, 514, 771): one
    digest under every scheduler, list-arena, wave and reduction setting, and
    the sample's two root ranges sum to it."""
    kw = dict(root_begin=0, root_end=1024, root_stride=257, root_phase=0)
    digests = {}
    for v in VARIANTS:
        print(f"hub sample {v}", flush=True)
        digests[str(v)] = run(c2, **kw, **v).digest()
    assert len(set(digests.values())) == 1, digests
    parts = [run(c2, root_begin=a, root_end=b, root_stride=257, root_phase=0).digest()
             for a, b in ((0, 300), (300, 1024))]
    assert combine(parts) == next(iter(digests.values()))


@pytest.mark.parametrize("stride,phase", [(997, 40000), (1999, 30000)])
def test_full_size_config3_root_sample(gpu, c3, stride, phase):
    """Lighter top-level vertices of the planted-block graph against the
    oracle, bit for bit (samples the oracle finishes in seconds: a root of
    processing rank below 20,000 can hold millions of bicliques)."""
    res = oracle.gmbe(c3.oracle_graph(), threads=8, stride=stride, phase=phase)
    f = IterFinderGpu6(c3, workloads.BOUND_HEIGHT, workloads.BOUND_SIZE, root_begin=phase,
                       root_stride=stride, root_phase=phase % stride)
    r = f.execute()
    f.close()
    assert r.overflow == 0
    assert r.digest() == res.digest()


def test_full_size_config3_ranges_partition(gpu, c3):
    """Root ranges of the planted-block graph partition the result (ranks
    40,000-49,999: about 2 x 10^7 bicliques, a second on the GPU; a rank
    below 20,000 can hold millions of bicliques by itself)."""
    whole = run(c3, root_begin=40000, root_end=50000)
    assert whole.maximal > 10 ** 6
    parts = [run(c3, root_begin=a, root_end=b).digest()
             for a, b in ((40000, 42000), (42000, 45000), (45000, 50000))]
    assert combine(parts) == whole.digest()
