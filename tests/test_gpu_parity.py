"""Parity of the HIP enumerator (through the C ABI) with the CPU oracle.

Bit-exact: the count and the (sum, xor) hash digest of every maximal biclique
must equal the oracle's; on small graphs the full list of (L, R) pairs is
compared.  Full-size graphs are checked on root samples (the digest over a
sample of top-level vertices is exact and additive) and through invariances.
"""
import json
import os

import numpy as np
import pytest

import graphs
import oracle
from gmbe_amd import BiGraph, GmbeError, IterFinderGpu6, digest_of, enumerate_graph, workloads

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def gpu_digest(g, **kw):
    r = enumerate_graph(g, **kw)
    assert r.overflow == 0
    return r


def gpu_list(g, cap=1 << 22, **kw):
    f = IterFinderGpu6(g, list_cap=cap, **kw)
    r = f.execute()
    s = f.bicliques()
    f.close()
    return r, s


@pytest.mark.parametrize("name", sorted(graphs.KATS))
def test_closed_forms(gpu, name):
    g, cnt = graphs.kat_graph(name, uvtrans=1, order=1, fast_mode=1)
    r = gpu_digest(g)
    assert r.maximal == cnt
    assert r.digest() == oracle.gmbe(g.oracle_graph()).digest()


@pytest.mark.parametrize("n", [2, 5, 9, 12])
def test_crown_list(gpu, n):
    g = BiGraph.from_rows(*graphs.crown(n)[:2], uvtrans=1, order=1, fast_mode=1)
    r, s = gpu_list(g)
    assert s == graphs.crown_list(n)
    assert r.digest() == digest_of(s)


@pytest.mark.parametrize("seed", range(10))
def test_random_small_lists(gpu, seed):
    rng = np.random.default_rng(seed)
    nU, nV = int(rng.integers(4, 50)), int(rng.integers(3, 20))
    g = graphs.random_graph(nU, nV, float(rng.uniform(0.1, 0.6)), 2000 + seed)
    g = g.derive(uvtrans=1, order=1, fast_mode=1)
    res, lst = oracle.brute(g.oracle_graph(), want_list=True) if nV <= 24 and nU <= 64 \
        else oracle.gmbe(g.oracle_graph(), want_list=True)
    from gmbe_amd import records_to_set
    expect = records_to_set(lst)
    for sel, bh, bs in [(2, 20, 1500), (0, 20, 1500), (2, 0, 0)]:
        r, s = gpu_list(g, sel=sel, bound_height=bh, bound_size=bs)
        assert s == expect, (sel, bh, bs)
        assert r.digest() == res.digest()


def test_config1_full_list(gpu):
    g = workloads.run_graph(1)
    res, lst = oracle.gmbe(g.oracle_graph(), want_list=True)
    from gmbe_amd import records_to_set
    expect = records_to_set(lst)
    assert len(expect) == 3988  # reference CPU count (tests/golden/ref_counts.json)
    r, s = gpu_list(g)
    assert s == expect
    assert r.digest() == res.digest() == digest_of(expect)


@pytest.mark.parametrize("sel,bh,bs,mp", [(2, 20, 1500, 4), (0, 20, 1500, 4), (2, 0, 0, 4),
                                          (2, 0, 0, 8), (2, 5, 50, 3)])
def test_config1_schedulers(gpu, sel, bh, bs, mp):
    g = workloads.run_graph(1)
    base = oracle.gmbe(g.oracle_graph()).digest()
    r = gpu_digest(g, sel=sel, bound_height=bh, bound_size=bs, max_prefix=mp)
    assert r.digest() == base


@pytest.mark.parametrize("lds_kb", [0, 3, 7, 16])
def test_lds_tiers(gpu, lds_kb):
    g = workloads.run_graph(2, 0.01)
    base = oracle.gmbe(g.oracle_graph(), threads=8).digest()
    assert gpu_digest(g, lds_kb=lds_kb).digest() == base


REF = json.load(open(os.path.join(HERE, "golden", "ref_counts.json")))["graphs"]


@pytest.mark.parametrize("name", ["config2_s0.02", "config3_s0.002"])
def test_scaled_configs_vs_reference_counts(gpu, name):
    rec = REF[name]["recipe"]
    g = workloads.run_graph(rec["config"], rec["scale"])
    r = gpu_digest(g)
    assert r.maximal == REF[name]["count"]
    assert r.digest() == oracle.gmbe(g.oracle_graph(), threads=8).digest()


def test_wide_bitsets(gpu):
    """|L| > 256 without splitting goes through the wide-bitset pass."""
    g = graphs.random_graph(700, 24, 0.45, 31).derive(uvtrans=1, order=1, fast_mode=1)
    assert g.stats()["V_size"] == 24
    base = oracle.gmbe(g.oracle_graph(), threads=8).digest()
    for sel in (0, 2):
        assert gpu_digest(g, sel=sel).digest() == base


def test_very_wide_l_must_split(gpu):
    """|L| > 1024 is split into children whatever the bounds."""
    g = graphs.random_graph(1500, 12, 0.7, 32).derive(uvtrans=1, order=1, fast_mode=1)
    base = oracle.gmbe(g.oracle_graph(), threads=8).digest()
    assert gpu_digest(g, sel=2, bound_height=1 << 20, bound_size=1 << 30).digest() == base


def test_edge_cases(gpu):
    empty = BiGraph.from_rows(3, [[], [], []], uvtrans=1, order=1, fast_mode=1)
    assert gpu_digest(empty).maximal == 0
    single = BiGraph.from_rows(1, [[0]], uvtrans=1, order=1, fast_mode=1)
    assert gpu_digest(single).maximal == 1
    twins = BiGraph.from_rows(4, [[0, 1], [0, 1], [1, 2, 3], [0, 1]], uvtrans=0, order=1,
                              fast_mode=1)
    assert gpu_digest(twins).digest() == oracle.brute(twins.oracle_graph()).digest()
    isolated = BiGraph.from_rows(3, [[0, 1], [], [1, 2]], uvtrans=0, order=0, fast_mode=1)
    assert gpu_digest(isolated).digest() == oracle.brute(isolated.oracle_graph()).digest()


def test_capacity_error_is_reported(gpu):
    """A per-wave arena too small for a task is an error, never a silent miss:
    on the 2% power-law graph some top-level vertex has more than 64 KB of
    candidate records (over 2,730 candidates at 24 bytes)."""
    g = workloads.run_graph(2, 0.02)
    f = IterFinderGpu6(g, arena_kb=64, lds_kb=0)
    with pytest.raises(GmbeError) as e:
        f.execute()
    assert e.value.code == -4
    assert f.result.overflow > 0
    f.close()


def test_full_queue_falls_back_to_local_runs(gpu):
    """The smallest queue (two slots per shard, 64 in all) with every task
    split and no shared contexts: pushes that find no room run the task
    locally instead, and the result is the oracle's.  (A shard of one slot
    would make a slot's "holds position p" and "free for position p + 1"
    sequence numbers equal, so a pusher could overwrite an entry before its
    popper read it; the finder rounds the capacity up to two.)"""
    g = workloads.run_graph(1)
    base = oracle.gmbe(g.oracle_graph()).digest()
    r = gpu_digest(g, queue_log2=0, share=0, bound_height=0, bound_size=0)
    assert r.digest() == base


@pytest.mark.parametrize("waves", [2, 4, 8])
def test_waves_per_block(gpu, waves):
    """Two, four or eight waves per block (the reference's warps-per-block
    builds, src/CMakeLists.txt:57-78) give the oracle's result."""
    g = workloads.run_graph(2, 0.01)
    base = oracle.gmbe(g.oracle_graph(), threads=8).digest()
    assert gpu_digest(g, waves_per_block=waves).digest() == base


def test_repeated_execute_is_identical(gpu):
    g = workloads.run_graph(2, 0.01)
    f = IterFinderGpu6(g)
    a = f.execute().digest()
    b = f.execute().digest()
    f.close()
    assert a == b


@pytest.mark.parametrize("stride,phase", [(2000, 1999), (2000, 1500), (5000, 4000)])
def test_full_size_config2_root_sample(gpu, stride, phase):
    g = workloads.run_graph(2)
    res = oracle.gmbe(g.oracle_graph(), threads=8, stride=stride, phase=phase)
    r = gpu_digest(g, root_stride=stride, root_phase=phase)
    assert r.digest() == res.digest()


def test_root_ranges_partition_the_result(gpu):
    """Rank ranges, alone and combined with a stride, partition the roots."""
    g = workloads.run_graph(2, 0.02)
    full = gpu_digest(g).digest()
    nV = g.stats()["V_size"]
    cuts = [0, 7, 500, nV // 2, nV]
    parts = [gpu_digest(g, root_begin=a, root_end=b).digest() for a, b in zip(cuts, cuts[1:])]
    parts += [gpu_digest(g, root_begin=a, root_end=b, root_stride=3, root_phase=p).digest()
              for a, b in [(0, 1000), (1000, nV)] for p in range(3)]
    for group in (parts[:4], parts[4:]):
        cnt = sum(d[0] for d in group)
        hs = sum(d[1] for d in group) % (1 << 64)
        hx = 0
        for d in group:
            hx ^= d[2]
        assert (cnt, hs, hx) == full


def test_deadline_cuts_the_run_short(gpu):
    """A deadline ends the kernel with partial counts and GMBE_ERR_DEADLINE;
    a generous one leaves the result exact."""
    g = workloads.run_graph(2, 0.1)
    f = IterFinderGpu6(g, deadline_ms=50)
    with pytest.raises(GmbeError) as e:
        f.execute()
    assert e.value.code == -10
    part = f.result
    assert 0 < part.kernel_time_s < 5.0
    assert part.maximal > 0
    f.close()
    small = workloads.run_graph(2, 0.01)
    exact = oracle.gmbe(small.oracle_graph(), threads=8).digest()
    assert gpu_digest(small, deadline_ms=60000).digest() == exact


@pytest.mark.parametrize("knobs", [
    dict(share=0), dict(share=1), dict(group=0), dict(group=1, merge_min=2),
    dict(share=1, share_work=1, share_queue=1 << 30),       # share every node it can
    dict(share=1, bound_height=0, bound_size=0, merge_min=2),
    dict(prune=0),
])
def test_scheduler_knobs_keep_the_result(gpu, knobs):
    """Sharing, grouping, twin merging and pruning change the traversal only."""
    g = workloads.run_graph(2, 0.01)
    base = oracle.gmbe(g.oracle_graph(), threads=8)
    r = gpu_digest(g, **knobs)
    assert r.digest() == base.digest()


def test_unsplit_matches_reference_traversal(gpu):
    """Without splitting, the pushes, non-maximal verdicts and algorithmic
    bytes are those of the reference's traversal (the oracle restates it).
    Dropping dominated excluded entries keeps every verdict (dom_k on); the
    skip marks leave out pushes known to be non-maximal, so they are off."""
    g = workloads.run_graph(1)
    o = oracle.gmbe(g.oracle_graph())
    for group, dom_k in ((0, 0), (1, 0), (0, 8), (1, 8)):
        r = gpu_digest(g, sel=0, group=group, dom_k=dom_k, dom_skip=0, count_bytes=1)
        assert r.digest() == o.digest()
        assert r.nodes == o.nodes
        assert r.non_maximal == o.non_maximal
        assert r.alg_bytes == o.b_alg


def test_exit_times_of_the_blocks(gpu):
    """exit_stamps=1 (the CLI's -p): one exit time per block of the persistent
    grid, none negative or past the run, and the result unchanged."""
    g = workloads.run_graph(1)
    f = IterFinderGpu6(g, exit_stamps=1)
    r = f.execute()
    t = f.exit_times()
    f.close()
    assert len(t) > 0
    assert all(0.0 <= x <= r.exe_time_s + 0.1 for x in t)
    assert r.digest() == gpu_digest(g).digest()
