"""Pins the CPU oracle (test infrastructure) before it is trusted.

* counts of the reference's own CPU code (baselines/MBE.zip built into
  oracle/_ref), frozen in tests/golden/ref_counts.json;
* closed forms (complete, matching, crown = 2^n - 2 with its full list, chain,
  stars, disjoint unions);
* the definition itself: brute force over every subset of V on small graphs;
* agreement of the GMBE-order restatement, the MBEA, iMBEA and PMBE
  restatements and brute force, with full lists;
* PMBE's containment skip on graphs built of nested neighbourhoods, with two
  planted mistakes in the order's ranges shown to fail this check, and its
  processing-node counts of the PMBE and iMBEA restatements equal to the
  reference binary's on small graphs (tests/golden/mbe_zip_nodes.json);
* invariance of the digest under side swap, vertex order, task splitting and
  PRUNE_EN (all change the traversal, none may change the result).
"""
import json
import os

import numpy as np
import pytest

import graphs
import oracle
from gmbe_amd import BiGraph, digest_of, hash_biclique, workloads

HERE = os.path.dirname(os.path.abspath(__file__))
M64 = (1 << 64) - 1


def _mix64(x):
    z = (x + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def _hash_py(A, B):
    sa = sum(_mix64((a << 1) | 0) for a in A) & M64
    sb = sum(_mix64((b << 1) | 1) for b in B) & M64
    rot = ((sb << 31) | (sb >> 33)) & M64
    return _mix64(sa ^ rot)


def test_mix64_is_splitmix64():
    """mix64(x) is splitmix64's output for state x + golden gamma: the first
    three outputs of splitmix64 seeded with 0 are the published
    0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4, 0x06c45d188009454f."""
    g = 0x9E3779B97F4A7C15
    assert _mix64(0) == 0xE220A8397B1DCDAF
    assert _mix64(g) == 0x6E789E6AA1B965F4
    assert _mix64((2 * g) & M64) == 0x06C45D188009454F
    # the engine and the oracle use the same function: a one-element
    # biclique's hash is mix64(mix64(a << 1) ^ rotl64(mix64(b << 1 | 1), 31))
    for a, b in [(0, 0), (7, 3), (123456, 654321)]:
        sb = _mix64((b << 1) | 1)
        want = _mix64(_mix64(a << 1) ^ (((sb << 31) | (sb >> 33)) & M64))
        assert oracle.hash_pair([a], [b]) == want
        assert hash_biclique([a], [b]) == want


def test_hash_known_answers():
    kat = json.load(open(os.path.join(HERE, "golden", "hash_kat.json")))
    for case in kat:
        A, B, h = case["A"], case["B"], case["h"]
        assert oracle.hash_pair(A, B) == h
        assert hash_biclique(A, B) == h
        assert _hash_py(A, B) == h
    # order independence within a side, side sensitivity
    assert hash_biclique([3, 1, 2], [9, 8]) == hash_biclique([1, 2, 3], [8, 9])
    assert hash_biclique([1], [2]) != hash_biclique([2], [1])


def build_recipe(rec):
    k = rec["kind"]
    if k == "kat":
        g, _ = graphs.kat_graph(rec["name"])
        return g
    if k == "random":
        return graphs.random_graph(rec["nU"], rec["nV"], rec["p"], rec["seed"])
    if k == "config":
        return workloads.raw_graph(rec["config"], rec["scale"])
    raise ValueError(k)


REF = json.load(open(os.path.join(HERE, "golden", "ref_counts.json")))["graphs"]


@pytest.mark.parametrize("name", sorted(REF))
def test_oracle_matches_reference_cpu_counts(name):
    """Counts of the reference's own MbeaAdvFinder / ImbeaAdvFinder / PmbeAdvFinder
    (its -s 0 / 1 / 2, frozen per graph) against the four restatements."""
    ent = REF[name]
    g = build_recipe(ent["recipe"]).derive(uvtrans=1, order=1, fast_mode=1)
    og = g.oracle_graph()
    assert oracle.gmbe(og, threads=8).maximal == ent["count"]
    if ent["count"] < 5_000_000:
        assert oracle.mbea(og, threads=8).maximal == ent["count"]
        assert oracle.imbea(og, threads=8).maximal == ent["count"]
    if ent["count"] < 1_000_000:
        assert oracle.pmbe(og, threads=8).maximal == ent["ref_finders"]["s2"]


@pytest.mark.parametrize("name", sorted(graphs.KATS))
def test_closed_forms(name):
    g, cnt = graphs.kat_graph(name)
    og = g.oracle_graph()
    for run in (oracle.gmbe, oracle.mbea, oracle.imbea, oracle.pmbe):
        assert run(og).maximal == cnt
    if g.stats()["V_size"] <= 24 and g.stats()["U_size"] <= 64:
        assert oracle.brute(og).maximal == cnt


@pytest.mark.parametrize("n", [2, 3, 6, 10])
def test_crown_graph_full_list(n):
    g = BiGraph.from_rows(*graphs.crown(n)[:2])
    expect = graphs.crown_list(n)
    for run in (oracle.gmbe, oracle.mbea):
        res, lst = run(g.oracle_graph(), want_list=True)
        from gmbe_amd import records_to_set
        got = records_to_set(lst)
        assert got == expect
        assert res.digest() == digest_of(expect)


@pytest.mark.parametrize("seed", range(12))
def test_three_way_agreement_small(seed):
    rng = np.random.default_rng(seed)
    nU, nV = int(rng.integers(4, 40)), int(rng.integers(3, 17))
    p = float(rng.uniform(0.1, 0.6))
    g = graphs.random_graph(nU, nV, p, 1000 + seed)
    og = g.oracle_graph()
    from gmbe_amd import records_to_set
    rb, lb = oracle.brute(og, want_list=True)
    rg, lg = oracle.gmbe(og, want_list=True)
    rm, lm = oracle.mbea(og, want_list=True)
    ri, li = oracle.imbea(og, want_list=True)
    rp, lp = oracle.pmbe(og, want_list=True)
    sb = records_to_set(lb)
    assert records_to_set(lg) == sb
    assert records_to_set(lm) == sb
    assert records_to_set(li) == sb
    assert records_to_set(lp) == sb
    assert rb.digest() == rg.digest() == rm.digest() == ri.digest() == rp.digest() == \
        digest_of(sb)


MBE_ZIP_NODES = json.load(open(os.path.join(HERE, "golden", "mbe_zip_nodes.json")))["graphs"]


def test_pmbe_walks_the_reference_traversal():
    """Maximal AND processing-node counts of the reference's own PmbeAdvFinder
    (-s 2, frozen per graph) on 120 small random and nested graphs, each
    oriented as its loader orients it.  With at most 16 root-side vertices
    the reference's containment order is fixed, so equal node counts pin the
    order, the released ranges, the skip rule and the maximality test, not
    only the result."""
    skipped = 0
    for name, ent in sorted(MBE_ZIP_NODES.items()):
        nU, rows = graphs.pmbe_recipe_rows(ent["recipe"])
        r = oracle.pmbe(graphs.mbe_orient(nU, rows).oracle_graph())
        assert (r.maximal, r.nodes) == (ent["s2"]["maximal"], ent["s2"]["nodes"]), name
        skipped += r.skipped
    assert skipped > 0


def test_imbea_walks_the_reference_traversal():
    """The same for ImbeaAdvFinder (-s 1): the graph numbered as its
    Reorder(2) numbers it, candidates in overlap order with ties kept, and
    P minus the twins computed by the reference's merge walk."""
    for name, ent in sorted(MBE_ZIP_NODES.items()):
        nU, rows = graphs.pmbe_recipe_rows(ent["recipe"])
        r = oracle.imbea(graphs.mbe_reorder2(nU, rows).oracle_graph())
        assert (r.maximal, r.nodes) == (ent["s1"]["maximal"], ent["s1"]["nodes"]), name


def test_pmbe_containment_skip_on_nested_graphs():
    """PMBE skips a candidate inside the range of vertices a done candidate
    released; the full list must still equal brute force.  Nested rows make
    the skip frequent; a range one id too wide, or stored at the neighbouring
    id, changes the list on about half of these graphs."""
    from gmbe_amd import records_to_set
    skipped = with_skips = 0
    for seed in range(60):
        rng = np.random.default_rng(seed)
        nU, nV = int(rng.integers(4, 30)), int(rng.integers(4, 20))
        og = BiGraph.from_rows(nU, graphs.nested_rows(nU, nV, seed)).oracle_graph()
        rb, lb = oracle.brute(og, want_list=True)
        rp, lp = oracle.pmbe(og, want_list=True)
        assert records_to_set(lp) == records_to_set(lb), seed
        assert rp.digest() == rb.digest()
        skipped += rp.skipped
        with_skips += rp.skipped > 0
    assert with_skips >= 20 and skipped >= 200


@pytest.mark.parametrize("opt", [(0, 0), (1, 1), (2, 2), (1, 0), (0, 3)])
def test_digest_invariant_to_orientation_and_order(opt):
    raw = graphs.random_graph(60, 45, 0.12, 77)
    base = oracle.gmbe(raw.derive(uvtrans=0, order=0, fast_mode=0).oracle_graph()).digest()
    g = raw.derive(uvtrans=opt[0], order=opt[1], fast_mode=0 if opt[1] == 3 else 1)
    assert oracle.gmbe(g.oracle_graph()).digest() == base
    assert oracle.mbea(g.oracle_graph()).digest() == base


@pytest.mark.parametrize("bh,bs,mp", [(0, 0, 4), (0, 0, 8), (3, 10, 4), (20, 1500, 4)])
def test_split_rule_keeps_the_result(bh, bs, mp):
    g = workloads.run_graph(1)
    og = g.oracle_graph()
    base = oracle.gmbe(og).digest()
    split = oracle.gmbe(og, bound_height=bh, bound_size=bs, max_prefix=mp)
    assert split.digest() == base


def test_prune_keeps_the_result_and_saves_pushes():
    og = workloads.run_graph(1).oracle_graph()
    a = oracle.gmbe(og, prune=1)
    b = oracle.gmbe(og, prune=0)
    assert a.digest() == b.digest()
    assert a.nodes <= b.nodes and a.non_maximal <= b.non_maximal


@pytest.mark.parametrize("algo", ["gmbe", "mbea", "imbea", "pmbe"])
def test_root_sample_partition(algo):
    """Digests over disjoint root samples add up to the full digest (PMBE's
    ranks follow its own containment order)."""
    run = getattr(oracle, algo)
    og = workloads.run_graph(2, 0.01).oracle_graph()
    full = run(og, threads=8)
    parts = [run(og, threads=8, stride=3, phase=p) for p in range(3)]
    assert sum(p.maximal for p in parts) == full.maximal
    assert sum(p.hash_sum for p in parts) % (1 << 64) == full.hash_sum
    x = 0
    for p in parts:
        x ^= p.hash_xor
    assert x == full.hash_xor


@pytest.mark.parametrize("algo", ["mbea", "imbea", "pmbe"])
def test_cpu_finder_time_limit_cuts_within_a_root(algo):
    """The CPU finders' time limit also abandons the subtree they are in: a
    tiny limit returns promptly with a partial count, a generous one the full
    count (the benchmark's CPU legs rely on both)."""
    run = getattr(oracle, algo)
    og = workloads.run_graph(2, 0.01).oracle_graph()
    full = run(og, threads=8)
    cut = run(og, threads=8, time_limit=0.01)
    assert cut.seconds < 2.0
    assert cut.maximal < full.maximal
    again = run(og, threads=8, time_limit=600.0)
    assert again.digest() == full.digest()


def test_rank_offset_samples_partition_the_roots():
    """Samples that start at a later rank (phase >= stride) select the ranks
    phase, phase + stride, ...: two offsets of one stride add up to the
    stride's sample from the earlier rank."""
    og = workloads.run_graph(2, 0.01).oracle_graph()
    a = oracle.gmbe(og, threads=4, stride=2, phase=10)   # ranks 10, 12, 14, ...
    b = oracle.gmbe(og, threads=4, stride=2, phase=11)   # ranks 11, 13, 15, ...
    c = oracle.gmbe(og, threads=4, stride=1, phase=10)   # ranks 10, 11, 12, ...
    assert a.maximal + b.maximal == c.maximal
    assert (a.hash_sum + b.hash_sum) % (1 << 64) == c.hash_sum
    assert a.hash_xor ^ b.hash_xor == c.hash_xor
