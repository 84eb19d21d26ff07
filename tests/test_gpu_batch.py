"""Batched form: per-graph counts and hashes equal the plain finder's (which
the finder parity tests tie to the oracle), across replays of the captured
launch, the looped baseline, size-class thresholds and the autotuner."""
import pytest

import graphs
import oracle
from gmbe_amd import Batch, BiGraph, GmbeError, enumerate_graph, workloads

pytestmark = pytest.mark.gpu


def plain(g):
    r = enumerate_graph(g)
    assert r.overflow == 0
    return r.digest()


@pytest.fixture(scope="module")
def small_set():
    gs = workloads.batch_graphs(48)
    # edge cases: closed-form graphs and a single edge
    for name in ("crown_5", "K_3x4", "matching_7", "star_v_10"):
        gs.append(graphs.kat_graph(name, uvtrans=1, order=1, fast_mode=1)[0])
    gs.append(BiGraph.from_rows(1, [[0]], uvtrans=1, order=1, fast_mode=1))
    return gs


@pytest.fixture(scope="module")
def expected(small_set):
    return [plain(g) for g in small_set]


def test_batch_matches_plain_finder(gpu, small_set, expected):
    b = Batch(small_set)
    per, tot = b.execute()
    assert [r.digest() for r in per] == expected
    assert tot.maximal == sum(e[0] for e in expected)
    b.close()


def test_oracle_spot_check(gpu, small_set):
    b = Batch(small_set[:6])
    per, _ = b.execute()
    for g, r in zip(small_set[:6], per):
        assert r.digest() == oracle.gmbe(g.oracle_graph(), threads=8).digest()
    b.close()


def test_replays_of_the_captured_launch_agree(gpu, small_set, expected):
    b = Batch(small_set)
    for i in range(4):
        per, _ = b.execute()
        assert [r.digest() for r in per] == expected, f"replay {i}"
        # the program's own eager launches and host waits between replays
        assert plain(small_set[i]) == expected[i]
    b.close()


def test_looped_baseline_matches(gpu, small_set, expected):
    b = Batch(small_set, per_graph_launch=1)
    per, _ = b.execute()
    assert [r.digest() for r in per] == expected
    b.close()


@pytest.mark.parametrize("knobs", [
    {"bound_height.0": 0, "bound_size.0": 0, "bound_height.1": 0, "bound_size.1": 0,
     "bound_height.2": 0, "bound_size.2": 0, "bound_height.3": 0, "bound_size.3": 0},
    {"share": 0}, {"lds_kb": 0}, {"lds_kb": 16}, {"waves_per_block": 8},
    {"waves_per_block": 2}, {"group": 0},
])
def test_knobs_keep_per_graph_results(gpu, small_set, expected, knobs):
    b = Batch(small_set)
    for k, v in knobs.items():
        b.set(k, v)
    per, _ = b.execute()
    assert [r.digest() for r in per] == expected
    b.close()


def _light(small_set):
    # graphs without hub roots: unsplit, a hub root's subtree runs for
    # minutes on one wave, since without splitting nothing hands it to others
    return [i for i, g in enumerate(small_set) if g.stats()["edge_size"] < 16000]


def test_unsplit_batch_keeps_per_graph_results(gpu, small_set, expected):
    pick = _light(small_set)
    b = Batch([small_set[i] for i in pick])
    b.set("split", 0)
    per, _ = b.execute()
    assert [r.digest() for r in per] == [expected[i] for i in pick]
    b.close()


def test_autotuner_keeps_results(gpu, small_set, expected):
    # graphs of every size class but the hub-heavy ones, so that the tuning
    # trials (unsplit trials among them) stay short
    pick = _light(small_set)
    gs = [small_set[i] for i in pick]
    b = Batch(gs)
    b.set("autotune", 1)
    for c in range(b.get("n_classes")):
        assert b.get(f"waves_per_block.{c}") in (2, 4, 8)
        assert b.get(f"lds_kb.{c}") in (0, 3, 7, 16)
        assert b.get(f"bound_height.{c}") >= 0
    per, _ = b.execute()
    assert [r.digest() for r in per] == [expected[i] for i in pick]
    b.close()


def test_per_class_kernel_variants(gpu, small_set, expected):
    """Each size class with its own kernel variant (one launch pair per class
    in the captured run): the per-graph results stay those of the plain
    finder, in the captured run and in the looped one."""
    b = Batch(small_set)
    variants = [(2, 7), (8, 7), (4, 0), (4, 16)]
    for c in range(b.get("n_classes")):
        w, l = variants[c % len(variants)]
        b.set(f"waves_per_block.{c}", w)
        b.set(f"lds_kb.{c}", l)
        assert (b.get(f"waves_per_block.{c}"), b.get(f"lds_kb.{c}")) == (w, l)
    per, _ = b.execute()
    assert [r.digest() for r in per] == expected
    per, _ = b.execute()  # a replay of the captured run
    assert [r.digest() for r in per] == expected
    light = _light(small_set)
    lb = Batch([small_set[i] for i in light])
    for c in range(lb.get("n_classes")):
        w, l = variants[c % len(variants)]
        lb.set(f"waves_per_block.{c}", w)
        lb.set(f"lds_kb.{c}", l)
    lb.set("per_graph_launch", 1)
    per, _ = lb.execute()
    assert [r.digest() for r in per] == [expected[i] for i in light]
    b.close()
    lb.close()


def test_light_graphs_of_the_first_1024_match_the_oracle(gpu):
    """Config 4's first 1,024 graphs, all those under 16,000 edges (778 of
    them, 8.3 x 10^7 bicliques): one batch, every graph's count and digest
    equal to the oracle's."""
    gs = [g for g in workloads.batch_graphs(1024) if g.stats()["edge_size"] < 16000]
    assert len(gs) > 700
    b = Batch(gs)
    per, _ = b.execute()
    b.close()
    for i, (g, r) in enumerate(zip(gs, per)):
        assert r.digest() == oracle.gmbe(g.oracle_graph(), threads=8).digest(), i


def test_bad_knobs_are_errors(gpu, small_set):
    b = Batch(small_set[:2])
    with pytest.raises(GmbeError):
        b.set("waves_per_block", 3)
    with pytest.raises(GmbeError):
        b.set("bound_height.9", 1)
    with pytest.raises(GmbeError):
        b.set("nonsense", 1)
    b.close()
