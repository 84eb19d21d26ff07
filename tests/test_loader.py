"""CSRBiGraph::Read semantics (src/BiGraph.cu:13-278) on hand-written .adj files."""
import numpy as np
import pytest

from gmbe_amd import BiGraph


def write(tmp_path, text, name="g.adj"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def rows_of(g):
    v_off, v_adj, u_off, u_adj, u_orig, v_orig = g.csr()
    rows = [list(v_adj[v_off[i]:v_off[i + 1]]) for i in range(len(v_off) - 1)]
    cols = [list(u_adj[u_off[i]:u_off[i + 1]]) for i in range(len(u_off) - 1)]
    return rows, cols, list(u_orig), list(v_orig)


def test_plain_read(tmp_path):
    # line i = V vertex i; tokens = U ids; |U| = max id + 1 (src/BiGraph.cu:27-37)
    g = BiGraph.read_adj(write(tmp_path, "0 2\n1\n0 1 2\n"), uvtrans=0, order=0, fast_mode=1)
    s = g.stats()
    assert (s["U_size"], s["V_size"], s["edge_size"]) == (3, 3, 6)
    rows, cols, _, _ = rows_of(g)
    assert rows == [[0, 2], [1], [0, 1, 2]]
    assert cols == [[0, 2], [1, 2], [0, 2]]  # U lists ascending


def test_rows_sorted_and_deduplicated(tmp_path):
    g = BiGraph.read_adj(write(tmp_path, "2 0 2\n1 1\n"), order=0, fast_mode=1)
    rows, _, _, _ = rows_of(g)
    assert rows == [[0, 2], [1]]
    assert g.stats()["edge_size"] == 3


def test_empty_line_is_degree_zero_vertex(tmp_path):
    g = BiGraph.read_adj(write(tmp_path, "0\n\n1\n"), order=0, fast_mode=1)
    rows, _, _, _ = rows_of(g)
    assert rows == [[0], [], [1]]


def test_uvtrans(tmp_path):
    text = "0 1\n1\n0\n"  # |U| = 2 < |V| = 3
    g0 = BiGraph.read_adj(write(tmp_path, text), uvtrans=0, order=0, fast_mode=1)
    g1 = BiGraph.read_adj(write(tmp_path, text), uvtrans=1, order=0, fast_mode=1)
    g2 = BiGraph.read_adj(write(tmp_path, text), uvtrans=2, order=0, fast_mode=1)
    assert (g0.stats()["V_size"], g0.stats()["swapped"]) == (3, 0)
    assert (g1.stats()["V_size"], g1.stats()["swapped"]) == (2, 1)  # swap when |U| < |V|
    assert (g2.stats()["V_size"], g2.stats()["swapped"]) == (3, 0)  # swap only when |U| > |V|
    rows, _, _, _ = rows_of(g1)
    assert rows == [[0, 2], [0, 1]]  # old U lists become the rows


def test_order_degree_then_lexicographic(tmp_path):
    text = "0 1 2\n3\n1 2\n0 2\n"
    g = BiGraph.read_adj(write(tmp_path, text), order=1, fast_mode=1)
    rows, _, _, v_orig = rows_of(g)
    assert rows == [[3], [0, 2], [1, 2], [0, 1, 2]]  # degree asc, then lexicographic
    assert v_orig == [1, 3, 2, 0]
    g = BiGraph.read_adj(write(tmp_path, text), order=2, fast_mode=1)
    rows, _, _, _ = rows_of(g)
    assert rows == [[0, 1, 2], [0, 2], [1, 2], [3]]


def test_fast_mode_order3_sorts_descending(tmp_path):
    # the reference's fast path treats every order other than 0 and 1 as
    # descending (src/BiGraph.cu:236-237)
    text = "0\n0 1 2\n1 2\n"
    g = BiGraph.read_adj(write(tmp_path, text), order=3, fast_mode=1)
    rows, _, _, _ = rows_of(g)
    assert rows == [[0, 1, 2], [1, 2], [0]]


def test_two_hop_statistics(tmp_path):
    # V0:{0,1} V1:{1,2} V2:{3}; d(U)=2 (U1), d(V)=2,
    # d2(V): V0 reaches {V0,V1} -> 1; d2(U): U1 reaches {U0,U1,U2} -> 2
    g = BiGraph.read_adj(write(tmp_path, "0 1\n1 2\n3\n"), order=0, fast_mode=0)
    s = g.stats()
    assert (s["U_degree"], s["V_degree"], s["U_2degree"], s["V_2degree"]) == (2, 2, 2, 1)
    g = BiGraph.read_adj(write(tmp_path, "0 1\n1 2\n3\n"), order=0, fast_mode=1)
    assert g.stats()["U_2degree"] == -1


def test_unilateral_order_is_a_permutation(tmp_path):
    rng = np.random.default_rng(5)
    lines = []
    for _ in range(40):
        r = sorted(set(int(x) for x in rng.integers(0, 30, size=rng.integers(1, 6))))
        lines.append(" ".join(map(str, r)))
    path = write(tmp_path, "\n".join(lines) + "\n")
    g = BiGraph.read_adj(path, order=3, fast_mode=0)
    rows, _, _, v_orig = rows_of(g)
    assert sorted(v_orig) == list(range(40))
    src = [sorted(set(int(x) for x in l.split())) for l in lines]
    assert [src[o] for o in v_orig] == rows


def test_unilateral_order_known_answer(tmp_path):
    """Order 3 on a graph whose 2-hop graph is a star (centre V0, leaves
    V1-V3) beside a triangle (V4-V6), each 2-hop edge one U vertex of degree
    two.  Worked by hand through the reference's bucket peeling
    (src/BiGraph.cu:93-210): 2-hop degrees 3,1,1,1,2,2,2 give the initial
    order V1 V2 V3 | V4 V5 V6 | V0.  Taking V1 lowers V0 to 2 (it already sits
    first in its new bucket); taking V2 lowers V0 to 1, which swaps it with
    V4, the first of bucket 2.  The order is V1 V2 V3 V0 V5 V6 V4: a plain
    sort by 2-hop degree would keep V4 before V5 and V6."""
    rows = ["0 1 2", "0", "1", "2", "3 4", "3 5", "4 5"]
    g = BiGraph.read_adj(write(tmp_path, "\n".join(rows) + "\n"), uvtrans=0, order=3,
                         fast_mode=0)
    _, _, _, v_orig = rows_of(g)
    assert list(v_orig) == [1, 2, 3, 0, 5, 6, 4]


def test_derive_keeps_original_ids(tmp_path):
    text = "0 1 2\n3\n1 2\n0 2\n2 3 4\n"
    raw = BiGraph.read_adj(write(tmp_path, text), order=0, fast_mode=1)
    g = raw.derive(uvtrans=1, order=1, fast_mode=1)
    rows, _, u_orig, v_orig = rows_of(g)
    s = g.stats()
    # every internal edge maps back to an original edge
    src = [set(int(x) for x in l.split()) for l in text.strip().split("\n")]
    for v, r in enumerate(rows):
        for u in r:
            a, b = (v_orig[v], u_orig[u]) if s["swapped"] else (u_orig[u], v_orig[v])
            assert a in src[b]


def test_write_adj_roundtrip(tmp_path):
    g = BiGraph.read_adj(write(tmp_path, "0 2\n1\n0 1 2\n"), order=0, fast_mode=1)
    out = tmp_path / "out.adj"
    g.write_adj(str(out))
    assert out.read_text() == "0 2\n1\n0 1 2\n"


def test_write_edges_matches_the_adj_rows(tmp_path):
    """The edge-list writer (preprocess/adj_to_graph.cpp:31-39): header
    "rows cols", then the edges row by row in original ids; a derived graph
    (side swap, reordering) writes the same file as its source."""
    rows = [[3, 1], [0], [], [2, 5, 1]]  # |U| = 6 > |V| = 4: uvtrans=2 swaps
    p = tmp_path / "g.adj"
    p.write_text("\n".join(" ".join(map(str, r)) for r in rows) + "\n")
    g = BiGraph.read_adj(str(p))
    out = tmp_path / "g.graph"
    g.write_edges(str(out))
    lines = out.read_text().splitlines()
    assert lines[0] == "4 6"
    want = [f"{v} {u}" for v, r in enumerate(rows) for u in sorted(r)]
    assert lines[1:] == want
    for opt in (dict(uvtrans=1, order=1, fast_mode=1), dict(uvtrans=2, order=2, fast_mode=0)):
        d = g.derive(**opt)
        assert d.stats()["swapped"] == (1 if opt["uvtrans"] == 2 else 0)
        out2 = tmp_path / "d.graph"
        d.write_edges(str(out2))
        assert out2.read_text() == out.read_text()
