"""Device generators equal the oracle's restatement element by element, and
have the properties the workloads need (no loops, no duplicates, degrees)."""
import numpy as np
import pytest

import pyoracle as O
import zgraph as Z

pytestmark = pytest.mark.gpu


def _check(dev, ref_s, ref_d):
    s, d, _ = dev.download()
    assert (s == ref_s).all() and (d == ref_d).all()
    assert (s < d).all()
    key = s.astype(np.uint64) << np.uint64(32) | d
    assert len(np.unique(key)) == len(key)
    return s, d


def test_erdos_renyi_equals_oracle():
    ref = O.gen_er(10000, 100000, 1)
    _check(Z.erdos_renyi(10000, 100000, seed=1), *ref)


def test_chung_lu_equals_oracle():
    ref = O.gen_chung_lu(200000, 2_000_000, 2.1, 4000.0, 3)
    s, d = _check(Z.chung_lu(200000, 2_000_000, 2.1, 4000.0, 3), *ref)
    deg = np.bincount(np.concatenate([s, d]))
    assert 0.7 * 4000 < deg.max() <= 1.1 * 4000
    assert deg.sum() == 4_000_000


def test_rmat_equals_oracle():
    ref = O.gen_rmat(16, 16 << 16, 0.57, 0.19, 0.19, 5)
    s, d = _check(Z.rmat(16, 16 << 16, 0.57, 0.19, 0.19, seed=5), *ref)
    assert d.max() < (1 << 16)


def test_weights_uniform():
    e = Z.erdos_renyi(1000, 20000, seed=2).add_uniform_weights(7)
    _, _, w = e.download()
    assert w.dtype == np.float32 and (w >= 0).all() and (w < 1).all()
    assert abs(w.mean() - 0.5) < 0.01


@pytest.mark.parametrize("gname", ["wiki-vote", "snap-wiki-Vote", "simple-30", "powerlaw", "K9", "multi"])
def test_device_exact_counts_equal_oracle(gname):
    import helpers as H
    if gname == "powerlaw":
        n, s, d = H.power_law_graph(3000, 20000, 2.1, 7)
    elif gname == "multi":
        n, s, d = H.multigraph(24, 0.4, 5)
    elif gname == "K9":
        n, s, d = H.complete_graph(9)
    else:
        n, s, d = H.golden_graph(gname)
    g = Z.Graph.from_edges(s, d, n, seed=4)
    og = O.OGraph(n, s, d, seed=4)
    assert g.exact_count("triangle") == og.triangles()
    assert g.exact_count("k4") == og.k4()
    assert g.exact_count("wedge") == og.wedges()
    assert g.exact_count("path3") == og.paths3()
    assert g.exact_count("star5") == og.stars(4)
    if n <= 400:
        assert g.exact_count("house") == og.house_brute()
    assert g.exact_count("c4") == og.c4()
    assert g.exact_count("path4") == og.paths4()
    if n <= 400:
        assert g.exact_count("path4") == og.paths_brute(5)
    if gname == "snap-wiki-Vote":
        assert g.exact_count("triangle") == 608389


@pytest.mark.parametrize("n,p,seed", [(40, 0.3, 1), (60, 0.15, 2), (25, 0.6, 3)])
def test_device_house_count_equals_brute_force(n, p, seed):
    import helpers as H
    n, s, d = H.random_graph(n, p, seed)
    g = Z.Graph.from_edges(s, d, n, seed=1)
    og = O.OGraph(n, s, d, seed=1)
    assert g.exact_count("house") == og.house_brute()


def test_device_house_count_matches_the_reference_ground_truth():
    """applications/FiveHouse.cpp:68 hard-codes 2,134,981,304 houses for its
    graph: the simple graph of the reference's wiki-vote fixture has exactly
    that many."""
    import helpers as H
    n, s, d = H.golden_graph("wiki-vote")
    g = Z.Graph.from_edges(s, d, n, seed=1)
    assert g.exact_count("house") == 2134981304


def test_device_house_count_agrees_with_the_house_estimator():
    """G(90, 0.4): the exact count (new algorithm) lies in the 99.9 % interval
    of 2^26 unbiased SamplerGenerator house estimators."""
    import helpers as H
    n, s, d = H.random_graph(90, 0.4, 21)
    g = Z.Graph.from_edges(s, d, n, seed=1)
    H_exact = g.exact_count("house")
    N = 1 << 26
    r = Z.run(g, Z.Plan(Z.PLAN_PATTERN, pattern=Z.FIVE_HOUSE_PATTERN), 3, 0, N)[0]
    hw = Z.clt_half_width(r.mean, r.sumsq, N, 0.999)
    print(f"[house] exact {H_exact} estimate {r.mean:.6g} +- {hw:.3g} ({r.successes} successes)", flush=True)
    assert r.successes > 100
    assert abs(r.mean - H_exact) <= hw


@pytest.mark.parametrize("n", [66, 130, 200])
def test_device_clique_counts_of_complete_graphs(n):
    """K_n: C(n, 3) triangles and C(n, 4) 4-cliques.  Ids break the degree
    ties, so the common out-list of edge (0, 1) holds n - 2 members: more
    than one group of 64 for the 4-clique counter's per-member work."""
    from math import comb
    import helpers as H
    n, s, d = H.complete_graph(n)
    g = Z.Graph.from_edges(s, d, n, seed=2)
    assert g.exact_count("triangle") == comb(n, 3)
    assert g.exact_count("k4") == comb(n, 4)


@pytest.mark.parametrize("nh,nl", [(2200, 2300), (4200, 4300)])
def test_device_k4_count_with_long_out_lists(nh, nl):
    """Two vertices u, v share nh common out-neighbours H, so u's bit rows are
    ceil((nh + 1) / 64) words long (35: the wavefront path with one word per
    lane; 66: more words than lanes) and its rows are written both by
    ballots and through the LDS copy.  H holds a matching of 100 pairs, and
    every H vertex also links to all of nl >= nh low-degree vertices L, which
    keeps H above u and v in degree order.
    Closed forms: 4-cliques {u, v, h1, h2} per matched pair = 100; triangles
    {u, v, h} nh, {u | v, h1, h2} 200, {h1, h2, l} 100 * nl."""
    pairs = 100
    u, v = 0, 1
    H0, L0 = 2, 2 + nh
    hs = np.arange(H0, H0 + nh, dtype=np.uint32)
    ls = np.arange(L0, L0 + nl, dtype=np.uint32)
    src = [np.array([u], np.uint32), np.full(nh, u, np.uint32), np.full(nh, v, np.uint32),
           hs[0:2 * pairs:2], np.repeat(hs, nl)]
    dst = [np.array([v], np.uint32), hs, hs, hs[1:2 * pairs:2], np.tile(ls, nh)]
    s, d = np.concatenate(src), np.concatenate(dst)
    g = Z.Graph.from_edges(s, d, L0 + nl, seed=3)
    assert g.exact_count("k4") == pairs
    assert g.exact_count("triangle") == nh + 2 * pairs + pairs * nl


@pytest.mark.parametrize("n", [600, 1500, 4097])
def test_device_k4_count_wide_rows_on_complete_graphs(n):
    """K_n with n - 1 >= 513: the high vertices' out-lists (513..4096 entries,
    in the orientation's order) have triangular bit rows of 9..64 words,
    counted one wavefront per row with the lanes packed in groups; K_4097's
    vertex 0 has 4096 out-neighbours, the longest sorted list.  C(n, 4)
    4-cliques."""
    from math import comb
    import helpers as H
    n, s, d = H.complete_graph(n)
    g = Z.Graph.from_edges(s, d, n, seed=2)
    assert g.exact_count("k4") == comb(n, 4)


_DENSE_K4 = {}


@pytest.mark.parametrize("scratch_words", [None, "300000"])
def test_device_k4_count_wide_rows_equals_oracle(monkeypatch, scratch_words):
    """A dense random graph (1200 vertices, p = 0.6): out-lists of up to ~700
    entries (wide triangular rows); the count equals the oracle's, in one
    pass and in many vertex ranges."""
    import helpers as H
    if scratch_words:
        monkeypatch.setenv("ZGRAPH_K4_SCRATCH_WORDS", scratch_words)
    n, s, d = H.random_graph(1200, 0.6, 31)
    if "k4" not in _DENSE_K4:   # the oracle's count once for both scratch sizes
        _DENSE_K4["k4"] = O.OGraph(n, s, d, seed=1).k4()
    g = Z.Graph.from_edges(s, d, n, seed=1)
    assert g.exact_count("k4") == _DENSE_K4["k4"] == 4007890405


@pytest.mark.parametrize("scratch_words", [None, "20000"])
def test_device_k4_count_rmat16_in_one_and_in_many_passes(monkeypatch, scratch_words):
    """R-MAT scale 16 (1 M edges): the 4-clique count equals the oracle's,
    with the default scratch and with a scratch of 20 000 words, which
    splits the vertices into many ranges counted one after another."""
    if scratch_words:
        monkeypatch.setenv("ZGRAPH_K4_SCRATCH_WORDS", scratch_words)
    s, d = O.gen_rmat(16, 16 << 16, 0.57, 0.19, 0.19, 1)
    og = O.OGraph(1 << 16, s, d, seed=1)
    g = Z.Graph.from_edges(s, d, 1 << 16, seed=1)
    assert g.exact_count("k4") == og.k4() == 456481728


@pytest.mark.parametrize("n,sort_wedges,wide", [(9, None, False), (70, None, False), (1000, None, False),
                                                (1000, "5000", False), (1000, "1", False), (1000, None, True),
                                                (1000, "5000", True)])
def test_device_c4_and_five_chains_of_complete_graphs(monkeypatch, n, sort_wedges, wide):
    """K_n: 3 C(n, 4) 4-cycles and n!/(n-5)!/2 5-vertex chains.  In K_1000
    the top vertices see up to 998 distinct opposite vertices, more than the
    LDS hash holds, so they are counted again by sorting their wedges; with
    the sort batches limited to 5 000 wedges the larger of them take the dense
    counters, and with 1 every one of them does."""
    from math import comb, perm
    import helpers as H
    if sort_wedges:
        monkeypatch.setenv("ZGRAPH_C4_SORT_WEDGES", sort_wedges)
    if wide:   # 64-bit sort keys, batches of at most 2 vertices
        monkeypatch.setenv("ZGRAPH_C4_WIDE_KEYS", "1")
        monkeypatch.setenv("ZGRAPH_C4_BATCH_VERTICES", "2")
    n, s, d = H.complete_graph(n)
    g = Z.Graph.from_edges(s, d, n, seed=2)
    assert g.exact_count("c4") == 3 * comb(n, 4)
    assert g.exact_count("path4") == perm(n, 5) // 2
    # houses: 5! / 2 per 5 vertices, through the same LDS, sorted (weighted) and dense paths
    assert g.exact_count("house") == 60 * comb(n, 5)


@pytest.mark.parametrize("a,b,sort_wedges", [(2, 60000, None), (300, 300, None), (300, 300, "5000")])
def test_device_c4_complete_bipartite(monkeypatch, a, b, sort_wedges):
    """K_{a,b}: C(a, 2) C(b, 2) 4-cycles.  K_{2,60000} puts 1.8e9 of them on
    one opposite pair (one LDS slot); in K_{300,300} the top vertices overflow
    the LDS hash into the sorted batches (and, with batches limited to 5 000
    wedges, into the dense counters), each summed in 128 bits."""
    from math import comb
    if sort_wedges:
        monkeypatch.setenv("ZGRAPH_C4_SORT_WEDGES", sort_wedges)
    u, v = np.meshgrid(np.arange(a), np.arange(a, a + b), indexing="ij")
    g = Z.Graph.from_edges(u.ravel().astype(np.uint32), v.ravel().astype(np.uint32), a + b, seed=3)
    assert g.exact_count("c4") == comb(a, 2) * comb(b, 2)
    assert g.exact_count("house") == 0   # no triangles


@pytest.mark.parametrize("env", [{}, {"ZGRAPH_C4_WIDE_KEYS": "1"}, {"ZGRAPH_C4_BATCH_VERTICES": "3"},
                                 {"ZGRAPH_C4_WIDE_KEYS": "1", "ZGRAPH_C4_BATCH_VERTICES": "2"},
                                 {"ZGRAPH_HOUSE_WIDE_VALUES": "1"},
                                 {"ZGRAPH_HOUSE_WIDE_VALUES": "1", "ZGRAPH_C4_WIDE_KEYS": "1"}],
                         ids=["default", "wide-keys", "3-vertex-batches", "wide-keys-2-vertex-batches",
                              "house-wide-values", "house-wide-values-wide-keys"])
def test_device_c4_and_five_chains_rmat14(monkeypatch, env):
    """R-MAT scale 14 (262 k edges, hubs of a few thousand): device counts
    equal the oracle's pair-counting 4-cycles and closed-form chains -- also
    with the hubs' wedges sorted under 64-bit keys (the path of graphs past
    2^30 vertices) and with batches split at a few vertices (the split that
    32-bit keys force at 2^(32 - ceil(log2 n)) vertices)."""
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    s, d = O.gen_rmat(14, 16 << 14, 0.57, 0.19, 0.19, 3)
    og = O.OGraph(1 << 14, s, d, seed=1)
    g = Z.Graph.from_edges(s, d, 1 << 14, seed=1)
    assert g.exact_count("c4") == og.c4()
    assert g.exact_count("path4") == og.paths4()
    assert g.exact_count("house") == og.house()


def test_device_house_count_through_the_dense_weighted_counters(monkeypatch):
    """R-MAT 14 with the hub batches limited to 2 000 wedges: the larger hubs'
    weighted wedges go to the dense counters, the rest to the sorted batches;
    the count equals the oracle's closed form."""
    monkeypatch.setenv("ZGRAPH_C4_SORT_WEDGES", "2000")
    s, d = O.gen_rmat(14, 16 << 14, 0.57, 0.19, 0.19, 3)
    og = O.OGraph(1 << 14, s, d, seed=1)
    g = Z.Graph.from_edges(s, d, 1 << 14, seed=1)
    assert g.exact_count("house") == og.house()


@pytest.mark.parametrize("order,what", [(4, "path3"), (5, "path4")])
def test_chain_mining_estimate_covers_the_exact_chain_count(order, what):
    """ChainMining(order) (applications/ChainMining.cpp) on R-MAT scale 14:
    the exact count of order-vertex chains lies in the 99.9 % CLT interval of
    2^26 estimators."""
    s, d = O.gen_rmat(14, 16 << 14, 0.57, 0.19, 0.19, 3)
    g = Z.Graph.from_edges(s, d, 1 << 14, seed=1)
    exact = g.exact_count(what)
    N = 1 << 26
    r = Z.run(g, Z.Plan(Z.PLAN_CHAIN, order=order), 5, 0, N)[0]
    hw = Z.clt_half_width(r.mean, r.sumsq, N, 0.999)
    print(f"[chain{order}] exact {exact} estimate {r.mean:.6g} +- {hw:.3g} ({r.successes} successes)", flush=True)
    assert r.successes > 1000
    assert abs(r.mean - exact) <= hw


def _ring(n, first=0):
    return [(first + i, first + (i + 1) % n) for i in range(n)]


@pytest.mark.parametrize("name,n,edges,c4,chains", [
    ("path P10", 10, [(i, i + 1) for i in range(9)], 0, 6),
    ("cycle C10", 10, _ring(10), 0, 10),
    ("cycle C5", 5, _ring(5), 0, 5),
    ("cycle C4", 4, _ring(4), 1, 0),
    ("two C4 and a K4", 12, _ring(4) + _ring(4, 4) + [(8, 9), (8, 10), (8, 11), (9, 10), (9, 11), (10, 11)], 5, 0),
    ("star S20", 21, [(0, i) for i in range(1, 21)], 0, 0),
    ("isolated vertices only", 7, [], 0, 0),
])
def test_device_c4_and_five_chains_of_small_structures(name, n, edges, c4, chains):
    """Closed forms: a path of 10 vertices holds 6 chains of 5, a cycle C_n
    (n >= 5) n of them and no 4-cycle, C4 one 4-cycle, K4 three 4-cycles and
    4!/2 = 12 chains of 4 vertices but none of 5, a star neither."""
    s = np.array([a for a, _ in edges], np.uint32)
    d = np.array([b for _, b in edges], np.uint32)
    g = Z.Graph.from_edges(s, d, n, seed=1)
    assert g.exact_count("c4") == c4, name
    assert g.exact_count("path4") == chains, name
    og = O.OGraph(n, s, d)
    assert og.c4() == c4 and og.paths4() == chains and og.paths_brute(5) == chains


def test_device_five_star_count_near_and_past_64_bits():
    """A star of 100 000 leaves holds C(100000, 4) ~ 4.2e18 5-stars (exact,
    near 2^64); one of 150 000 leaves holds ~2.1e19, past 2^64, which the call
    reports as an overflow instead of wrapping."""
    from math import comb
    for leaves, ok in ((100_000, True), (150_000, False)):
        s = np.zeros(leaves, np.uint32)
        d = np.arange(1, leaves + 1, dtype=np.uint32)
        g = Z.Graph.from_edges(s, d, leaves + 1, seed=1)
        if ok:
            assert g.exact_count("star5") == comb(leaves, 4)
        else:
            with pytest.raises(Z.ZGraphError):
                g.exact_count("star5")
