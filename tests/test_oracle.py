"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each check is independent of the formula it checks: published known-answer
vectors, an independent implementation (rocRAND), closed forms, brute-force
enumeration, exhaustive expectation over every random choice, the
reference's own naive counters built from source, and the reference tests'
printed values.
"""
import json
import math
import os
import subprocess

import numpy as np
import pytest

import helpers as H
import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ---- Philox -----------------------------------------------------------------
def test_philox_random123_kat_and_rocrand():
    with open(os.path.join(H.GOLDEN, "philox_vectors.json")) as f:
        vecs = json.load(f)["vectors"]
    assert len(vecs) >= 67
    for v in vecs:
        assert O.philox(v["ctr"], v["key"]) == v["out"]


def _stream_words(seed, tag, eid, k):
    out = []
    for b in range((k + 3) // 4):
        out += O.philox([b, tag, eid & 0xffffffff, eid >> 32], [seed & 0xffffffff, seed >> 32])
    return out[:k]


# ---- bounded draws ------------------------------------------------------------
def _draws(n, count, seed=7):
    """count draws of U[0, n) from the estimator stream of ids 0.. (one per id)"""
    L = O.lib()
    import ctypes as C
    st = (C.c_uint8 * 64)()
    res, words = [], []
    L.og_stream_init.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint64]
    L.og_uniform.argtypes = [C.c_void_p, C.c_uint64]
    L.og_uniform.restype = C.c_uint64
    for i in range(count):
        L.og_stream_init(st, seed, 0, i)
        res.append(L.og_uniform(st, n))
        words.append(int.from_bytes(bytes(st[24:32]), "little"))   # og_stream.word (after key, tag, id)
    return np.array(res, dtype=np.uint64), np.array(words)


def test_uniform_power_of_two_is_top_bits():
    # n = 2^k never rejects and returns the top k bits of the first word
    for k in (1, 5, 17, 32):
        r, w = _draws(1 << k, 200)
        assert (w == 1).all()
        first = np.array([_stream_words(7, 0, i, 1)[0] for i in range(200)], dtype=np.uint64)
        assert (r == (first >> np.uint64(32 - k))).all()


def test_uniform_rejection_rate_and_range():
    # n = 2^31 + 1 rejects a word with probability (2^32 mod n)/2^32 ~ 1/2
    n = (1 << 31) + 1
    r, w = _draws(n, 4000)
    assert (r < n).all()
    extra = (w - 1).mean()            # expected words beyond the first: p/(1-p) ~ 1
    assert 0.85 < extra < 1.15
    # n > 2^32 uses two words per attempt
    r, w = _draws((1 << 40) + 3, 500)
    assert (r < (1 << 40) + 3).all() and (w % 2 == 0).all()


@pytest.mark.parametrize("n", [3, 7, 1000])
def test_uniform_chi_square(n):
    r, _ = _draws(n, 30000, seed=11)
    cnt = np.bincount(r.astype(np.int64), minlength=n)
    exp = len(r) / n
    chi = ((cnt - exp) ** 2 / exp).sum()
    dof = n - 1
    assert chi < dof + 6 * math.sqrt(2 * dof)


# ---- CSR construction ----------------------------------------------------------
@pytest.mark.parametrize("name", ["simple-5", "simple-10", "simple-30", "small-10", "wiki-vote"])
def test_csr_invariants(name):
    n, s, d = H.golden_graph(name)
    g = O.OGraph(n, s, d, seed=3)
    a = g.arrays()
    m = len(s)
    assert a["off"][-1] == 2 * m
    deg = np.diff(a["off"].astype(np.int64))
    assert (deg == np.bincount(np.concatenate([s, d]).astype(np.int64), minlength=n)).all()
    # each rank appears twice, once in each endpoint's list
    assert (np.bincount(a["adj_rank"].astype(np.int64), minlength=m) == 2).all()
    for v in range(min(n, 2000)):
        lo, hi = a["off"][v], a["off"][v + 1]
        rk = a["adj_rank"][lo:hi]
        assert (np.diff(rk.astype(np.int64)) > 0).all()      # ascending by rank (src/Graph.cpp:324-334)
        assert (a["adj_src"][lo:hi] == v).all()
        for r, nb in zip(rk, a["adj_dst"][lo:hi]):
            assert {a["src"][r], a["dst"][r]} == {v, nb}
    # the rank order is a permutation of the input edges
    inp = sorted(zip(s.tolist(), d.tolist()))
    assert sorted(zip(a["src"].tolist(), a["dst"].tolist())) == inp


def test_csr_rejects_self_loops():
    with pytest.raises(ValueError):
        O.OGraph(3, np.array([0, 1]), np.array([1, 1]))


def test_subsample_and_filter():
    n, s, d = H.golden_graph("wiki-vote")
    w = H.golden_weights("wiki-vote")
    g = O.OGraph(n, s, d, w=w, seed=5)
    sg = g.subsample(0.1, 9)
    frac = sg.m / g.m
    assert abs(frac - 0.1) < 0.01
    f = g.filter(50.0, True, 4)
    assert f.m == int((w >= 50).sum())
    f2 = g.filter(50.0, False, 4)
    assert f.m + f2.m == g.m


# ---- primitives by brute force ---------------------------------------------------
def _brute_entries(a, min_valid, verts):
    out = []
    for v in verts:
        for i in range(a["off"][v], a["off"][v + 1]):
            if a["adj_rank"][i] >= min_valid:
                out.append((a["adj_rank"][i], a["adj_src"][i], a["adj_dst"][i]))
    return out


def test_conditional_pick_matches_enumeration():
    import ctypes as C
    n, s, d = H.random_graph(9, 0.6, 4)
    # add a hub so that both the linear (deg <= 32) and binary paths run
    hub_s = np.full(40, 9, np.uint32)
    hub_d = np.arange(10, 50, dtype=np.uint32)
    s, d = np.concatenate([s, hub_s]), np.concatenate([d, hub_d])
    g = O.OGraph(50, s, d, seed=2)
    a = g.arrays()
    rng = np.random.default_rng(1)
    L = O.lib()
    for trial in range(300):
        nv = int(rng.integers(1, 4))
        verts = rng.choice(50, nv, replace=False).astype(np.uint32)
        if trial % 3 == 0:
            verts[0] = 9
        mv = int(rng.integers(0, g.m + 2))
        want = _brute_entries(a, mv, verts)
        pos = np.zeros(16, np.uint64)
        cnt = L.og_cond_valid_count(C.byref(g.g), mv, verts.ctypes.data, nv, pos.ctypes.data)
        assert cnt == len(want)
        for r in range(len(want)):
            rk, sr, ds = C.c_uint64(), C.c_uint32(), C.c_uint32()
            assert L.og_cond_pick(C.byref(g.g), mv, verts.ctypes.data, nv, r, C.byref(rk), C.byref(sr), C.byref(ds))
            assert (rk.value, sr.value, ds.value) == tuple(int(x) for x in want[r])


def test_close_matches_edge_set():
    import ctypes as C
    n, s, d = H.random_graph(12, 0.4, 8)
    g = O.OGraph(n, s, d, seed=6)
    a = g.arrays()
    rank = {}
    for r in range(g.m):
        rank[(min(a["src"][r], a["dst"][r]), max(a["src"][r], a["dst"][r]))] = r
    L = O.lib()
    rng = np.random.default_rng(3)
    for _ in range(500):
        x, y = rng.choice(n, 2, replace=False)
        mv = int(rng.integers(0, g.m + 1))
        verts = np.array([x, y, rng.integers(0, n)], np.uint32)
        ws, wd = np.array([x], np.uint32), np.array([y], np.uint32)
        got = L.og_close(C.byref(g.g), mv, verts.ctypes.data, 3, ws.ctypes.data, wd.ctypes.data, 1)
        key = (min(x, y), max(x, y))
        assert bool(got) == (key in rank and rank[key] >= mv)


def test_close_on_parallel_edges_matches_any_copy():
    # with parallel edges the awaited edge is closed when some copy has rank >= min_valid
    import ctypes as C
    n, s, d = H.multigraph(10, 0.5, 4, max_copies=5, heavy_copies=12)
    g = O.OGraph(n, s, d, seed=2)
    a = g.arrays()
    ranks = {}
    for r in range(g.m):
        ranks.setdefault((min(a["src"][r], a["dst"][r]), max(a["src"][r], a["dst"][r])), []).append(r)
    assert max(len(v) for v in ranks.values()) == 12
    L = O.lib()
    rng = np.random.default_rng(5)
    for _ in range(800):
        x, y = rng.choice(n, 2, replace=False)
        mv = int(rng.integers(0, g.m + 1))
        verts = np.array([x, y], np.uint32)
        ws, wd = np.array([x], np.uint32), np.array([y], np.uint32)
        got = L.og_close(C.byref(g.g), mv, verts.ctypes.data, 2, ws.ctypes.data, wd.ctypes.data, 1)
        key = (min(x, y), max(x, y))
        assert bool(got) == any(r >= mv for r in ranks.get(key, []))


def test_exact_counters_collapse_parallel_edges():
    # the naive_implementation counts are over the simple graph under the edge list
    n, s, d = H.multigraph(12, 0.5, 6)
    multi = O.OGraph(n, s, d)
    pairs = sorted(set((min(x, y), max(x, y)) for x, y in zip(s.tolist(), d.tolist())))
    ss = np.array([p[0] for p in pairs], np.uint32)
    sd = np.array([p[1] for p in pairs], np.uint32)
    simple = O.OGraph(n, ss, sd)
    assert multi.m == len(s) > simple.m == len(pairs)
    assert multi.triangles() == simple.pattern_brute(*O.TRIANGLE)
    assert multi.k4() == simple.pattern_brute(*O.K4)
    assert multi.wedges() == simple.paths_brute(3)
    assert multi.paths3() == simple.paths_brute(4)


# ---- estimators: exhaustive expectation equals the exact count ---------------------
def _exact(g, name):
    if name in ("triangle", "pattern_triangle"):
        return g.triangles()
    if name in ("k4_close", "k4_split", "pattern_k4"):
        return g.k4()
    if name == "three_chain":
        return g.wedges()
    if name in ("chain4", "four_chain"):
        return g.paths_brute(4)
    if name == "chain5":
        return g.paths_brute(5)
    if name == "pattern_5star":
        return g.pattern_brute(*O.FIVE_STAR)
    if name == "pattern_house":
        return g.house_brute()
    raise KeyError(name)


TINY = [H.complete_graph(5), H.random_graph(7, 0.6, 1), H.random_graph(7, 0.5, 2),
        H.random_graph(8, 0.45, 3)]


@pytest.mark.parametrize("gi", range(len(TINY)))
@pytest.mark.parametrize("name,oargs", [(p[0], p[1]) for p in H.PLANS if p[0] not in ("k4_split", "house_split")])
def test_unbiased_by_exhaustive_expectation(gi, name, oargs):
    n, s, d = TINY[gi]
    for seed in (1, 2):
        g = O.OGraph(n, s, d, seed=seed)
        plan = H.oracle_plan(O, **oargs)
        if name == "three_motif":
            assert O.expectation(g, plan, 0) == pytest.approx(g.wedges(), abs=1e-7)
            assert O.expectation(g, plan, 1) == pytest.approx(g.triangles(), abs=1e-7)
            continue
        e = O.expectation(g, plan, 0)
        assert e >= 0, "too many paths"
        assert e == pytest.approx(_exact(g, name), rel=1e-9, abs=1e-7)


@pytest.mark.parametrize("graph,J", [(H.complete_graph(4), 1), (H.complete_graph(4), 2), (H.complete_graph(5), 2),
                                     (H.random_graph(7, 0.6, 1), 2)] + [(t, 1) for t in TINY])
def test_split_k4_unbiased_by_exhaustive_expectation(graph, J):
    # every draw of the first stage and of all J completions enumerated: the
    # mean of J completions over J is the 4-clique count for each rank order
    n, s, d = graph
    for seed in (1, 2):
        g = O.OGraph(n, s, d, seed=seed)
        e = O.expectation(g, O.make_plan(O.PLAN_K4_SPLIT, order=J), 0)
        assert e >= 0, "too many paths"
        assert e == pytest.approx(g.k4(), rel=1e-9, abs=1e-7)


def test_split_k4_first_stage_is_the_k4_plans():
    # same first edge per id; no completion unless the 4-clique plan's first
    # stage closed its triangle (its ncond reaches 4: the three shapes' samples)
    n, s, d = H.power_law_graph(400, 3000, 2.1, 5)
    g = O.OGraph(n, s, d, seed=4)
    k4 = O.run_outcomes(g, O.make_plan(O.PLAN_K4_CLOSE), 9, 100, 60000)
    for J, nc in ((1, 4), (16, 49), (128, 255)):
        sp = O.run_outcomes(g, O.make_plan(O.PLAN_K4_SPLIT, order=J), 9, 100, 60000)
        assert (sp["e0"] == k4["e0"]).all()
        closed = k4["ncond"] == 4
        assert closed.sum() > 100
        assert (sp["status"][~closed] == 0).all() and (sp["ncond"][~closed] == k4["ncond"][~closed]).all()
        assert (sp["ncond"][closed] == nc).all()
        ok = sp["status"] == 1
        assert ok.any() and (sp["log2den"][ok] == int(math.log2(J))).all()


def test_split_k4_variance_falls_with_completions():
    # the same estimator ids with 1, 16 and 64 completions: the mean stays at
    # the count and the per-estimator variance falls about J-fold until the
    # first stage's share of it dominates (here between 16 and 64 completions:
    # 22x and 26x below one completion); one shared completion stream per
    # estimator, or an unscaled sum, would fail this
    n, s, d = H.power_law_graph(2000, 30000, 2.1, 11)
    g = O.OGraph(n, s, d, seed=1)
    K4 = g.k4()
    N = 200_000
    var = {}
    for J in (1, 16, 64):
        o = O.run_outcomes(g, O.make_plan(O.PLAN_K4_SPLIT, order=J), 3, 0, N)
        v = (o["num_lo"].astype(np.float64) + o["num_hi"].astype(np.float64) * 2.0 ** 64) / 2.0 ** o["log2den"]
        v[o["status"] != 1] = 0
        var[J] = v.var()
        assert abs(v.mean() - K4) <= 4.5 * math.sqrt(var[J] / N), (J, v.mean(), K4)
    assert var[16] < var[1] / 6 and var[64] < var[1] / 6


def _houses_copy_weighted(n, s, d):
    """Brute force, independent of the estimator's decomposition: every
    injective map of the house pattern (applications/FiveHouse.cpp:18-27) into
    the graph, weighted by the product of its six pairs' multiplicities, over
    the pattern's two automorphisms."""
    from collections import Counter
    from itertools import permutations
    mult = Counter()
    for a, b in zip(s.tolist(), d.tolist()):
        mult[(a, b)] += 1
        mult[(b, a)] += 1
    pat = [(0, 1), (0, 2), (1, 2), (1, 3), (2, 4), (3, 4)]
    tot = 0
    for phi in permutations(range(n), 5):
        w = 1
        for u, v in pat:
            w *= mult[(phi[u], phi[v])]
            if not w:
                break
        tot += w
    return tot // 2


def test_house_brute_force_closed_forms():
    # K_n holds C(n, 5) * 5! / 2 houses (the pattern has two automorphisms);
    # doubling every edge of K5 multiplies each house's weight by 2^6
    for k in (5, 6, 7):
        n, s, d = H.complete_graph(k)
        assert _houses_copy_weighted(n, s, d) == math.comb(k, 5) * 60
        assert O.OGraph(n, s, d).house_brute() == math.comb(k, 5) * 60
    n, s, d = H.complete_graph(5)
    assert _houses_copy_weighted(n, np.concatenate([s, d]), np.concatenate([d, s])) == 60 * 64
    for gi in range(1, len(TINY)):
        n, s, d = TINY[gi]
        assert _houses_copy_weighted(n, s, d) == O.OGraph(n, s, d).house_brute()


HOUSE_GRAPHS = [H.complete_graph(5), H.random_graph(7, 0.6, 1), H.random_graph(8, 0.5, 2),
                H.random_graph(8, 0.45, 3), H.multigraph(6, 0.7, 4, max_copies=2, heavy_copies=3)]


@pytest.mark.parametrize("gi", range(len(HOUSE_GRAPHS)))
@pytest.mark.parametrize("J", [1, 2])
def test_split_house_unbiased_by_exhaustive_expectation(gi, J):
    # every draw of the base edge, the roof and all J completions enumerated:
    # the expectation is the copy-weighted house count (the simple count
    # without parallel edges), for either rank order
    n, s, d = HOUSE_GRAPHS[gi]
    want = _houses_copy_weighted(n, s, d)
    assert want > 0
    for seed in (1, 2):
        g = O.OGraph(n, s, d, seed=seed)
        e = O.expectation(g, O.make_plan(O.PLAN_HOUSE_SPLIT, order=J), 0, max_paths=200_000_000)
        assert e >= 0, "too many paths"
        assert e == pytest.approx(want, rel=1e-9)


def test_split_house_outcome_fields():
    # the first stage's words and ncond: 1 after the roof, 1 + 2 J once it
    # closes; successes carry denominator J
    n, s, d = H.power_law_graph(400, 3000, 2.1, 5)
    g = O.OGraph(n, s, d, seed=4)
    base = O.run_outcomes(g, O.make_plan(O.PLAN_HOUSE_SPLIT, order=1), 9, 100, 60000)
    for J in (4, 128):
        o = O.run_outcomes(g, O.make_plan(O.PLAN_HOUSE_SPLIT, order=J), 9, 100, 60000)
        assert (o["e0"] == base["e0"]).all() and (o["nwords"] == base["nwords"]).all()
        closed = o["ncond"] > 1
        assert closed.sum() > 100 and (closed == (base["ncond"] > 1)).all()
        assert (o["ncond"][closed] == min(1 + 2 * J, 255)).all() and (o["ncond"][~closed] == 1).all()
        ok = o["status"] == 1
        assert ok.any() and (o["log2den"][ok] == int(math.log2(J))).all()


def test_split_house_reference_ground_truth_on_wiki_vote():
    # applications/FiveHouse.cpp:68 records 2,134,981,304 houses for the
    # reference's wiki-vote fixture; 2^20 split estimators put it inside
    # their 99.9 % interval
    n, s, d = H.golden_graph("wiki-vote")
    g = O.OGraph(n, s, d, seed=1)
    N = 1 << 20
    r = O.run(g, O.make_plan(O.PLAN_HOUSE_SPLIT, order=64), 5, 0, N, threads=8)[0]
    mean = O.mean(r)
    hw = 3.29 * math.sqrt(max(r.sumsq / N - mean * mean, 0.0) / N)
    assert r.successes > 1000
    assert abs(mean - 2134981304) <= hw, (mean, hw)
    assert hw < 0.05 * 2134981304


def test_reference_fourclique_shape_alone_is_biased():
    # Shape A alone is what applications/FourClique.cpp captures: averaged over
    # all 720 rank orders of K4 it finds 3/5 of the clique (FourClique.cpp:35 TODO).
    from itertools import permutations
    n, s, d = H.complete_graph(4)
    # enumerate rank orders by building the graph with each input order and a
    # fixed shuffle seed, and count orders whose 3 lowest ranks form a spanning
    # tree whose 2nd edge touches the 1st
    good, tot = 0, 0
    edges = list(zip(s.tolist(), d.tolist()))
    for perm in permutations(range(6)):
        ranked = [edges[i] for i in perm]
        e1, e2, e3 = ranked[0], ranked[1], ranked[2]
        adj12 = len(set(e1) & set(e2)) == 1
        span = len(set(e1) | set(e2) | set(e3)) == 4
        good += adj12 and span
        tot += 1
    assert good * 5 == tot * 3


# ---- SamplerGenerator pattern machinery (tests/TestPattern2.cpp, TestPattern.cpp) ----
def test_pattern_unique_relabelings():
    assert O.pattern_unique(*O.FOUR_CHAIN) == (12, 24)
    assert O.pattern_unique(*O.K4) == (1, 24)
    assert O.pattern_unique(*O.FIVE_STAR) == (5, 120)
    assert O.pattern_unique(*O.FIVE_HOUSE) == (60, 120)   # n!/|Aut| = 120/2


def test_pattern_edge_order_classes():
    assert O.pattern_edge_order_classes(*O.FIVE_CHAIN) == (12, 24)
    k4 = [(i, j) for i in range(4) for j in range(i)]
    assert O.pattern_edge_order_classes(4, k4) == (30, 720)   # 720 / |Aut(K4)| = 720/24


# ---- exact counters ---------------------------------------------------------------
def test_snap_wiki_vote_triangles():
    n, s, d = H.golden_graph("snap-wiki-Vote")
    g = O.OGraph(n, s, d)
    assert g.triangles() == 608389            # SNAP's published count for wiki-Vote


@pytest.mark.parametrize("gi", range(3))
def test_exact_counters_agree_with_brute_force(gi):
    n, s, d = [H.random_graph(12, 0.5, 10), H.random_graph(14, 0.35, 11), H.complete_graph(7)][gi]
    g = O.OGraph(n, s, d)
    assert g.triangles() == g.pattern_brute(*O.TRIANGLE)
    assert g.k4() == g.pattern_brute(*O.K4)
    assert g.paths3() == g.paths_brute(4)
    assert g.wedges() == g.paths_brute(3)
    if gi == 2:   # K7 closed forms
        assert g.triangles() == math.comb(7, 3) and g.k4() == math.comb(7, 4)
        assert g.house_brute() == math.comb(7, 5) * 60   # every 5 vertices hold 5!/2 houses


@pytest.mark.parametrize("gi", range(6))
def test_c4_and_five_chain_counts_agree_with_brute_force(gi):
    """4-cycles by pair counting equal the brute-force count of the 4-cycle
    pattern, and the closed form for 5-vertex chains equals the enumeration of
    simple 4-edge paths (FiveChain.cpp's count), on graphs with triangles,
    hubs and parallel edges."""
    n, s, d = [H.random_graph(12, 0.5, 10), H.random_graph(14, 0.35, 11), H.complete_graph(7),
               H.random_graph(30, 0.2, 12), H.multigraph(20, 0.4, 13), H.hub_graph(40, 3, 14)][gi]
    g = O.OGraph(n, s, d)
    assert g.c4() == g.pattern_brute(*O.C4)
    assert g.paths4() == g.paths_brute(5)
    if gi == 2:   # K7: 3 C(7, 4) 4-cycles, 7!/2!/2 chains of 5 vertices
        assert g.c4() == 3 * math.comb(7, 4) and g.paths4() == math.perm(7, 5) // 2


def _complete_bipartite(a, b):
    u, v = np.meshgrid(np.arange(a), np.arange(a, a + b), indexing="ij")
    return a + b, u.ravel().astype(np.uint32), v.ravel().astype(np.uint32)


@pytest.mark.parametrize("a,b", [(2, 2), (2, 9), (3, 5), (6, 7), (1, 12)])
def test_c4_complete_bipartite_closed_form_and_bound(a, b):
    """K_{a,b} holds C(a, 2) C(b, 2) 4-cycles (two vertices on each side), and
    every simple graph with m edges at most C(m, 2): each 4-cycle holds two
    pairs of disjoint edges and each such pair closes at most two 4-cycles --
    the bound that keeps the device's 4-cycle count below 2^61 for m < 2^31."""
    n, s, d = _complete_bipartite(a, b)
    g = O.OGraph(n, s, d)
    assert g.c4() == math.comb(a, 2) * math.comb(b, 2)
    assert g.c4() == g.pattern_brute(*O.C4)
    m = a * b
    assert g.c4() <= math.comb(m, 2)


@pytest.mark.parametrize("gi", range(4))
def test_c4_bound_on_random_graphs(gi):
    n, s, d = [H.random_graph(16, 0.9, 20), H.random_graph(40, 0.3, 21), H.complete_graph(9),
               H.power_law_graph(200, 900, 2.1, 22)][gi]
    g = O.OGraph(n, s, d)
    assert 0 < g.c4() <= math.comb(len(s), 2)


def test_five_chain_closed_form_misses_a_dropped_term():
    """Each term of the closed form matters: on G(30, 0.2) leaving out the
    4-cycle term, or the triangle terms, changes the count."""
    n, s, d = H.random_graph(30, 0.2, 12)
    g = O.OGraph(n, s, d)
    assert g.c4() > 0 and g.triangles() > 0
    assert g.paths4() + 4 * g.c4() != g.paths_brute(5)


NAIVE = {"TriangleCounting": lambda g: g.triangles(), "FourClique": lambda g: g.k4(),
         "ThreeChain": lambda g: g.wedges(), "FourChain": lambda g: g.paths3(),
         "FiveChain": lambda g: g.paths_brute(5), "FiveStar": lambda g: g.pattern_brute(*O.FIVE_STAR),
         "FiveHouse": lambda g: g.house_brute()}


@pytest.mark.parametrize("prog", sorted(NAIVE))
def test_reference_naive_counters(prog, tmp_path):
    exe = os.path.join(ROOT, "oracle", "_ref", prog)
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref not built (needs the reference sources: make -C oracle ref)")
    n, s, d = H.golden_graph("simple-30")
    p = tmp_path / "g.edgelist"
    with open(p, "w") as f:
        f.write(f"{n} {len(s)}\n")
        for a, b in zip(s, d):
            f.write(f"{a} {b}\n")
    out = subprocess.run([exe, str(p)], capture_output=True, timeout=120).stdout.decode(errors="replace")
    got = int([ln for ln in out.replace("\r", "\n").split("\n") if "count:" in ln][-1].split(":")[1])
    assert got == NAIVE[prog](O.OGraph(n, s, d))


# ---- regression ELP ------------------------------------------------------------------
def _ld(arr):
    import ctypes as C
    return (C.c_longdouble * len(arr))(*arr)


def test_regression_known_answer():
    # tests/TestRegression.cpp: k^2 printed as 609.8565
    x = [v / 100 for v in (0.09, 0.26, 0.15, 0.42, 0.57, 0.10, 2.62)]
    y = [v / 100 for v in (1, 2, 4, 8, 16, 32, 64)]
    k = O.lib().og_regression_k(_ld(x), _ld(y), 7)
    assert f"{k * k:.4f}" == "609.8565"


def test_elp_single_point_closed_form():
    import ctypes as C
    F, e0 = 1000.0, 1100.0
    a, b, dmax_s, m_s, dmax, m, eps, conf = 1.0, 1.0, 50, 10000, 400, 100000, 0.05, 0.95
    err = abs(e0 - F) / F * 5
    ema = 100 * 0.96 + err * 0.04
    bound = lambda e, d, est, D, M: (1 / e ** 2) * (D ** b) * (M ** a) / est * math.log(2 / d)
    X = math.sqrt(bound(ema, 1 - conf, F, dmax_s, m_s) * ema ** 2 / 16)
    K = ema / X
    want = bound(eps, 1 - conf, F / 0.1 ** (a + b), dmax, m) * K * K
    Kout = C.c_longdouble()
    got = O.lib().og_elp_num_estimators(_ld([e0]), 1, F, a, b, dmax_s, m_s, dmax, m, eps, conf, C.byref(Kout))
    assert got == pytest.approx(want, rel=1e-12)
    assert float(Kout.value) == pytest.approx(K, rel=1e-12)


def test_star_and_house_closed_forms():
    """K_n: 5-stars n*C(n-1,4); houses n!/(n-5)!/2 (two automorphisms)."""
    from math import comb, perm
    for k in (5, 6, 8):
        n, s, d = H.complete_graph(k)
        og = O.OGraph(n, s, d, seed=1)
        assert og.stars(4) == k * comb(k - 1, 4)
        assert og.house_brute() == perm(k, 5) // 2


def test_mt19937_engine_matches_the_standard(tmp_path):
    """The cost-faithful baseline's MT19937: the C++ standard's check value
    ([rand.predef]: the 10000th output of a default-seeded mt19937 is
    4123659995) and std::mt19937's first outputs for seed 1 (g++)."""
    assert O.lib().og_mt19937_nth(5489, 10000) == 4123659995
    src = tmp_path / "mt.cpp"
    src.write_text("#include <random>\n#include <cstdio>\nint main(){std::mt19937 g(1u);"
                   "for(int i=0;i<5;++i)std::printf(\"%u\\n\",(unsigned)g());}\n")
    exe = tmp_path / "mt"
    subprocess.check_call(["g++", "-O1", "-o", str(exe), str(src)])
    want = [int(x) for x in subprocess.check_output([str(exe)]).split()]
    assert [O.lib().og_mt19937_nth(1, k) for k in range(1, 6)] == want


@pytest.mark.parametrize("gi", range(3))
def test_line_model_basic_costs(gi):
    """The CPU line model of the device kernels: a 3-chain estimator on a graph
    without parallel edges reads only its first edge's record (one line); a
    triangle estimator reads at most record + entry + the buckets of one
    closing check; parallel edges add the 3-chain's entry read."""
    n, s, d = [H.golden_graph("wiki-vote"), H.power_law_graph(2000, 12000, 2.2, 3), H.multigraph(24, 0.4, 5)][gi]
    g = O.OGraph(n, s, d, seed=4)
    k = 1 << 14
    chain = O.device_lines(g, O.PLAN_THREE_CHAIN, 1, 0, k)
    tri = O.device_lines(g, O.PLAN_TRIANGLE, 1, 0, k)
    motif = O.device_lines(g, O.PLAN_THREE_MOTIF, 1, 0, k)
    if gi < 2:
        assert chain == k
    else:
        assert k < chain <= 2 * k
    assert tri == motif          # the fused plan reads exactly the triangle plan's lines
    assert k < tri < (4 if gi < 2 else 8) * k   # the multigraph's copies overflow buckets: longer walks
    with pytest.raises(ValueError):
        O.device_lines(g, O.PLAN_FOUR_CHAIN, 1, 0, 16)


def test_house_closed_form_equals_brute_force():
    # houses = sum_e t_e c4_e - 2 sum_e t_e (t_e - 1): equal to the brute force
    # over pattern maps on random graphs, multigraphs (their simple graphs) and
    # K_n, and to FiveHouse.cpp:68's count for the reference's wiki-vote fixture
    for k in (5, 6, 8):
        n, s, d = H.complete_graph(k)
        assert O.OGraph(n, s, d).house() == math.comb(k, 5) * 60
    graphs = [H.random_graph(20 + 5 * i, 0.3, i) for i in range(6)]
    graphs += [H.multigraph(25, 0.3, 7), H.power_law_graph(60, 300, 2.1, 3), H.random_graph(40, 0.8, 9)]
    for n, s, d in graphs:
        g = O.OGraph(n, s, d)
        assert g.house() == g.house_brute()
    n, s, d = H.golden_graph("wiki-vote")
    assert O.OGraph(n, s, d).house() == 2134981304
