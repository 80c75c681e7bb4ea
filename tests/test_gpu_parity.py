"""Device vs oracle parity through the C ABI: the device CSR equals the
oracle's CSR array for array, and every estimator's outcome (status, exact
numerator, denominator exponent, first rank, conditional calls, Philox words
used) is bit-identical to the CPU restatement on the same Philox streams."""
import sys

import numpy as np
import pytest

import helpers as H
import pyoracle as O
import zgraph as Z

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def both(n, s, d, seed, w=None):
    say(f"[graph] n={n} m={len(s)} seed={seed}")
    return Z.Graph.from_edges(s, d, n, weights=w, seed=seed), O.OGraph(n, s, d, w=w, seed=seed)


def assert_same_csr(zg, og):
    a = zg.export()
    b = og.arrays()
    assert zg.num_local_edges == og.m and zg.num_vertices == og.n
    assert zg.max_degree == og.g.max_degree
    assert (a["src"] == b["src"]).all() and (a["dst"] == b["dst"]).all()
    assert (a["off"] == b["off"]).all()
    assert (a["adj_rank"] == b["adj_rank"].astype(np.uint32)).all()
    assert (a["adj_nbr"] == b["adj_dst"]).all()
    if "w" in b:
        assert (a["w"] == b["w"]).all()


GRAPHS = {
    "simple-30": lambda: H.golden_graph("simple-30"),
    "small-10": lambda: H.golden_graph("small-10"),
    "wiki-vote": lambda: H.golden_graph("wiki-vote"),
    "K7": lambda: H.complete_graph(7),
    "gnp-60": lambda: H.random_graph(60, 0.3, 5),
    "powerlaw": lambda: H.power_law_graph(3000, 20000, 2.1, 7),
    "multi": lambda: H.multigraph(24, 0.4, 5),
    "hub": lambda: H.hub_graph(6000, 3000, 2),
    "bipartite": lambda: H.bipartite_graph(40, 60, 0.3, 3),
    "sparse-ids": lambda: H.sparse_ids_graph(1_000_000, 3000, 4),
}


@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_parallel_edge_flag_and_three_chain_reads(name):
    """The build flags a graph with parallel edges exactly when some pair
    repeats; on a graph without them the 3-chain plan reads one line per
    estimator (the sampled edge's record: the sampled list entry can only set
    the parallel-copy denominator, which is then 0), with outcomes unchanged
    (the parity tests cover both cases)."""
    n, s, d = GRAPHS[name]()
    g = Z.Graph.from_edges(s, d, n, seed=11)
    pairs = np.minimum(s, d).astype(np.uint64) << np.uint64(32) | np.maximum(s, d).astype(np.uint64)
    assert g.has_parallel_edges == (len(np.unique(pairs)) < len(pairs))
    k = 1 << 14
    reads = Z.count_reads(g, Z.Plan(Z.PLAN_THREE_CHAIN), 3, 0, k)
    if g.has_parallel_edges:
        assert reads > k
    else:
        assert reads == k


@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_read_count_equals_the_cpu_line_model(name):
    """The algorithmic bytes of the roofline come from the device's
    read-counting launch (zg_count_reads); the CPU line model (oracle/lines.c:
    the device layout's cost of each step along the oracle's draws -- records,
    entries, search probes, hash buckets walked) gives the same count for the
    triangle, 3-chain, 3-motif and 4-clique plans, estimator set by estimator
    set."""
    n, s, d = GRAPHS[name]()
    zg, og = both(n, s, d, seed=11)
    k = 1 << 15
    for kind in (Z.PLAN_TRIANGLE, Z.PLAN_THREE_CHAIN, Z.PLAN_THREE_MOTIF, Z.PLAN_K4_CLOSE):
        dev = Z.count_reads(zg, Z.Plan(kind), 9, 1 << 20, k)
        cpu = O.device_lines(og, kind, 9, 1 << 20, k)
        assert dev == cpu, (name, kind, dev, cpu)
    for J in (1, 16):
        dev = Z.count_reads(zg, Z.Plan(Z.PLAN_HOUSE_SPLIT, completions=J), 9, 1 << 20, k)
        cpu = O.device_lines_house(og, J, 9, 1 << 20, k)
        assert dev == cpu, (name, "house", J, dev, cpu)


@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_device_csr_equals_oracle(name):
    n, s, d = GRAPHS[name]()
    zg, og = both(n, s, d, seed=11)
    assert_same_csr(zg, og)


def test_empty_and_rejected_graphs():
    g = Z.Graph.from_edges(np.zeros(0, np.uint32), np.zeros(0, np.uint32), 4)
    assert g.num_local_edges == 0 and g.max_degree == 0
    sums = Z.run(g, Z.Plan(Z.PLAN_TRIANGLE), 1, 0, 1000)[0]
    assert sums.count == 1000 and sums.successes == 0 and sums.total() == 0
    with pytest.raises(Z.ZGraphError):
        Z.Graph.from_edges(np.array([0, 2], np.uint32), np.array([1, 2], np.uint32), 3)   # self-loop
    with pytest.raises(Z.ZGraphError):
        Z.Graph.from_edges(np.array([0], np.uint32), np.array([5], np.uint32), 3)         # id >= n


def test_estimator_id_ranges_past_2_64_are_rejected():
    """Ids are 64-bit: a run whose first_id + count exceeds 2^64 - 1 fails
    with an error instead of reusing ids from 0; the last range within it
    still runs."""
    n, s, d = H.golden_graph("simple-30")
    g = Z.Graph.from_edges(s, d, n, seed=1)
    plan = Z.Plan(Z.PLAN_TRIANGLE)
    top = (1 << 64) - 1 - 1000
    assert Z.run(g, plan, 1, top, 1000)[0].count == 1000
    with pytest.raises(Z.ZGraphError):
        Z.run(g, plan, 1, top, 1001)


def test_subsample_and_filter_equal_oracle():
    n, s, d = H.golden_graph("wiki-vote")
    w = H.golden_weights("wiki-vote").astype(np.float32)
    zg, og = both(n, s, d, seed=5, w=w)
    assert_same_csr(zg, og)
    assert_same_csr(zg.subsample(0.1, 9), og.subsample(0.1, 9))
    assert_same_csr(zg.filter(50.0, True, 4), og.filter(50.0, True, 4))
    assert_same_csr(zg.filter(50.0, False, 4), og.filter(50.0, False, 4))


def test_subsample_and_filter_on_parallel_edges_equal_oracle():
    n, s, d = H.multigraph(40, 0.3, 9)
    w = np.random.default_rng(9).integers(0, 79, size=len(s)).astype(np.float32)
    zg, og = both(n, s, d, seed=6, w=w)
    assert_same_csr(zg, og)
    assert_same_csr(zg.subsample(0.3, 2), og.subsample(0.3, 2))
    assert_same_csr(zg.filter(50.0, True, 3), og.filter(50.0, True, 3))
    zf, of = zg.filter(50.0, False, 3), og.filter(50.0, False, 3)
    assert_same_csr(zf, of)
    compare_outcomes(zf, of, "triangle", dict(kind=1), seed=4, first=0, count=20000)


def compare_outcomes(zg, og, name, oargs, seed, first, count):
    zplan = H.device_plan(Z, **oargs)
    oplan = H.oracle_plan(O, **oargs)
    say(f"[parity] plan={name} seed={seed} ids=[{first},{first + count})")
    zo, zs = Z.run_outcomes(zg, zplan, seed, first, count)
    oo = O.run_outcomes(og, oplan, seed, first, count)
    zo = zo.reshape(count, -1)
    oo = oo.reshape(count, -1)
    for f in ("status", "num_lo", "num_hi", "log2den", "ncond", "nwords", "e0"):
        bad = np.argwhere(zo[f] != oo[f])
        if len(bad):
            i, c = (int(x) for x in bad[0])
            say(f"[parity] first mismatch in {f} at id {first + i} channel {c}: "
                f"device {zo[i, c].tolist()} oracle {oo[i, c].tolist()}")
        assert len(bad) == 0, f"{len(bad)} estimators differ in {f}"
    osums = O.run(og, oplan, seed, first, count)
    for ch in range(zplan.channels):
        assert zs[ch].total() == osums[ch].total()
        assert zs[ch].successes == osums[ch].successes
        assert zs[ch].sumsq == pytest.approx(osums[ch].sumsq, rel=1e-9)
    return zo


@pytest.mark.parametrize("gname", ["wiki-vote", "powerlaw", "gnp-60", "K7", "simple-30", "multi", "hub", "bipartite",
                                   "sparse-ids"])
@pytest.mark.parametrize("name,oargs", [(p[0], p[1]) for p in H.PLANS])
def test_outcomes_bit_exact(gname, name, oargs):
    n, s, d = GRAPHS[gname]()
    zg, og = both(n, s, d, seed=3)
    count = 40000 if gname in ("wiki-vote", "powerlaw") else 20000
    zo = compare_outcomes(zg, og, name, oargs, seed=17, first=123, count=count)
    if gname == "K7" and name not in ("pattern_house", "pattern_k4", "pattern_5star"):
        assert (zo["status"] == 1).any()
    if gname == "bipartite" and name in ("triangle", "k4_close", "pattern_triangle", "pattern_k4"):
        assert not (zo["status"] == 1).any()   # no odd cycles


@pytest.mark.parametrize("gname", ["wiki-vote", "powerlaw", "gnp-60", "K7", "multi"])
@pytest.mark.parametrize("fused,completions", [(Z.PLAN_TRIANGLE_K4, 0), (Z.PLAN_TRIANGLE_K4_SPLIT, 64)])
def test_fused_triangle_k4_channels_equal_the_single_plans(gname, fused, completions):
    # channel 0 is the triangle plan's outcome and channel 1 the 4-clique plan's
    # (three-shape or split), id for id
    n, s, d = GRAPHS[gname]()
    zg, og = both(n, s, d, seed=3)
    count = 40000 if gname in ("wiki-vote", "powerlaw") else 20000
    say(f"[parity] fused triangle + 4-clique (plan {fused}) on {gname}, {count} ids")
    zo, zs = Z.run_outcomes(zg, Z.Plan(fused, completions=completions), 17, 123, count)
    zo = zo.reshape(count, 2)
    k4 = Z.PLAN_K4_CLOSE if fused == Z.PLAN_TRIANGLE_K4 else Z.PLAN_K4_SPLIT
    for ch, kind in ((0, Z.PLAN_TRIANGLE), (1, k4)):
        c = completions if kind == Z.PLAN_K4_SPLIT else 0
        oo = O.run_outcomes(og, H.oracle_plan(O, kind=kind, completions=c), 17, 123, count).reshape(count, -1)[:, 0]
        for f in ("status", "num_lo", "num_hi", "log2den", "e0"):
            assert (zo[:, ch][f] == oo[f]).all(), f"channel {ch} differs from the oracle in {f}"
        single = Z.run(zg, Z.Plan(kind, completions=c), 17, 123, count)[0]
        assert zs[ch].total() == single.total() and zs[ch].successes == single.successes
        assert zs[ch].sumsq == pytest.approx(single.sumsq, rel=1e-12)
    if gname == "K7":
        assert (zo[:, 1]["status"] == 1).any()


def test_er_10k_100k_triangle_one_million_bit_exact():
    # Erdos-Renyi 10,000 vertices / 100,000 edges, 10^6 estimators, Philox seed 1
    e = Z.erdos_renyi(10000, 100000, seed=1)
    s, d, _ = e.download()
    assert len(s) == 100000 and (s < d).all()
    assert len(np.unique(s.astype(np.uint64) << np.uint64(32) | d)) == 100000
    zg = Z.Graph.from_device(e, seed=1)
    og = O.OGraph(10000, s, d, seed=1)
    assert_same_csr(zg, og)
    zo = compare_outcomes(zg, og, "triangle", dict(kind=1), seed=1, first=0, count=1_000_000)
    assert (zo["status"] == 1).sum() > 0


def test_ids_not_launch_shape_define_outcomes():
    n, s, d = H.power_law_graph(3000, 20000, 2.1, 7)
    zg = Z.Graph.from_edges(s, d, n, seed=2)
    plan = Z.Plan(Z.PLAN_K4_CLOSE)
    a = Z.run(zg, plan, 5, 0, 300000, block_size=64, blocks_per_cu=1)[0]
    b = Z.run(zg, plan, 5, 0, 300000, block_size=256, blocks_per_cu=32)[0]
    c1 = Z.run(zg, plan, 5, 0, 100000)[0]
    c2 = Z.run(zg, plan, 5, 100000, 200000)[0]
    assert a.total() == b.total() == c1.total() + c2.total()
    assert a.successes == b.successes == c1.successes + c2.successes


def _random_case(k):
    """Seeded random graph k: G(n, p) with n in [5, 200] and p in [0.03, 0.9],
    sometimes with parallel edges, sometimes spread over a large id space."""
    rng = np.random.default_rng(9000 + k)
    n = int(rng.integers(5, 201))
    p = float(rng.uniform(0.03, 0.9))
    kind = k % 3
    if kind == 1:
        n, s, d = H.multigraph(n, p, 9000 + k, max_copies=int(rng.integers(2, 13)),
                               heavy_copies=int(rng.integers(9, 30)))
    else:
        n, s, d = H.random_graph(n, p, 9000 + k)
    if kind == 2 and len(s):   # spread the vertices over an id space 1000x larger
        ids = np.sort(rng.choice(n * 1000, size=n, replace=False)).astype(np.uint32)
        s, d, n = ids[s], ids[d], n * 1000
    first = int(rng.integers(0, 1 << 44))
    seed = int(rng.integers(1, 1 << 62))
    return n, s, d, seed, first


@pytest.mark.parametrize("k", range(12))
def test_random_graphs_all_plans_bit_exact(k):
    n, s, d, seed, first = _random_case(k)
    if len(s) == 0:
        pytest.skip("empty draw")
    zg, og = both(n, s, d, seed=seed % 1000)
    assert_same_csr(zg, og)
    for name, oargs in [(p[0], p[1]) for p in H.PLANS]:
        compare_outcomes(zg, og, name, oargs, seed=seed, first=first, count=4000)
