"""Batched queries (zg_batch_*): one replayable HIP graph around one
persistent work-queue launch over every query's chunks must give every query
exactly the sums of its own zg_run, for any launch shape, across replays and
seed updates."""
import numpy as np
import pytest

import zgraph as Z
import helpers as H

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def graphs():
    n, s, d = H.golden_graph("wiki-vote")
    g1 = Z.Graph.from_edges(s, d, n, seed=2)
    e = Z.chung_lu(20000, 120000, 2.2, 800.0, 4)
    g2 = Z.Graph.from_device(e, seed=4)
    e.free()
    e = Z.erdos_renyi(300, 9000, 6)
    g3 = Z.Graph.from_device(e, seed=6)
    e.free()
    return [g1, g2, g3]


@pytest.fixture(scope="module")
def plans():
    return {"tri": Z.Plan(Z.PLAN_TRIANGLE), "k4": Z.Plan(Z.PLAN_K4_CLOSE), "c4": Z.Plan(Z.PLAN_FOUR_CHAIN),
            "motif": Z.Plan(Z.PLAN_THREE_MOTIF), "chain5": Z.Plan(Z.PLAN_CHAIN, order=5),
            "star": Z.Plan(Z.PLAN_PATTERN, pattern=Z.FIVE_STAR_PATTERN),
            "house": Z.Plan(Z.PLAN_PATTERN, pattern=Z.FIVE_HOUSE_PATTERN),
            "w3": Z.Plan(Z.PLAN_THREE_CHAIN), "tk4": Z.Plan(Z.PLAN_TRIANGLE_K4),
            "k4s": Z.Plan(Z.PLAN_K4_SPLIT, completions=128),
            "tk4s": Z.Plan(Z.PLAN_TRIANGLE_K4_SPLIT, completions=32),
            "hs": Z.Plan(Z.PLAN_HOUSE_SPLIT, completions=64)}


def _queries(graphs, plans):
    rng = np.random.default_rng(5)
    qs = []
    names = list(plans)
    for i in range(40):
        g = graphs[i % 3]
        p = plans[names[i % len(names)]]
        count = int(rng.choice([0, 1, 255, 16384, 16385, 100000, 1 << 18]))
        qs.append((g, p, int(rng.integers(1, 1 << 40)), int(rng.integers(0, 1 << 34)), count))
    return qs


def _same(a, b, sumsq_rel=1e-9):
    assert a.total() == b.total()
    assert a.log2_den == b.log2_den
    assert a.successes == b.successes
    assert a.count == b.count
    assert a.sumsq == pytest.approx(b.sumsq, rel=sumsq_rel, abs=1e-300)


def test_batch_equals_single_runs(graphs, plans):
    qs = _queries(graphs, plans)
    say(f"[batch] {len(qs)} queries, default config")
    b = Z.Batch(qs)
    res = b.run()
    for (g, p, seed, first, count), r in zip(qs, res):
        ref = Z.run(g, p, seed, first, count)
        assert len(r) == len(ref)
        for x, y in zip(r, ref):
            _same(x, y)
    # replays with eager work in between give the same records every time
    for rep in range(3):
        Z.run(graphs[0], plans["tri"], 9, 0, 1 << 16)
        again = b.run()
        for r, a in zip(res, again):
            for x, y in zip(r, a):
                _same(x, y, sumsq_rel=0)


@pytest.mark.parametrize("cfg", [dict(block_size=64, blocks_per_cu=2, chunk_estimators=256),
                                 dict(block_size=128, blocks_per_cu=16, chunk_estimators=65536),
                                 dict(block_size=256, blocks_per_cu=1, chunk_estimators=4096),
                                 dict(block_size=192, blocks_per_cu=3, chunk_estimators=64)])
def test_batch_launch_shapes_agree(graphs, plans, cfg):
    qs = _queries(graphs, plans)[:16]
    say(f"[batch] config {cfg}")
    res = Z.Batch(qs, **cfg).run()
    for (g, p, seed, first, count), r in zip(qs, res):
        ref = Z.run(g, p, seed, first, count)
        for x, y in zip(r, ref):
            _same(x, y)


def test_batch_set_ids_and_autotune(graphs, plans):
    qs = _queries(graphs, plans)[:12]
    b = Z.Batch(qs)
    seeds = [q[2] + 1 for q in qs]
    firsts = [q[3] + 7 for q in qs]
    b.set_ids(seeds, firsts)
    res = b.run()
    for (g, p, _, _, count), s, f, r in zip(qs, seeds, firsts, res):
        ref = Z.run(g, p, s, f, count)
        for x, y in zip(r, ref):
            _same(x, y)
    cfg = b.autotune(trials=1)
    say(f"[batch] autotuned {cfg}")
    assert cfg["block_size"] in (64, 128, 256) and cfg["blocks_per_cu"] in (2, 4, 8, 16)
    assert cfg["chunk_estimators"] in (256, 1024, 4096, 16384)
    assert b.config() == {k: cfg[k] for k in ("block_size", "blocks_per_cu", "chunk_estimators")}
    res2 = b.run()
    for r, a in zip(res, res2):
        for x, y in zip(r, a):
            _same(x, y)


def test_batch_rejects_bad_queries(graphs, plans):
    with pytest.raises(Z.ZGraphError):
        Z.Batch([(graphs[0], plans["tri"], 1, 0, (1 << 31) + 1)])
    with pytest.raises(Z.ZGraphError):
        Z.Batch([(graphs[0], plans["tri"], 1, 0, 10)], block_size=96)
    with pytest.raises(Z.ZGraphError):   # estimators per claim: a multiple of 64
        Z.Batch([(graphs[0], plans["tri"], 1, 0, 10)], chunk_estimators=1000)
    b = Z.Batch([(graphs[0], plans["tri"], 1, 0, 10), (graphs[1], plans["k4"], 2, 0, 10)])
    with pytest.raises(ValueError):
        b.set_ids(seeds=[1])
    with pytest.raises(ValueError):
        b.set_ids(first_ids=5)
