"""The batched-query workload at its full shape: 256 mixed-pattern queries
(triangle, 4-clique, 5-star, 5-house, 4-chain) over 64 seeded Chung-Lu /
R-MAT graphs of 1e4..1e7 edges with budgets 2^16..2^26, served by one
persistent work-queue launch (the workload bench.py times).  Every query's
sums equal its own zg_run exactly; on the graphs small enough for the CPU
oracle, the batch's exact sums and success counts also equal the oracle's
run of the same (graph, plan, seed, ids) -- the restatement of
ZGraphInstance::calculate (src/ZGraphInstance.cpp:256-297) over the
reference's estimate() bodies."""
import os
import sys

import pytest

import zgraph as Z
import pyoracle as O
import helpers as H

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
import bench  # noqa: E402  (the workload definition)


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def workload():
    graphs, plans, queries, build_s = bench.batched_workload(Z, seed=1)
    say(f"[cfg4] 64 graphs built in {build_s:.1f}s, {len(queries)} queries, "
        f"{sum(q[4] for q in queries):.3e} estimators")
    yield graphs, plans, queries
    for g in graphs:
        g.close()


def _oracle_plan(p):
    pats = {id(Z.FIVE_STAR_PATTERN): O.FIVE_STAR, id(Z.FIVE_HOUSE_PATTERN): O.FIVE_HOUSE}
    if p.kind == Z.PLAN_PATTERN:
        return O.make_plan(p.kind, pattern=pats[id(p.pattern)])
    return O.make_plan(p.kind, order=p.order)


def test_cfg4_shape(workload):
    graphs, plans, queries = workload
    assert len(graphs) == 64 and len(queries) == 256
    ms = [g.num_local_edges for g in graphs]
    assert min(ms) >= 10_000 and max(ms) <= 10_000_000
    counts = [q[4] for q in queries]
    assert min(counts) >= 1 << 16 and max(counts) <= 1 << 26
    assert {q[1].kind for q in queries} == {Z.PLAN_TRIANGLE, Z.PLAN_K4_CLOSE, Z.PLAN_PATTERN, Z.PLAN_FOUR_CHAIN}


def test_cfg4_batch_equals_single_runs_and_oracle(workload):
    graphs, plans, queries = workload
    b = Z.Batch(queries)
    res, ms = b.run(timed=True)
    say(f"[cfg4] batch replay {ms:.1f} ms, config {b.config()}")
    for i, ((g, p, seed, first, count), r) in enumerate(zip(queries, res)):
        ref = Z.run(g, p, seed, first, count)
        for x, y in zip(r, ref):
            assert x.total() == y.total(), i
            assert x.successes == y.successes and x.count == y.count == count and x.log2_den == y.log2_den, i
            assert x.sumsq == pytest.approx(y.sumsq, rel=1e-9), i
    # the oracle on the graphs and budgets it finishes in seconds
    small = [(i, q) for i, q in enumerate(queries) if q[0].num_local_edges <= 150_000 and q[4] <= 1 << 20]
    assert len(small) >= 10
    cache = {}
    for i, (g, p, seed, first, count) in small:
        if id(g) not in cache:
            a = g.export()
            cache[id(g)] = O.OGraph.from_csr(g.num_vertices, a["src"], a["dst"], a["off"], a["adj_rank"],
                                             a["adj_nbr"])
        osum = O.run(cache[id(g)], _oracle_plan(p), seed, first, count)[0]
        x = res[i][0]
        assert x.total() == osum.total(), i
        assert x.successes == osum.successes and x.log2_den == osum.K, i
    say(f"[cfg4] {len(small)} queries on {len(cache)} small graphs equal the oracle")
    # replays give the same records; a second shape (smallest claims) the same sums
    again = b.run()
    for r, a in zip(res, again):
        for x, y in zip(r, a):
            assert x.total() == y.total() and x.sumsq == y.sumsq
    b.close()
    b2 = Z.Batch(queries, block_size=128, blocks_per_cu=4, chunk_estimators=256)
    for r, a in zip(res, b2.run()):
        for x, y in zip(r, a):
            assert x.total() == y.total() and x.successes == y.successes
    b2.close()
