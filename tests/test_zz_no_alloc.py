"""zg_run_noalloc: the run entry point that allocates nothing (block partials
and the result record in the library's fixed device scratch) must return
exactly zg_run's sums, on consecutive calls whose sizes shrink (the scratch
keeps the previous call's records), and must reject launches whose grid
needs more partial records than the scratch holds."""
import ctypes as C

import pytest

import zgraph as Z
from zgraph import _capi
import helpers as H

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _noalloc(graph, plan, seed, first, count, block_size=0, blocks_per_cu=0, order=0):
    sums = (_capi.Sums * plan.channels)()
    launch = _capi.Launch(block_size, blocks_per_cu, order)
    rc = _capi.lib().zg_run_noalloc(graph.h, plan.h, seed, first, count, C.byref(launch), sums, None)
    return rc, [sums[i] for i in range(plan.channels)]


def _same(a, b):
    assert a.total() == b.total()
    assert a.log2_den == b.log2_den
    assert a.successes == b.successes
    assert a.count == b.count
    assert a.status == b.status == 0
    assert a.sumsq == b.sumsq   # same launch shape: same reduction order, same bits


@pytest.fixture(scope="module")
def graphs():
    e = Z.chung_lu(200000, 2000000, 2.1, 3000.0, 7)
    big = Z.Graph.from_device(e, seed=7)
    e.free()
    n, s, d = H.golden_graph("wiki-vote")
    small = Z.Graph.from_edges(s, d, n, seed=3)
    return big, small


@pytest.mark.parametrize("plan_args", [dict(kind=Z.PLAN_TRIANGLE), dict(kind=Z.PLAN_TRIANGLE_K4),
                                       dict(kind=Z.PLAN_PATTERN, pattern=Z.FIVE_STAR_PATTERN),
                                       dict(kind=Z.PLAN_THREE_MOTIF),
                                       dict(kind=Z.PLAN_TRIANGLE_K4_SPLIT, completions=64),
                                       dict(kind=Z.PLAN_HOUSE_SPLIT, completions=16)])
def test_noalloc_equals_run_on_shrinking_calls(graphs, plan_args):
    big, small = graphs
    plan = Z.Plan(**plan_args)
    calls = [(big, 11, 5, 1 << 22, 0, 0), (small, 12, 1 << 40, 1 << 16, 0, 0),
             (small, 13, 9, 3000, 128, 2), (big, 14, 0, 1, 64, 1)]
    for g, seed, first, count, bs, bpc in calls:
        say(f"[noalloc] plan {plan_args['kind']} m={g.num_local_edges} seed={seed} first={first} count={count} "
            f"block={bs} bpc={bpc}")
        rc, got = _noalloc(g, plan, seed, first, count, bs, bpc)
        assert rc == 0, _capi.last_error()
        ref = Z.run(g, plan, seed, first, count, block_size=bs, blocks_per_cu=bpc)
        for x, y in zip(got, ref):
            _same(x, y)


def test_noalloc_rejects_oversized_grids_and_ordering(graphs):
    big, _ = graphs
    plan = Z.Plan(Z.PLAN_TRIANGLE_K4)
    cus = Z.device_info()["num_cus"]
    bpc = 16384 // (2 * cus) + 1   # one workgroup per CU too many for two channels
    say(f"[noalloc] rejection: {cus} CUs x {bpc} workgroups x 2 channels")
    rc, _ = _noalloc(big, plan, 1, 0, 1 << 24, 256, bpc)
    assert rc == _capi.ZG_EINVAL
    assert "scratch" in _capi.last_error()
    rc, _ = _noalloc(big, plan, 1, 0, 1 << 16, 256, 0, order=4)
    assert rc == _capi.ZG_EINVAL
    # a valid call right after the rejected ones is unaffected
    rc, got = _noalloc(big, plan, 2, 0, 1 << 18)
    assert rc == 0, _capi.last_error()
    for x, y in zip(got, Z.run(big, plan, 2, 0, 1 << 18)):
        _same(x, y)
