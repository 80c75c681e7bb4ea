"""First-edge ordered launches (zg_launch.order): the same estimators with
the same outcomes, so exact sums equal the estimator-id-order launch."""
import pytest

import zgraph as Z
import helpers as H

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _plans():
    return [Z.Plan(Z.PLAN_TRIANGLE), Z.Plan(Z.PLAN_THREE_CHAIN), Z.Plan(Z.PLAN_THREE_MOTIF),
            Z.Plan(Z.PLAN_CHAIN, order=5), Z.Plan(Z.PLAN_FOUR_CHAIN), Z.Plan(Z.PLAN_K4_CLOSE),
            Z.Plan(Z.PLAN_PATTERN, pattern=Z.FIVE_STAR_PATTERN), Z.Plan(Z.PLAN_TRIANGLE_K4)]


@pytest.mark.parametrize("order", [2, 8, 16])
def test_ordered_equals_id_order(order):
    n, s, d = H.golden_graph("wiki-vote")
    g = Z.Graph.from_edges(s, d, n, seed=3)
    for plan in _plans():
        say(f"[order] plan {plan.kind} order {order}")
        a = Z.run(g, plan, 9, 12345, 1_000_003, order=1)
        b = Z.run(g, plan, 9, 12345, 1_000_003, order=order)
        for x, y in zip(a, b):
            assert x.total() == y.total() and x.successes == y.successes and x.count == y.count
            assert x.sumsq == pytest.approx(y.sumsq, rel=1e-9)


def test_ordered_launch_on_a_large_graph():
    e = Z.chung_lu(600_000, 5_000_000, 2.2, 3000.0, 8)
    g = Z.Graph.from_device(e, seed=8)
    e.free()
    for plan in (Z.Plan(Z.PLAN_TRIANGLE), Z.Plan(Z.PLAN_K4_CLOSE), Z.Plan(Z.PLAN_FOUR_CHAIN)):
        N = (1 << 22) + 77
        (a,), ta = Z.run(g, plan, 4, 0, N, order=1, timed=True)
        (b,), tb = Z.run(g, plan, 4, 0, N, order=8, timed=True)
        say(f"[order] plan {plan.kind}: id order {ta:.3f} ms, first-edge order {tb:.3f} ms")
        assert a.total() == b.total() and a.successes == b.successes


def test_order_argument_checked():
    n, s, d = H.golden_graph("simple-30")
    g = Z.Graph.from_edges(s, d, n, seed=3)
    with pytest.raises(Z.ZGraphError):
        Z.run(g, Z.Plan(Z.PLAN_TRIANGLE), 1, 0, 10, order=17)
