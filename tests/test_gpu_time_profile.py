"""TimeProfile (src/ErrorLatencyProfile.cpp:24-47) on the device, and the
budget-less at-least-one predicate triangle application that runs two of them
(applications/AtLeastOnePredicateWeightedTriangle.cpp:135-152), on the
reference's weighted wiki-vote (u32 weights in [0, 79), predicate >= 50).

* The batches are consecutive estimator-id ranges from 0 whose sizes double
  from the base while `used + 2 * last <= limit` holds (checked against the
  batch times the profile recorded), then at most one batch sized to the time
  left.
* The returned estimate is the running mean over those batches and equals one
  zg_run over the union of their ids: the exact sums word for word, the
  estimate to double rounding.
* at-least-one = all - none, each term over the ids its own profile ran."""
import os
import re
import subprocess

import numpy as np
import pytest

import zgraph as Z
from helpers import golden_graph, golden_weights

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "sampling-graph-mining_amd", "host", "_bin")


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def weighted():
    n, s, d = golden_graph("wiki-vote")
    w = golden_weights("wiki-vote").astype(np.uint32)
    g = Z.Graph.from_edges(s, d, n, weights=w.astype(np.float32), seed=1)
    return n, s, d, w, g


def _check_schedule(tp, limit, base):
    b = tp.batches
    assert b and b[0][0] == 0 and b[0][1] == base
    for (f0, c0, _), (f1, _, _) in zip(b, b[1:]):
        assert f1 == f0 + c0                       # consecutive id ranges
    doubling = b[:-1] if tp.remainder else b
    used = 0
    for i, (_, c, ms) in enumerate(doubling):
        assert c == base << i                      # sizes double from the base
        used += ms
        cur = ms
        if i + 1 < len(doubling):
            assert used + 2 * cur <= limit         # the loop went on
    assert used + 2 * cur > limit                  # ... and stopped where the rule says
    if tp.remainder:
        nxt = base << len(doubling)
        assert b[-1][1] == nxt * (limit - used) // (cur * 2)
    return sum(c for _, c, _ in b)


@pytest.mark.parametrize("limit", [40, 150])
def test_time_profile_batches_and_union(weighted, limit):
    _, _, _, _, g = weighted
    inst = Z.TriangleCoutingInstance(g, seed=7)
    tp = Z.TimeProfile(basic_estimators=1 << 16)
    est = tp.run(inst, 3, limit)
    total = _check_schedule(tp, limit, 1 << 16)
    say(f"[tp] limit {limit} ms: {len(tp.batches)} batches, {total} estimators, estimate {est:.6g}")
    ref = Z.run(g, inst.plan, 7, 0, total)[0]
    assert tp.estimation.exact_sum == ref.total()
    assert tp.estimation.get_num_finished_estimators() == total
    assert est == pytest.approx(ref.mean, rel=1e-12)


def test_time_profile_at_least_one(weighted):
    n, s, d, w, g = weighted
    all_i = Z.TriangleCoutingInstance(g, seed=2)
    none_i = Z.AllPredicateZGraphInstance(g, 50.0, keep_ge=False, seed=2, filter_seed=1)
    tp = Z.TimeProfile()
    a = tp.run(all_i, 3, 60)
    na = sum(c for _, c, _ in tp.batches)
    b = tp.run(none_i, 3, 60)
    nb = sum(c for _, c, _ in tp.batches)
    ra = Z.run(g, all_i.plan, 2, 0, na)[0]
    rb = Z.run(none_i.graph, none_i.plan, 2, 0, nb)[0]
    say(f"[tp] at-least-one: all over {na} ids, none over {nb} ids: {a - b:.6g}")
    assert a - b == pytest.approx(ra.mean - rb.mean, rel=1e-9)


def _report(out):
    rows = {}
    for line in out.splitlines():
        m = re.match(r"^(\S+) (\S+) (\d+)$", line.strip())
        if m:
            rows[m.group(1)] = (float(m.group(2)), int(m.group(3)))
    return rows


def test_at_least_one_app_without_budget(tmp_path, weighted):
    """The C++ main with no budget: two 500 ms TimeProfiles, as the reference
    runs it; each term equals a zg_run over the ids its profile reports."""
    n, s, d, w, g = weighted
    prefix = str(tmp_path / "wvw")
    with open(prefix + ".csv", "w") as f:
        f.write("num_vertices,num_edges\n%d,%d\n" % (n, len(s)))
    Z.save_biedgelist(prefix + ".biedgelist", s, d, w)
    app = os.path.join(BIN, "AtLeastOnePredicateWeightedTriangle")
    if not os.path.exists(app):
        pytest.fail(f"{app} is not built")
    say("[tp] AtLeastOnePredicateWeightedTriangle without a budget (2 x 500 ms)")
    r = subprocess.run([app, prefix], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = _report(r.stdout)
    est_a, na = rows["time-profile-all"]
    est_b, nb = rows["time-profile-none"]
    cnt = rows["at-least-one-predicate-triangle"][0]
    assert na >= 1 << 16 and nb >= 1 << 16
    ra = Z.run(g, Z.Plan(Z.PLAN_TRIANGLE), 1, 0, na)[0]          # the app's default estimator seed is 1
    gn = g.filter(50.0, False, 1)
    rb = Z.run(gn, Z.Plan(Z.PLAN_TRIANGLE), 1, 0, nb)[0]
    assert est_a == pytest.approx(ra.mean, rel=1e-12)
    assert est_b == pytest.approx(rb.mean, rel=1e-12)
    assert cnt == pytest.approx(est_a - est_b, rel=1e-12)
    say(f"[tp] app: all {est_a:.6g} over {na} ids, none {est_b:.6g} over {nb} ids, at-least-one {cnt:.6g}")
