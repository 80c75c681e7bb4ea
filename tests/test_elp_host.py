"""Host logic of the learned error-latency profile: features, CLT sizing, the
judge (sampling/IEStop.cpp:47-56) and the top-up loop, driven by a synthetic
instance whose estimator values are a fixed function of the estimator id (no
GPU)."""
import math

import numpy as np
import pytest

import zgraph as Z
from zgraph import elp as L


class _Sums:
    def __init__(self, vals):
        self.log2_den = 0
        self._tot = int(vals.sum())
        self.sumsq = float((vals.astype(np.float64) ** 2).sum())
        self.successes = int((vals != 0).sum())
        self.count = len(vals)
        self.mean = self._tot / max(len(vals), 1)

    def total(self):
        return self._tot


class _Plan:
    def __init__(self, kind):
        self.kind = kind


class _Graph:
    _elp_features = (1e4, 1e5, 3e6, 5e8, 300.0)


class FakeInstance:
    """Estimator id i has value v(i) in {0, 10, 1000}: mean 35.5."""
    channel = 0

    def __init__(self, kind=Z.PLAN_TRIANGLE):
        self.plan = _Plan(kind)
        self.graph = _Graph()
        self.seed = 1
        self.calls = []

    @staticmethod
    def values(lo, hi):
        i = np.arange(lo, hi, dtype=np.uint64)
        h = (i * np.uint64(0x9E3779B97F4A7C15)) >> np.uint64(40)
        u = (h % np.uint64(1000)).astype(np.int64)
        return np.where(u < 25, 1000, np.where(u < 550, 10, 0)).astype(np.int64)

    def calculate(self, n, k=None, est=None):
        done = est.get_num_finished_estimators()
        if n > done:
            s = _Sums(self.values(done, n))
            est.absorb(s)
            est.update(n - done, s.mean)
            self.calls.append((done, n))
        return est.get_currnet_estimation()


class FixedProfile:
    def __init__(self, n):
        self.n = n

    def get_num_estimators(self, instance, k, eps, conf):
        return self.n


class FakeModel:
    def __init__(self, r, val_rmse=(0.0, 0.0, 0.0, 0.0)):
        self.r = r
        self.val_rmse = list(val_rmse)

    def predict(self, X):
        return np.full(np.atleast_2d(X).shape[0], self.r, np.float32)


def test_feature_vector_layout():
    x = L.feature_vector((1e4, 1e5, 3e6, 5e8, 300.0), Z.PLAN_K4_CLOSE, 65536, 12.0, 0.25)
    assert x.shape == (L.IN_DIM,)
    assert x[0] == pytest.approx(4) and x[1] == pytest.approx(5)
    assert x[2] == pytest.approx(math.log10(20))          # mean degree 2m/n
    assert x[3] == pytest.approx(math.log10(3e6 / 2e5))
    assert x[4] == pytest.approx(math.log10(5e8 / 3e6))
    assert x[5] == pytest.approx(math.log10(300))
    assert list(x[6:10]) == [0, 1, 0, 0]                   # 4-clique slot
    assert x[10] == pytest.approx(math.log10(12.0))
    assert x[11] == pytest.approx(math.log10(0.25))
    assert x[12] == pytest.approx(math.log10(65536))
    assert x[13] == pytest.approx(math.log10(144 * 0.25))
    assert x[15] == 0


def test_pilot_stats_closed_form():
    v = np.array([0, 0, 4, 8], np.int64)
    mean, cv, succ = L.pilot_stats(_Sums(v))
    assert mean == 3 and succ == 0.5
    assert cv == pytest.approx(math.sqrt(80 / 4 - 9) / 3)


def test_judge_matches_iestop_rule():
    est = Z.Estimation()
    inst = FakeInstance()
    inst.calculate(1 << 14, None, est)
    v = inst.values(0, 1 << 14).astype(np.float64)
    sd = v.std(ddof=1)
    rel = 1.959963984540054 * sd / (math.sqrt(len(v)) * v.mean())
    assert est.half_width(0.95) / est.exact_mean() == pytest.approx(rel, rel=1e-9)
    assert L.judge(est, rel * 1.001, 0.95)
    assert not L.judge(est, rel * 0.999, 0.95)
    small = Z.Estimation()
    FakeInstance().calculate(100, None, small)
    assert not L.judge(small, 10.0, 0.95)                  # below the minimum sample


def test_topup_reaches_the_judged_error_and_equals_one_run():
    inst = FakeInstance()
    rec = L.estimate_to_error(inst, 0.05, 0.95, FixedProfile(2000))
    assert rec["accepted"] and rec["topup_rounds"] >= 1
    assert rec["rel_half_width"] <= 0.05
    n = rec["estimators"]
    assert rec["estimate"] == pytest.approx(inst.values(0, n).mean(), rel=1e-12)
    # the runs tile the id range [0, n) without gaps or overlaps
    assert inst.calls[0][0] == 0 and inst.calls[-1][1] == n
    assert all(a[1] == b[0] for a, b in zip(inst.calls, inst.calls[1:]))


def test_unjudged_run_uses_the_sized_count():
    inst = FakeInstance()
    rec = L.estimate_to_error(inst, 0.05, 0.95, FixedProfile(3000), judged=False)
    assert rec["estimators"] == 3000 and rec["topup_rounds"] == 0


@pytest.mark.parametrize("aim", [0.0, 1.0, 2.0])
def test_learned_sizing_is_the_clt_formula(aim):
    """N = (z * CV_pilot * 10^(r - aim * s) / eps)^2 with r the network's
    correction and s its validation RMSE for the query's pattern slot."""
    prof = L.LearnedErrorProfile(weights=None, pilot=1 << 14, fallback=FixedProfile(777), aim_sigmas=aim)
    prof.model = FakeModel(0.1, val_rmse=(0.03, 0.05, 0.0, 0.02))
    inst = FakeInstance()
    est = Z.Estimation()
    n, how = prof._size(inst, 0.01, 0.95, est)
    assert how == "learned"
    v = inst.values(0, 1 << 14).astype(np.float64)
    cv = math.sqrt(max((v ** 2).mean() - v.mean() ** 2, 0)) / v.mean()
    s = (0.03, 0.05, 0.0, 0.02)[L.PATTERN_SLOTS[inst.plan.kind]]
    expect = math.ceil((1.959963984540054 * cv * 10 ** (0.1 - aim * s) / 0.01) ** 2)
    assert n == expect
    assert est.get_num_finished_estimators() == 1 << 14    # the pilot is part of the query


def test_fallback_for_untrained_plans_and_missing_weights():
    prof = L.LearnedErrorProfile(weights=None, fallback=FixedProfile(777))
    assert prof.get_num_estimators(FakeInstance(), 3, 0.01, 0.95) == 777   # no weights
    # a plan the network has no slot for: sized from its pilot's CV (r = 0), never the network
    prof.model = FakeModel(0.7)
    est = Z.Estimation()
    n, how = prof._size(FakeInstance(Z.PLAN_PATTERN), 0.01, 0.95, est)
    v = FakeInstance.values(0, 1 << 16).astype(np.float64)
    cv = math.sqrt(max((v ** 2).mean() - v.mean() ** 2, 0)) / v.mean()
    assert how == "pilot" and n == math.ceil((1.959963984540054 * cv / 0.01) ** 2)
    rec = L.estimate_to_error(FakeInstance(Z.PLAN_K4_SPLIT), 0.05, 0.95, prof)
    assert rec["sizing"] == "pilot" and rec["accepted"]


def test_split_model_fit_recovers_the_model_and_minimises_time_to_error():
    # pilots drawn from known Vb, Vw, c0, c1: the fit returns them, and the
    # chosen completions minimise Var(J) t(J) over every power of two in
    # [1, 4096] (brute force); edge cases clamp to 1 and 4096
    rng = np.random.default_rng(5)
    for _ in range(200):
        vb, vw = 10 ** rng.uniform(-2, 4), 10 ** rng.uniform(-1, 7)
        c0, c1 = 10 ** rng.uniform(-8, -5), 10 ** rng.uniform(-11, -7)
        stats = {j: {"var": vb + vw / j, "ms_per_id": c0 + c1 * j} for j in (64, 1024)}
        r = L.fit_split_model(stats, 64, 1024)
        assert r["v_between"] == pytest.approx(vb, rel=1e-6, abs=1e-9 * vw)
        assert r["v_within"] == pytest.approx(vw, rel=1e-9)
        assert r["c1_ms_per_id_per_completion"] == pytest.approx(c1, rel=1e-6)
        cost = {1 << k: (vb + vw / (1 << k)) * (c0 + c1 * (1 << k)) for k in range(13)}
        best = min(cost, key=cost.get)
        assert cost[r["completions"]] <= cost[best] * (1 + 1e-9)
    flat = {64: {"var": 5.0, "ms_per_id": 1e-6}, 1024: {"var": 5.0, "ms_per_id": 2e-6}}
    assert L.fit_split_model(flat, 64, 1024)["completions"] == 1          # completions add no information
    pure = {64: {"var": 1.0 / 64, "ms_per_id": 1e-6}, 1024: {"var": 1.0 / 1024, "ms_per_id": 2e-6}}
    assert L.fit_split_model(pure, 64, 1024)["completions"] == 4096       # all variance in the completions
