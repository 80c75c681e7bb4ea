"""The learned error-latency profile on the device: the deployed network
evaluated by the BF16-MFMA forward kernel matches the CPU reference of the
same network, and judged queries reach the requested error against exact
counts."""
import math
import os

import numpy as np
import pytest

import zgraph as Z
from zgraph import elp as L
import helpers as H
import mlp_ref

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_deployed_weights_forward_matches_reference():
    assert os.path.exists(L.DEFAULT_WEIGHTS), "trained weights are part of the package"
    m = L.LearnedModel()
    rng = np.random.default_rng(4)
    gf = (7115.0, 100762.0, 3.0e7, 3.0e10, 1065.0)
    X = np.stack([L.feature_vector(gf, k, 1 << int(rng.integers(14, 20)), float(10 ** rng.uniform(0, 3)),
                                   float(10 ** rng.uniform(-4, 0)))
                  for k in list(L.PATTERN_SLOTS) * 8])
    got = m.predict(X)
    Xn = ((X - m.mu) / m.sd).astype(np.float32)
    want = mlp_ref.forward_backward(m.params, Xn) * np.float32(m.y_scale)
    slot = np.argmax(X[:, 6:10], axis=1)
    want = np.where(np.array(m.use_model)[slot], want, 0.0)
    np.testing.assert_allclose(got, want, rtol=2e-3, atol=2e-4)


@pytest.mark.parametrize("kind,what", [(Z.PLAN_TRIANGLE, "triangle"), (Z.PLAN_FOUR_CHAIN, "path3"),
                                       (Z.PLAN_THREE_CHAIN, "wedge")])
def test_judged_queries_reach_the_error(kind, what):
    n, s, d = H.golden_graph("wiki-vote")
    g = Z.Graph.from_edges(s, d, n, seed=1)
    truth = g.exact_count(what)
    prof = L.LearnedErrorProfile()
    eps = 0.02
    errs = []
    for q in range(6):
        inst = Z.ZGraphInstance(g, Z.Plan(kind), seed=300 + q)
        r = L.estimate_to_error(inst, eps, 0.95, prof)
        assert r["accepted"] and r["sizing"] == "learned"
        assert r["rel_half_width"] <= eps
        errs.append(abs(r["estimate"] - truth) / truth)
        say(f"[elp] kind {kind} q{q}: N={r['estimators']} sized={r['sized_estimators']} "
            f"top-ups={r['topup_rounds']} err={errs[-1]:.4f}")
    # each query misses with probability about 5 %; five or more of six hit
    assert sum(e <= eps for e in errs) >= 5


def test_sizing_failure_falls_back_to_the_judge():
    # 4-cliques on a sparse graph: the pilot stays below 32 successes and the
    # regression profile is undefined; a judged query still answers
    e = Z.erdos_renyi(20000, 60000, 3)
    g = Z.Graph.from_device(e, seed=3)
    e.free()
    prof = L.LearnedErrorProfile(max_pilot=1 << 16)
    inst = Z.ZGraphInstance(g, Z.Plan(Z.PLAN_K4_CLOSE), seed=5)
    r = L.estimate_to_error(inst, 0.5, 0.95, prof, max_estimators=1 << 26, max_rounds=2)
    assert r["sizing"] in ("judge-only", "regression")
    assert r["estimators"] > 0


@pytest.mark.parametrize("inst_cls,whats", [(Z.ThreeMotifInstance, ("wedge", "triangle")),
                                            (Z.TriangleFourCliqueInstance, ("triangle", "k4")),
                                            (Z.TriangleFourCliqueSplitInstance, ("triangle", "k4"))])
def test_two_channel_query_reaches_the_error_on_both(inst_cls, whats):
    # one id range serves both channels; each channel's answer equals its own plan over those ids
    n, s, d = H.golden_graph("wiki-vote")
    g = Z.Graph.from_edges(s, d, n, seed=1)
    truths = [g.exact_count(w) for w in whats]
    prof = L.LearnedErrorProfile()
    eps = 0.05
    inst = inst_cls(g, seed=41)
    r = L.estimate_both_to_error(inst, eps, 0.95, prof, max_estimators=1 << 32)
    say(f"[elp] {inst_cls.__name__}: N={r['estimators']} rounds={r['topup_rounds']} "
        + ", ".join(f"{c['sizing']} {c['estimate']:.6g} +- {c['half_width']:.3g}" for c in r["channels"]))
    kinds = L.CHANNEL_KINDS[inst.plan.kind]
    for c, (chan, truth) in enumerate(zip(r["channels"], truths)):
        assert chan["accepted"] and chan["rel_half_width"] <= eps
        assert abs(chan["estimate"] - truth) <= 2 * chan["half_width"]
        single = Z.run(g, Z.Plan(kinds[c], completions=inst.plan.completions if c else 0), 41, 0,
                       r["estimators"])[0]
        assert chan["estimate"] == pytest.approx(single.mean, rel=1e-12)


@pytest.mark.parametrize("kind", [Z.PLAN_K4_SPLIT, Z.PLAN_TRIANGLE_K4_SPLIT, Z.PLAN_HOUSE_SPLIT])
def test_split_completions_choice(kind):
    # two pilots fit Var(J) = Vb + Vw / J and t(J) = c0 + c1 J; the choice is a
    # power of two in [1, 4096] and the pilots' ids stay clear of the query's
    n, s, d = H.golden_graph("wiki-vote")
    g = Z.Graph.from_edges(s, d, n, seed=1)
    J, rec = L.choose_split_completions(g, seed=7, kind=kind, eps=None)   # both pilots
    say(f"[elp] split choice kind={kind}: J={J} J*={rec['j_star']:.1f} Vb={rec['v_between']:.4g} "
        f"Vw={rec['v_within']:.4g} c0={rec['c0_ms_per_id']:.3g} c1={rec['c1_ms_per_id_per_completion']:.3g} "
        f"{rec['seconds'] * 1e3:.1f} ms")
    assert J in [1 << k for k in range(13)]
    lo, hi = (rec["pilots"][j] for j in sorted(rec["pilots"]))
    assert lo["ids"] == hi["ids"] and lo["successes"] >= 32 and hi["successes"] >= lo["successes"]
    assert lo["ids"] * lo["ms_per_id"] >= 1.0
    # on wiki-vote part of a closed first stage's variance is in its completions
    assert rec["v_within"] > 0
    if kind != Z.PLAN_TRIANGLE_K4_SPLIT:
        # 4-cliques against the exact count, houses against the reference's own (FiveHouse.cpp:68)
        truth = g.exact_count("k4") if kind == Z.PLAN_K4_SPLIT else 2134981304
        r = L.split_to_error(g, 0.02, 0.95, L.LearnedErrorProfile(), kind=kind, seed=7, max_estimators=1 << 32)
        say(f"[elp] split auto kind={kind}: J={r['completions']} N={r['estimators']} err="
            f"{abs(r['estimate'] - truth) / truth:.4f} hw={r['rel_half_width']:.4f}")
        assert r["accepted"] and r["rel_half_width"] <= 0.02
        assert abs(r["estimate"] - truth) <= 2 * r["half_width"]
        direct = Z.run(g, Z.Plan(kind, completions=r["completions"]), 7, 0, r["estimators"])[0]
        assert r["estimate"] == pytest.approx(direct.mean, rel=1e-12)
