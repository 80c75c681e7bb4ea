"""The R-MAT workload (edge factor 16, a=.57, b=c=.19, float weights U[0,1),
predicate threshold 0.5) at scale.

Size-independent properties at scales 20 and 26 (1.07 B edges; runs whenever
the device has 200 GiB of HBM free, as an MI355X does):
  * the fused 3-motif plan's two channels equal the separate 3-chain and
    triangle plans exactly (same draws per id);
  * 3-chain, triangle and 4-chain (FourChain) estimates lie in the 99.9 %
    intervals of the exact wedge, triangle and 3-edge path counts, and the
    split 4-clique plan's (fused with the triangle channel, which equals the
    triangle plan) in that of the exact 4-clique count;
  * predicate triangles: the all-edges estimate on the w >= 0.5 graph lies in
    the interval of that graph's exact triangle count, and the at-least-one
    estimate (all triangles minus those of the w < 0.5 graph,
    applications/AtLeastOnePredicateWeightedTriangle.cpp:135-152) in the
    interval of the exact difference;
  * a top-up equals one run; the two complementary filters partition the edges.
Per-estimator parity at scale 18 (4.2 M edges): the device CSR, both weight
filters and every estimator of the four R-MAT plans (3-motif, 3-chain,
4-chain, predicate triangles) equal the CPU oracle's."""
import numpy as np
import pytest

import zgraph as Z
import pyoracle as O

pytestmark = pytest.mark.gpu

A, B, C_ = 0.57, 0.19, 0.19
HBM_SCALE26 = 200 << 30   # peak HBM at scale 26: the graph (~100 GiB) and one filtered graph being built


def say(*a):
    print(*a, flush=True)


def _in_interval(name, s, n, exact):
    hw = Z.clt_half_width(s.mean, s.sumsq, n, 0.999)
    say(f"[scale] {name}: exact {exact} estimate {s.mean:.6g} +- {hw:.3g}")
    assert abs(s.mean - exact) <= hw, name


def _check(scale, N):
    m = 16 << scale
    say(f"[scale] R-MAT scale {scale} m={m}: generating")
    e = Z.rmat(scale, m, A, B, C_, 3).add_uniform_weights(4)
    g = Z.Graph.from_device(e, seed=3)
    e.free()
    assert g.num_local_edges == m and g.num_vertices == 1 << scale
    say(f"[scale] built: maxdeg {g.max_degree}, {g.device_bytes / 2**30:.1f} GiB")
    motif = Z.run(g, Z.Plan(Z.PLAN_THREE_MOTIF), 11, 0, N)
    chain = Z.run(g, Z.Plan(Z.PLAN_THREE_CHAIN), 11, 0, N)[0]
    tri = Z.run(g, Z.Plan(Z.PLAN_TRIANGLE), 11, 0, N)[0]
    assert motif[0].total() == chain.total() and motif[0].successes == chain.successes
    assert motif[1].total() == tri.total() and motif[1].successes == tri.successes
    _in_interval("wedges", chain, N, g.exact_count("wedge"))
    T = g.exact_count("triangle")
    _in_interval("triangles", tri, N, T)
    four = Z.run(g, Z.Plan(Z.PLAN_FOUR_CHAIN), 11, 0, N)[0]
    _in_interval("3-edge paths (4-chain)", four, N, g.exact_count("path3"))
    # 4-cliques by the split plan (256 completions per closed triangle): the first stage's
    # triangles are the triangle plan's, and the estimate lies in the interval of the exact count
    split = Z.run(g, Z.Plan(Z.PLAN_TRIANGLE_K4_SPLIT, completions=256), 11, 0, N // 4)
    assert split[0].total() == Z.run(g, Z.Plan(Z.PLAN_TRIANGLE), 11, 0, N // 4)[0].total()
    _in_interval("4-cliques (split plan)", split[1], N // 4, g.exact_count("k4"))
    if scale <= 22:
        # 5-houses by the split plan (16 completions per closed roof) against the exact count
        # (per-edge triangles weighting the 4-cycle listing; hubs use dense counters of n entries,
        # so the count stays off the 1.07 B-edge graph)
        house = Z.run(g, Z.Plan(Z.PLAN_HOUSE_SPLIT, completions=16), 11, 0, N // 4)[0]
        _in_interval("5-houses (split plan)", house, N // 4, g.exact_count("house"))
    half = Z.run(g, Z.Plan(Z.PLAN_FOUR_CHAIN), 11, 0, N // 2)[0]
    rest = Z.run(g, Z.Plan(Z.PLAN_FOUR_CHAIN), 11, N // 2, N - N // 2)[0]
    assert half.total() + rest.total() == four.total()
    # all-edges predicate triangles (AllPredicateWeightedTriangle.cpp): the plan on the filtered graph
    ge = g.filter(0.5, True, 5)
    m_ge = ge.num_local_edges
    all_ge = Z.run(ge, Z.Plan(Z.PLAN_TRIANGLE), 11, 0, N)[0]
    _in_interval("all-edges-predicate triangles", all_ge, N, ge.exact_count("triangle"))
    ge.close()   # one filtered graph resident at a time
    # at-least-one = all - (all edges failing the predicate), independent runs
    lt = g.filter(0.5, False, 5)
    assert m_ge + lt.num_local_edges == m
    assert abs(m_ge - m / 2) < 6 * (m ** 0.5)
    none = Z.run(lt, Z.Plan(Z.PLAN_TRIANGLE), 12, 0, N)[0]
    T_lt = lt.exact_count("triangle")
    lt.close()
    est = tri.mean - none.mean
    hw = np.hypot(Z.clt_half_width(tri.mean, tri.sumsq, N, 0.999), Z.clt_half_width(none.mean, none.sumsq, N, 0.999))
    say(f"[scale] at-least-one-predicate triangles: exact {T - T_lt} estimate {est:.6g} +- {hw:.3g}")
    assert abs(est - (T - T_lt)) <= hw
    g.close()


def test_rmat_scale20_properties():
    _check(20, 1 << 24)


def test_rmat_scale26_properties():
    free, total = Z.device_memory()
    if free < HBM_SCALE26:
        pytest.skip(f"needs {HBM_SCALE26 >> 30} GiB of free HBM, the device has {free >> 30} GiB free")
    _check(26, 1 << 27)


def test_rmat_scale18_oracle_parity():
    scale, seed = 18, 3
    m = 16 << scale
    e = Z.rmat(scale, m, A, B, C_, seed).add_uniform_weights(4)
    s, d, w = e.download()
    g = Z.Graph.from_device(e, seed=seed)
    e.free()
    og = O.OGraph(1 << scale, s, d, w, seed=seed)
    say(f"[scale18] m={m} maxdeg={g.max_degree}")

    def same_csr(dev, ora):
        a, b = dev.export(), ora.arrays()
        for k in ("src", "dst", "off", "adj_rank"):
            assert (a[k] == b[k]).all(), k
        assert (a["w"] == b["w"]).all()

    same_csr(g, og)
    n_ids = 1 << 18
    cases = [("3-motif", g, og, 3), ("3-chain", g, og, 2), ("4-chain", g, og, 5)]
    ge, oge = g.filter(0.5, True, 5), og.filter(0.5, True, 5)
    lt, olt = g.filter(0.5, False, 5), og.filter(0.5, False, 5)
    same_csr(ge, oge)
    same_csr(lt, olt)
    cases += [("predicate triangles (w >= 0.5)", ge, oge, 1), ("predicate triangles (w < 0.5)", lt, olt, 1)]
    for name, dg, dog, kind in cases:
        zo, zs = Z.run_outcomes(dg, Z.Plan(kind), 11, 7, n_ids)
        oo = O.run_outcomes(dog, O.make_plan(kind), 11, 7, n_ids)
        for f in ("status", "num_lo", "num_hi", "log2den", "e0", "ncond", "nwords"):
            assert (zo[f] == oo[f]).all(), (name, f)
        osums = O.run(dog, O.make_plan(kind), 11, 7, n_ids)
        for c, o in zip(zs, osums):
            assert c.total() == o.total() and c.successes == o.successes, name
        say(f"[scale18] {name}: {n_ids} estimators bit-exact")
    for x in (ge, lt, g):
        x.close()
