"""The C++ host API (sampling-graph-mining_amd/host/ZGraph.hpp): its own
host-side test program, the application mains failing loudly without a GPU,
and (GPU) every application main against the Python API on the same graph,
seed and estimator ids."""
import os
import re
import subprocess

import numpy as np
import pytest

import zgraph as Z
from helpers import golden_graph, golden_weights

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "sampling-graph-mining_amd", "host")
BIN = os.path.join(HOST, "_bin")


def _bin(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        pytest.fail(f"{path} is not built (make -C sampling-graph-mining_amd/host)")
    return path


def _write_graph(tmp_path, name, n, src, dst, weights=None):
    prefix = str(tmp_path / name)
    with open(prefix + ".csv", "w") as f:
        f.write("num_vertices,num_edges\n%d,%d\n" % (n, len(src)))
    Z.save_biedgelist(prefix + ".biedgelist", src, dst, weights)
    return prefix


def _report(out):
    rows = {}
    for line in out.splitlines():
        m = re.match(r"^(\S+) (\S+) (\d+)$", line.strip())
        if m:
            rows[m.group(1)] = (float(m.group(2)), int(m.group(3)))
    return rows


def test_host_api_program(tmp_path):
    r = subprocess.run([_bin("test_host_api"), str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "regression k^2 = 609.8565" in r.stdout
    assert r.stdout.strip().endswith("OK")


def test_apps_fail_loudly_without_gpu(tmp_path):
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    n, s, d = golden_graph("simple-30")
    prefix = _write_graph(tmp_path, "g", n, s, d)
    r = subprocess.run([_bin("Triangle"), prefix, "1000"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "error" in r.stderr


APPS = [
    # (binary, extra args, report name, python instance factory)
    ("Triangle", [], "triangle", lambda g: Z.Plan(Z.PLAN_TRIANGLE)),
    ("FourClique", [], "4-clique", lambda g: Z.Plan(Z.PLAN_K4_CLOSE)),
    ("ThreeChain", [], "3-chain", lambda g: Z.Plan(Z.PLAN_THREE_CHAIN)),
    ("FourChain", [], "4-chain", lambda g: Z.Plan(Z.PLAN_FOUR_CHAIN)),
    ("ChainMining", ["5"], "chain", lambda g: Z.Plan(Z.PLAN_CHAIN, order=5)),
    ("FiveStar", [], "5-star", lambda g: Z.Plan(Z.PLAN_PATTERN, pattern=Z.FIVE_STAR_PATTERN)),
    ("FiveHouse", [], "5-house", lambda g: Z.Plan(Z.PLAN_PATTERN, pattern=Z.FIVE_HOUSE_PATTERN)),
]


@pytest.mark.gpu
@pytest.mark.parametrize("app,extra,name,plan", APPS, ids=[a[0] for a in APPS])
def test_app_matches_python_api(tmp_path, app, extra, name, plan):
    n, s, d = golden_graph("wiki-vote")
    prefix = _write_graph(tmp_path, "wv", n, s, d)
    N, seed = 1 << 20, 5
    print(f"[cpp] {app} N={N} seed={seed}", flush=True)
    r = subprocess.run([_bin(app), prefix, str(N), str(seed)] + extra, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    est, cnt = _report(r.stdout)[name]
    assert cnt == N
    g = Z.Graph.from_edges(s, d, n, seed=1)          # the loader's default shuffle seed
    ref = Z.run(g, plan(g), seed, 0, N)[0]
    exact = ref.total() / 2 ** ref.log2_den / N
    assert est == pytest.approx(exact, rel=1e-13, abs=1e-9)


@pytest.mark.gpu
def test_three_motif_app(tmp_path):
    n, s, d = golden_graph("wiki-vote")
    prefix = _write_graph(tmp_path, "wv", n, s, d)
    N, seed = 1 << 20, 3
    r = subprocess.run([_bin("ThreeMotif"), prefix, str(N), str(seed)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    rows = _report(r.stdout)
    g = Z.Graph.from_edges(s, d, n, seed=1)
    ref = Z.run(g, Z.Plan(Z.PLAN_THREE_MOTIF), seed, 0, N)
    assert rows["3-chain"][0] == pytest.approx(ref[0].mean, rel=1e-12)
    assert rows["triangle"][0] == pytest.approx(ref[1].mean, rel=1e-12)


@pytest.mark.gpu
def test_triangle_four_clique_app(tmp_path):
    n, s, d = golden_graph("wiki-vote")
    prefix = _write_graph(tmp_path, "wv", n, s, d)
    N, seed = 1 << 20, 5
    r = subprocess.run([_bin("TriangleFourClique"), prefix, str(N), str(seed)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    rows = _report(r.stdout)
    g = Z.Graph.from_edges(s, d, n, seed=1)
    tri = Z.run(g, Z.Plan(Z.PLAN_TRIANGLE), seed, 0, N)[0]
    k4 = Z.run(g, Z.Plan(Z.PLAN_K4_CLOSE), seed, 0, N)[0]
    assert rows["triangle"][0] == pytest.approx(tri.mean, rel=1e-12)
    assert rows["4-clique"][0] == pytest.approx(k4.mean, rel=1e-12)


@pytest.mark.gpu
def test_predicate_apps(tmp_path):
    """All-predicate and at-least-one weighted triangles on the reference's
    weighted wiki-vote (u32 weights in [0, 79)): the C++ mains equal the
    Python API on the filtered graphs, and at-least-one = all - none."""
    n, s, d = golden_graph("wiki-vote")
    w = golden_weights("wiki-vote").astype(np.uint32)
    prefix = _write_graph(tmp_path, "wvw", n, s, d, w)
    N, seed = 1 << 20, 2
    r = subprocess.run([_bin("AllPredicateWeightedTriangle"), prefix, str(N), str(seed)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    est_all = _report(r.stdout)["all-predicate-triangle"][0]
    g = Z.Graph.from_edges(s, d, n, weights=w.astype(np.float32), seed=1)
    gf = g.filter(50.0, True, 1)
    ref = Z.run(gf, Z.Plan(Z.PLAN_TRIANGLE), seed, 0, N)[0]
    assert est_all == pytest.approx(ref.total() / 2 ** ref.log2_den / N, rel=1e-13)
    r = subprocess.run([_bin("AtLeastOnePredicateWeightedTriangle"), prefix, str(N), str(seed)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    est_one = _report(r.stdout)["at-least-one-predicate-triangle"][0]
    ga = Z.run(g, Z.Plan(Z.PLAN_TRIANGLE), seed, 0, N)[0]
    gn = Z.run(g.filter(50.0, False, 1), Z.Plan(Z.PLAN_TRIANGLE), seed, 0, N)[0]
    assert est_one == pytest.approx(ga.mean - gn.mean, rel=1e-9, abs=1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("pattern,kind,completions,sizing", [("triangle", Z.PLAN_TRIANGLE, 0, "learned"),
                                                             ("4-chain", Z.PLAN_FOUR_CHAIN, 0, "learned"),
                                                             ("4-clique-split", Z.PLAN_K4_SPLIT, 256, "pilot"),
                                                             ("5-house-split", Z.PLAN_HOUSE_SPLIT, 256, "pilot")])
def test_learned_profile_cpp_equals_python(tmp_path, pattern, kind, completions, sizing):
    """The C++ LearnedErrorProfile sizes and judges a query as zgraph.elp does
    (the split 4-clique plan, which the network has no slot for, from its
    pilot's CV)."""
    from zgraph import elp as L
    n, s, d = golden_graph("wiki-vote")
    prefix = _write_graph(tmp_path, "wv", n, s, d)
    eps, seed = 0.02, 11
    print(f"[cpp] EstimateToError {pattern}", flush=True)
    r = subprocess.run([_bin("EstimateToError"), prefix, pattern, str(eps), str(seed)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    sized = int(re.search(r"^sized (\d+)$", r.stdout, re.M).group(1))
    judged = re.search(r"^judged (\d) (\S+)$", r.stdout, re.M)
    est, cnt = _report(r.stdout)[pattern]
    assert judged.group(1) == "1" and float(judged.group(2)) <= eps
    g = Z.Graph.from_edges(s, d, n, seed=1)
    inst = Z.ZGraphInstance(g, Z.Plan(kind, completions=completions), seed=seed)
    rec = L.estimate_to_error(inst, eps, 0.95, L.LearnedErrorProfile())
    assert rec["sizing"] == sizing
    assert abs(rec["sized_estimators"] - sized) <= 1
    if rec["estimators"] == cnt:
        assert est == pytest.approx(rec["estimate"], rel=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("pattern,kind", [("4-clique-split-auto", Z.PLAN_K4_SPLIT),
                                          ("5-house-split-auto", Z.PLAN_HOUSE_SPLIT)])
def test_split_completions_choice_cpp_equals_python(tmp_path, pattern, kind):
    """zgraph::choose_split_completions fits the same variance model from the
    same pilots as zgraph.elp.choose_split_completions (the kernel times, and
    so the choice itself, may differ between runs), and the query it sizes
    reaches the error."""
    from zgraph import elp as L
    n, s, d = golden_graph("wiki-vote")
    prefix = _write_graph(tmp_path, "wv", n, s, d)
    eps, seed = 0.02, 13
    print(f"[cpp] EstimateToError {pattern}", flush=True)
    r = subprocess.run([_bin("EstimateToError"), prefix, pattern, str(eps), str(seed)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    m = re.search(r"^completions (\d+) (\d) (\S+) (\S+) (\d+)$", r.stdout, re.M)
    J, skipped, vb, vw, ids = int(m.group(1)), m.group(2) == "1", float(m.group(3)), float(m.group(4)), int(m.group(5))
    assert J in [1 << k for k in range(13)]
    judged = re.search(r"^judged (\d) (\S+)$", r.stdout, re.M)
    assert judged.group(1) == "1" and float(judged.group(2)) <= eps
    g = Z.Graph.from_edges(s, d, n, seed=1)
    # the pilots' sizes (and whether the second one runs) depend on kernel
    # times; refit the C++ pilots' ids in Python
    stats = {j: L.split_pilot_record(Z.run(g, Z.Plan(kind, completions=j), seed, 1 << 60, ids)[0], ids, 1.0)
             for j in (64, 1024)}
    assert stats[64]["successes"] >= 32
    if skipped:
        assert J == 64
    else:
        rec = L.fit_split_model(stats, 64, 1024)
        assert vb == pytest.approx(rec["v_between"], rel=1e-9)
        assert vw == pytest.approx(rec["v_within"], rel=1e-9)
    # 4-cliques: the exact count; houses: the reference's own (FiveHouse.cpp:68)
    truth = g.exact_count("k4") if kind == Z.PLAN_K4_SPLIT else 2134981304
    est, _ = _report(r.stdout)[pattern]
    assert abs(est - truth) <= 2 * float(judged.group(2)) * est


@pytest.mark.gpu
@pytest.mark.parametrize("pattern,cls,names", [("triangle+4-clique", "TriangleFourCliqueInstance", ("triangle", "4-clique")),
                                               ("triangle+4-clique-split", "TriangleFourCliqueSplitInstance",
                                                ("triangle", "4-clique")),
                                               ("3-motif", "ThreeMotifInstance", ("3-chain", "triangle"))])
def test_learned_two_channel_cpp_equals_python(tmp_path, pattern, cls, names):
    """The C++ LearnedErrorProfile::run_both sizes, runs and judges a
    two-channel query as zgraph.elp.estimate_both_to_error does."""
    from zgraph import elp as L
    n, s, d = golden_graph("wiki-vote")
    prefix = _write_graph(tmp_path, "wv", n, s, d)
    eps, seed = 0.05, 41
    print(f"[cpp] EstimateToError {pattern}", flush=True)
    r = subprocess.run([_bin("EstimateToError"), prefix, pattern, str(eps), str(seed)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    sized = int(re.search(r"^sized (\d+)$", r.stdout, re.M).group(1))
    rows = _report(r.stdout)
    g = Z.Graph.from_edges(s, d, n, seed=1)
    rec = L.estimate_both_to_error(getattr(Z, cls)(g, seed=seed), eps, 0.95, L.LearnedErrorProfile())
    assert abs(max(c["sized_estimators"] for c in rec["channels"]) - sized) <= 1
    for c, name in enumerate(names):
        judged = re.search(r"^judged " + re.escape(name) + r" (\d) (\S+)$", r.stdout, re.M)
        assert judged.group(1) == "1" and float(judged.group(2)) <= eps
        est, cnt = rows[name]
        if rec["estimators"] == cnt:
            assert est == pytest.approx(rec["channels"][c]["estimate"], rel=1e-12)


@pytest.mark.gpu
def test_triangle_app_error_profile_equals_python(tmp_path):
    """Triangle main without a budget sizes by ErrorProfile (the reference's
    flow, +-5 % at 95 %): the C++ ErrorProfile and the Python one give the
    same budget and the same estimate on the same graph and seed."""
    n, s, d = golden_graph("wiki-vote")
    prefix = _write_graph(tmp_path, "wv", n, s, d)
    r = subprocess.run([_bin("Triangle"), prefix], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    est, cnt = _report(r.stdout)["triangle"]
    g = Z.Graph.from_edges(s, d, n, seed=1)
    inst = Z.TriangleCoutingInstance(g, seed=1)
    N = Z.ErrorProfile().get_num_estimators(inst, 3, 0.05, 0.95)
    print(f"[cpp] ErrorProfile N={N} (C++ {cnt})", flush=True)
    assert cnt == N
    ref = Z.run(g, Z.Plan(Z.PLAN_TRIANGLE), 1, 0, N)[0]
    assert est == pytest.approx(ref.total() / 2 ** ref.log2_den / N, rel=1e-12)


@pytest.mark.gpu
def test_triangle2_and_three_motif_without_cache_apps(tmp_path):
    n, s, d = golden_graph("wiki-vote")
    prefix = _write_graph(tmp_path, "wv", n, s, d)
    g = Z.Graph.from_edges(s, d, n, seed=1)
    N, seed = 1 << 20, 6
    r = subprocess.run([_bin("Triangle2"), prefix, str(N), str(seed)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    ref = Z.run(g, Z.Plan(Z.PLAN_PATTERN, pattern=Z.TRIANGLE_PATTERN), seed, 0, N)[0]
    assert _report(r.stdout)["triangle-pattern"][0] == pytest.approx(ref.total() / 2 ** ref.log2_den / N,
                                                                     rel=1e-13)
    nc, nt = 1 << 19, 3 << 18
    r = subprocess.run([_bin("ThreeMotifWithoutCache"), prefix, str(nc), str(nt), str(seed)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = _report(r.stdout)
    rc = Z.run(g, Z.Plan(Z.PLAN_THREE_CHAIN), seed, 0, nc)[0]
    rt = Z.run(g, Z.Plan(Z.PLAN_TRIANGLE), seed, 0, nt)[0]
    assert rows["3-chain"] == (pytest.approx(rc.total() / 2 ** rc.log2_den / nc, rel=1e-13), nc)
    assert rows["triangle"] == (pytest.approx(rt.total() / 2 ** rt.log2_den / nt, rel=1e-13), nt)


EXACT = {"TriangleCounting": "triangle", "FourClique": "k4", "ThreeChain": "wedge", "FourChain": "path3",
         "FiveChain": "path4", "FiveStar": "star5", "FiveHouse": "house", "FourCycle": "c4"}


@pytest.mark.gpu
@pytest.mark.parametrize("gname", ["simple-30", "wiki-vote"])
def test_exact_count_app_matches_the_naive_programs_output(tmp_path, gname):
    """host/apps/ExactCount reads the naive_implementation programs' text
    edge list and prints their "count: N" line; every program's count equals
    the Python API's and the oracle's (which the CPU tests pin against the
    reference's programs built from source)."""
    import pyoracle as O
    n, s, d = golden_graph(gname)
    p = tmp_path / "g.txt"
    with open(p, "w") as f:
        f.write(f"{n} {len(s)}\n")
        for a, b in zip(s, d):
            f.write(f"{a} {b}\n")
    g = Z.Graph.from_edges(s, d, n, seed=1)
    og = O.OGraph(n, s, d)
    oracle = {"triangle": og.triangles, "k4": og.k4, "wedge": og.wedges, "path3": og.paths3, "path4": og.paths4,
              "star5": lambda: og.stars(4), "house": og.house_brute if n <= 400 else None, "c4": og.c4}
    for prog, what in EXACT.items():
        print(f"[cpp] ExactCount {gname} {prog}", flush=True)
        r = subprocess.run([_bin("ExactCount"), str(p), prog], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        got = int(r.stdout.strip().split("count:")[1])
        assert got == g.exact_count(what)
        if oracle[what] is not None:
            assert got == oracle[what]()


def test_exact_count_app_rejects_unknown_programs(tmp_path):
    p = tmp_path / "g.txt"
    p.write_text("3 2\n0 1\n1 2\n")
    r = subprocess.run([_bin("ExactCount"), str(p), "SixClique"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "unknown program" in r.stderr


def test_exact_count_app_rejects_bad_input(tmp_path):
    """A missing file, a header without counts and an edge naming a vertex
    outside [0, n) are reported before any device work."""
    exe = _bin("ExactCount")
    r = subprocess.run([exe, str(tmp_path / "missing.txt"), "FiveChain"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "cannot read the header" in r.stderr
    p = tmp_path / "bad.txt"
    p.write_text("3 2\n0 1\n1 7\n")
    r = subprocess.run([exe, str(p), "TriangleCounting"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "bad edge 1" in r.stderr
