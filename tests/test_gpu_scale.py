"""Size-independent properties at R-MAT scale (edge factor 16, a=.57,
b=c=.19, float weights U[0,1)): the fused 3-motif plan's two channels equal
the separate 3-chain and triangle plans exactly (same draws per id); the
3-chain and triangle estimates lie in the 99.9 % intervals of the exact wedge
and triangle counts; a top-up
equals one run; the two complementary weight filters partition the edges.
Scale 20 (16.8 M edges) always runs; scale 26 (1.07 B edges, about 150 GB of
HBM while building) runs when ZG_FULL_SCALE=1."""
import os

import pytest

import zgraph as Z

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _check(scale, N):
    m = 16 << scale
    say(f"[scale] R-MAT scale {scale} m={m}: generating")
    e = Z.rmat(scale, m, 0.57, 0.19, 0.19, 3).add_uniform_weights(4)
    g = Z.Graph.from_device(e, seed=3)
    e.free()
    assert g.num_local_edges == m and g.num_vertices == 1 << scale
    say(f"[scale] built: maxdeg {g.max_degree}, {g.device_bytes / 2**30:.1f} GiB")
    motif = Z.run(g, Z.Plan(Z.PLAN_THREE_MOTIF), 11, 0, N)
    chain = Z.run(g, Z.Plan(Z.PLAN_THREE_CHAIN), 11, 0, N)[0]
    tri = Z.run(g, Z.Plan(Z.PLAN_TRIANGLE), 11, 0, N)[0]
    assert motif[0].total() == chain.total() and motif[0].successes == chain.successes
    assert motif[1].total() == tri.total() and motif[1].successes == tri.successes
    W = g.exact_count("wedge")
    hw = Z.clt_half_width(chain.mean, chain.sumsq, N, 0.999)
    say(f"[scale] wedges exact {W} estimate {chain.mean:.6g} +- {hw:.3g}")
    assert abs(chain.mean - W) <= hw
    T = g.exact_count("triangle")
    hwt = Z.clt_half_width(tri.mean, tri.sumsq, N, 0.999)
    say(f"[scale] triangles exact {T} estimate {tri.mean:.6g} +- {hwt:.3g}")
    assert abs(tri.mean - T) <= hwt
    half = Z.run(g, Z.Plan(Z.PLAN_FOUR_CHAIN), 11, 0, N // 2)[0]
    rest = Z.run(g, Z.Plan(Z.PLAN_FOUR_CHAIN), 11, N // 2, N - N // 2)[0]
    whole = Z.run(g, Z.Plan(Z.PLAN_FOUR_CHAIN), 11, 0, N)[0]
    assert half.total() + rest.total() == whole.total()
    ge = g.filter(0.5, True, 5)
    lt = g.filter(0.5, False, 5)
    assert ge.num_local_edges + lt.num_local_edges == m
    assert abs(ge.num_local_edges - m / 2) < 6 * (m ** 0.5)
    for f in (ge, lt):
        r = Z.run(f, Z.Plan(Z.PLAN_TRIANGLE), 11, 0, N // 4)[0]
        assert r.count == N // 4
        f.close()
    g.close()


def test_rmat_scale20_properties():
    _check(20, 1 << 24)


@pytest.mark.skipif(os.environ.get("ZG_FULL_SCALE") != "1", reason="needs ~150 GB of HBM; set ZG_FULL_SCALE=1")
def test_rmat_scale26_properties():
    _check(26, 1 << 27)
