"""CPU-side checks of the product: the C-ABI library loads and exports every
symbol include/zgraph_amd.h declares, the host-only ELP arithmetic matches the
oracle, the file-format readers round-trip, and compute entry points fail
loudly without a GPU."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

import pyoracle as O
import zgraph as Z
from zgraph import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    with open(os.path.join(ROOT, "include", "zgraph_amd.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(zg_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_are_exported():
    names = _declared()
    assert len(names) >= 25
    assert sorted(_capi.EXPORTED) == names
    lib = _capi.lib()
    for nm in names:
        assert hasattr(lib, nm), nm


def test_library_is_gfx950():
    with open(_capi.LIB_PATH, "rb") as f:
        blob = f.read()
    assert b"gfx950" in blob


def _gpu_present():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(_gpu_present(), reason="checks the no-GPU error path")
def test_no_gpu_fails_loudly():
    with pytest.raises(Z.ZGraphError) as ei:
        Z.Graph.from_edges(np.array([0], np.uint32), np.array([1], np.uint32), 2)
    assert ei.value.code == _capi.ZG_ESTATE


def _ld(arr):
    return (C.c_longdouble * len(arr))(*arr)


def test_regression_known_answer_product():
    x = [v / 100 for v in (0.09, 0.26, 0.15, 0.42, 0.57, 0.10, 2.62)]
    y = [v / 100 for v in (1, 2, 4, 8, 16, 32, 64)]
    k = _capi.lib().zg_regression_k(_ld(x), _ld(y), 7)
    assert f"{k * k:.4f}" == "609.8565"   # tests/TestRegression.cpp


def test_elp_arithmetic_matches_oracle():
    rng = np.random.default_rng(3)
    for trial in range(20):
        npts = int(rng.integers(1, 3000))
        F = float(rng.uniform(1e3, 1e7))
        est = (F * (1 + rng.normal(0, 0.3, npts) / np.sqrt(np.arange(1, npts + 1)))).tolist()
        args = (npts, F, 1.0, float(rng.integers(1, 5)), int(rng.integers(10, 5000)),
                int(rng.integers(1000, 10 ** 7)), int(rng.integers(100, 50000)),
                int(rng.integers(10 ** 4, 10 ** 9)), 0.05, 0.95)
        k1, k2 = C.c_longdouble(), C.c_longdouble()
        a = _capi.lib().zg_elp_from_pilots(_ld(est), *args, C.byref(k1))
        b = O.lib().og_elp_num_estimators(_ld(est), *args, C.byref(k2))
        assert a == b and k1.value == k2.value
    assert _capi.lib().zg_elp_bound(0.05, 0.05, 1, 2, 1e6, 100, 10 ** 6) == \
        O.lib().og_elp_bound(0.05, 0.05, 1, 2, 1e6, 100, 10 ** 6)


def test_biedgelist_and_meta_roundtrip(tmp_path):
    s = np.array([0, 2, 1, 4], np.uint32)
    d = np.array([1, 3, 4, 0], np.uint32)
    w = np.array([3, 75, 50, 0], np.uint32)
    p = str(tmp_path / "g")
    Z.save_biedgelist(p + ".biedgelist", s, d)
    with open(p + ".csv", "w") as f:
        f.write("vertices, edges\n5, 4\n")       # datasets/simple-*/dataset/*.csv layout
    assert Z.load_graph_meta_data(p + ".csv") == 5
    s2, d2, w2 = Z.load_biedgelist(p + ".biedgelist")
    assert (s2 == s).all() and (d2 == d).all() and w2 is None
    Z.save_biedgelist(p + ".w.biedgelist", s, d, w)
    s3, d3, w3 = Z.load_biedgelist(p + ".w.biedgelist", weighted=True)
    assert (w3 == w.astype(np.float32)).all()
    with open(p + ".edgelist", "w") as f:
        f.write("5 4\n" + "".join(f"{a} {b}\n" for a, b in zip(s, d)))
    n, s4, d4, _ = Z.load_edgelist(p + ".edgelist")
    assert n == 5 and (s4 == s).all() and (d4 == d).all()


def test_clt_half_width():
    # Bernoulli(1/2) scaled by 2: mean 1, variance 1
    hw = Z.clt_half_width(1.0, 2.0 * 100, 100, 0.95)
    assert hw == pytest.approx(1.959964 * math.sqrt(100 / 99) / 10, rel=1e-5)


@pytest.mark.parametrize("pattern,expect", [
    (Z.Pattern(4, [(0, 1), (1, 2), (2, 3)]), (12, 24)),      # tests/TestPattern2.cpp:100-108 (4-chain)
    (Z.FOUR_CLIQUE_PATTERN, (1, 24)),                           # tests/TestPattern2.cpp:111-122
    (Z.FIVE_STAR_PATTERN, (5, 120)),                            # tests/TestPattern2.cpp:127-137
    (Z.FIVE_HOUSE_PATTERN, (60, 120)),
    (Z.TRIANGLE_PATTERN, (1, 6)),
])
def test_pattern_unique_relabelings_product(pattern, expect):
    """The product's host-side dedupe (zg_pattern_unique) reproduces the
    reference's TestPattern2 counts and agrees with the oracle."""
    assert pattern.unique_relabelings() == expect
    assert O.pattern_unique(pattern.num_vertices, pattern.edges) == expect


def test_time_profile_doubling_rule(monkeypatch):
    """TimeProfile.run follows src/ErrorLatencyProfile.cpp:24-47: doubling
    batches while used + 2*last <= limit, then one batch of
    next*(limit-used)/(2*last); batches take consecutive disjoint id ranges.
    Driven by a fake clock (1 ms per 2^16 estimators) and a fake run."""
    clock = {"t": 0.0}
    calls = []

    class FakeTime:
        @staticmethod
        def perf_counter():
            return clock["t"]

    class S:   # a zg_sums stand-in: integer values first % 7, so the exact sum is known
        def __init__(self, v, count):
            self.mean, self.log2_den, self.sumsq, self.successes = float(v), 0, float(v * v * count), count
            self._total = v * count

        def total(self):
            return self._total

    def fake_run(graph, plan, seed, first, count, *a, **k):
        calls.append((first, count))
        clock["t"] += count / 65536 * 1e-3
        return [S(first % 7, count)]

    monkeypatch.setattr(Z, "time", FakeTime)
    monkeypatch.setattr(Z, "run", fake_run)

    class Inst:
        graph = plan = None
        seed = 1
        channel = 0

    tp = Z.TimeProfile(1 << 16)
    est = tp.run(Inst(), 3, 100)
    sizes = [c for _, c in calls]
    assert [(f, c) for f, c, _ in tp.batches] == calls and tp.remainder
    assert [ms for _, _, ms in tp.batches[:6]] == [1, 2, 4, 8, 16, 32]
    assert tp.estimation.exact_sum == sum((f % 7) * c for f, c in calls)
    assert sizes[:6] == [1 << (16 + i) for i in range(6)]          # 1, 2, 4, 8, 16, 32 ms
    assert sizes[6] == (1 << 22) * (100 - 63) // (32 * 2)          # the remaining-time batch
    assert len(calls) == 7
    assert all(a[0] + a[1] == b[0] for a, b in zip(calls, calls[1:])) and calls[0][0] == 0
    want = sum((f % 7) * c for f, c in calls) / sum(sizes)
    assert est == pytest.approx(want)


def test_abi_argument_errors_without_gpu():
    """Argument checks of the C ABI that run before any device work return
    ZG_EINVAL with a message (host logic; no GPU needed)."""
    L = _capi.lib()
    assert L.zg_plan_create(None, None) == _capi.ZG_EINVAL
    assert L.zg_run(None, None, 1, 0, 10, None, None, None) == _capi.ZG_EINVAL
    assert L.zg_exact_counts(None, 0, None) == _capi.ZG_EINVAL
    assert L.zg_batch_create(None, 0, None, None) == _capi.ZG_EINVAL
    assert L.zg_mlp_forward(None, None, 0, None) == _capi.ZG_EINVAL
    # pattern descriptions are validated on the host
    for verts, edges in ((3, [(0, 1), (0, 1)]), (3, [(0, 3)]), (2, [(0, 1), (1, 0)]), (7, [(0, 1)]),
                         (3, [(1, 1)])):
        P = Z.Pattern(verts, edges)
        with pytest.raises(Z.ZGraphError):
            P.unique_relabelings()
    assert "pattern" in _capi.last_error()


def test_plan_kinds_and_channels_without_gpu():
    """Plan creation is host logic (pattern plans aside, which upload their
    acceptance table): every kind has its channel count and conditional-call
    constant; an unknown kind is rejected."""
    want = {Z.PLAN_TRIANGLE: (1, 1), Z.PLAN_THREE_CHAIN: (1, 1), Z.PLAN_THREE_MOTIF: (2, 1),
            Z.PLAN_FOUR_CHAIN: (1, 2), Z.PLAN_K4_CLOSE: (1, 2), Z.PLAN_TRIANGLE_K4: (2, 2)}
    for kind, (nch, calls) in want.items():
        p = Z.Plan(kind)
        assert (p.channels, p.cond_calls) == (nch, calls)
    assert Z.Plan(Z.PLAN_CHAIN, order=6).cond_calls == 4
    for kind, nch, calls in ((Z.PLAN_K4_SPLIT, 1, 2), (Z.PLAN_TRIANGLE_K4_SPLIT, 2, 2), (Z.PLAN_HOUSE_SPLIT, 1, 3)):
        for J in (1, 64, 4096):
            p = Z.Plan(kind, completions=J)
            assert (p.channels, p.cond_calls) == (nch, calls)
        for J in (0, 3, 8192):   # completions: a power of two in [1, 4096]
            with pytest.raises(Z.ZGraphError):
                Z.Plan(kind, completions=J)
    for bad in (0, 12, 100):
        with pytest.raises(Z.ZGraphError):
            Z.Plan(bad)
    with pytest.raises(Z.ZGraphError):
        Z.Plan(Z.PLAN_CHAIN, order=2)
