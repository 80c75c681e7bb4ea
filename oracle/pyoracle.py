"""ctypes binding of the CPU oracle (oracle/oracle.c).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

PLAN_TRIANGLE = 1
PLAN_THREE_CHAIN = 2
PLAN_THREE_MOTIF = 3
PLAN_CHAIN = 4
PLAN_FOUR_CHAIN = 5
PLAN_PATTERN = 6
PLAN_K4_CLOSE = 7
PLAN_K4_SPLIT = 9     # order = completions (a power of two)
PLAN_HOUSE_SPLIT = 11 # order = completions (a power of two)

MAX_PE = 16


class Graph(C.Structure):
    _fields_ = [("n", C.c_uint32), ("m", C.c_uint64), ("num_edges", C.c_uint64),
                ("max_degree", C.c_uint64),
                ("src", C.POINTER(C.c_uint32)), ("dst", C.POINTER(C.c_uint32)),
                ("w", C.POINTER(C.c_float)), ("off", C.POINTER(C.c_uint64)),
                ("adj_rank", C.POINTER(C.c_uint64)), ("adj_src", C.POINTER(C.c_uint32)),
                ("adj_dst", C.POINTER(C.c_uint32))]


class Plan(C.Structure):
    _fields_ = [("kind", C.c_int), ("order", C.c_int), ("pk", C.c_int), ("pe", C.c_int),
                ("pat_src", C.c_uint32 * MAX_PE), ("pat_dst", C.c_uint32 * MAX_PE)]


class Outcome(C.Structure):
    _fields_ = [("num_lo", C.c_uint64), ("num_hi", C.c_uint64), ("e0", C.c_uint32),
                ("status", C.c_uint8), ("log2den", C.c_uint8), ("ncond", C.c_uint8),
                ("nwords", C.c_uint8)]


OUTCOME_DTYPE = np.dtype([("num_lo", "<u8"), ("num_hi", "<u8"), ("e0", "<u4"), ("status", "u1"),
                          ("log2den", "u1"), ("ncond", "u1"), ("nwords", "u1")])
assert OUTCOME_DTYPE.itemsize == C.sizeof(Outcome) == 24


class Sums(C.Structure):
    _fields_ = [("sum", C.c_uint64 * 3), ("sumsq", C.c_double), ("successes", C.c_uint64),
                ("count", C.c_uint64), ("K", C.c_int)]

    def total(self):
        """exact sum of values times 2^K, as a Python int"""
        return self.sum[0] | (self.sum[1] << 64) | (self.sum[2] << 128)


assert C.sizeof(Sums) == 56


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(_LIB_PATH)
        P = C.POINTER
        L.og_philox4x32_10.argtypes = [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]
        L.og_key64.restype = C.c_uint64
        L.og_key64.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64]
        L.og_graph_build.argtypes = [P(Graph), C.c_uint32, C.c_uint64, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_uint64]
        L.og_graph_free.argtypes = [P(Graph)]
        L.og_graph_from_csr.argtypes = [P(Graph), C.c_uint32, C.c_uint64] + [C.c_void_p] * 5
        L.og_graph_subsample.argtypes = [P(Graph), P(Graph), C.c_double, C.c_uint64]
        L.og_device_lines.argtypes = [P(Graph), C.c_int, C.c_uint64, C.c_uint64, C.c_uint64]
        L.og_device_lines.restype = C.c_uint64
        L.og_device_lines_house.argtypes = [P(Graph), C.c_int, C.c_uint64, C.c_uint64, C.c_uint64]
        L.og_device_lines_house.restype = C.c_uint64
        L.og_graph_filter.argtypes = [P(Graph), P(Graph), C.c_float, C.c_int, C.c_uint64]
        L.og_estimate.argtypes = [P(Graph), P(Plan), C.c_uint64, C.c_uint64, P(Outcome)]
        L.og_run.argtypes = [P(Graph), P(Plan), C.c_uint64, C.c_uint64, C.c_uint64, C.c_int, P(Sums)]
        L.og_run_outcomes.argtypes = [P(Graph), P(Plan), C.c_uint64, C.c_uint64, C.c_uint64,
                                      C.c_int, C.c_void_p]
        L.og_sums_mean.restype = C.c_double
        L.og_sums_mean.argtypes = [P(Sums)]
        L.og_expectation.restype = C.c_double
        L.og_expectation.argtypes = [P(Graph), P(Plan), C.c_int, C.c_uint64]
        L.og_plan_channels.argtypes = [P(Plan)]
        L.og_plan_max_log2den.argtypes = [P(Plan)]
        L.og_plan_cond_calls.restype = C.c_double
        L.og_plan_cond_calls.argtypes = [P(Plan)]
        L.og_pattern_unique_count.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, P(C.c_int)]
        L.og_pattern_edge_order_classes.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, P(C.c_int)]
        L.og_cond_valid_count.restype = C.c_uint64
        L.og_cond_valid_count.argtypes = [P(Graph), C.c_uint64, C.c_void_p, C.c_int, C.c_void_p]
        L.og_cond_pick.argtypes = [P(Graph), C.c_uint64, C.c_void_p, C.c_int, C.c_uint64,
                                   P(C.c_uint64), P(C.c_uint32), P(C.c_uint32)]
        L.og_close.argtypes = [P(Graph), C.c_uint64, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int]
        for nm in ("og_exact_triangles", "og_exact_k4", "og_exact_paths3"):
            getattr(L, nm).restype = C.c_uint64
            getattr(L, nm).argtypes = [P(Graph), C.c_int]
        L.og_exact_wedges.restype = C.c_uint64
        L.og_exact_wedges.argtypes = [P(Graph)]
        for nm in ("og_exact_c4", "og_exact_paths4"):
            getattr(L, nm).restype = C.c_uint64
            getattr(L, nm).argtypes = [P(Graph)]
        L.og_exact_paths_brute.restype = C.c_uint64
        L.og_exact_paths_brute.argtypes = [P(Graph), C.c_int]
        L.og_set_rng_mode.restype = None
        L.og_set_rng_mode.argtypes = [C.c_int]
        L.og_mt19937_nth.restype = C.c_uint32
        L.og_mt19937_nth.argtypes = [C.c_uint32, C.c_uint32]
        L.og_exact_stars64.restype = None
        L.og_exact_stars64.argtypes = [P(Graph), C.c_int, P(C.c_uint64), P(C.c_uint64)]
        L.og_exact_house_brute.restype = C.c_uint64
        L.og_exact_house_brute.argtypes = [P(Graph)]
        L.og_exact_house.restype = C.c_uint64
        L.og_exact_house.argtypes = [P(Graph)]
        L.og_exact_pattern_brute.restype = C.c_uint64
        L.og_exact_pattern_brute.argtypes = [P(Graph), C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        L.og_regression_k.restype = C.c_longdouble
        L.og_regression_k.argtypes = [C.c_void_p, C.c_void_p, C.c_uint]
        L.og_elp_bound.restype = C.c_double
        L.og_elp_bound.argtypes = [C.c_double] * 5 + [C.c_uint64, C.c_uint64]
        L.og_elp_num_estimators.restype = C.c_double
        L.og_elp_num_estimators.argtypes = [C.c_void_p, C.c_uint64, C.c_longdouble, C.c_double,
                                            C.c_double, C.c_uint64, C.c_uint64, C.c_uint64,
                                            C.c_uint64, C.c_double, C.c_double,
                                            P(C.c_longdouble)]
        P2 = P(P(C.c_uint32))
        L.og_gen_er.argtypes = [C.c_uint32, C.c_uint64, C.c_uint64, C.c_uint64, P2, P2]
        L.og_gen_chung_lu.argtypes = [C.c_uint32, C.c_uint64, C.c_double, C.c_double, C.c_uint64,
                                      C.c_uint64, P2, P2]
        L.og_gen_rmat.argtypes = [C.c_int, C.c_uint64, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                  C.c_uint64, P2, P2]
        L.og_free.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def _take(ps, pd, m):
    s = np.ctypeslib.as_array(ps, (max(m, 1),))[:m].copy()
    d = np.ctypeslib.as_array(pd, (max(m, 1),))[:m].copy()
    lib().og_free(ps)
    lib().og_free(pd)
    return s, d


def gen_er(n, m, seed, max_candidates=None):
    ps, pd = C.POINTER(C.c_uint32)(), C.POINTER(C.c_uint32)()
    if lib().og_gen_er(n, m, seed, max_candidates or 64 * m + 1024, C.byref(ps), C.byref(pd)):
        raise ValueError("not enough candidates")
    return _take(ps, pd, m)


def gen_chung_lu(n, m, beta, max_deg, seed, max_candidates=None):
    ps, pd = C.POINTER(C.c_uint32)(), C.POINTER(C.c_uint32)()
    if lib().og_gen_chung_lu(n, m, beta, max_deg, seed, max_candidates or 64 * m + 1024, C.byref(ps),
                             C.byref(pd)):
        raise ValueError("not enough candidates")
    return _take(ps, pd, m)


def gen_rmat(scale, m, a, b, c, seed, max_candidates=None):
    ps, pd = C.POINTER(C.c_uint32)(), C.POINTER(C.c_uint32)()
    if lib().og_gen_rmat(scale, m, a, b, c, seed, max_candidates or 64 * m + 1024, C.byref(ps),
                         C.byref(pd)):
        raise ValueError("not enough candidates")
    return _take(ps, pd, m)


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().og_philox4x32_10(c, k, o)
    return list(o)


class OGraph:
    """Owns an og_graph built from an input edge list."""

    def __init__(self, n, src, dst, w=None, seed=1, _raw=None):
        self.g = Graph()
        if _raw is not None:
            return
        src = np.ascontiguousarray(src, dtype=np.uint32)
        dst = np.ascontiguousarray(dst, dtype=np.uint32)
        wp = None
        if w is not None:
            w = np.ascontiguousarray(w, dtype=np.float32)
            wp = w.ctypes.data
        rc = lib().og_graph_build(C.byref(self.g), n, len(src), src.ctypes.data, dst.ctypes.data,
                                  wp, seed)
        if rc != 0:
            raise ValueError(f"og_graph_build failed ({rc})")

    def __del__(self):
        try:
            lib().og_graph_free(C.byref(self.g))
        except Exception:
            pass

    @property
    def n(self):
        return self.g.n

    @property
    def m(self):
        return self.g.m

    def arrays(self):
        g = self.g
        m, n = g.m, g.n
        out = {"src": np.ctypeslib.as_array(g.src, (max(m, 1),))[:m].copy(),
               "dst": np.ctypeslib.as_array(g.dst, (max(m, 1),))[:m].copy(),
               "off": np.ctypeslib.as_array(g.off, (n + 1,)).copy(),
               "adj_rank": np.ctypeslib.as_array(g.adj_rank, (max(2 * m, 1),))[:2 * m].copy(),
               "adj_src": np.ctypeslib.as_array(g.adj_src, (max(2 * m, 1),))[:2 * m].copy(),
               "adj_dst": np.ctypeslib.as_array(g.adj_dst, (max(2 * m, 1),))[:2 * m].copy()}
        if g.w:
            out["w"] = np.ctypeslib.as_array(g.w, (max(m, 1),))[:m].copy()
        return out

    @classmethod
    def from_csr(cls, n, src, dst, off, adj_rank, adj_nbr):
        o = cls(0, None, None, _raw=True)
        arrs = [np.ascontiguousarray(a) for a in (src, dst, off, adj_rank, adj_nbr)]
        rc = lib().og_graph_from_csr(C.byref(o.g), n, len(arrs[0]), *[a.ctypes.data for a in arrs])
        if rc:
            raise ValueError(rc)
        return o

    def subsample(self, rate, seed):
        o = OGraph(0, None, None, _raw=True)
        rc = lib().og_graph_subsample(C.byref(self.g), C.byref(o.g), rate, seed)
        if rc:
            raise ValueError(rc)
        return o

    def filter(self, thr, keep_ge, seed):
        o = OGraph(0, None, None, _raw=True)
        rc = lib().og_graph_filter(C.byref(self.g), C.byref(o.g), thr, 1 if keep_ge else 0, seed)
        if rc:
            raise ValueError(rc)
        return o

    # exact counts
    def triangles(self, threads=0):
        return lib().og_exact_triangles(C.byref(self.g), threads)

    def k4(self, threads=0):
        return lib().og_exact_k4(C.byref(self.g), threads)

    def wedges(self):
        return lib().og_exact_wedges(C.byref(self.g))

    def paths3(self, threads=0):
        return lib().og_exact_paths3(C.byref(self.g), threads)

    def c4(self):
        return lib().og_exact_c4(C.byref(self.g))

    def paths4(self):
        return lib().og_exact_paths4(C.byref(self.g))

    def paths_brute(self, nverts):
        return lib().og_exact_paths_brute(C.byref(self.g), nverts)

    def stars(self, leaves=4):
        lo, hi = C.c_uint64(), C.c_uint64()
        lib().og_exact_stars64(C.byref(self.g), leaves, C.byref(lo), C.byref(hi))
        return lo.value | (hi.value << 64)

    def house_brute(self):
        return lib().og_exact_house_brute(C.byref(self.g))

    def house(self):
        return lib().og_exact_house(C.byref(self.g))

    def pattern_brute(self, k, edges):
        s = np.array([e[0] for e in edges], np.uint32)
        d = np.array([e[1] for e in edges], np.uint32)
        return lib().og_exact_pattern_brute(C.byref(self.g), k, len(edges), s.ctypes.data, d.ctypes.data)


def make_plan(kind, order=0, pattern=None):
    p = Plan()
    p.kind = kind
    p.order = order
    if pattern is not None:
        k, edges = pattern
        p.pk = k
        p.pe = len(edges)
        for i, (a, b) in enumerate(edges):
            p.pat_src[i] = a
            p.pat_dst[i] = b
    return p


def channels(plan):
    return lib().og_plan_channels(C.byref(plan))


def run_outcomes(g, plan, seed, first, count, threads=0):
    nch = channels(plan)
    out = np.zeros(count * nch, dtype=OUTCOME_DTYPE)
    lib().og_run_outcomes(C.byref(g.g), C.byref(plan), seed, first, count, threads, out.ctypes.data)
    return out.reshape(count, nch) if nch > 1 else out


def run(g, plan, seed, first, count, threads=0):
    nch = channels(plan)
    sums = (Sums * nch)()
    lib().og_run(C.byref(g.g), C.byref(plan), seed, first, count, threads, sums)
    return list(sums)


def device_lines(g, kind, seed, first, count):
    """128-byte lines the device kernel of `kind` reads for these estimators
    (lines.c: the device layout's costs along the oracle's draws)."""
    r = lib().og_device_lines(C.byref(g.g), kind, seed, first, count)
    if r == (1 << 64) - 1:
        raise ValueError("no line model for this plan kind")
    return r


def device_lines_house(g, completions, seed, first, count):
    """device_lines for the split 5-house plan with `completions` per closed roof."""
    return lib().og_device_lines_house(C.byref(g.g), completions, seed, first, count)


def mean(s):
    return lib().og_sums_mean(C.byref(s))


def expectation(g, plan, channel=0, max_paths=50_000_000):
    return lib().og_expectation(C.byref(g.g), C.byref(plan), channel, max_paths)


def pattern_unique(k, edges):
    s = np.array([e[0] for e in edges], np.uint32)
    d = np.array([e[1] for e in edges], np.uint32)
    tot = C.c_int()
    u = lib().og_pattern_unique_count(k, len(edges), s.ctypes.data, d.ctypes.data, C.byref(tot))
    return u, tot.value


def pattern_edge_order_classes(k, edges):
    s = np.array([e[0] for e in edges], np.uint32)
    d = np.array([e[1] for e in edges], np.uint32)
    tot = C.c_int()
    u = lib().og_pattern_edge_order_classes(k, len(edges), s.ctypes.data, d.ctypes.data, C.byref(tot))
    return u, tot.value


# patterns used by the reference applications
TRIANGLE = (3, [(0, 1), (0, 2), (1, 2)])                          # Triangle2.cpp:18-24
K4 = (4, [(1, 0), (2, 0), (2, 1), (3, 0), (3, 1), (3, 2)])          # FourClique2 / TestPattern2
FIVE_STAR = (5, [(0, 1), (0, 2), (0, 3), (0, 4)])                  # FiveStar.cpp:18-26
FIVE_HOUSE = (5, [(0, 1), (0, 2), (1, 2), (1, 3), (2, 4), (3, 4)])  # FiveHouse.cpp:18-27
C4 = (4, [(0, 1), (1, 2), (2, 3), (3, 0)])                         # 4-cycle
FOUR_CHAIN = (4, [(0, 1), (1, 2), (2, 3)])                         # TestPattern2.cpp:101-108
FIVE_CHAIN = (5, [(0, 1), (1, 2), (2, 3), (3, 4)])                 # TestPattern.cpp:92-99
