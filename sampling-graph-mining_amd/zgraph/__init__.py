"""zgraph -- Python host mirror of ZGraph's estimator API over the MI355X C ABI.

Names follow the reference (jy-yuan/sampling-graph-mining):
  ZGraphInstance.calculate        include/ZGraphInstance.hpp:125, src/ZGraphInstance.cpp:256-297
  Estimation                      include/ZGraphInstance.hpp:18-28, src/ZGraphInstance.cpp:15-35
  SamplerGenerator / Pattern      include/SamplerGenerator.hpp:9-80
  ErrorProfile / TimeProfile      include/ErrorLatencyProfile.hpp:13-32
  load_graph_meta_data            src/Utilities.cpp:24-35
The estimate() bodies of the reference applications are device plans; each
application class below selects its plan.  Every call goes to the HIP
library; nothing here computes estimates on the CPU.
"""
import ctypes as C
import math
import time

import numpy as np

from . import _capi
from ._capi import (PLAN_CHAIN, PLAN_FOUR_CHAIN, PLAN_HOUSE_SPLIT, PLAN_K4_CLOSE, PLAN_K4_SPLIT, PLAN_PATTERN,
                    PLAN_THREE_CHAIN,
                    PLAN_THREE_MOTIF, PLAN_TRIANGLE, PLAN_TRIANGLE_K4, PLAN_TRIANGLE_K4_SPLIT, ZGraphError, check,
                    lib)
from .formats import load_biedgelist, load_edgelist, load_graph_meta_data, save_biedgelist

__all__ = [
    "Graph", "DeviceEdges", "erdos_renyi", "chung_lu", "rmat", "Plan", "Pattern", "Estimation",
    "ZGraphInstance", "TriangleCoutingInstance", "ThreeChainCountingInstance", "ThreeMotifInstance",
    "TriangleFourCliqueInstance", "TriangleFourCliqueSplitInstance",
    "ChainMiningInstance", "FourChainCountingInstance", "FourCliqueCountingInstance", "FourCliqueSplitCountingInstance",
    "SamplerGenerator", "AllPredicateZGraphInstance", "ErrorProfile", "TimeProfile",
    "load_graph_meta_data", "load_biedgelist", "load_edgelist", "save_biedgelist", "ZGraphError",
    "device_info", "device_memory", "clt_half_width", "run", "run_outcomes", "Batch", "random_gather_bandwidth",
    "random_line_bandwidth", "random_chain_bandwidth", "random_region_bandwidth",
]

OUTCOME_DTYPE = np.dtype([("num_lo", "<u8"), ("num_hi", "<u8"), ("e0", "<u4"), ("status", "u1"),
                          ("log2den", "u1"), ("ncond", "u1"), ("nwords", "u1")])
assert OUTCOME_DTYPE.itemsize == C.sizeof(_capi.Outcome)


def device_info():
    cu, lds, ws = C.c_int(), C.c_int(), C.c_int()
    check(lib().zg_device_info(C.byref(cu), C.byref(lds), C.byref(ws)))
    return {"num_cus": cu.value, "lds_per_cu": lds.value, "warp_size": ws.value}


def device_memory():
    """(free, total) bytes of device memory"""
    f, t = C.c_uint64(), C.c_uint64()
    check(lib().zg_device_memory(C.byref(f), C.byref(t)))
    return f.value, t.value


# ---------------------------------------------------------------------------
# synthetic inputs (device generators)
# ---------------------------------------------------------------------------
class DeviceEdges:
    """An edge list in HBM produced by a seeded generator (src < dst)."""

    def __init__(self):
        self.e = _capi.Edges()

    @property
    def m(self):
        return self.e.m

    @property
    def n(self):
        return self.e.n

    def add_uniform_weights(self, seed):
        check(lib().zg_edges_add_uniform_weights(C.byref(self.e), seed))
        return self

    def download(self):
        m = self.e.m
        s = np.empty(m, np.uint32)
        d = np.empty(m, np.uint32)
        w = np.empty(m, np.float32) if self.e.d_w else None
        check(lib().zg_edges_download(C.byref(self.e), s.ctypes.data, d.ctypes.data,
                                      w.ctypes.data if w is not None else None))
        return s, d, w

    def free(self):
        if self.e.d_src or self.e.d_dst:
            lib().zg_edges_free(C.byref(self.e))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def erdos_renyi(n, m, seed):
    out = DeviceEdges()
    check(lib().zg_gen_erdos_renyi(n, m, seed, C.byref(out.e)))
    return out


def chung_lu(n, m, beta, max_degree, seed):
    out = DeviceEdges()
    check(lib().zg_gen_chung_lu(n, m, beta, max_degree, seed, C.byref(out.e)))
    return out


def rmat(scale, m, a=0.57, b=0.19, c=0.19, seed=1):
    out = DeviceEdges()
    check(lib().zg_gen_rmat(scale, m, a, b, c, seed, C.byref(out.e)))
    return out


# ---------------------------------------------------------------------------
# graphs
# ---------------------------------------------------------------------------
class Graph:
    """Device-resident CSR (the reference's Graph<E> at one worker)."""

    def __init__(self, handle):
        self.h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle
        info = _capi.GraphInfo()
        check(lib().zg_graph_info_get(self.h, C.byref(info)))
        self.num_vertices = info.n
        self.num_local_edges = info.m
        self.num_edges = info.num_edges
        self.max_degree = info.max_degree
        self.device_bytes = info.device_bytes
        self.has_weights = bool(info.has_weights)
        self.has_parallel_edges = bool(info.has_parallel_edges)

    @classmethod
    def from_edges(cls, src, dst, num_vertices, weights=None, seed=1):
        src = np.ascontiguousarray(src, np.uint32)
        dst = np.ascontiguousarray(dst, np.uint32)
        if src.shape != dst.shape:
            raise ValueError("src/dst length mismatch")
        wp = None
        if weights is not None:
            weights = np.ascontiguousarray(weights, np.float32)
            if weights.shape != src.shape:
                raise ValueError("weights length mismatch")
            wp = weights.ctypes.data
        h = C.c_void_p()
        check(lib().zg_graph_create(src.ctypes.data, dst.ctypes.data, wp, len(src), num_vertices, seed,
                                    C.byref(h)))
        return cls(h)

    @classmethod
    def from_device(cls, edges, seed=1):
        h = C.c_void_p()
        e = edges.e
        check(lib().zg_graph_create_device(e.d_src, e.d_dst, e.d_w, e.m, e.n, seed, C.byref(h)))
        return cls(h)

    @classmethod
    def load(cls, prefix, weighted=False, seed=1):
        """`<prefix>.csv` + `<prefix>.biedgelist` (DistGraphLoader::load_graph's files)."""
        n = load_graph_meta_data(prefix + ".csv")
        src, dst, w = load_biedgelist(prefix + ".biedgelist", weighted=weighted)
        return cls.from_edges(src, dst, n, weights=w, seed=seed)

    def export(self):
        m, n = self.num_local_edges, self.num_vertices
        out = {"src": np.empty(m, np.uint32), "dst": np.empty(m, np.uint32),
               "off": np.empty(n + 1, np.uint64), "adj_rank": np.empty(2 * m, np.uint32),
               "adj_nbr": np.empty(2 * m, np.uint32)}
        if self.has_weights:
            out["w"] = np.empty(m, np.float32)
        check(lib().zg_graph_export(self.h, out["src"].ctypes.data, out["dst"].ctypes.data,
                                    out["w"].ctypes.data if self.has_weights else None,
                                    out["off"].ctypes.data, out["adj_rank"].ctypes.data,
                                    out["adj_nbr"].ctypes.data))
        return out

    def subsample(self, rate, seed):
        """GraphSampler::sample_graph (src/Graph.cpp:628-780)."""
        h = C.c_void_p()
        check(lib().zg_graph_subsample(self.h, rate, seed, C.byref(h)))
        return Graph(h)

    def filter(self, threshold, keep_ge=True, seed=1):
        """GraphFilter::generate_graph (src/GraphFilters.cpp:16-177) on `w >= thr` (or `<`)."""
        h = C.c_void_p()
        check(lib().zg_graph_filter(self.h, threshold, 1 if keep_ge else 0, seed, C.byref(h)))
        return Graph(h)

    def exact_count(self, what):
        """Exact non-induced count: 'triangle', 'k4', 'wedge', 'path3', 'star5', 'house',
        'path4' (5-vertex chains) or 'c4' (4-cycles)."""
        code = {"triangle": 0, "k4": 1, "wedge": 2, "path3": 3, "star5": 4, "house": 5, "path4": 6, "c4": 7}[what]
        r = C.c_uint64()
        check(lib().zg_exact_counts(self.h, code, C.byref(r)))
        return r.value

    def degree_moments(self):
        """(sum d^2, sum d^3) over all vertices"""
        a, b = C.c_double(), C.c_double()
        check(lib().zg_graph_degree_moments(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def close(self):
        if self.h:
            lib().zg_graph_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# plans and patterns
# ---------------------------------------------------------------------------
class Pattern:
    """Pattern(num_vertices, num_edges, edges) (include/SamplerGenerator.hpp:9-47)."""

    def __init__(self, num_vertices, edges):
        self.num_vertices = num_vertices
        self.edges = [(int(a), int(b)) for a, b in edges]

    @property
    def num_edges(self):
        return len(self.edges)

    def desc(self):
        d = _capi.PlanDesc()
        d.kind = PLAN_PATTERN
        if self.num_edges > _capi.MAX_PATTERN_EDGES:
            raise ValueError("pattern too large")
        d.pattern_vertices = self.num_vertices
        d.pattern_edges = self.num_edges
        for i, (a, b) in enumerate(self.edges):
            d.pattern_src[i] = a
            d.pattern_dst[i] = b
        return d

    def unique_relabelings(self):
        """(unique, total) relabelings after is_automorphic deduplication
        (src/SamplerGenerator.cpp:311-363); host only."""
        u, t = C.c_int(), C.c_int()
        check(lib().zg_pattern_unique(C.byref(self.desc()), C.byref(u), C.byref(t)))
        return u.value, t.value


TRIANGLE_PATTERN = Pattern(3, [(0, 1), (0, 2), (1, 2)])                              # Triangle2.cpp:18-24
FOUR_CLIQUE_PATTERN = Pattern(4, [(1, 0), (2, 0), (2, 1), (3, 0), (3, 1), (3, 2)])    # TestPattern2.cpp:111-122
FIVE_STAR_PATTERN = Pattern(5, [(0, 1), (0, 2), (0, 3), (0, 4)])                      # FiveStar.cpp:18-26
FIVE_HOUSE_PATTERN = Pattern(5, [(0, 1), (0, 2), (1, 2), (1, 3), (2, 4), (3, 4)])     # FiveHouse.cpp:18-27


class Plan:
    """A device plan (an estimate() body).  order: ZG_PLAN_CHAIN's path
    vertices; pattern: ZG_PLAN_PATTERN's pattern; completions: the split plans'
    completions per closed triangle (ZG_PLAN_K4_SPLIT, ZG_PLAN_TRIANGLE_K4_SPLIT)
    or per closed roof (ZG_PLAN_HOUSE_SPLIT), a power of two in [1, 4096]."""

    def __init__(self, kind, order=0, pattern=None, completions=0):
        d = _capi.PlanDesc()
        d.kind = kind
        d.order = order
        d.completions = completions
        if pattern is not None:
            if pattern.num_edges > _capi.MAX_PATTERN_EDGES:
                raise ValueError("pattern too large")
            d.pattern_vertices = pattern.num_vertices
            d.pattern_edges = pattern.num_edges
            for i, (a, b) in enumerate(pattern.edges):
                d.pattern_src[i] = a
                d.pattern_dst[i] = b
        self.kind = kind
        self.order = order
        self.pattern = pattern
        self.completions = completions
        self.h = C.c_void_p()
        check(lib().zg_plan_create(C.byref(d), C.byref(self.h)))
        self.channels = lib().zg_plan_channels(self.h)
        self.cond_calls = lib().zg_plan_cond_calls(self.h)

    def unique_patterns(self):
        u, t = C.c_int(), C.c_int()
        check(lib().zg_plan_unique_patterns(self.h, C.byref(u), C.byref(t)))
        return u.value, t.value

    def __del__(self):
        try:
            if self.h:
                lib().zg_plan_destroy(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass


def run(graph, plan, seed, first_id, count, block_size=0, blocks_per_cu=0, timed=False, order=0):
    """One launch of the estimator loop; returns the per-channel sums (and kernel
    ms).  order: 0 automatic, 1 estimator-id order, 2..16 first-edge order."""
    sums = (_capi.Sums * plan.channels)()
    launch = _capi.Launch(block_size, blocks_per_cu, order)
    ms = C.c_float(0)
    check(lib().zg_run(graph.h, plan.h, seed, first_id, count, C.byref(launch), sums,
                       C.byref(ms) if timed else None))
    out = [sums[i] for i in range(plan.channels)]
    return (out, ms.value) if timed else out


class Batch:
    """Many (graph, plan, seed, first_id, count) queries served by one
    replayable HIP graph: one persistent work-queue launch over the chunks of
    every query, whatever its plan.  chunk_estimators is the number of
    estimator ids one wavefront claims at a time (a multiple of 64).
    run() returns, per query, the list of per-channel sums -- equal to
    run(graph, plan, seed, first_id, count) for that query."""

    def __init__(self, queries, block_size=0, blocks_per_cu=0, chunk_estimators=0):
        self._keep = [(q[0], q[1]) for q in queries]   # graphs and plans outlive the batch
        self.nq = len(queries)
        arr = (_capi.Query * self.nq)()
        for i, (g, p, seed, first, count) in enumerate(queries):
            arr[i] = _capi.Query(g.h, p.h, seed, first, count)
        self.channels = [q[1].channels for q in queries]
        cfg = _capi.BatchConfig(block_size, blocks_per_cu, chunk_estimators)
        self.h = C.c_void_p()
        check(lib().zg_batch_create(arr, self.nq, C.byref(cfg), C.byref(self.h)))
        self.nrec = lib().zg_batch_records(self.h)

    def run(self, timed=False):
        out = (_capi.Sums * self.nrec)()
        ms = C.c_float(0)
        check(lib().zg_batch_run(self.h, out, C.byref(ms)))
        res, k = [], 0
        for ch in self.channels:
            res.append([out[k + c] for c in range(ch)])
            k += ch
        return (res, ms.value) if timed else res

    def set_ids(self, seeds=None, first_ids=None):
        """New seeds and/or first ids for later runs: 1-D sequences of one
        entry per query (the library reads exactly nq entries)."""
        s = np.ascontiguousarray(seeds, np.uint64) if seeds is not None else None
        f = np.ascontiguousarray(first_ids, np.uint64) if first_ids is not None else None
        for name, a in (("seeds", s), ("first_ids", f)):
            if a is not None and (a.ndim != 1 or a.shape[0] != self.nq):
                raise ValueError(f"{name} must be a 1-D array of {self.nq} entries, got shape {a.shape}")
        check(lib().zg_batch_set_ids(self.h, s.ctypes.data if s is not None else None,
                                     f.ctypes.data if f is not None else None))

    def autotune(self, trials=3):
        cfg = _capi.BatchConfig()
        ms = C.c_float(0)
        check(lib().zg_batch_autotune(self.h, trials, C.byref(cfg), C.byref(ms)))
        return {"block_size": cfg.block_size, "blocks_per_cu": cfg.blocks_per_cu,
                "chunk_estimators": cfg.chunk_estimators, "ms": ms.value}

    def config(self):
        cfg = _capi.BatchConfig()
        check(lib().zg_batch_config_get(self.h, C.byref(cfg)))
        return {"block_size": cfg.block_size, "blocks_per_cu": cfg.blocks_per_cu,
                "chunk_estimators": cfg.chunk_estimators}

    def close(self):
        if self.h:
            lib().zg_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_outcomes(graph, plan, seed, first_id, count):
    """Per-estimator records (count x channels) for parity checks."""
    out = np.zeros(count * plan.channels, OUTCOME_DTYPE)
    sums = (_capi.Sums * plan.channels)()
    check(lib().zg_run_outcomes(graph.h, plan.h, seed, first_id, count, out.ctypes.data, sums))
    return (out.reshape(count, plan.channels) if plan.channels > 1 else out), [sums[i] for i in
                                                                               range(plan.channels)]


def count_reads(graph, plan, seed, first_id, count):
    """Dependent global reads (128-byte lines) issued by these estimators."""
    r = C.c_uint64()
    check(lib().zg_count_reads(graph.h, plan.h, seed, first_id, count, C.byref(r)))
    return r.value


def random_gather_bandwidth(buffer_bytes=16 << 30, chains=1, loads=1 << 28, block_size=0, blocks_per_cu=0):
    """Achievable random-gather rate of this GPU in GB/s (one 128-byte line per
    random 8-byte read from a buffer far larger than the caches)."""
    gbs, ms = C.c_double(), C.c_double()
    check(lib().zg_random_gather_bw(buffer_bytes, chains, loads, block_size, blocks_per_cu, C.byref(gbs),
                                    C.byref(ms)))
    return gbs.value, ms.value


def random_line_bandwidth(buffer_bytes=16 << 30, loads_per_line=1, cooperative=False, lines=1 << 28,
                          block_size=0, blocks_per_cu=0):
    """Random 128-byte line rate in GB/s when each line is read with
    `loads_per_line` 16-byte loads by one lane, or (cooperative) by 8 lanes
    with one 16-byte load each in a single instruction."""
    gbs, ms = C.c_double(), C.c_double()
    check(lib().zg_random_line_bw(buffer_bytes, loads_per_line, int(bool(cooperative)), lines, block_size,
                                  blocks_per_cu, C.byref(gbs), C.byref(ms)))
    return gbs.value, ms.value


def random_region_bandwidth(region_bytes, xcd_matched=True, buffer_bytes=8 << 30, rounds_per_phase=256,
                            lines=1 << 30, block_size=0, blocks_per_cu=0):
    """Random 128-byte line rate in GB/s when each workgroup's reads stay in
    one region of `region_bytes` (changing every `rounds_per_phase` reads),
    the regions matched to the XCDs (each XCD's workgroups share one) or not."""
    gbs, ms = C.c_double(), C.c_double()
    check(lib().zg_random_region_bw(buffer_bytes, region_bytes, int(bool(xcd_matched)), rounds_per_phase, lines,
                                    block_size, blocks_per_cu, C.byref(gbs), C.byref(ms)))
    return gbs.value, ms.value


def random_chain_bandwidth(m, estimators=1 << 28, block_size=0, blocks_per_cu=0):
    """Line rate in GB/s of `estimators` triangle-estimator access chains (a
    random 32-byte record, an 8-byte entry, a 16-byte bucket header, each
    address taken from the previous read) over arrays sized for m edges."""
    gbs, ms = C.c_double(), C.c_double()
    check(lib().zg_random_chain_bw(m, estimators, block_size, blocks_per_cu, C.byref(gbs), C.byref(ms)))
    return gbs.value, ms.value


def clt_half_width(mean, sumsq, count, confidence=0.95):
    """CLT half-width of the mean at the given two-sided confidence."""
    if count < 2:
        return float("inf")
    var = max(sumsq / count - mean * mean, 0.0) * count / (count - 1)
    z = _z(confidence)
    return z * math.sqrt(var / count)


def _z(conf):
    # inverse normal CDF by bisection on erf (no scipy dependency)
    p = 0.5 + conf / 2
    lo, hi = 0.0, 10.0
    for _ in range(100):
        mid = (lo + hi) / 2
        if 0.5 * (1 + math.erf(mid / math.sqrt(2))) < p:
            lo = mid
        else:
            hi = mid
    return (lo + hi) / 2


# ---------------------------------------------------------------------------
# Estimation / ZGraphInstance
# ---------------------------------------------------------------------------
class Estimation:
    """Running estimator-weighted mean (src/ZGraphInstance.cpp:15-35), plus the
    exact integer sum and the sum of squares that the device reduction returns."""

    def __init__(self):
        self.reset()

    def reset(self):
        self.num_finished_estimators = 0
        self.current_estimation = 0.0
        self.exact_sum = 0        # sum of values * 2^log2_den
        self.log2_den = None
        self.sumsq = 0.0
        self.successes = 0

    def update(self, new_num_estimators, new_estimation):
        n0, n1 = self.num_finished_estimators, new_num_estimators
        self.current_estimation = (1.0 * n0 / (n0 + n1)) * self.current_estimation + \
                                  (1.0 * n1 / (n0 + n1)) * new_estimation
        self.num_finished_estimators += n1

    def get_num_finished_estimators(self):
        return self.num_finished_estimators

    def get_currnet_estimation(self):   # spelling kept from include/ZGraphInstance.hpp:27
        return self.current_estimation

    get_current_estimation = get_currnet_estimation

    def absorb(self, sums):
        if self.log2_den is None:
            self.log2_den = sums.log2_den
        self.exact_sum += sums.total()
        self.sumsq += sums.sumsq
        self.successes += sums.successes

    def exact_mean(self):
        if not self.num_finished_estimators:
            return 0.0
        return self.exact_sum / (2 ** self.log2_den) / self.num_finished_estimators

    def half_width(self, confidence=0.95):
        return clt_half_width(self.exact_mean(), self.sumsq, self.num_finished_estimators, confidence)

    def sums_view(self, channel=0):
        """The accumulated sums as a zg_sums-like record (count, successes,
        sumsq, mean)."""
        return _SumsView(self.num_finished_estimators, self.successes, self.sumsq, self.exact_mean())


class _SumsView:
    __slots__ = ("count", "successes", "sumsq", "mean")

    def __init__(self, count, successes, sumsq, mean):
        self.count, self.successes, self.sumsq, self.mean = count, successes, sumsq, mean


class ZGraphInstance:
    """Base estimator instance (include/ZGraphInstance.hpp:107-140).

    Estimator ids are global: calculate(N, k, init) runs ids
    [init.finished, N), so a top-up equals one run of N estimators."""

    plan_kind = None

    def __init__(self, graph, plan=None, seed=1):
        self.graph = graph
        self.plan = plan if plan is not None else Plan(self.plan_kind)
        self.seed = seed
        self.channel = 0

    def get_graph(self):
        return self.graph

    def set_graph(self, graph):
        self.graph = graph

    def calculate(self, num_estimators, num_vertices_in_pattern=None, init_estimation=None):
        est = init_estimation if init_estimation is not None else Estimation()
        done = est.get_num_finished_estimators()
        if done >= num_estimators:
            return est.get_currnet_estimation()
        new = num_estimators - done
        sums = run(self.graph, self.plan, self.seed, done, new)[self.channel]
        est.absorb(sums)
        est.update(new, sums.mean)   # W = 1: get_scaling_factor == 1 (src/ZGraphInstance.cpp:117-125)
        return est.get_currnet_estimation()

    def outcomes(self, first_id, count):
        return run_outcomes(self.graph, self.plan, self.seed, first_id, count)[0]

    def get_num_sample_edge_calls(self):
        return 1.0   # src/ZGraphInstance.cpp:452

    def get_num_conditional_sample_edge_calls(self):
        return self.plan.cond_calls


class TriangleCoutingInstance(ZGraphInstance):
    """applications/Triangle.cpp:42-74 (class name spelled as the reference)."""
    plan_kind = PLAN_TRIANGLE


TriangleCountingInstance = TriangleCoutingInstance


class ThreeChainCountingInstance(ZGraphInstance):
    """applications/ThreeChain.cpp:16-38"""
    plan_kind = PLAN_THREE_CHAIN


class _TwoChannelInstance(ZGraphInstance):
    """A plan whose every estimator id yields two patterns' outcomes."""

    def calculate_both(self, n0, n1):
        """Channel 0 over ids [0, n0) and channel 1 over ids [0, n1), from
        max(n0, n1) ids: two Estimations."""
        n = max(n0, n1)
        e0, e1 = Estimation(), Estimation()
        if n0 == n1:
            s = run(self.graph, self.plan, self.seed, 0, n)
            e0.absorb(s[0]); e0.update(n, s[0].mean)
            e1.absorb(s[1]); e1.update(n, s[1].mean)
        else:
            lo = min(n0, n1)
            s = run(self.graph, self.plan, self.seed, 0, lo)
            s2 = run(self.graph, self.plan, self.seed, lo, n - lo)
            for ch, e, cnt in ((0, e0, n0), (1, e1, n1)):
                e.absorb(s[ch]); e.update(lo, s[ch].mean)
                if cnt > lo:
                    e.absorb(s2[ch]); e.update(n - lo, s2[ch].mean)
        return e0, e1


class ThreeMotifInstance(_TwoChannelInstance):
    """applications/ThreeMotif.cpp: 3-chain (channel 0) and triangle (channel 1)
    from shared wedges, the device analogue of CachedSubPatterns."""
    plan_kind = PLAN_THREE_MOTIF


class TriangleFourCliqueInstance(_TwoChannelInstance):
    """Triangles (channel 0, TriangleCoutingInstance's estimator) and 4-cliques
    (channel 1, the three-shape plan) from one first edge, conditional sample
    and closing check per id: the CachedSubPatterns sharing of ThreeMotif
    (src/ZGraphInstance.cpp:596-833) applied to triangle -> 4-clique.  Each
    channel equals its own plan's run on the same ids."""
    plan_kind = PLAN_TRIANGLE_K4


class TriangleFourCliqueSplitInstance(_TwoChannelInstance):
    """Triangles (channel 0) and 4-cliques by the split plan (channel 1,
    FourCliqueSplitCountingInstance's estimator with `completions`) from one
    first stage per id; each channel equals its own plan's run on the same ids."""

    def __init__(self, graph, completions=256, seed=1):
        super().__init__(graph, Plan(PLAN_TRIANGLE_K4_SPLIT, completions=completions), seed)


class ChainMiningInstance(ZGraphInstance):
    """applications/ChainMining.cpp:18-85 (order = vertices in the path)."""

    def __init__(self, graph, order=4, seed=1):
        super().__init__(graph, Plan(PLAN_CHAIN, order=order), seed)
        self.order = order


class FourChainCountingInstance(ZGraphInstance):
    """applications/FourChain.cpp:18-125"""
    plan_kind = PLAN_FOUR_CHAIN


class FourCliqueCountingInstance(ZGraphInstance):
    """4-cliques by three spanning shapes + closing checks (unbiased; the
    reference's applications/FourClique.cpp:17-77 misses two of the shapes)."""
    plan_kind = PLAN_K4_CLOSE


class FourCliqueSplitCountingInstance(ZGraphInstance):
    """4-cliques, split: FourCliqueCountingInstance's first stage, then
    `completions` independent draws of its three shapes per closed triangle,
    averaged (same mean, lower variance per estimator)."""

    def __init__(self, graph, completions=64, seed=1):
        super().__init__(graph, Plan(PLAN_K4_SPLIT, completions=completions), seed)


class FiveHouseSplitCountingInstance(ZGraphInstance):
    """5-houses (the pattern of applications/FiveHouse.cpp:18-27), split: a
    base edge and its roof apex, closed by one pair lookup, then `completions`
    draws of the square's two other vertices per closed roof, averaged.  Its
    mean is the house count that SamplerGenerator(FIVE_HOUSE_PATTERN) estimates
    (each house once per choice of copies of its edges), with far less
    variance per estimator."""

    def __init__(self, graph, completions=256, seed=1):
        super().__init__(graph, Plan(PLAN_HOUSE_SPLIT, completions=completions), seed)


class SamplerGenerator(ZGraphInstance):
    """SamplerGenerator<Empty>(graph, pattern) (src/SamplerGenerator.cpp:311-476)."""

    def __init__(self, graph, pattern, seed=1):
        super().__init__(graph, Plan(PLAN_PATTERN, pattern=pattern), seed)
        self.pattern = pattern

    def get_num_conditional_sample_edge_calls(self):
        return self.pattern.num_edges - 1   # src/SamplerGenerator.cpp:471-474


class AllPredicateZGraphInstance(ZGraphInstance):
    """AllPredicateZGraphInstance (src/ZGraphInstance.cpp:490-525): runs the plan
    on the sub-graph of edges whose weight satisfies the predicate."""

    def __init__(self, graph, threshold, keep_ge=True, plan_kind=PLAN_TRIANGLE, seed=1, filter_seed=1):
        self.original_graph = graph
        filtered = graph.filter(threshold, keep_ge, filter_seed)
        super().__init__(filtered, Plan(plan_kind), seed)

    def get_graph(self):
        return self.original_graph


# ---------------------------------------------------------------------------
# error-latency profiles
# ---------------------------------------------------------------------------
class ErrorProfile:
    """ErrorProfile<E> (src/ErrorLatencyProfile.cpp:49-207) on the device."""

    def __init__(self, pilot_scale=0):
        self.pilot_scale = pilot_scale
        self.last = None

    @staticmethod
    def calculate_bound(epsilon, delta, num_sample_edge_calls, num_conditional_sample_edge_calls,
                        estimation, max_degree, num_edges):
        return lib().zg_elp_bound(epsilon, delta, num_sample_edge_calls,
                                  num_conditional_sample_edge_calls, estimation, max_degree, num_edges)

    def get_num_estimators(self, instance, num_vertices_in_pattern, required_error_rate,
                           required_confidence):
        r = _capi.ElpResult()
        check(lib().zg_elp_regression(instance.graph.h, instance.plan.h, required_error_rate,
                                      required_confidence, instance.seed, self.pilot_scale, C.byref(r)))
        self.last = r
        return int(r.num_estimators) if math.isfinite(r.num_estimators) else 0

    def run(self, instance, num_vertices_in_pattern, required_error_rate, required_confidence,
            init_estimation=None):
        n = self.get_num_estimators(instance, num_vertices_in_pattern, required_error_rate,
                                    required_confidence)
        return instance.calculate(n, num_vertices_in_pattern, init_estimation)


class TimeProfile:
    """TimeProfile<E>::run (src/ErrorLatencyProfile.cpp:24-47): doubling batches
    while used + 2 * last <= limit, then one batch sized to the rest.  Batches
    take disjoint estimator-id ranges so they stay independent."""

    def __init__(self, basic_estimators=1 << 16):
        self.basic = basic_estimators
        self.batches = []          # (first id, count, whole ms) of the last run's batches
        self.remainder = False     # whether the last run ended with a batch sized to the time left
        self.estimation = None     # the last run's Estimation (running mean + exact sums)

    def run(self, instance, num_vertices_in_pattern, time_limit_ms):
        used = 0
        cur = 0
        nxt = self.basic
        first = 0
        est = Estimation()
        self.batches, self.remainder = [], False
        while used + cur * 2 <= time_limit_ms:
            t = time.perf_counter()
            s = run(instance.graph, instance.plan, instance.seed, first, nxt)[instance.channel]
            cur = int((time.perf_counter() - t) * 1000)
            used += cur
            est.absorb(s)
            est.update(nxt, s.mean)
            self.batches.append((first, nxt, cur))
            first += nxt
            nxt <<= 1
        if cur > 0 and time_limit_ms > used:
            rest = nxt * (time_limit_ms - used) // (cur * 2)
            if rest > 0:
                s = run(instance.graph, instance.plan, instance.seed, first, rest)[instance.channel]
                est.absorb(s)
                est.update(rest, s.mean)
                self.batches.append((first, rest, None))
                self.remainder = True
        self.estimation = est
        return est.get_currnet_estimation()
