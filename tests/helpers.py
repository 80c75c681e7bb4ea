"""Shared fixtures for the tests: golden graphs, seeded small graphs, and the
mapping between oracle plans and device plans."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
_graphs = None


def golden_graph(name):
    global _graphs
    if _graphs is None:
        _graphs = np.load(os.path.join(GOLDEN, "graphs.npz"))
    n = int(_graphs[name + "/n"])
    return n, _graphs[name + "/src"], _graphs[name + "/dst"]


def golden_weights(name):
    golden_graph(name)
    return _graphs[name + "/w"]


def random_graph(n, p, seed, flip=True):
    """G(n, p) edge list with random orientation (simple graph)."""
    rng = np.random.default_rng(seed)
    iu = np.triu_indices(n, 1)
    keep = rng.random(len(iu[0])) < p
    s, d = iu[0][keep].astype(np.uint32), iu[1][keep].astype(np.uint32)
    if flip:
        f = rng.random(len(s)) < 0.5
        s, d = np.where(f, d, s), np.where(f, s, d)
    perm = rng.permutation(len(s))
    return n, s[perm].copy(), d[perm].copy()


def complete_graph(n):
    iu = np.triu_indices(n, 1)
    return n, iu[0].astype(np.uint32), iu[1].astype(np.uint32)


def power_law_graph(n, m, beta, seed):
    """Small Chung-Lu-like multigraph-free edge list for CPU tests (numpy)."""
    rng = np.random.default_rng(seed)
    w = (np.arange(n) + 1.0) ** (-1.0 / (beta - 1.0))
    w /= w.sum()
    pairs = set()
    while len(pairs) < m:
        u = rng.choice(n, size=2 * m, p=w)
        v = rng.choice(n, size=2 * m, p=w)
        for a, b in zip(u, v):
            if a != b:
                pairs.add((min(a, b), max(a, b)))
                if len(pairs) >= m:
                    break
    e = np.array(sorted(pairs), dtype=np.uint32)
    perm = rng.permutation(len(e))
    return n, e[perm, 0].copy(), e[perm, 1].copy()


def hub_graph(leaves, extra, seed):
    """Vertex 0 joined to every leaf 1..leaves, plus `extra` random leaf-leaf
    edges: one very long list (the long-list search paths) and triangles that
    all pass through the hub."""
    rng = np.random.default_rng(seed)
    pairs = {(0, i) for i in range(1, leaves + 1)}
    while len(pairs) < leaves + extra:
        a, b = (int(x) for x in rng.integers(1, leaves + 1, size=2))
        if a != b:
            pairs.add((min(a, b), max(a, b)))
    e = np.array(sorted(pairs), dtype=np.uint32)
    e = e[rng.permutation(len(e))]
    return leaves + 1, e[:, 0].copy(), e[:, 1].copy()


def bipartite_graph(a, b, p, seed):
    """Random bipartite graph between {0..a-1} and {a..a+b-1}: no odd cycles."""
    rng = np.random.default_rng(seed)
    u, v = np.meshgrid(np.arange(a), np.arange(a, a + b), indexing="ij")
    keep = rng.random(u.shape) < p
    s, d = u[keep].astype(np.uint32), v[keep].astype(np.uint32)
    perm = rng.permutation(len(s))
    return a + b, s[perm].copy(), d[perm].copy()


def sparse_ids_graph(n, m, seed):
    """m random edges among a few hundred vertices of an id space of size n
    (most vertices isolated)."""
    rng = np.random.default_rng(seed)
    ids = rng.choice(n, size=300, replace=False).astype(np.uint32)
    pairs = set()
    while len(pairs) < m:
        a, b = rng.choice(ids, size=2, replace=False)
        pairs.add((min(a, b), max(a, b)))
    e = np.array(sorted(pairs), dtype=np.uint32)
    e = e[rng.permutation(len(e))]
    return n, e[:, 0].copy(), e[:, 1].copy()


def multigraph(n, p, seed, max_copies=12, heavy_copies=40):
    """G(n, p) with every pair repeated 1..max_copies times (each copy randomly
    oriented) and one pair repeated heavy_copies times: parallel edges, which
    the reference's loader accepts, beyond what one hash bucket holds."""
    rng = np.random.default_rng(seed)
    _, s, d = random_graph(n, p, seed)
    reps = rng.integers(1, max_copies + 1, size=len(s))
    reps[0] = heavy_copies
    s, d = np.repeat(s, reps), np.repeat(d, reps)
    f = rng.random(len(s)) < 0.5
    s, d = np.where(f, d, s), np.where(f, s, d)
    perm = rng.permutation(len(s))
    return n, s[perm].astype(np.uint32), d[perm].astype(np.uint32)


# (name, oracle plan args, device plan args)
PLANS = [
    ("triangle", dict(kind=1), dict(kind=1)),
    ("three_chain", dict(kind=2), dict(kind=2)),
    ("three_motif", dict(kind=3), dict(kind=3)),
    ("chain4", dict(kind=4, order=4), dict(kind=4, order=4)),
    ("chain5", dict(kind=4, order=5), dict(kind=4, order=5)),
    ("four_chain", dict(kind=5), dict(kind=5)),
    ("k4_close", dict(kind=7), dict(kind=7)),
    ("k4_split", dict(kind=9, completions=64), dict(kind=9, completions=64)),
    ("house_split", dict(kind=11, completions=16), dict(kind=11, completions=16)),
    ("pattern_triangle", dict(kind=6, pattern="TRIANGLE"), dict(kind=6, pattern="TRIANGLE")),
    ("pattern_k4", dict(kind=6, pattern="K4"), dict(kind=6, pattern="K4")),
    ("pattern_5star", dict(kind=6, pattern="FIVE_STAR"), dict(kind=6, pattern="FIVE_STAR")),
    ("pattern_house", dict(kind=6, pattern="FIVE_HOUSE"), dict(kind=6, pattern="FIVE_HOUSE")),
]


def oracle_plan(O, kind, order=0, pattern=None, completions=0):
    # the oracle's split plans take their completions in `order`
    return O.make_plan(kind, order=completions if kind in (O.PLAN_K4_SPLIT, O.PLAN_HOUSE_SPLIT) else order,
                       pattern=getattr(O, pattern) if pattern else None)


def device_plan(Z, kind, order=0, pattern=None, completions=0):
    pats = {"TRIANGLE": Z.TRIANGLE_PATTERN, "K4": Z.FOUR_CLIQUE_PATTERN, "FIVE_STAR": Z.FIVE_STAR_PATTERN,
            "FIVE_HOUSE": Z.FIVE_HOUSE_PATTERN}
    return Z.Plan(kind, order=order, pattern=pats[pattern] if pattern else None, completions=completions)
