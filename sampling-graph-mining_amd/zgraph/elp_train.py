"""Training data and training of the learned error-latency profile (GPU).

    python -m zgraph.elp_train --graphs 256 --pilots 64 --out DIR
    python -m zgraph.elp_train --first-graph 0 --num-graphs 128 --collect-only --out DIR/a
    python -m zgraph.elp_train --merge DIR/a/elp_data.npz DIR/b/elp_data.npz --out DIR

1. Seeded graphs from the device generators (Erdos-Renyi, Chung-Lu with
   degree exponents 2.05..3, R-MAT), 2e4..1e9 edges, built in HBM; graph i is
   drawn from its own seeded stream, so index ranges can be collected
   separately and merged.  Every 16th graph is large: R-MAT scale 24..26
   (edge factor 8..16, up to 1.07e9 edges) or Chung-Lu with 1.5e8..6e8 edges.
2. For every (graph, pattern) with pattern in {triangle, 4-clique, 3-chain,
   4-chain}: the exact count T from the device exact counters and a large run
   (2^26 estimators, extended up to 2^30 until 4096 succeed) whose sum of
   squares gives the estimator's CV, CV_true = sqrt(E[X^2] / T^2 - 1).
3. `pilots` pilot runs on disjoint id ranges, each sized as the learned
   profile sizes its pilot (2^14..2^18 estimators, widened x8 until 32
   successes, at most 2^28), give one tuple each: the 16 features of zgraph.elp.feature_vector and the target
   r = log10(CV_true / CV_pilot).
4. The device MLP is trained on the tuples of 70 % of the graphs (Adam,
   MSE), the parameters kept from the training segment with the lowest error
   on 15 % of the graphs, and scored on the last 15 %.  The weights, input normalisation and the hold-out
   errors are written to DIR/elp_weights.npz, the tuples to DIR/elp_data.npz.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import (PLAN_FOUR_CHAIN, PLAN_K4_CLOSE, PLAN_THREE_CHAIN, PLAN_TRIANGLE, Graph, Plan, chung_lu, erdos_renyi,
               rmat, run)
from .elp import (IN_DIM, MAX_PILOT, MIN_PILOT_SUCCESSES, export_binary, feature_vector, graph_features,
                  pilot_stats)
from .mlp import MLP

PATTERNS = [("triangle", PLAN_TRIANGLE, 0), ("4-clique", PLAN_K4_CLOSE, 1), ("3-chain", PLAN_THREE_CHAIN, 2),
            ("4-chain", PLAN_FOUR_CHAIN, 3)]


def log(*a):
    print("[elp_train]", *a, file=sys.stderr, flush=True)


def random_graph(rng, i):
    """One seeded synthetic graph description and its device edges."""
    fam = ["er", "cl", "cl", "rmat"][i % 4]
    seed = 1000 + i
    if i % 16 == 11:   # the large graphs: R-MAT 24..26 or Chung-Lu up to 6e8 edges
        if rng.random() < 0.5:
            scale = int(rng.integers(24, 27))
            ef = float(rng.uniform(8, 16))
            m = int(ef * (1 << scale))
            a = float(rng.uniform(0.5, 0.6))
            b = c = float((1 - a) * rng.uniform(0.35, 0.45))
            desc = {"family": "rmat", "scale": scale, "m": m, "a": a, "b": b, "c": c, "seed": seed}
            return desc, rmat(scale, m, a, b, c, seed)
        m = int(10 ** rng.uniform(8.2, 8.78))
        n = int(2 * m / 10 ** rng.uniform(1.0, 1.6))
        beta = float(rng.uniform(2.05, 2.8))
        dmax = float(10 ** rng.uniform(3.5, 4.6))
        desc = {"family": "chung_lu", "n": n, "m": m, "beta": beta, "max_degree": dmax, "seed": seed}
        return desc, chung_lu(n, m, beta, dmax, seed)
    if fam == "rmat":
        scale = int(rng.integers(10, 21))
        ef = float(10 ** rng.uniform(0.5, 1.25))
        m = int(ef * (1 << scale))
        a = float(rng.uniform(0.45, 0.6))
        b = c = float((1 - a) * rng.uniform(0.3, 0.45))
        desc = {"family": "rmat", "scale": scale, "m": m, "a": a, "b": b, "c": c, "seed": seed}
        return desc, rmat(scale, m, a, b, c, seed)
    m = int(10 ** rng.uniform(4.3, 7.0))
    if i % 37 == 5:
        m = int(10 ** rng.uniform(7.0, 7.5))
    if i % 16 == 6:   # a few graphs at the scale of the 69 M-edge benchmark graph and above
        m = int(10 ** rng.uniform(7.5, 8.2))
    avg = 10 ** rng.uniform(0.5, 1.8)
    n = max(int(2 * m / avg), 64)
    while n * (n - 1) / 2 < 3 * m:
        n *= 2
    if fam == "er":
        desc = {"family": "er", "n": n, "m": m, "seed": seed}
        return desc, erdos_renyi(n, m, seed)
    beta = float(rng.uniform(2.05, 3.0))
    dmax = float(min(10 ** rng.uniform(1.5, 4.4), n / 4))
    desc = {"family": "chung_lu", "n": n, "m": m, "beta": beta, "max_degree": dmax, "seed": seed}
    return desc, chung_lu(n, m, beta, dmax, seed)


def collect(first_graph, num_graphs, pilots, seed, big=1 << 26):
    plans = {k: Plan(k) for _, k, _ in PATTERNS}
    X, Y, G, P, NP = [], [], [], [], []
    graphs = []
    t0 = time.perf_counter()
    for i in range(first_graph, first_graph + num_graphs):
        rng = np.random.default_rng([seed, i])
        try:
            desc, edges = random_graph(rng, i)
        except RuntimeError as e:   # e.g. a generator that could not find m distinct pairs
            log(f"graph {i}: skipped ({e})")
            continue
        try:
            g = Graph.from_device(edges, seed=desc["seed"])
        finally:
            edges.free()
        gf = graph_features(g)
        desc.update({"vertices": g.num_vertices, "edges": g.num_local_edges, "max_degree_built": g.max_degree})
        truths = {"triangle": g.exact_count("triangle"), "4-clique": g.exact_count("k4"),
                  "3-chain": g.exact_count("wedge"), "4-chain": g.exact_count("path3")}
        desc["truth"] = truths
        desc["cv"] = {}
        for name, kind, slot in PATTERNS:
            T = truths[name]
            if T == 0:
                continue
            # the estimator's second moment from `big` estimators, extended
            # (x4, at most 2^30) until 4096 of them succeed
            n_big, sumsq, succ = 0, 0.0, 0
            step = big
            while True:
                s = run(g, plans[kind], 7, n_big, step)[0]
                n_big += step
                sumsq += s.sumsq
                succ += s.successes
                if succ >= 4096 or n_big >= (1 << 30):
                    break
                step = min(3 * n_big, (1 << 30) - n_big)
            ex2 = sumsq / n_big
            cv2 = ex2 / (float(T) ** 2) - 1.0
            if cv2 <= 0:
                continue
            cv_true = math.sqrt(cv2)
            desc["cv"][name] = cv_true
            for j in range(pilots):
                # the pilot procedure of LearnedErrorProfile: start at 2^14..2^18
                # estimators, widen x8 until 32 successes (at most 2^28)
                n_p = 1 << int(rng.integers(14, 19))
                first = (j + 1) << 32
                while True:
                    ps = run(g, plans[kind], 7, first, n_p)[0]
                    if ps.successes >= MIN_PILOT_SUCCESSES or n_p >= MAX_PILOT:
                        break
                    n_p = min(n_p * 8, MAX_PILOT)
                if ps.successes < MIN_PILOT_SUCCESSES:
                    continue
                mean, cv, succ = pilot_stats(ps)
                if not math.isfinite(cv) or cv <= 0:
                    continue
                X.append(feature_vector(gf, kind, n_p, cv, succ))
                Y.append(math.log10(cv_true) - math.log10(cv))
                G.append(i)
                P.append(slot)
                NP.append(n_p)
        desc["index"] = i
        graphs.append(desc)
        g.close()
        log(f"graph {i}: {desc['family']} n={desc['vertices']} m={desc['edges']} dmax={desc['max_degree_built']} "
            f"tuples={len(Y)} t={time.perf_counter() - t0:.1f}s")
    return (np.array(X, np.float32), np.array(Y, np.float32), np.array(G, np.int32), np.array(P, np.int32),
            np.array(NP, np.int64), graphs)


def train(X, Y, G, steps=20000, segment=500, batch=256, lr=1e-3, seed=3, init_seed=None):
    """Adam on the tuples of 70 % of the graphs, targets scaled to unit
    variance; the parameters after the segment with the lowest validation
    error (15 % of the graphs) are kept; the remaining 15 % of the graphs are
    the test set.  `seed` fixes the split, `init_seed` (default: seed) the
    initial weights and minibatches."""
    init_seed = seed if init_seed is None else init_seed
    graphs = np.unique(G)
    rng = np.random.default_rng(seed)
    rng.shuffle(graphs)
    k = max(1, len(graphs) * 15 // 100)
    te = np.isin(G, graphs[:k])
    va = np.isin(G, graphs[k:2 * k])
    tr = ~(te | va)
    mu = X[tr].mean(0)
    sd = X[tr].std(0)
    sd[sd < 1e-6] = 1.0
    Xn = ((X - mu) / sd).astype(np.float32)
    y_scale = float(Y[tr].std()) or 1.0
    Ys = (Y / y_scale).astype(np.float32)
    rmse = lambda a, b: float(np.sqrt(np.mean((a - b) ** 2))) if len(b) else float("nan")
    net = MLP(init_seed)
    t = time.perf_counter()
    best = (float("inf"), None, 0)
    curve = []
    for s in range(max(1, steps // segment)):
        net.train(Xn[tr], Ys[tr], segment, batch, lr, init_seed * 1000 + s)
        v = rmse(net.forward(Xn[va]) * y_scale, Y[va])
        curve.append(v)
        if v < best[0]:
            best = (v, net.params(), (s + 1) * segment)
    dt = time.perf_counter() - t
    net.set_params(best[1])
    pred = net.forward(Xn) * y_scale
    # per pattern slot, keep the network's correction only where it beats
    # no correction on the validation graphs
    use_model = []
    for slot in range(4):
        sel = va & (X[:, 6 + slot] > 0.5)
        use_model.append(bool(sel.any() and rmse(pred[sel], Y[sel]) < rmse(0 * Y[sel], Y[sel])))
    # the deployed correction's error on the validation graphs, per pattern slot (log10 CV):
    # the learned profile aims its first run this far below the predicted budget
    val_rmse_slot = []
    for slot in range(4):
        sel = va & (X[:, 6 + slot] > 0.5)
        used = pred[sel] if use_model[slot] else 0 * Y[sel]
        val_rmse_slot.append(rmse(used, Y[sel]) if sel.any() else 0.0)
    stats = {"train_rmse": rmse(pred[tr], Y[tr]), "validation_rmse": rmse(pred[va], Y[va]),
             "test_rmse": rmse(pred[te], Y[te]), "test_rmse_no_correction": rmse(0 * Y[te], Y[te]),
             "test_rmse_mean_correction": rmse(np.full(te.sum(), Y[tr].mean()), Y[te]),
             "test_rmse_by_pattern": {}, "chosen_steps": best[2], "validation_curve": curve,
             "use_model": use_model, "y_scale": y_scale, "validation_rmse_by_slot": val_rmse_slot,
             "train_seconds": dt, "steps": steps, "batch": batch, "lr": lr, "init_seed": init_seed,
             "train_tuples": int(tr.sum()),
             "validation_tuples": int(va.sum()), "test_tuples": int(te.sum()),
             "test_graphs": sorted(int(g) for g in graphs[:k])}
    return net.params(), mu.astype(np.float32), sd.astype(np.float32), stats, pred, te


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--graphs", type=int, default=256, help="graphs 0..graphs-1 (without --num-graphs)")
    ap.add_argument("--first-graph", type=int, default=0)
    ap.add_argument("--num-graphs", type=int, default=0, help="collect graphs first-graph .. +num-graphs-1")
    ap.add_argument("--collect-only", action="store_true", help="write the tuples, do not train")
    ap.add_argument("--pilots", type=int, default=64)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--out", required=True)
    ap.add_argument("--data", default=None, help="train on this saved dataset instead of collecting")
    ap.add_argument("--merge", nargs="+", default=None, help="train on the union of these saved datasets")
    ap.add_argument("--search", action="store_true",
                    help="train over learning rate x batch x initial seed and keep the lowest validation error")
    a = ap.parse_args()
    os.makedirs(a.out, exist_ok=True)
    if a.data or a.merge:
        parts = []
        for path in ([a.data] if a.data else a.merge):
            with np.load(path, allow_pickle=False) as z:
                parts.append((z["X"], z["Y"], z["G"], z["P"], z["NP"], json.loads(bytes(z["graphs"]).decode())))
        X, Y, G, P, NP = (np.concatenate([p[k] for p in parts]) for k in range(5))
        graphs = [d for p in parts for d in p[5]]
        if len({d.get("index", -1) for d in graphs}) < len(graphs) and a.merge:
            raise SystemExit("merged datasets hold the same graph twice")
    else:
        n = a.num_graphs or a.graphs
        X, Y, G, P, NP, graphs = collect(a.first_graph, n, a.pilots, a.seed)
        np.savez_compressed(os.path.join(a.out, "elp_data.npz"), X=X, Y=Y, G=G, P=P, NP=NP,
                            graphs=np.frombuffer(json.dumps(graphs).encode(), np.uint8))
        if a.collect_only:
            log(f"{len(Y)} tuples from {len(graphs)} graphs written")
            return
    log(f"{len(Y)} tuples from {len(graphs)} graphs")
    if a.search:
        # model selection on the validation graphs only; the test graphs stay unseen
        grid, best = [], None
        for lr in (3e-4, 1e-3, 3e-3):
            for batch in (256, 1024):
                for init in (3, 4, 5):
                    r = train(X, Y, G, steps=a.steps, lr=lr, batch=batch, init_seed=init)
                    v = r[3]["validation_rmse"]
                    grid.append({"lr": lr, "batch": batch, "init_seed": init, "validation_rmse": v})
                    log(f"search lr={lr:g} batch={batch} init={init}: validation RMSE {v:.5f}")
                    if best is None or v < best[3]["validation_rmse"]:
                        best = r
        params, mu, sd, stats, pred, te = best
        stats["search"] = grid
    else:
        params, mu, sd, stats, pred, te = train(X, Y, G, steps=a.steps)
    names = ["triangle", "4-clique", "3-chain", "4-chain"]
    for slot in range(4):
        sel = te & (P == slot)
        if sel.any():
            used = pred[sel] if stats["use_model"][slot] else 0 * Y[sel]
            stats["test_rmse_by_pattern"][names[slot]] = {
                "tuples": int(sel.sum()), "model": float(np.sqrt(np.mean((pred[sel] - Y[sel]) ** 2))),
                "no_correction": float(np.sqrt(np.mean(Y[sel] ** 2))),
                "as_deployed": float(np.sqrt(np.mean((used - Y[sel]) ** 2)))}
    dep = np.where(np.array(stats["use_model"])[P], pred, 0)
    stats["test_rmse_as_deployed"] = float(np.sqrt(np.mean((dep[te] - Y[te]) ** 2)))
    stats.update({"tuples": int(len(Y)), "graphs": len(graphs), "in_dim": IN_DIM,
                  "feature_min": X.min(0).tolist(), "feature_max": X.max(0).tolist()})
    log(json.dumps({k: v for k, v in stats.items() if not isinstance(v, list)}))
    np.savez(os.path.join(a.out, "elp_weights.npz"), params=params, mu=mu, sd=sd,
             meta=np.frombuffer(json.dumps(stats).encode(), np.uint8))
    export_binary(os.path.join(a.out, "elp_weights.npz"), os.path.join(a.out, "elp_weights.bin"))
    print(json.dumps({k: v for k, v in stats.items() if not isinstance(v, list)}), flush=True)


if __name__ == "__main__":
    main()
