"""Learned error-latency profile, CLT judge and adaptive top-up.

The reference sizes a query with ErrorProfile::get_num_estimators
(src/ErrorLatencyProfile.cpp:71-207): pilots on a 10 % edge sample, a
through-origin regression of pilot error on the Hoeffding-style bound, and the
bound scaled by K^2.  The learned profile replaces that sizing:

  1. graph features from the device CSR (n, m, degree moments, max degree);
  2. a pilot: the first `pilot` estimators of the query itself (ids
     [0, pilot)), whose exact sums give the pilot mean, coefficient of
     variation (CV) and success fraction;
  3. the MLP (16-64-64-64-1, BF16 MFMA kernels, fp32 master weights) predicts
     r = log10(CV_true / CV_pilot), the correction from the pilot's CV to the
     CV of the estimator's full distribution;
  4. N = (z * CV_pilot * 10^(r - s) / eps)^2 estimators (CLT sizing), s the
     correction's RMSE on validation graphs: the first run aims one typical
     error low, since a short run costs one top-up and a long one its excess;
  5. the judge (the rule of sampling/IEStop.cpp:47-56: stop once
     z * sd / (sqrt(t) * mean) <= eps) checks the exact running sums and tops
     the same id sequence up until it holds.

Plans the network was not trained on are sized from the pilot's own CV (r =
0) and judged the same way.  The regression profile stays as the fallback: for
pilots with too few successes, or without a weights file.
Because estimator ids are global and every run continues the id sequence,
the final answer equals one run of N_final estimators.
"""
import json
import math
import os
import time

import numpy as np

from . import (PLAN_FOUR_CHAIN, PLAN_HOUSE_SPLIT, PLAN_K4_CLOSE, PLAN_K4_SPLIT, PLAN_THREE_CHAIN, PLAN_THREE_MOTIF,
               PLAN_TRIANGLE, PLAN_TRIANGLE_K4, PLAN_TRIANGLE_K4_SPLIT, ErrorProfile, Estimation, ZGraphError, _z, run)

IN_DIM = 16
PATTERN_SLOTS = {PLAN_TRIANGLE: 0, PLAN_K4_CLOSE: 1, PLAN_THREE_CHAIN: 2, PLAN_FOUR_CHAIN: 3}
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "elp_weights.npz")
MIN_PILOT_SUCCESSES = 32
# the single plan whose estimator each channel of a two-channel plan carries
CHANNEL_KINDS = {PLAN_THREE_MOTIF: (PLAN_THREE_CHAIN, PLAN_TRIANGLE),
                 PLAN_TRIANGLE_K4: (PLAN_TRIANGLE, PLAN_K4_CLOSE),
                 PLAN_TRIANGLE_K4_SPLIT: (PLAN_TRIANGLE, PLAN_K4_SPLIT)}
MAX_PILOT = 1 << 28


def graph_features(graph):
    """(n, m, sum d^2, sum d^3, max degree) of the graph, cached on it."""
    f = getattr(graph, "_elp_features", None)
    if f is None:
        s2, s3 = graph.degree_moments()
        f = (float(graph.num_vertices), float(graph.num_local_edges), s2, s3, float(graph.max_degree))
        graph._elp_features = f
    return f


def pilot_stats(sums):
    """(mean, cv, success fraction) of one channel's sums."""
    n = sums.count
    mean = sums.mean
    if n == 0 or mean <= 0 or sums.successes == 0:
        return mean, float("inf"), sums.successes / max(n, 1)
    var = max(sums.sumsq / n - mean * mean, 0.0)
    return mean, math.sqrt(var) / mean, sums.successes / n


def feature_vector(gf, plan_kind, pilot_n, pilot_cv, pilot_succ):
    """The 16 network inputs (log10 scales; one-hot pattern slot)."""
    n, m, s2, s3, dmax = gf
    x = np.zeros(IN_DIM, np.float32)
    lg = lambda v: math.log10(max(v, 1e-30))
    x[0] = lg(n)
    x[1] = lg(m)
    x[2] = lg(2 * m / max(n, 1))
    x[3] = lg(s2 / max(2 * m, 1))
    x[4] = lg(s3 / max(s2, 1))
    x[5] = lg(dmax)
    x[6 + PATTERN_SLOTS[plan_kind]] = 1.0
    x[10] = lg(pilot_cv)
    x[11] = lg(pilot_succ)
    x[12] = lg(pilot_n)
    x[13] = lg(pilot_cv * pilot_cv * pilot_succ)   # CV^2 of the successes' part
    x[14] = lg(s2 * s2 / max(m * s3, 1))            # tail-heaviness of the degree sequence
    return x


BIN_MAGIC = b"ZGELP002"


def export_binary(npz_path, bin_path):
    """Write the weights as the flat little-endian file the C++ host API reads:
    magic "ZGELP002", u32 nparam, u32 in_dim, f32 mu[in_dim], f32 sd[in_dim],
    u8 use_model[4], f32 y_scale, f32 validation RMSE per pattern slot [4],
    f32 params[nparam]."""
    with np.load(npz_path, allow_pickle=False) as z:
        params = z["params"].astype("<f4")
        mu = z["mu"].astype("<f4")
        sd = z["sd"].astype("<f4")
        meta = json.loads(bytes(z["meta"]).decode()) if "meta" in z else {}
    use = np.array([1 if u else 0 for u in meta.get("use_model", [True] * 4)], np.uint8)
    with open(bin_path, "wb") as f:
        f.write(BIN_MAGIC)
        f.write(np.array([len(params), len(mu)], "<u4").tobytes())
        f.write(mu.tobytes())
        f.write(sd.tobytes())
        f.write(use.tobytes())
        f.write(np.array([meta.get("y_scale", 1.0)], "<f4").tobytes())
        f.write(np.array(meta.get("validation_rmse_by_slot", [0.0] * 4), "<f4").tobytes())
        f.write(params.tobytes())


class LearnedModel:
    """Weights + input normalisation, evaluated by the device MLP."""

    def __init__(self, path=DEFAULT_WEIGHTS):
        from .mlp import MLP
        with np.load(path, allow_pickle=False) as z:
            self.params = z["params"].astype(np.float32)
            self.mu = z["mu"].astype(np.float32)
            self.sd = z["sd"].astype(np.float32)
            self.meta = json.loads(bytes(z["meta"]).decode()) if "meta" in z else {}
        # pattern slots where the network beat "no correction" on validation graphs
        self.use_model = list(self.meta.get("use_model", [True] * len(PATTERN_SLOTS)))
        # the deployed correction's RMSE on the validation graphs (log10 CV), per slot
        # (rounded to fp32, as the C++ API reads them from the binary weights)
        self.val_rmse = [float(np.float32(v)) for v in
                         self.meta.get("validation_rmse_by_slot", [0.0] * len(PATTERN_SLOTS))]
        self.y_scale = float(self.meta.get("y_scale", 1.0))   # the network predicts r / y_scale
        self.net = MLP()
        self.net.set_params(self.params)

    def predict(self, X):
        X = np.atleast_2d(np.asarray(X, np.float32))
        r = self.net.forward(((X - self.mu) / self.sd).astype(np.float32)) * np.float32(self.y_scale)
        slot = np.argmax(X[:, 6:10], axis=1)
        return np.where(np.array(self.use_model)[slot], r, 0.0).astype(np.float32)


class LearnedErrorProfile:
    """Drop-in for ErrorProfile with learned sizing and a CLT judge.
    With correction=False the network is left out (r = 0): the pilot's own CV
    sizes the query -- the baseline that shows what the network adds.

    get_num_estimators(instance, k, eps, conf) sizes the query; run(...)
    sizes, runs, judges and tops up, returning the estimate.  `last` holds the
    record of the latest call (sizes, rounds, times, which sizing was used).
    """

    def __init__(self, weights=DEFAULT_WEIGHTS, pilot=1 << 16, max_pilot=MAX_PILOT, max_rounds=8,
                 fallback=None, min_sample=1 << 12, correction=True, aim_sigmas=1.0):
        self.model = None
        self.correction = correction   # False: size from the pilot's CV alone (r = 0), no network
        # the first run aims `aim_sigmas` validation RMSEs below the predicted CV: a run that
        # falls short costs one top-up of the id sequence, one that overshoots costs its excess
        self.aim_sigmas = aim_sigmas
        if correction and weights and os.path.exists(weights):
            self.model = LearnedModel(weights)
        self.pilot = pilot
        self.max_pilot = max_pilot
        self.max_rounds = max_rounds
        self.fallback = fallback if fallback is not None else ErrorProfile()
        self.min_sample = min_sample
        self.last = None

    # -- sizing --------------------------------------------------------------
    def _size(self, instance, eps, conf, est):
        """Runs the pilot into `est`; returns (N, method): "learned" (pilot +
        network), "pilot" (the pilot's own CV: correction=False, or a plan the
        network has no pattern slot for), or (None, "regression") when the
        pilot cannot size the query or the weights are missing."""
        kind = instance.plan.kind
        if self.correction and self.model is None:
            return None, "regression"
        use_net = self.correction and kind in PATTERN_SLOTS
        n_p = self.pilot
        while True:
            instance.calculate(n_p, None, est)
            s = est.sums_view(instance.channel)
            if s.successes >= MIN_PILOT_SUCCESSES or n_p >= self.max_pilot:
                break
            n_p = min(n_p * 8, self.max_pilot)
        mean, cv, succ = pilot_stats(s)
        if s.successes < MIN_PILOT_SUCCESSES or not math.isfinite(cv) or cv <= 0:
            return None, "regression"
        if use_net:
            x = feature_vector(graph_features(instance.graph), kind, n_p, cv, succ)
            r = float(self.model.predict(x)[0]) - self.aim_sigmas * self.model.val_rmse[PATTERN_SLOTS[kind]]
        else:
            r = 0.0
        cv_hat = cv * 10.0 ** r
        z = _z(conf)
        return max(int(math.ceil((z * cv_hat / eps) ** 2)), n_p), "learned" if use_net else "pilot"

    def get_num_estimators(self, instance, num_vertices_in_pattern, required_error_rate, required_confidence):
        est = Estimation()
        n, how = self._size(instance, required_error_rate, required_confidence, est)
        if n is None:
            n = self.fallback.get_num_estimators(instance, num_vertices_in_pattern, required_error_rate,
                                                 required_confidence)
        return n

    # -- sizing + judge + top-up ---------------------------------------------
    def run(self, instance, num_vertices_in_pattern, required_error_rate, required_confidence,
            init_estimation=None):
        rec = estimate_to_error(instance, required_error_rate, required_confidence, self,
                                init_estimation=init_estimation, num_vertices_in_pattern=num_vertices_in_pattern)
        self.last = rec
        return rec["estimate"]


def judge(est, eps, conf, min_sample=1 << 12):
    """True when the CLT interval of the running mean is within eps (relative),
    the stopping rule of sampling/IEStop.cpp:47-56."""
    t = est.get_num_finished_estimators()
    mean = est.exact_mean()
    if t < min_sample or mean <= 0:
        return False
    return est.half_width(conf) <= eps * mean


def estimate_to_error(instance, eps, conf, profile, init_estimation=None, num_vertices_in_pattern=None,
                      max_rounds=None, max_estimators=1 << 40, judged=True, fallback_budget=1 << 24):
    """Size the query with `profile` (LearnedErrorProfile or ErrorProfile),
    run it, and (judged=True) top up until the CLT judge accepts.  When the
    sizing is undefined, a judged query starts from `fallback_budget`
    estimators (0: report the failure).  Returns a record with the estimate,
    estimator counts and wall times."""
    t0 = time.perf_counter()
    est = init_estimation if init_estimation is not None else Estimation()
    how = "regression"
    n = None
    if isinstance(profile, LearnedErrorProfile):
        n, how = profile._size(instance, eps, conf, est)
        if n is None:
            profile = profile.fallback
    error = None
    if n is None:
        try:
            n = profile.get_num_estimators(instance, num_vertices_in_pattern, eps, conf)
            how = "regression"
        except ZGraphError as e:
            # the regression profile is undefined (no success among its pilots,
            # or a zero estimate on the sub-sampled graph)
            if not judged or not fallback_budget:
                return {"estimate": float("nan"), "sizing": "regression", "error_message": str(e),
                        "sized_estimators": 0, "estimators": 0, "topup_rounds": 0, "half_width": float("inf"),
                        "rel_half_width": float("inf"), "accepted": False, "capped": False,
                        "sizing_s": time.perf_counter() - t0, "total_s": time.perf_counter() - t0}
            n, how, error = fallback_budget, "judge-only", str(e)
    t_size = time.perf_counter() - t0
    n0 = n
    capped = n > max_estimators
    n = min(max(n, 1), max_estimators)
    instance.calculate(n, num_vertices_in_pattern, est)
    rounds = 0
    rmax = max_rounds if max_rounds is not None else getattr(profile, "max_rounds", 8)
    while judged and not judge(est, eps, conf) and rounds < rmax and n < max_estimators:
        mean = est.exact_mean()
        hw = est.half_width(conf)
        if mean > 0 and math.isfinite(hw) and hw > 0:
            grow = (hw / (eps * mean)) ** 2 * 1.05
        else:
            grow = 4.0
        n = min(int(math.ceil(n * min(max(grow, 1.25), 64.0))), max_estimators)
        instance.calculate(n, num_vertices_in_pattern, est)
        rounds += 1
    t1 = time.perf_counter()
    mean = est.exact_mean()
    hw = est.half_width(conf)
    rec = {"estimate": mean, "sizing": how, "sized_estimators": n0, "estimators": est.get_num_finished_estimators(),
            "topup_rounds": rounds, "half_width": hw, "rel_half_width": hw / mean if mean > 0 else float("inf"),
            "accepted": judge(est, eps, conf), "capped": capped, "sizing_s": t_size, "total_s": t1 - t0}
    if error:
        rec["error_message"] = error
    return rec


def estimate_both_to_error(instance, eps, conf, profile, max_estimators=1 << 40, max_rounds=8,
                           fallback_budget=1 << 24):
    """Both channels of a two-channel plan (3-motif; fused triangle + 4-clique,
    with the three-shape or the split 4-clique plan) to +-eps at `conf`: one
    pilot gives both channels' statistics, each channel is sized by the network
    with its own pattern slot (from its pilot's CV for a channel the network
    has no slot for; judge-only start where the pilot holds too few successes),
    the query runs to the larger size over one id range, and the
    judge tops up until both channels' intervals hold.  Returns one record per
    channel (estimate, estimators, half-width, accepted) and the shared times."""
    kinds = CHANNEL_KINDS.get(instance.plan.kind)
    if kinds is None:
        raise ValueError("not a two-channel plan")
    t0 = time.perf_counter()
    g, plan, seed = instance.graph, instance.plan, instance.seed
    ests = [Estimation(), Estimation()]
    done = 0

    def extend(n):
        nonlocal done
        if n > done:
            sums = run(g, plan, seed, done, n - done)
            for c in (0, 1):
                ests[c].absorb(sums[c])
                ests[c].update(n - done, sums[c].mean)
            done = n

    n_p = profile.pilot
    while True:
        extend(n_p)
        if min(e.successes for e in ests) >= MIN_PILOT_SUCCESSES or n_p >= profile.max_pilot:
            break
        n_p = min(n_p * 8, profile.max_pilot)
    gf = graph_features(g)
    z = _z(conf)
    sizes, how = [], []
    for c in (0, 1):
        mean, cv, succ = pilot_stats(ests[c].sums_view())
        if (profile.model is not None and ests[c].successes >= MIN_PILOT_SUCCESSES and math.isfinite(cv)
                and cv > 0):
            if kinds[c] in PATTERN_SLOTS:
                r = float(profile.model.predict(feature_vector(gf, kinds[c], n_p, cv, succ))[0]) - \
                    profile.aim_sigmas * profile.model.val_rmse[PATTERN_SLOTS[kinds[c]]]
                how.append("learned")
            else:   # no network slot: the pilot's own CV
                r = 0.0
                how.append("pilot")
            sizes.append(max(int(math.ceil((z * cv * 10.0 ** r / eps) ** 2)), n_p))
        else:
            sizes.append(max(fallback_budget, n_p))
            how.append("judge-only")
    t_size = time.perf_counter() - t0
    n = min(max(sizes), max_estimators)
    extend(n)
    rounds = 0
    while not all(judge(e, eps, conf) for e in ests) and rounds < max_rounds and n < max_estimators:
        grow = 1.25
        for e in ests:
            mean, hw = e.exact_mean(), e.half_width(conf)
            g_c = (hw / (eps * mean)) ** 2 * 1.05 if mean > 0 and math.isfinite(hw) and hw > 0 else 4.0
            grow = max(grow, g_c)
        n = min(int(math.ceil(n * min(grow, 64.0))), max_estimators)
        extend(n)
        rounds += 1
    t1 = time.perf_counter()
    chans = []
    for c in (0, 1):
        mean, hw = ests[c].exact_mean(), ests[c].half_width(conf)
        chans.append({"estimate": mean, "sizing": how[c], "sized_estimators": sizes[c],
                      "half_width": hw, "rel_half_width": hw / mean if mean > 0 else float("inf"),
                      "accepted": judge(ests[c], eps, conf)})
    return {"channels": chans, "estimators": done, "topup_rounds": rounds, "capped": max(sizes) > max_estimators,
            "sizing_s": t_size, "total_s": t1 - t0}


def choose_split_completions(graph, seed=1, kind=PLAN_K4_SPLIT, j_lo=64, j_hi=1024, min_successes=32,
                             min_ms=1.0, pilot=1 << 18, max_pilot=1 << 30, first_id=1 << 60, eps=0.01, conf=0.95):
    """Completions for a split plan (ZG_PLAN_K4_SPLIT, the 4-clique channel of
    ZG_PLAN_TRIANGLE_K4_SPLIT, or ZG_PLAN_HOUSE_SPLIT), from two pilots over
    the same ids with j_lo and j_hi completions.  The j_lo pilot starts at
    `pilot` ids and grows (at least 4x, more when far short) until it holds
    `min_successes` successes and its kernel runs `min_ms` milliseconds --
    long enough for a launch-independent time -- or reaches `max_pilot` ids.
    If a +-eps answer at `conf` with j_lo completions, projected from that
    pilot, would take less time than the j_hi pilot can (j_hi / j_lo times the
    j_lo pilot's), the choice is j_lo: it could not repay its own cost
    (eps=None turns this off).  Otherwise the j_hi pilot runs the same ids
    (its completions include the j_lo pilot's).  Per estimator
    Var(J) = Vb + Vw / J (Vb: the first stage's share, Vw: the completions')
    and the kernel time t(J) = c0 + c1 J; the time to a given error is
    proportional to Var(J) t(J), least at J* = sqrt(Vw c0 / (Vb c1)), returned
    as the nearest power of two in [1, 4096] (1 when Vw <= 0, 4096 when
    Vb <= 0 or c1 <= 0).  The pilots run ids from `first_id` on, away from the
    query's own.  Returns (J, record)."""
    from . import Plan
    if kind not in (PLAN_K4_SPLIT, PLAN_TRIANGLE_K4_SPLIT, PLAN_HOUSE_SPLIT):
        raise ValueError("not a split plan kind")
    ch = 1 if kind == PLAN_TRIANGLE_K4_SPLIT else 0
    t0 = time.perf_counter()
    plans = {J: Plan(kind, completions=J) for J in (j_lo, j_hi)}
    for p in plans.values():
        run(graph, p, seed, first_id, 1 << 12)   # code objects loaded before timing
    n = pilot
    while True:
        sums, ms = run(graph, plans[j_lo], seed, first_id, n, timed=True)
        s = sums[ch]
        if (s.successes >= min_successes and ms >= min_ms) or n >= max_pilot:
            break
        short = max(min_successes / max(s.successes, 1), min_ms / max(ms, 1e-3))
        n = min(int(n * min(max(1.5 * short, 4.0), 256.0)), max_pilot)
    stats = {j_lo: split_pilot_record(s, n, ms)}
    if eps is not None and s.mean > 0:
        z = _z(conf)
        projected_ms = (z * z * stats[j_lo]["var"] / (eps * s.mean) ** 2) * stats[j_lo]["ms_per_id"]
        if projected_ms < ms * j_hi / j_lo:
            return j_lo, {"completions": j_lo, "hi_pilot_skipped": True, "projected_ms": projected_ms,
                          "pilots": stats, "seconds": time.perf_counter() - t0}
    sums, ms = run(graph, plans[j_hi], seed, first_id, n, timed=True)
    stats[j_hi] = split_pilot_record(sums[ch], n, ms)
    rec = fit_split_model(stats, j_lo, j_hi)
    rec["hi_pilot_skipped"] = False
    rec["pilots"] = stats
    rec["seconds"] = time.perf_counter() - t0
    return rec["completions"], rec


def split_pilot_record(sums, n, ms):
    """A split pilot's record from its channel sums over n ids and its kernel ms."""
    mean = sums.mean
    return {"ids": n, "successes": sums.successes, "mean": mean, "var": max(sums.sumsq / n - mean * mean, 0.0),
            "ms_per_id": ms / n}


def fit_split_model(stats, j_lo, j_hi):
    """Var(J) = Vb + Vw / J and t(J) = c0 + c1 J through the two pilots
    (stats[j]: split_pilot_record), and the completions that minimise Var(J) t(J):
    the power of two nearest J* = sqrt(Vw c0 / (Vb c1)) in [1, 4096]."""
    a, b = stats[j_lo], stats[j_hi]
    vw = (a["var"] - b["var"]) / (1.0 / j_lo - 1.0 / j_hi)
    vb = b["var"] - vw / j_hi
    c1 = (b["ms_per_id"] - a["ms_per_id"]) / (j_hi - j_lo)
    c0 = a["ms_per_id"] - c1 * j_lo
    if vw <= 0:
        jstar = 1.0
    elif vb <= 0 or c1 <= 0:
        jstar = 4096.0
    else:
        jstar = math.sqrt(vw * max(c0, 1e-12) / (vb * c1))
    J = int(min(4096, max(1, 2 ** round(math.log2(max(jstar, 1.0))))))
    return {"completions": J, "j_star": jstar, "v_between": vb, "v_within": vw, "c0_ms_per_id": c0,
            "c1_ms_per_id_per_completion": c1}


def split_to_error(graph, eps, conf, profile, kind=PLAN_K4_SPLIT, seed=1, max_estimators=1 << 40, **choose):
    """A count to +-eps at `conf` by a one-channel split plan (4-cliques:
    ZG_PLAN_K4_SPLIT; 5-houses: ZG_PLAN_HOUSE_SPLIT) with its completions
    chosen by choose_split_completions (pilots on ids away from the query's),
    then sized by `profile` and judged as estimate_to_error does.  The record
    adds the completions, the choice's record and its time (in total_s)."""
    from . import Plan, ZGraphInstance
    if kind not in (PLAN_K4_SPLIT, PLAN_HOUSE_SPLIT):
        raise ValueError("not a one-channel split plan kind")
    t0 = time.perf_counter()
    J, choice = choose_split_completions(graph, seed=seed, kind=kind, eps=eps, conf=conf, **choose)
    inst = ZGraphInstance(graph, Plan(kind, completions=J), seed=seed)
    rec = estimate_to_error(inst, eps, conf, profile, max_estimators=max_estimators)
    rec["completions"] = J
    rec["choice"] = choice
    rec["choice_s"] = choice["seconds"]
    rec["total_s"] = time.perf_counter() - t0
    return rec
