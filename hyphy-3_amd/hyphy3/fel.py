"""FEL (fixed-effects likelihood) per-site refits on the MI355X engine.

Follows FELAnalyzer with every branch tested (include/hyphy/analyses/fel.hpp):
``prepare_branch_lengths`` (:227-272) converts the fitted tree's lengths by
3 / Q_scale of the global MG94 model, then ``analyze_pattern`` (:326-445)
refits each pattern's synonymous and non-synonymous rates: a 12-point start
grid, ``NelderMead2D`` over (alpha, beta) in [1e-4, 1000] (nelder_mead.hpp:
111-205), the null alpha = beta from a 7-point grid and ``Brent1D`` on
[1e-4, 50] (nelder_mead.hpp:15-108), the likelihood-ratio statistic and
p-value erfc(sqrt(LRT / 2)); ``run`` (:778-834) maps patterns to sites.

The reference runs one pattern at a time per thread, each candidate a fresh
61 x 61 eigendecomposition and a single-pattern pruning.  Here every
pattern's refit is a coroutine that asks for one candidate log-likelihood at
a time, and the driver advances all patterns in lockstep: one
``hy_fel_eval`` call evaluates the current candidate of every pattern still
being refitted (spectra of all candidates in one launch, warm-started from the
same pattern's previous candidate, then one workgroup per candidate prunes
its pattern).  Each pattern sees the reference's sequence of evaluations, so
results differ from a sequential run only through the candidate values.
"""
from __future__ import annotations

import math
import time

import numpy as np

START_GRID = ((0.01, 0.1), (1.0, 0.1), (1.0, 0.5), (1.0, 1.0), (1.0, 5.0), (10.0, 0.1), (0.01, 0.5), (0.01, 5.0),
              (10.0, 0.5), (10.0, 1.0), (10.0, 50.0), (100.0, 1.0))  # fel.hpp:387-391
NULL_GRID = (0.01, 0.1, 0.5, 1.0, 2.0, 5.0, 10.0)  # fel.hpp:418


def brent_1d(ax, bx, cx, tol=1e-4, max_iter=60):
    """Brent1D::minimize_point (nelder_mead.hpp:24-101) as a coroutine: yields
    each abscissa, receives the objective there; returns (x, f(x), iterations,
    converged)."""
    cgold, zeps = 0.3819660, 1e-10
    a, b = min(ax, cx), max(ax, cx)
    v = w = x = bx
    e = 0.0
    fx = yield x
    fv = fw = fx
    iters, converged = 0, False
    for it in range(max_iter):
        iters = it + 1
        xm = 0.5 * (a + b)
        tol1 = tol * abs(x) + zeps
        tol2 = 2.0 * tol1
        if abs(x - xm) <= tol2 - 0.5 * (b - a):
            converged = True
            break
        d = 0.0
        if abs(e) > tol1:
            # parabola through (v, fv), (w, fw), (x, fx)
            r = (x - w) * (fx - fv)
            q = (x - v) * (fx - fw)
            p = (x - v) * q - (x - w) * r
            q = 2.0 * (q - r)
            if q > 0.0:
                p = -p
            q = abs(q)
            etemp, e = e, d
            if abs(p) >= abs(0.5 * q * etemp) or p <= q * (a - x) or p >= q * (b - x):
                e = (a - x) if x >= xm else (b - x)
                d = cgold * e
            else:
                d = p / q
                if (x + d) - a < tol2 or b - (x + d) < tol2:
                    d = abs(tol1) if xm - x >= 0.0 else -abs(tol1)
        else:
            e = (a - x) if x >= xm else (b - x)
            d = cgold * e
        u = x + d if abs(d) >= tol1 else x + (abs(tol1) if d >= 0.0 else -abs(tol1))
        fu = yield u
        if fu <= fx:
            if u >= x:
                a = x
            else:
                b = x
            v, fv, w, fw, x, fx = w, fw, x, fx, u, fu
        else:
            if u < x:
                a = u
            else:
                b = u
            if fu <= fw or w == x:
                v, fv, w, fw = w, fw, u, fu
            elif fu <= fv or v == x or v == w:
                v, fv = u, fu
    return x, fx, iters, converged


def nelder_mead_2d(init_a, init_b, lb=1e-4, ub=1000.0, tol=1e-4, max_iter=100):
    """NelderMead2D::minimize (nelder_mead.hpp:121-204) as a coroutine: yields
    each (clamped) vertex, receives the objective there; returns (a, b, f,
    iterations, converged).  Vertices are kept sorted by a stable sort, as the
    reference's insertion sort of three points keeps ties in order."""

    def vertex(a, b):
        ca, cb = min(max(a, lb), ub), min(max(b, lb), ub)
        f = yield (ca, cb)
        return (ca, cb, f)

    step_a, step_b = max(0.1, 0.2 * init_a), max(0.1, 0.2 * init_b)
    s = [(yield from vertex(init_a, init_b)), (yield from vertex(init_a + step_a, init_b)),
         (yield from vertex(init_a, init_b + step_b))]
    iters, converged = 0, False
    key = lambda t: t[2]  # noqa: E731
    for it in range(max_iter):
        iters = it + 1
        s.sort(key=key)
        best, mid, worst = s
        if abs(worst[2] - best[2]) < tol and abs(worst[0] - best[0]) < tol and abs(worst[1] - best[1]) < tol:
            converged = True
            break
        ca, cb = 0.5 * (best[0] + mid[0]), 0.5 * (best[1] + mid[1])
        r = yield from vertex(ca + (ca - worst[0]), cb + (cb - worst[1]))
        if r[2] < best[2]:
            e = yield from vertex(ca + 2.0 * (r[0] - ca), cb + 2.0 * (r[1] - cb))
            s[2] = e if e[2] < r[2] else r
        elif r[2] < mid[2]:
            s[2] = r
        else:
            outside = r[2] < worst[2]
            toward = r if outside else worst
            c = yield from vertex(ca + 0.5 * (toward[0] - ca), cb + 0.5 * (toward[1] - cb))
            if c[2] < toward[2]:
                s[2] = c
            else:
                s[1] = yield from vertex(best[0] + 0.5 * (mid[0] - best[0]), best[1] + 0.5 * (mid[1] - best[1]))
                s[2] = yield from vertex(best[0] + 0.5 * (worst[0] - best[0]), best[1] + 0.5 * (worst[1] - best[1]))
    s.sort(key=key)
    return s[0][0], s[0][1], s[0][2], iters, converged


def invariant(states_row):
    """All non-gap states of the pattern identical (fel.hpp:333-350)."""
    seen = [int(s) for s in states_row if s >= 0]
    return all(s == seen[0] for s in seen)


def refit_pattern(states_row, tested_length):
    """FELAnalyzer::analyze_pattern with every branch tested (fel.hpp:326-445)
    as a coroutine: yields (alpha, beta) candidates, receives their pattern
    log-likelihoods; returns the site result."""
    res = dict(alpha=0.0, beta=0.0, beta_nuisance=0.0, alpha_null=0.0, lrt=0.0, p_value=1.0, total_branch_length=0.0,
               log_l_alt=0.0, log_l_null=0.0)
    if invariant(states_row):
        return res
    best_lnl, best_a, best_b = -1e20, 1.0, 1.0
    for a, b in START_GRID:
        lnl = yield (a, b)
        if lnl > best_lnl:
            best_lnl, best_a, best_b = lnl, a, b
    nm = nelder_mead_2d(best_a, best_b, 1e-4, 1000.0, 1e-4, 100)
    pa, pb, pf, _, _ = yield from _negated(nm, lambda v: v)
    res.update(alpha=pa, beta=pb, log_l_alt=-pf)
    init_null = (min(pa, 100.0) + 3.0 * min(pb, 100.0)) / 4.0
    best_null_lnl = yield (init_null, init_null)
    best_null_a = init_null
    for c in NULL_GRID:
        lnl = yield (c, c)
        if lnl > best_null_lnl:
            best_null_lnl, best_null_a = lnl, c
    br = brent_1d(1e-4, min(max(best_null_a, 0.01), 45.0), 50.0, 1e-4, 60)
    xa, fa, _, _ = yield from _negated(br, lambda x: (x, x))
    res.update(alpha_null=xa, log_l_null=-fa)
    lrt = 0.0
    if res["log_l_alt"] > -1e10 and res["log_l_null"] > -1e10:
        lrt = max(2.0 * (res["log_l_alt"] - res["log_l_null"]), 0.0)
    res.update(lrt=lrt, p_value=math.erfc(math.sqrt(lrt / 2.0)), total_branch_length=xa * tested_length)
    return res


def _negated(opt, to_candidate):
    """Runs a minimiser coroutine on -lnL: forwards its points as candidates
    and the negated log-likelihoods back; returns the minimiser's result."""
    try:
        pt = next(opt)
        while True:
            lnl = yield to_candidate(pt)
            pt = opt.send(-lnl)
    except StopIteration as stop:
        return stop.value


def drive(coroutines, evaluate):
    """Advances every coroutine in lockstep.  ``coroutines`` maps a pattern to
    its refit; each round, ``evaluate(patterns, alphas, betas)`` returns the
    log-likelihoods of every pending candidate (one per pattern).  Returns
    ({pattern: result}, number of evaluate calls, number of candidates)."""
    results, pending = {}, {}
    for p, co in coroutines.items():
        try:
            pending[p] = next(co)
        except StopIteration as stop:
            results[p] = stop.value
    calls = cands = 0
    while pending:
        ps = np.fromiter(pending.keys(), dtype=np.int32, count=len(pending))
        ab = np.array([pending[int(p)] for p in ps], dtype=np.float64).reshape(-1, 2)
        lnl = evaluate(ps, ab[:, 0].copy(), ab[:, 1].copy())
        calls += 1
        cands += len(ps)
        nxt = {}
        for p, v in zip(ps.tolist(), np.asarray(lnl, dtype=np.float64).tolist()):
            try:
                nxt[p] = coroutines[p].send(v)
            except StopIteration as stop:
                results[p] = stop.value
        pending = nxt
    return results, calls, cands


class FEL:
    """FEL on one alignment, every branch tested.

    ``tree`` carries the fitted global MG94 branch lengths (expected
    substitutions per codon), ``theta`` the five exchangeabilities (AC, AT,
    CG, CT, GT) and ``alpha``/``beta`` (with ``delta``/``psi``) the global
    rates, as FELAnalyzer's constructor receives them (fel.hpp:167-200).
    ``engine``: an ``Engine`` already holding this problem and model
    (``hy_set_problem``/``hy_set_model``), or None to build one from
    ``tree``/``aln`` (hyphy3.Tree, hyphy3.Alignment) on ``device``.
    """

    def __init__(self, tree, aln, theta, alpha=1.0, beta=1.0, delta=0.0, psi=0.0, p_value_threshold=0.1,
                 engine=None, device=0, arrays=None):
        from hyphy3.engine import Engine
        if engine is None:
            import hyphy3
            arrays = hyphy3.problem_arrays(tree, aln)
            engine = Engine(device)
            n = len(arrays["parent"])
            cp, ci = arrays["child_ptr"], arrays["child_idx"]
            engine.set_problem(arrays["parent"], [list(ci[cp[i]:cp[i + 1]]) for i in range(n)], arrays["leaf_taxon"],
                               arrays["states"], arrays["masks"], arrays["weights"], S=int(arrays["S"]))
            engine.set_model(int(arrays["code_id"]), arrays["pos_freqs"], arrays["pi"])
            self.states = np.asarray(arrays["states"])
            self.site_to_pattern = np.asarray(aln.site_to_pattern)
            bl = np.asarray(tree.get_branch_lengths(), dtype=np.float64)
            root = int(tree.root_id)
        else:
            self.states = np.asarray(arrays["states"])
            self.site_to_pattern = np.asarray(arrays["site_to_pattern"])
            bl = np.asarray(arrays["bl"], dtype=np.float64)
            root = int(arrays["root"])
        self.engine = engine
        self.theta = np.asarray(theta, dtype=np.float64)
        self.delta, self.psi = float(delta), float(psi)
        self.p_value_threshold = p_value_threshold
        # prepare_branch_lengths (fel.hpp:240-262): Q_scale of the global
        # model = alpha * mean rate at (1, beta / alpha), the generator being
        # linear in (alpha, beta)
        a_safe = max(float(alpha), 1e-12)
        _, mu = engine.fel_eval([0], [a_safe], [float(beta)], self.theta, np.where(np.arange(bl.size) == root, 0.0, bl),
                                self.delta, self.psi, scale=True)
        self.q_scale = float(alpha) * float(mu[0])
        self.conversion_factor = 3.0 / self.q_scale if self.q_scale > 1e-12 else 1.0
        mask = np.arange(bl.size) != root
        self.baseline_tree_length = float(np.sum(bl[mask]))
        self.bl = np.where(mask, bl * self.conversion_factor, 0.0)
        self.stats = {}

    def evaluate(self, patterns, alphas, betas):
        return self.engine.fel_eval(patterns, alphas, betas, self.theta, self.bl, self.delta, self.psi)

    def run(self, patterns=None, evaluate=None):
        """Refits every pattern (or ``patterns``) and returns per-site results
        (FELAnalyzer::run, fel.hpp:778-834) when all patterns are refitted,
        else the per-pattern results.  ``evaluate`` replaces the engine's
        candidate evaluator (same contract as ``FEL.evaluate``)."""
        pats = range(self.states.shape[0]) if patterns is None else patterns
        cos = {int(p): refit_pattern(self.states[int(p)], self.baseline_tree_length) for p in pats}
        t0 = time.perf_counter()
        res, calls, cands = drive(cos, evaluate or self.evaluate)
        self.stats = {"seconds": time.perf_counter() - t0, "calls": calls, "candidates": cands,
                      "patterns": len(cos)}
        if patterns is not None:
            return res
        out = []
        for s, p in enumerate(self.site_to_pattern.tolist()):
            r = dict(res[p])
            r["site_index"] = s
            out.append(r)
        return out
