"""FP64 restatement of MEME's per-site likelihood.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  The product
never imports this module.

Restated reference semantics (paths relative to the reference root), no
bridges:
  * MEMEAnalyzer::analyze_pattern's eval_pattern_meme_lnl
        include/hyphy/analyses/meme.hpp:219-277
    MG94(alpha = a, beta = a w1), MG94(a, b_plus), MG94(a, b_bg) built as the
    reference builds them (MG94Matrix::update, rate_matrix.hpp:165-248);
    test branches P = p1 P_1(t) + (1 - p1) P_+(t), background P_bg(t)
  * FELAnalyzer::compute_pattern_log_likelihood  fel.hpp:276-308 (the
    oracle's pruning of the one-pattern problem)
  * NelderMeadND::minimize                        include/hyphy/opt/nelder_mead.hpp:208-360
  * MEMEAnalyzer::analyze_pattern, every branch tested (no background, no
    bridges): start grids :279-320, alternative 4-D fit :351-367, null 3-D
    fit and warm alternative refit :370-453, the test :455-482, branch EBFs
    :484-600, MEME vs FEL :602-611, total branch length :614-638
  * MEMEAnalyzer::prepare_branch_lengths          meme.hpp:150-163 (FEL's
    conversion, fel_ref.converted_lengths)
"""
from __future__ import annotations

import math
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from oracle import fel_ref as fr
from oracle import hyoracle as ho
from oracle.fel_ref import single_pattern


class MemeCandidateOracle:
    def __init__(self, prob, code, theta, pos, pi, bl, p, is_test=None, delta=0.0, psi=0.0):
        self.sub = single_pattern(prob, p)
        self.code, self.theta, self.pos, self.pi = code, list(theta), pos, np.asarray(pi, dtype=np.float64)
        self.delta, self.psi = delta, psi
        self.bl = np.asarray(bl, dtype=np.float64)
        self.root = prob.root
        n = len(self.bl)
        self.is_test = np.ones(n, dtype=bool) if is_test is None else np.asarray(is_test, dtype=bool)
        self.count = 0
        self._cache = {}

    def _spec(self, alpha, beta):
        key = (alpha, beta)
        if key not in self._cache:
            if len(self._cache) > 64:
                self._cache.clear()
            Q, _ = ho.mg94_Q(self.code, ho.params(alpha, beta, self.theta, self.delta, self.psi), self.pos, self.pi)
            self._cache[key] = ho.decompose(Q, self.pi)
        return self._cache[key]

    def __call__(self, a, w1, p1, b_plus, b_bg=0.0, cond_node=-1, cond_class=0):
        """lnL of the pattern; with cond_node >= 0 that (test) branch takes
        P_1 (cond_class 0) or P_+ (1) alone, the others the mixture
        (meme.hpp:552-580)."""
        self.count += 1
        s1 = self._spec(a, a * w1)
        sp = self._spec(a, b_plus)
        has_bg = not bool(np.all(self.is_test[[n for n in range(len(self.bl)) if n != self.root]]))
        sbg = self._spec(a, b_bg) if has_bg else None
        S = len(self.pi)
        P = np.zeros((len(self.bl), S, S))
        for n in range(len(self.bl)):
            if n == self.root:
                continue
            if n == cond_node:
                P[n] = ho.transition(*(s1 if cond_class == 0 else sp), self.bl[n])
            elif self.is_test[n]:
                P[n] = p1 * ho.transition(*s1, self.bl[n]) + (1.0 - p1) * ho.transition(*sp, self.bl[n])
            else:
                P[n] = ho.transition(*sbg, self.bl[n])
        ll = ho.loglik_from_P(self.sub, P, self.pi)
        return ll if math.isfinite(ll) else -1e20


def nelder_mead_nd(f, init, lb, ub, tol=1e-4, max_iter=150, grid=()):
    """NelderMeadND::minimize (nelder_mead.hpp:219-360), one evaluation at a
    time: points clamped to [lb, ub] before evaluation; the best of init and
    the grid (first strictly lower) starts the simplex, vertex d + 1 moving
    coordinate d by max(0.05, 0.2 |x_d|) (up if that stays within ub, else
    down); reflection 1, expansion 2, contraction 0.5, shrink 0.5; stops when
    best and worst differ by less than tol in f and every coordinate.
    Returns (x, f, iterations, converged); ties keep the vertices' order
    (Python's stable sort; the reference's std::sort of at most six points
    is an insertion sort)."""
    D = len(init)

    def ev(x):
        c = [min(max(x[d], lb[d]), ub[d]) for d in range(D)]
        return (c, f(c))

    best = ev(init)
    for g in grid:
        p = ev(g)
        if p[1] < best[1]:
            best = p
    s = [best]
    for d in range(D):
        x = list(best[0])
        step = max(0.05, 0.2 * abs(x[d]))
        x[d] = x[d] + step if x[d] + step <= ub[d] else x[d] - step
        s.append(ev(x))
    iters, converged = 0, False
    for it in range(max_iter):
        iters = it + 1
        s.sort(key=lambda v: v[1])
        dx = max(abs(s[D][0][d] - s[0][0][d]) for d in range(D))
        if abs(s[D][1] - s[0][1]) < tol and dx < tol:
            converged = True
            break
        c = [0.0] * D
        for i in range(D):
            for d in range(D):
                c[d] += s[i][0][d]
        c = [v / D for v in c]
        r = ev([c[d] + 1.0 * (c[d] - s[D][0][d]) for d in range(D)])
        if r[1] < s[0][1]:
            e = ev([c[d] + 2.0 * (r[0][d] - c[d]) for d in range(D)])
            s[D] = e if e[1] < r[1] else r
        elif r[1] < s[D - 1][1]:
            s[D] = r
        else:
            outside = r[1] < s[D][1]
            tw = r[0] if outside else s[D][0]
            cv = ev([c[d] + 0.5 * (tw[d] - c[d]) for d in range(D)])
            if cv[1] < (r[1] if outside else s[D][1]):
                s[D] = cv
            else:
                for i in range(1, D + 1):
                    s[i] = ev([s[0][0][d] + 0.5 * (s[i][0][d] - s[0][0][d]) for d in range(D)])
    s.sort(key=lambda v: v[1])
    return s[0][0], s[0][1], iters, converged


LB_ALT, UB_ALT = (1e-4, 0.0, 1e-6, 1e-4), (10000.0, 1.0, 1.0 - 1e-6, 100000.0)
LB_NULL, UB_NULL = (1e-4, 0.0, 1e-6), (10000.0, 1.0, 1.0 - 1e-6)


def start_grid(fel_alpha, fel_beta):
    """The regime's start grid (meme.hpp:279-320)."""
    a = min(max(fel_alpha, 1e-4), 10000.0)
    b = min(max(fel_beta, 1e-4), 100000.0)
    w = min(max(b / max(a, 1e-4), 0.0), 1.0)
    if a >= b:
        g = [(a, w, 0.75, b), (a, 0.0, 0.75, b), (a, 0.0, 0.95, a * 0.5), (a, 0.25, 0.85, a * 0.9)]
        bps = (max(a * 1.5, b * 2.0), max(a * 2.5, b * 4.0), max(a * 4.0, b * 8.0))
        for pp in (0.05, 0.10, 0.20):
            for bp in bps:
                g.append((a, w, 1.0 - pp, bp))
    else:
        g = [(a, w, 0.75, b), (a, 0.0, 0.75, b), (a, 0.25, 0.50, b * 2.0), (a, 0.25, 0.75, b * 5.0),
             (a, 0.10, 0.90, b * 15.0), (a, 0.00, 0.95, b * 30.0), (a, 1.00, 0.01, max(b, a * 1.5))]
    return g


def p_value(lrt):
    """The asymptotic mixture p-value of meme.hpp:480."""
    return (2.0 / 3.0) * (0.45 * math.erfc(math.sqrt(lrt / 2.0)) + 0.55 * math.exp(-lrt / 2.0))


def analyze_pattern(evalf, invariant, fel, num_nodes, root):
    """MEMEAnalyzer::analyze_pattern, every branch tested (meme.hpp:166-648).
    evalf(a, w1, p1, bp, cond_node=-1, cond_class=0) is the pattern's lnL;
    fel the pattern's FEL result dict (fel_ref.analyze_pattern's keys).
    Returns (result dict, ebf list of num_nodes)."""
    res = dict(alpha=0.0, beta1=0.0, p1=1.0, beta_plus=0.0, p_plus=0.0, lrt=0.0, p_value=1.0,
               branches_under_selection=0, total_branch_length=0.0, log_l_meme=0.0, log_l_fel=0.0,
               p_value_meme_vs_fel=1.0, fel_alpha=fel["alpha"], fel_beta=fel["beta"], beta_bg=0.0)
    ebf = [1.0] * num_nodes
    if invariant:
        return res, ebf
    res["log_l_fel"] = fel["log_l_alt"]
    grid = start_grid(fel["alpha"], fel["beta"])
    x, fx, _, _ = nelder_mead_nd(lambda v: -evalf(*v), grid[0], LB_ALT, UB_ALT, 1e-5, 800, grid)
    alpha, omega1, p1, bp = x
    ll = -fx
    p_plus = 1.0 - p1
    if bp > alpha and p_plus > 1e-6:
        an = min(max(fel["alpha_null"], 1e-4), 50.0)
        ng = [(alpha, omega1, p1), (an, 1.0, 0.5), (an, omega1, p1), (min(max(alpha, 1e-4), 50.0), 1.0, 0.5)]
        for ac in (0.01, 0.1, 0.5, 1.0, 2.0, 5.0, 10.0, 20.0):
            ng += [(ac, 0.5, 0.5), (ac, 0.0, 0.9)]
        xn, fn, _, _ = nelder_mead_nd(lambda v: -evalf(v[0], v[1], v[2], v[0]), ng[0], LB_NULL, UB_NULL, 1e-5, 800,
                                      ng)
        ll_null = -fn
        a_n, w_n, p_n = xn
        if ll_null > ll and math.isfinite(ll_null):
            wg = [(a_n, w_n, p_n, a_n), (a_n, w_n, 0.99, a_n)]
            x2, f2, _, _ = nelder_mead_nd(lambda v: -evalf(*v), wg[0], LB_ALT, UB_ALT, 1e-5, 800, wg)
            if -f2 > ll:
                alpha, omega1, p1, bp = x2
                p_plus = 1.0 - p1
                ll = -f2
        if ll_null <= -1e10 or ll <= -1e10 or not math.isfinite(ll_null) or not math.isfinite(ll):
            lrt, pv = 0.0, 1.0
        elif ll_null > ll:
            alpha, omega1, p1, bp = a_n, w_n, p_n, a_n
            p_plus = 1.0 - p_n
            ll = ll_null
            lrt, pv = 0.0, 2.0 / 3.0
        elif abs(ll - ll_null) <= 1e-12:
            lrt, pv = 0.0, 2.0 / 3.0
        else:
            lrt = 2.0 * (ll - ll_null)
            pv = p_value(lrt)
        under = 0
        for n in range(num_nodes):
            if n == root:
                continue
            l1 = evalf(alpha, omega1, p1, bp, cond_node=n, cond_class=0)
            lp = evalf(alpha, omega1, p1, bp, cond_node=n, cond_class=1)
            if l1 > -1e10 and lp > -1e10:
                e = math.exp(min(max(lp - l1, -50.0), 50.0))
            elif l1 <= -1e10 and lp > -1e10:
                e = 1e6
            else:
                e = 0.0
            ebf[n] = e
            under += e >= 100.0
        res.update(lrt=lrt, p_value=pv, branches_under_selection=under)
    else:
        res.update(lrt=0.0, p_value=2.0 / 3.0)
    res.update(alpha=alpha, beta1=alpha * omega1, p1=p1, beta_plus=bp, p_plus=p_plus, log_l_meme=ll)
    d = 2.0 * (ll - res["log_l_fel"])
    res["p_value_meme_vs_fel"] = math.exp(-d / 2.0) if d > 0.0 else 1.0
    return res, ebf


def meme(prob, code, theta, pos, pi, bl, base_alpha=1.0, base_beta=1.0, delta=0.0, psi=0.0, patterns=None,
         threads=1):
    """MEME's per-pattern results for ``patterns`` (default all), every
    branch tested: FEL's fit of each pattern (fel_ref), then analyze_pattern
    with the candidate oracle, then the total branch length at the fitted
    rates (meme.hpp:614-638).  Returns a list of (result dict, ebf)."""
    scale = fr.mg94_scale(code, theta, pos, pi, base_alpha, base_beta, delta, psi)
    blc, tested = fr.converted_lengths(bl, prob.root, scale)
    pats = list(range(prob.num_patterns)) if patterns is None else list(patterns)

    def one(p):
        inv = fr.is_invariant(prob.states[p])
        fev = fr.CandidateOracle(prob, code, theta, pos, pi, blc, p, delta, psi)
        f = fr.analyze_pattern(fev, prob.states[p], tested)
        ev = MemeCandidateOracle(prob, code, theta, pos, pi, blc, p, delta=delta, psi=psi)
        r, ebf = analyze_pattern(ev, inv, f, prob.num_nodes, prob.root)
        if not inv:
            mu1 = fr.mg94_scale(code, theta, pos, pi, r["alpha"], r["beta1"], delta, psi)
            mup = fr.mg94_scale(code, theta, pos, pi, r["alpha"], r["beta_plus"], delta, psi)
            rate = (r["p1"] * mu1 + r["p_plus"] * mup) / 3.0
            tl = 0.0
            for n in range(prob.num_nodes):
                if n != prob.root:
                    tl += blc[n] * rate
            r["total_branch_length"] = tl
        r["evaluations"] = ev.count
        return r, ebf

    if threads > 1:
        with ThreadPoolExecutor(threads) as ex:
            return list(ex.map(one, pats))
    return [one(p) for p in pats]
