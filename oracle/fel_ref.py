"""Sequential FP64 restatement of FEL's per-site refits (all branches tested).

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's ``cpu_baseline``
leg of ``--fel``, as the checker and the CPU baseline.  The product never
imports this module.

Restated reference semantics (paths relative to the reference root):
  * Brent1D::minimize_point          include/hyphy/opt/nelder_mead.hpp:15-108
  * NelderMead2D::minimize           nelder_mead.hpp:111-205
  * FELAnalyzer::prepare_branch_lengths (no background branches)
                                     include/hyphy/analyses/fel.hpp:227-272
  * compute_pattern_log_likelihood   fel.hpp:276-308
  * analyze_pattern, fast path       fel.hpp:326-445
  * run (pattern loop, site mapping) fel.hpp:778-834

Each function is written as the reference's plain loop: one pattern at a
time, every candidate evaluated by a direct call.  The candidate evaluator
builds MG94 at alpha = 1, beta = omega with the oracle's generator and
eigensolver, P_b = max(U diag(e^{l a t_b}) V, 0), and prunes the single
pattern with the oracle's FP64 pruning.
"""
from __future__ import annotations

import math
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from oracle import hyoracle as ho

START_GRID = ((0.01, 0.1), (1.0, 0.1), (1.0, 0.5), (1.0, 1.0), (1.0, 5.0), (10.0, 0.1), (0.01, 0.5), (0.01, 5.0),
              (10.0, 0.5), (10.0, 1.0), (10.0, 50.0), (100.0, 1.0))  # fel.hpp:387-391
NULL_GRID = (0.01, 0.1, 0.5, 1.0, 2.0, 5.0, 10.0)  # fel.hpp:418


def brent_1d(f, ax, bx, cx, tol=1e-4, max_iter=60):
    """Brent1D::minimize_point (nelder_mead.hpp:24-101): bracketed parabolic
    interpolation / golden section.  Returns (x, f(x), iterations, converged)."""
    CGOLD, ZEPS = 0.3819660, 1e-10
    a, b = min(ax, cx), max(ax, cx)
    v = w = x = bx
    e = 0.0
    fx = f(x)
    fv = fw = fx
    iters, converged = 0, False
    for it in range(max_iter):
        iters = it + 1
        xm = 0.5 * (a + b)
        tol1 = tol * abs(x) + ZEPS
        tol2 = 2.0 * tol1
        if abs(x - xm) <= tol2 - 0.5 * (b - a):
            converged = True
            break
        d = 0.0
        if abs(e) > tol1:
            r = (x - w) * (fx - fv)
            q = (x - v) * (fx - fw)
            p = (x - v) * q - (x - w) * r
            q = 2.0 * (q - r)
            if q > 0.0:
                p = -p
            q = abs(q)
            etemp = e
            e = d
            if abs(p) >= abs(0.5 * q * etemp) or p <= q * (a - x) or p >= q * (b - x):
                e = a - x if x >= xm else b - x
                d = CGOLD * e
            else:
                d = p / q
                u = x + d
                if u - a < tol2 or b - u < tol2:
                    d = abs(tol1) if xm - x >= 0.0 else -abs(tol1)
        else:
            e = a - x if x >= xm else b - x
            d = CGOLD * e
        u = x + d if abs(d) >= tol1 else x + (abs(tol1) if d >= 0.0 else -abs(tol1))
        fu = f(u)
        if fu <= fx:
            if u >= x:
                a = x
            else:
                b = x
            v, fv = w, fw
            w, fw = x, fx
            x, fx = u, fu
        else:
            if u < x:
                a = u
            else:
                b = u
            if fu <= fw or w == x:
                v, fv = w, fw
                w, fw = u, fu
            elif fu <= fv or v == x or v == w:
                v, fv = u, fu
    return x, fx, iters, converged


def nelder_mead_2d(f, init_a, init_b, lb=1e-4, ub=1000.0, tol=1e-4, max_iter=100):
    """NelderMead2D::minimize (nelder_mead.hpp:121-204): box-clamped simplex
    with reflection 1, expansion 2, contraction 0.5, shrink 0.5.  Returns
    (a, b, f, iterations, converged); ties keep the vertices' order (the
    reference sorts three points by insertion)."""

    def ev(a, b):
        ca, cb = min(max(a, lb), ub), min(max(b, lb), ub)
        return (ca, cb, f(ca, cb))

    step_a, step_b = max(0.1, 0.2 * init_a), max(0.1, 0.2 * init_b)
    s = [ev(init_a, init_b), ev(init_a + step_a, init_b), ev(init_a, init_b + step_b)]
    iters, converged = 0, False
    for it in range(max_iter):
        iters = it + 1
        s.sort(key=lambda t: t[2])
        if abs(s[2][2] - s[0][2]) < tol and abs(s[2][0] - s[0][0]) < tol and abs(s[2][1] - s[0][1]) < tol:
            converged = True
            break
        ca, cb = 0.5 * (s[0][0] + s[1][0]), 0.5 * (s[0][1] + s[1][1])
        r = ev(ca + (ca - s[2][0]), cb + (cb - s[2][1]))
        if r[2] < s[0][2]:
            e = ev(ca + 2.0 * (r[0] - ca), cb + 2.0 * (r[1] - cb))
            s[2] = e if e[2] < r[2] else r
        elif r[2] < s[1][2]:
            s[2] = r
        else:
            outside = r[2] < s[2][2]
            c = (ev(ca + 0.5 * (r[0] - ca), cb + 0.5 * (r[1] - cb)) if outside
                 else ev(ca + 0.5 * (s[2][0] - ca), cb + 0.5 * (s[2][1] - cb)))
            ref_f = r[2] if outside else s[2][2]
            if c[2] < ref_f:
                s[2] = c
            else:
                s[1] = ev(s[0][0] + 0.5 * (s[1][0] - s[0][0]), s[0][1] + 0.5 * (s[1][1] - s[0][1]))
                s[2] = ev(s[0][0] + 0.5 * (s[2][0] - s[0][0]), s[0][1] + 0.5 * (s[2][1] - s[0][1]))
    s.sort(key=lambda t: t[2])
    return s[0][0], s[0][1], s[0][2], iters, converged


def single_pattern(prob, p):
    """The one-pattern problem of pattern p (weight 1)."""
    n = prob.num_nodes
    children = [list(prob.child_idx[prob.child_ptr[k]:prob.child_ptr[k + 1]]) for k in range(n)]
    return ho.Problem(prob.parent, children, prob.post_order, prob.is_leaf, prob.leaf_taxon,
                      prob.states[p:p + 1], prob.masks[p:p + 1], [1.0], S=prob.S)


class CandidateOracle:
    """eval_pattern_lnl (fel.hpp:367-385) for one pattern: MG94 at alpha = 1,
    beta = omega = b / max(a, 1e-12) (base theta, delta, psi), rebuilt when
    omega moves by more than 1e-12, P_b(max(a, 1e-12) t_b), and
    compute_pattern_log_likelihood (fel.hpp:276-308)."""

    def __init__(self, prob, code, theta, pos, pi, bl, p, delta=0.0, psi=0.0):
        self.sub = single_pattern(prob, p)
        self.code, self.theta, self.pos, self.pi = code, list(theta), pos, np.asarray(pi, dtype=np.float64)
        self.delta, self.psi = delta, psi
        self.bl = np.asarray(bl, dtype=np.float64)
        self.root = prob.root
        self.cached = -1.0
        self.spec = None
        self.count = 0

    def __call__(self, a, b):
        self.count += 1
        a_safe = max(a, 1e-12)
        omega = b / a_safe
        if abs(omega - self.cached) > 1e-12:
            Q, _ = ho.mg94_Q(self.code, ho.params(1.0, omega, self.theta, self.delta, self.psi), self.pos, self.pi)
            self.spec = ho.decompose(Q, self.pi)
            self.cached = omega
        S = len(self.pi)
        P = np.zeros((len(self.bl), S, S))
        for n in range(len(self.bl)):
            if n != self.root:
                P[n] = ho.transition(*self.spec, a_safe * self.bl[n])
        ll = ho.loglik_from_P(self.sub, P, self.pi)
        self.last = (a_safe, P, ll)
        return ll if math.isfinite(ll) else -1e20

    def rounding_bound(self, rel_eps=64 * 2.220446049250313e-16):
        """Bound on how far the last candidate's lnL can move when every
        entry of every P_b moves by at most d_b = rel_eps (1 + |l_min| a t_b)
        (the rounding of U diag(e^{l a t}) V in FP64, any eigensolver): the
        pattern likelihood is a polynomial in the entries with non-negative
        coefficients, so it moves by at most L(P + d) - L(P).  Sites whose
        likelihood rests on entries near zero (very short or very long times)
        get a wide bound: there the eigen route itself is that inexact."""
        a_safe, P, ll = self.last
        lmin = float(np.max(np.abs(self.spec[0])))
        Pd = P.copy()
        for n in range(len(self.bl)):
            if n != self.root:
                Pd[n] += rel_eps * (1.0 + lmin * a_safe * self.bl[n])
        return ho.loglik_from_P(self.sub, Pd, self.pi) - ll


def mg94_scale(code, theta, pos, pi, alpha=1.0, beta=1.0, delta=0.0, psi=0.0):
    """MG94Matrix::scale_factor: sum_i pi_i (-Q_ii) (rate_matrix.hpp:226-233)."""
    return ho.mg94_Q(code, ho.params(alpha, beta, theta, delta, psi), pos, pi)[1]


def converted_lengths(bl, root, scale):
    """prepare_branch_lengths with every branch tested (fel.hpp:240-262):
    lengths times 3 / Q_scale (1 when Q_scale <= 1e-12); returns the
    converted lengths and the unconverted tested tree length."""
    conv = 3.0 / scale if scale > 1e-12 else 1.0
    bl = np.asarray(bl, dtype=np.float64)
    mask = np.arange(len(bl)) != root
    out = np.where(mask, bl * conv, 0.0)
    return out, float(np.sum(bl[mask]))


def is_invariant(states_row):
    """All non-gap states identical (fel.hpp:333-350)."""
    first = -1
    for s in states_row:
        s = int(s)
        if s >= 0:
            if first == -1:
                first = s
            elif s != first:
                return False
    return True


def analyze_pattern(evalf, states_row, tested_length):
    """FELAnalyzer::analyze_pattern, all branches tested (fel.hpp:326-445)."""
    res = dict(alpha=0.0, beta=0.0, beta_nuisance=0.0, alpha_null=0.0, lrt=0.0, p_value=1.0, total_branch_length=0.0,
               log_l_alt=0.0, log_l_null=0.0)
    if is_invariant(states_row):
        return res
    best_lnl, best_a, best_b = -1e20, 1.0, 1.0
    for a, b in START_GRID:
        lnl = evalf(a, b)
        if lnl > best_lnl:
            best_lnl, best_a, best_b = lnl, a, b
    pa, pb, pf, _, _ = nelder_mead_2d(lambda a, b: -evalf(a, b), best_a, best_b, 1e-4, 1000.0, 1e-4, 100)
    res.update(alpha=pa, beta=pb, log_l_alt=-pf)
    init_null = (min(pa, 100.0) + 3.0 * min(pb, 100.0)) / 4.0
    best_null_lnl, best_null_a = evalf(init_null, init_null), init_null
    for c in NULL_GRID:
        lnl = evalf(c, c)
        if lnl > best_null_lnl:
            best_null_lnl, best_null_a = lnl, c
    bx = min(max(best_null_a, 0.01), 45.0)
    xa, fa, _, _ = brent_1d(lambda a: -evalf(a, a), 1e-4, bx, 50.0, 1e-4, 60)
    res.update(alpha_null=xa, log_l_null=-fa)
    lrt = 0.0
    if res["log_l_alt"] > -1e10 and res["log_l_null"] > -1e10:
        lrt = max(2.0 * (res["log_l_alt"] - res["log_l_null"]), 0.0)
    res.update(lrt=lrt, p_value=math.erfc(math.sqrt(lrt / 2.0)), total_branch_length=xa * tested_length)
    return res


def fel(prob, code, theta, pos, pi, bl, base_alpha=1.0, base_beta=1.0, delta=0.0, psi=0.0, patterns=None,
        threads=1):
    """FEL's per-pattern results for ``patterns`` (default all): a list of
    dicts in pattern order, plus the number of candidate evaluations.  The
    patterns are independent (the reference runs them in an OpenMP dynamic
    loop); ``threads`` runs them on a thread pool (the oracle's C calls
    release the interpreter lock)."""
    scale = mg94_scale(code, theta, pos, pi, base_alpha, base_beta, delta, psi)
    blc, tested = converted_lengths(bl, prob.root, scale)
    pats = list(range(prob.num_patterns)) if patterns is None else list(patterns)

    def one(p):
        ev = CandidateOracle(prob, code, theta, pos, pi, blc, p, delta, psi)
        r = analyze_pattern(ev, prob.states[p], tested)
        r["evaluations"] = ev.count
        return r

    if threads > 1:
        with ThreadPoolExecutor(threads) as ex:
            out = list(ex.map(one, pats))
    else:
        out = [one(p) for p in pats]
    return out
