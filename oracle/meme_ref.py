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
"""
from __future__ import annotations

import math

import numpy as np

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

    def _spec(self, alpha, beta):
        Q, _ = ho.mg94_Q(self.code, ho.params(alpha, beta, self.theta, self.delta, self.psi), self.pos, self.pi)
        return ho.decompose(Q, self.pi)

    def __call__(self, a, w1, p1, b_plus, b_bg=0.0):
        s1 = self._spec(a, a * w1)
        sp = self._spec(a, b_plus)
        has_bg = not bool(np.all(self.is_test[[n for n in range(len(self.bl)) if n != self.root]]))
        sbg = self._spec(a, b_bg) if has_bg else None
        S = len(self.pi)
        P = np.zeros((len(self.bl), S, S))
        for n in range(len(self.bl)):
            if n == self.root:
                continue
            if self.is_test[n]:
                P[n] = p1 * ho.transition(*s1, self.bl[n]) + (1.0 - p1) * ho.transition(*sp, self.bl[n])
            else:
                P[n] = ho.transition(*sbg, self.bl[n])
        ll = ho.loglik_from_P(self.sub, P, self.pi)
        return ll if math.isfinite(ll) else -1e20
