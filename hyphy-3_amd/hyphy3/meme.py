"""MEME (mixed-effects model of evolution) per-site refits on the MI355X engine.

Follows MEMEAnalyzer with every branch tested (include/hyphy/analyses/
meme.hpp): the fitted tree's lengths are converted by 3 / Q_scale of the
global MG94 model (``prepare_branch_lengths``, :150-163, FEL's conversion);
each pattern first gets FEL's fit (FELAnalyzer::analyze_pattern), then
``analyze_pattern`` (:166-648) fits the two-class branch mixture
p1 P(alpha, alpha omega1) + (1 - p1) P(alpha, beta_plus): a start grid chosen
by the FEL regime, a 4-D ``NelderMeadND`` over (alpha, omega1, p1, beta_plus)
(nelder_mead.hpp:208-360); when beta_plus > alpha and p_plus > 1e-6 the
null beta_plus = alpha as a 3-D fit, a warm alternative refit when the null
is better, the LRT and its mixture p-value 2/3 (0.45 erfc(sqrt(LRT/2)) +
0.55 e^{-LRT/2}), and each branch's empirical Bayes factor (the site's
likelihood with that branch in class + over class 1); the MEME-vs-FEL
p-value and the total branch length at the fitted rates.  ``run``
(:648-700) maps patterns to sites.

The refits are native coroutines (``hy_meme_drive``) advanced in lockstep;
``hy_meme_eval`` evaluates every pending candidate of every pattern in one
GPU launch (three pooled spectra per candidate, one workgroup per candidate
pruning its pattern, EBF conditionals as candidates with one branch held in
one class).
"""
from __future__ import annotations

import time

import numpy as np

from hyphy3.engine import MEME_FIELD_NAMES, meme_drive
from hyphy3.fel import FEL, invariant_patterns


def rows_to_dicts(rows):
    ints = ("evaluations", "branches_under_selection")
    return [{k: (int(v) if k in ints else float(v)) for k, v in zip(MEME_FIELD_NAMES, r)} for r in rows]


class MEME(FEL):
    """MEME on one alignment, every branch tested; the constructor takes
    FEL's arguments (the fitted tree, the global MG94 rates; MEMEAnalyzer's
    constructor, meme.hpp:79-86)."""

    def run(self, patterns=None, evaluate=None, fel_rows=None):
        """Refits every pattern and returns one dict per codon site, with
        "branch_ebf" the per-node EBFs (1 at the root).  With a custom
        ``evaluate(patterns, alphas, omega1s, p1s, beta_pluses, cond_nodes,
        cond_classes)`` over the converted lengths and ``fel_rows`` (the
        patterns' FEL results, [len(patterns), HY_FEL_FIELDS] in
        hy_fel_drive's layout) for a subset of ``patterns``, the native
        driver runs over that evaluator and the results come back as
        {pattern: dict}."""
        t0 = time.perf_counter()
        if patterns is None and evaluate is None:
            r = self.engine.meme_run(self.theta, self.alpha, self.beta, self.bl, self.delta, self.psi)
            self.stats = {"seconds": time.perf_counter() - t0, "calls": r["calls"], "candidates": r["candidates"],
                          "patterns": int(r["sites"].shape[0])}
            res = rows_to_dicts(r["sites"])
            out = []
            for s, p in enumerate(self.site_to_pattern.tolist()):
                d = dict(res[p])
                d["site_index"] = s
                d["branch_ebf"] = r["ebf"][p].tolist()
                out.append(d)
            return out
        if evaluate is None or fel_rows is None:
            raise ValueError("run: a pattern subset needs an evaluator on the converted lengths and FEL rows")
        pats = np.arange(self.states.shape[0]) if patterns is None else np.asarray(patterns, dtype=np.int64)
        inv = invariant_patterns(self.states[pats])

        def sub_evaluate(idx, a, w1, p1, bp, cn, cc):
            return evaluate(pats[idx].astype(np.int32), a, w1, p1, bp, cn, cc)

        sites, ebf, calls, cands = meme_drive(inv, np.asarray(fel_rows, dtype=np.float64), self.bl.size, self.root,
                                              sub_evaluate)
        self.stats = {"seconds": time.perf_counter() - t0, "calls": calls, "candidates": cands,
                      "patterns": int(pats.size)}
        out = {}
        for p, d, e in zip(pats.tolist(), rows_to_dicts(sites), ebf):
            d["branch_ebf"] = e.tolist()
            out[int(p)] = d
        return out
