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
61 x 61 eigendecomposition and a single-pattern pruning.  Here the refits are
native coroutines (``hy_fel_drive``) advanced in lockstep: one evaluator call
takes the pending candidates of every pattern still being refitted, and
``hy_fel_eval`` evaluates them on the GPU (spectra per distinct omega, pooled
and warm-started; one workgroup per candidate prunes its pattern).  Each
pattern sees the reference's sequence of evaluations, so results differ from
a sequential run only through the candidate values.
"""
from __future__ import annotations

import time

import numpy as np

from hyphy3.engine import FEL_FIELD_NAMES, fel_drive


def invariant_patterns(states):
    """Per pattern: all resolved codons the same (fel.hpp:333-350)."""
    st = np.asarray(states)
    out = np.ones(st.shape[0], dtype=np.uint8)
    for p in range(st.shape[0]):
        row = st[p][st[p] >= 0]
        out[p] = 1 if row.size == 0 or np.all(row == row[0]) else 0
    return out


def rows_to_dicts(rows):
    return [{k: (int(v) if k == "evaluations" else float(v)) for k, v in zip(FEL_FIELD_NAMES, r)} for r in rows]


class FEL:
    """FEL on one alignment, every branch tested.

    ``tree`` carries the fitted global MG94 branch lengths (expected
    substitutions per codon), ``theta`` the five exchangeabilities (AC, AT,
    CG, CT, GT) and ``alpha``/``beta`` (with ``delta``/``psi``) the global
    rates, as FELAnalyzer's constructor receives them (fel.hpp:167-200).
    ``engine``: an ``Engine`` already holding this problem and model, with
    ``arrays`` giving ``states``, ``site_to_pattern``, ``bl`` and ``root``;
    or None to build one from ``tree``/``aln`` (hyphy3.Tree,
    hyphy3.Alignment) on ``device``.
    """

    def __init__(self, tree, aln, theta, alpha=1.0, beta=1.0, delta=0.0, psi=0.0, p_value_threshold=0.1,
                 engine=None, device=0, arrays=None):
        if engine is None:
            import hyphy3
            from hyphy3.engine import Engine
            arrays = hyphy3.problem_arrays(tree, aln)
            engine = Engine(device)
            n = len(arrays["parent"])
            cp, ci = arrays["child_ptr"], arrays["child_idx"]
            engine.set_problem(arrays["parent"], [list(ci[cp[i]:cp[i + 1]]) for i in range(n)], arrays["leaf_taxon"],
                               arrays["states"], arrays["masks"], arrays["weights"], S=int(arrays["S"]))
            engine.set_model(int(arrays["code_id"]), arrays["pos_freqs"], arrays["pi"])
            self.site_to_pattern = np.asarray(aln.site_to_pattern)
            self.bl = np.asarray(tree.get_branch_lengths(), dtype=np.float64)
            self.root = int(tree.root_id)
        else:
            self.site_to_pattern = np.asarray(arrays["site_to_pattern"])
            self.bl = np.asarray(arrays["bl"], dtype=np.float64)
            self.root = int(arrays["root"])
        self.states = np.asarray(arrays["states"])
        self.engine = engine
        self.theta = np.asarray(theta, dtype=np.float64)
        self.alpha, self.beta = float(alpha), float(beta)
        self.delta, self.psi = float(delta), float(psi)
        self.p_value_threshold = p_value_threshold
        mask = np.arange(self.bl.size) != self.root
        self.baseline_tree_length = float(np.sum(self.bl[mask]))
        self.stats = {}

    def run(self, patterns=None, evaluate=None):
        """Refits every pattern and returns one dict per codon site
        (FELAnalyzer::run, fel.hpp:778-834).  With a custom
        ``evaluate(patterns, alphas, betas)`` (candidates on the converted
        lengths, e.g. a CPU evaluator) and optionally a subset of ``patterns``,
        the native driver runs over that evaluator instead and the
        per-pattern results come back as {pattern: dict}."""
        t0 = time.perf_counter()
        if patterns is None and evaluate is None:
            rows, st = self.engine.fel_run(self.theta, self.alpha, self.beta, self.bl, self.delta, self.psi)
            self.q_scale = st["q_scale"]
            self.stats = {"seconds": time.perf_counter() - t0, "calls": st["calls"], "candidates": st["candidates"],
                          "patterns": int(rows.shape[0])}
            res = rows_to_dicts(rows)
            out = []
            for s, p in enumerate(self.site_to_pattern.tolist()):
                r = dict(res[p])
                r["site_index"] = s
                out.append(r)
            return out
        pats = np.arange(self.states.shape[0]) if patterns is None else np.asarray(patterns, dtype=np.int64)
        if evaluate is None:
            raise ValueError("run: a pattern subset needs an evaluator built for the converted lengths")
        inv = invariant_patterns(self.states[pats])

        def sub_evaluate(idx, a, b):
            return evaluate(pats[idx].astype(np.int32), a, b)

        rows, calls, cands = fel_drive(inv, self.baseline_tree_length, sub_evaluate)
        self.stats = {"seconds": time.perf_counter() - t0, "calls": calls, "candidates": cands,
                      "patterns": int(pats.size)}
        return {int(p): r for p, r in zip(pats.tolist(), rows_to_dicts(rows))}
