"""FEL per-site refits, CPU side (no GPU):

  * the oracle's optimisers (Brent1D, NelderMead2D restatements) on
    functions whose minima are known, including minima outside the box;
  * the native lockstep driver (hy_fel_drive) on synthetic per-pattern
    objectives asks for exactly the candidates the oracle's plain-loop
    analyze_pattern evaluates, in the same order, and returns the same
    results;
  * the oracle's candidate evaluator against the oracle's whole-alignment
    MG94 pruning (site log-likelihoods at the same rates) and against a
    brute-force sum over ancestral codons with Taylor-series P;
  * the native driver fed the oracle's candidates reproduces the oracle's
    sequential FEL on CD2 bit for bit;
  * the oracle's FEL on CD2 against HyPhy 2.5's per-site table at the
    thresholds of the reference's own parity test
    (tests/test_fel_cd2_parity.cpp:67-120).
"""
import json
import math
import os

import numpy as np
import pytest

from helpers import GOLDEN, branch_lengths
from oracle import data_ref as dr
from oracle import fel_ref as fr
from oracle import hyoracle as ho
from hyphy3 import fel as pf
from hyphy3.engine import fel_drive


def recorded(f):
    pts = []

    def g(*x):
        pts.append(x if len(x) > 1 else x[0])
        return f(*x)
    return g, pts


BRENT_CASES = [
    (lambda x: (x - 2.0) ** 2 + 1.0, 2.0, (1e-4, 0.01, 50.0)),
    (lambda x: abs(x - 7.5), 7.5, (1e-4, 1.0, 50.0)),
    (lambda x: x, 1e-4, (1e-4, 10.0, 50.0)),  # minimum on the bracket's edge
    (lambda x: -math.log(x) + 0.3 * x, 1.0 / 0.3, (1e-4, 45.0, 50.0)),
]


@pytest.mark.parametrize("case", range(len(BRENT_CASES)))
def test_brent_minima(case):
    f, xmin, (ax, bx, cx) = BRENT_CASES[case]
    g, pts = recorded(f)
    x, fx, it, conv = fr.brent_1d(g, ax, bx, cx, 1e-4, 60)
    assert abs(x - xmin) <= 2e-4 * max(1.0, xmin) + 1e-9, (x, xmin)
    assert fx == f(x) and all(min(ax, cx) <= p <= max(ax, cx) for p in pts)


NM_CASES = [
    (lambda a, b: (a - 3.0) ** 2 + 2.0 * (b - 7.0) ** 2 + 0.5 * (a - 3.0) * (b - 7.0), (3.0, 7.0), (1.0, 1.0)),
    (lambda a, b: (a + 5.0) ** 2 + (b - 2.0) ** 2, (1e-4, 2.0), (1.0, 0.5)),  # a pinned at the lower bound
    (lambda a, b: 0.01 * (a - 40.0) ** 2 + (b - 3.0) ** 2, (40.0, 3.0), (10.0, 1.0)),
    (lambda a, b: abs(a - 0.4) + abs(b - 0.9), (0.4, 0.9), (0.01, 5.0)),
]


@pytest.mark.parametrize("case", range(len(NM_CASES)))
def test_nelder_mead_minima(case):
    f, (amin, bmin), (a0, b0) = NM_CASES[case]
    g, pts = recorded(f)
    a, b, fv, it, conv = fr.nelder_mead_2d(g, a0, b0, 1e-4, 1000.0, 1e-4, 100)
    assert conv
    assert abs(a - amin) < 5e-3 * max(1.0, amin) and abs(b - bmin) < 5e-3 * max(1.0, bmin), (a, b)
    assert all(1e-4 <= p[0] <= 1000.0 and 1e-4 <= p[1] <= 1000.0 for p in pts)
    assert fv == f(a, b)


def test_nelder_mead_collapsed_simplex():
    """A reflection clamped to the box can flatten the simplex onto a bound,
    after which the reference's method stays there (here b = 1e-4, not the
    minimum's 0.3)."""
    f = lambda a, b: 0.01 * (a - 40.0) ** 2 + (b - 0.3) ** 2  # noqa: E731
    res = fr.nelder_mead_2d(f, 10.0, 1.0, 1e-4, 1000.0, 1e-4, 100)
    assert res[1] == 1e-4


SYNTH = [  # per-pattern log-likelihood surfaces (alternative optimum, null curvature, floor)
    lambda a, b: -((math.log(a) - 0.7) ** 2 + 2.0 * (math.log(b) - 1.9) ** 2) - 20.0,
    lambda a, b: -0.01 * (a - 40.0) ** 2 - (b - 0.3) ** 2 - 5.0,  # collapses onto b = 1e-4
    lambda a, b: -abs(a - 0.4) - abs(b - 0.9) - 3.0,
    lambda a, b: -((a - 2000.0) ** 2) * 1e-6 - (b - 3.0) ** 2,  # alpha beyond the box
    lambda a, b: -1e20 if a > 50.0 else -((a - 0.02) ** 2 + (b - 0.02) ** 2) * 100.0,  # sentinel values
    lambda a, b: -(math.log(b + 1e-3) - math.log(a + 1e-3)) ** 2 - 0.3 * a - 7.0,
]


def test_native_driver_matches_oracle_sequence():
    """hy_fel_drive against the oracle's analyze_pattern on synthetic
    objectives: the same candidates in the same order, the same results."""
    asked = {p: [] for p in range(len(SYNTH) + 1)}

    def evaluate(ps, a, b):
        out = np.empty(len(ps))
        for i, p in enumerate(ps.tolist()):
            asked[p].append((float(a[i]), float(b[i])))
            out[i] = SYNTH[p](float(a[i]), float(b[i]))
        return out

    inv = np.zeros(len(SYNTH) + 1, dtype=np.uint8)
    inv[-1] = 1  # an invariant pattern: never evaluated
    rows, calls, cands = fel_drive(inv, 2.5, evaluate)
    assert asked[len(SYNTH)] == [] and rows[-1][4] == 0.0 and rows[-1][5] == 1.0
    assert cands == sum(len(v) for v in asked.values())
    for p, f in enumerate(SYNTH):
        seq = []

        def g(a, b, f=f, seq=seq):
            seq.append((a, b))
            return f(a, b)
        ref = fr.analyze_pattern(g, np.array([0, 1]), 2.5)
        assert asked[p] == seq, p
        r = dict(zip(("alpha", "beta", "beta_nuisance", "alpha_null", "lrt", "p_value", "total_branch_length",
                      "log_l_alt", "log_l_null"), rows[p][:9]))
        for k in ("alpha", "beta", "alpha_null", "lrt", "total_branch_length", "log_l_alt", "log_l_null"):
            assert r[k] == ref[k], (p, k, r[k], ref[k])
        assert abs(r["p_value"] - ref["p_value"]) <= 1e-15 * max(ref["p_value"], 1e-300), p
        assert rows[p][9] == len(seq)


def test_nelder_mead_converges_on_smooth_bowl():
    f = NM_CASES[0][0]
    a, b, fv, it, conv = fr.nelder_mead_2d(f, 1.0, 1.0, 1e-4, 1000.0, 1e-4, 100)
    assert conv and it < 100


def small_problem(seed, n_codons=12):
    rng = np.random.default_rng(seed)
    code = dr.Code()
    names = ["A", "B", "C", "D"]
    seqs = []
    for _ in names:
        cod = rng.integers(0, code.S, size=n_codons)
        s = "".join("ACGT"[(code.sense[c] >> 4) & 3] + "ACGT"[(code.sense[c] >> 2) & 3] + "ACGT"[code.sense[c] & 3]
                    for c in cod)
        seqs.append(s)
    seqs[1] = "NNN" + seqs[1][3:6] + "AYG" + seqs[1][9:]  # a gap and an ambiguous codon
    aln = dr.AlignmentRef(names, seqs)
    tree = dr.TreeRef.from_newick("((A:0.11,B:0.27):0.08,C:0.31,D:0.05);")
    return tree, aln


def test_candidate_matches_whole_alignment_pruning():
    tree, aln = small_problem(5)
    prob = ho.Problem.from_refs(tree, aln)
    code = ho.Code(0)
    theta = [0.6, 0.3, 0.5, 1.2, 0.35]
    bl = branch_lengths(tree) * 2.7
    for a, b in ((1.0, 1.0), (0.01, 5.0), (10.0, 0.1), (1e-4, 1000.0), (0.0, 0.4)):
        a_safe = max(a, 1e-12)
        _, site = ho.mg94_loglik(prob, code, ho.params(1.0, b / a_safe, theta), aln.pos_freqs, aln.cf3x4, bl,
                                 branch_scale=a_safe, site=True)
        for p in range(prob.num_patterns):
            ev = fr.CandidateOracle(prob, code, theta, aln.pos_freqs, aln.cf3x4, bl, p)
            v = ev(a, b)
            assert abs(v - site[p]) <= 1e-12 * max(1.0, abs(site[p])), (a, b, p, v, site[p])


def test_candidate_brute_force():
    """Four taxa, one internal node below the root: L = sum over the root
    codon r and the internal codon u of pi_r P_u[r, u] P_C[r, x_C] P_D[r, x_D]
    P_A[u, x_A] P_B[u, x_B], P = exp(Q a t) by Taylor series, Q the MG94
    generator at alpha = 1, beta = b / a."""
    tree, aln = small_problem(9, n_codons=6)
    prob = ho.Problem.from_refs(tree, aln)
    code = ho.Code(0)
    theta = [0.9, 0.4, 0.7, 1.8, 0.5]
    bl = branch_lengths(tree)
    pi = aln.cf3x4
    l2t = tree.leaf_to_taxon(aln.names)
    root = tree.nodes[tree.root]
    u = [c for c in root.children if not tree.nodes[c].is_leaf][0]
    for a, b in ((0.7, 2.3), (4.0, 0.05)):
        Q, _ = ho.mg94_Q(code, ho.params(1.0, b / a, theta), aln.pos_freqs, pi)
        P = {n.id: ho.expm_taylor(Q, a * bl[n.id]) for n in tree.nodes if n.id != tree.root}
        for p in range(prob.num_patterns):
            x = {}
            for n in list(tree.nodes[u].children) + [c for c in root.children if c != u]:
                st, m = int(aln.states[p, l2t[n]]), int(aln.pmasks[p, l2t[n]])
                x[n] = np.eye(code.S)[st] if st >= 0 else np.array([(m >> i) & 1 for i in range(code.S)], float)
            lu = tree.nodes[u].children
            lr = [c for c in root.children if c != u]
            a_u = (P[lu[0]] @ x[lu[0]]) * (P[lu[1]] @ x[lu[1]])
            c_r = (P[lr[0]] @ x[lr[0]]) * (P[lr[1]] @ x[lr[1]])
            L = math.fsum((pi[:, None] * c_r[:, None] * P[u] * a_u[None, :]).ravel())
            v = fr.CandidateOracle(prob, code, theta, aln.pos_freqs, pi, bl, p)(a, b)
            # eigen-based P carries ~1e-16 absolute error per entry, Taylor P
            # another; sites of likelihood ~1e-15 see ~1e-9 relative differences
            assert abs(v - math.log(L)) < 1e-8, (a, b, p)


def test_scale_is_mean_rate():
    tree, aln = small_problem(3)
    theta = [0.6, 0.3, 0.5, 1.2, 0.35]
    Q, _ = ho.mg94_Q(ho.Code(0), ho.params(1.0, 1.7, theta), aln.pos_freqs, aln.cf3x4)
    mu = fr.mg94_scale(ho.Code(0), theta, aln.pos_freqs, aln.cf3x4, 1.0, 1.7)
    assert abs(mu - float(-np.sum(aln.cf3x4 * np.diag(Q)))) < 1e-14 * mu
    # Q is linear in (alpha, beta): scale(alpha, beta) = alpha scale(1, beta / alpha)
    mu2 = fr.mg94_scale(ho.Code(0), theta, aln.pos_freqs, aln.cf3x4, 2.5, 1.7 * 2.5)
    assert abs(mu2 - 2.5 * mu) < 1e-13 * mu2


def load_cd2_fel():
    names, seqs = dr.parse_fasta(open(os.path.join(GOLDEN, "cd2.fna")).read())
    aln = dr.AlignmentRef(names, seqs)
    tree = dr.TreeRef.from_newick(open(os.path.join(GOLDEN, "cd2_fitted.nwk")).read())
    truth = json.load(open(os.path.join(GOLDEN, "cd2_fel_mle.json")))
    return tree, aln, truth


def oracle_evaluator(prob, code, theta, pos, pi, blc):
    cache = {}

    def evaluate(ps, alphas, betas):
        out = np.empty(len(ps))
        for i, p in enumerate(ps.tolist()):
            if p not in cache:
                cache[p] = fr.CandidateOracle(prob, code, theta, pos, pi, blc, p)
            out[i] = cache[p](float(alphas[i]), float(betas[i]))
        return out
    return evaluate


def test_lockstep_driver_reproduces_sequential_oracle():
    tree, aln, truth = load_cd2_fel()
    g = truth["global_mg94"]
    prob = ho.Problem.from_refs(tree, aln)
    code = ho.Code(0)
    bl = branch_lengths(tree)
    pats = [0, 1, 2, 3, 5, 8, 13, 40, 77, 120]
    seq = fr.fel(prob, code, g["theta"], aln.pos_freqs, aln.cf3x4, bl, g["alpha"], g["beta"], patterns=pats)
    arrays = dict(states=aln.states, site_to_pattern=aln.site_to_pattern, bl=bl, root=tree.root)
    f = pf.FEL(None, None, g["theta"], g["alpha"], g["beta"], engine="no engine: candidates come from the oracle",
               arrays=arrays)
    scale = fr.mg94_scale(code, g["theta"], aln.pos_freqs, aln.cf3x4, g["alpha"], g["beta"])
    blc, tested = fr.converted_lengths(bl, tree.root, scale)
    assert f.baseline_tree_length == tested
    np.testing.assert_array_equal(pf.invariant_patterns(aln.states[pats]),
                                  [fr.is_invariant(aln.states[p]) for p in pats])
    res = f.run(patterns=pats, evaluate=oracle_evaluator(prob, code, g["theta"], aln.pos_freqs, aln.cf3x4, blc))
    assert f.stats["candidates"] == sum(r["evaluations"] for r in seq)
    for p, r in zip(pats, seq):
        for k in ("alpha", "beta", "alpha_null", "lrt", "total_branch_length", "log_l_alt", "log_l_null"):
            assert res[p][k] == r[k], (p, k, res[p][k], r[k])
        assert abs(res[p]["p_value"] - r["p_value"]) <= 1e-15 * max(r["p_value"], 1e-300)
        assert res[p]["evaluations"] == r["evaluations"]


def correlations(x, y):
    from scipy.stats import pearsonr, spearmanr
    return pearsonr(x, y)[0], spearmanr(x, y)[0]


@pytest.mark.timeout(600)
def test_oracle_fel_cd2_matches_hyphy_ground_truth():
    """The oracle's FEL on CD2 with the fitted tree and global MG94 rates of
    the reference's matched-tree parity test, against HyPhy 2.5's per-site
    LRT and p-values at that test's thresholds."""
    tree, aln, truth = load_cd2_fel()
    g = truth["global_mg94"]
    prob = ho.Problem.from_refs(tree, aln)
    res = fr.fel(prob, ho.Code(0), g["theta"], aln.pos_freqs, aln.cf3x4, branch_lengths(tree), g["alpha"],
                 g["beta"], threads=8)
    mle = truth["mle"]
    assert len(mle) == aln.num_codons == 187
    lrt = np.array([res[aln.site_to_pattern[s]]["lrt"] for s in range(aln.num_codons)])
    pv = np.array([res[aln.site_to_pattern[s]]["p_value"] for s in range(aln.num_codons)])
    th = truth["thresholds"]
    r_lrt, rho_lrt = correlations(lrt, [m[3] for m in mle])
    r_p, rho_p = correlations(pv, [m[4] for m in mle])
    print(f"LRT r {r_lrt:.5f} rho {rho_lrt:.5f}, p r {r_p:.5f} rho {rho_p:.5f}")
    assert r_lrt > th["lrt_pearson"] and rho_lrt > th["lrt_spearman"]
    assert r_p > th["p_pearson"] and rho_p > th["p_spearman"]
    # invariant sites: zero rates, p = 1 (fel.hpp:352-361); HyPhy reports zeros too
    for s in range(aln.num_codons):
        r = res[aln.site_to_pattern[s]]
        if fr.is_invariant(prob.states[aln.site_to_pattern[s]]):
            assert r["alpha"] == r["beta"] == r["lrt"] == 0.0 and r["p_value"] == 1.0
            assert mle[s][0] == mle[s][1] == 0
