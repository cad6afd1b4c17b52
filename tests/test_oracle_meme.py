"""Pins of the oracle's MEME (oracle/meme_ref.py) and the native MEME driver:

  * the per-site likelihood (meme.hpp:219-277) against the FEL candidate
    oracle (fel.hpp:367-385, pinned in test_fel_cpu.py), which builds its
    generators at alpha = 1 and scales time instead: with p1 = 1 MEME is FEL
    at (a, a w1); with p1 = 0 it is FEL at (a, b_plus); background branches
    at b_bg = b_plus change nothing when p1 = 0;
  * NelderMeadND (nelder_mead.hpp:208-360) on functions with known minima,
    inside the box and beyond it, and its start (grid choice, first vertices,
    steps down at the upper bound);
  * the mixture p-value against scipy's chi-square tails;
  * the EBF conditionals against the linearity identity
    p1 L_1 + (1 - p1) L_+ = L for every branch;
  * the lockstep native driver (hy_meme_drive) against the oracle's plain-loop
    analyze_pattern: identical candidate sequences and results on synthetic
    sites, and bit-identical results with the oracle's real evaluator;
  * (opt-in, slow) the oracle's whole MEME on CD2 against HyPhy 2.5's table
    at the reference's parity thresholds."""
import numpy as np

from helpers import branch_lengths, sim_problem
from oracle import fel_ref as fr
from oracle import hyoracle as ho
from oracle import meme_ref as mr


def test_meme_reduces_to_fel():
    tree, aln, sim = sim_problem(9, 30, 91, gaps=0.1)
    prob = ho.Problem.from_refs(tree, aln)
    bl = branch_lengths(tree)
    code = ho.Code(0)
    rng = np.random.default_rng(4)
    is_test = rng.random(len(bl)) < 0.6
    for p in (0, 5, 11):
        m = mr.MemeCandidateOracle(prob, code, sim["theta"], aln.pos_freqs, aln.cf3x4, bl, p)
        mb = mr.MemeCandidateOracle(prob, code, sim["theta"], aln.pos_freqs, aln.cf3x4, bl, p, is_test=is_test)
        f = fr.CandidateOracle(prob, code, sim["theta"], aln.pos_freqs, aln.cf3x4, bl, p)
        for a, w1, bp in ((0.8, 0.3, 2.5), (1.7, 0.9, 0.4), (0.05, 0.0, 6.0)):
            fa = f(a, a * w1)
            assert abs(m(a, w1, 1.0, bp) - fa) < 1e-10 * abs(fa), (p, a)
            fb = f(a, bp)
            assert abs(m(a, w1, 0.0, bp) - fb) < 1e-10 * abs(fb), (p, a)
            assert abs(mb(a, w1, 0.0, bp, bp) - fb) < 1e-10 * abs(fb), (p, a)


# ---------------------------------------------------------------------------
# NelderMeadND, the p-value, the EBF conditionals and the MEME driver
# ---------------------------------------------------------------------------
import math  # noqa: E402
import os  # noqa: E402

import pytest  # noqa: E402

from hyphy3.engine import HY_FEL_FIELDS, MEME_FIELD_NAMES, meme_drive  # noqa: E402

NM_ND_CASES = [
    # (f, minimum inside the box or on its face, init, lb, ub)
    (lambda x: (x[0] - 3.0) ** 2 + 2.0 * (x[1] - 0.4) ** 2 + 3.0 * (x[2] - 0.7) ** 2 + 0.5 * (x[3] - 9.0) ** 2
     + 0.3 * (x[0] - 3.0) * (x[3] - 9.0), (3.0, 0.4, 0.7, 9.0), (1.0, 0.5, 0.5, 1.0), mr.LB_ALT, mr.UB_ALT),
    (lambda x: (x[0] + 1.0) ** 2 + (x[1] - 2.0) ** 2 + (x[2] - 0.5) ** 2, (1e-4, 1.0, 0.5), (2.0, 0.2, 0.2),
     mr.LB_NULL, mr.UB_NULL),  # minimum outside the box: the nearest face
    (lambda x: 100.0 * (x[1] - x[0] ** 2) ** 2 + (1.0 - x[0]) ** 2 + x[2] ** 2, (1.0, 1.0, 1e-6), (0.3, 0.3, 0.3),
     (1e-4, 0.0, 1e-6), (10.0, 10.0, 1.0)),
]


@pytest.mark.parametrize("case", range(len(NM_ND_CASES)))
def test_nelder_mead_nd_minima(case):
    f, xmin, init, lb, ub = NM_ND_CASES[case]
    pts = []

    def g(x):
        pts.append(tuple(x))
        return f(x)
    x, fx, it, conv = mr.nelder_mead_nd(g, init, lb, ub, 1e-10, 4000)
    assert conv, (case, it)
    assert all(lb[d] <= p[d] <= ub[d] for p in pts for d in range(len(lb)))
    assert max(abs(x[d] - xmin[d]) for d in range(len(x))) < 1e-3, (x, xmin)
    assert fx == f(x)


def test_nelder_mead_nd_grid_start():
    # the best grid point (the first strictly lower) starts the simplex: its
    # D perturbed vertices are the next D evaluations
    pts = []

    def g(x):
        pts.append(tuple(x))
        return (x[0] - 5.0) ** 2 + (x[1] - 0.5) ** 2 + (x[2] - 0.5) ** 2
    grid = [(1.0, 0.1, 0.1), (5.0, 0.5, 0.5), (5.0, 0.5, 0.5), (9.0, 0.9, 0.9)]
    mr.nelder_mead_nd(g, (1.0, 0.1, 0.1), mr.LB_NULL, mr.UB_NULL, 1e-5, 10, grid)
    assert pts[:5] == [(1.0, 0.1, 0.1)] + grid
    # vertex 1: 5 + max(0.05, 1.0); vertex 2: 0.5 + 0.1; vertex 3: 0.5 + 0.1
    assert pts[5] == (6.0, 0.5, 0.5) and pts[6] == (5.0, 0.6, 0.5) and pts[7] == (5.0, 0.5, 0.6)
    # at the upper bound the step goes down
    pts.clear()
    mr.nelder_mead_nd(g, (5.0, 1.0, 1.0 - 1e-6), mr.LB_NULL, mr.UB_NULL, 1e-5, 1)
    assert pts[2] == (5.0, 1.0 - 0.2, 1.0 - 1e-6) and pts[3][2] == pytest.approx(1.0 - 1e-6 - 0.2)


def test_p_value_is_the_chi2_mixture():
    # 2/3 [0.45 P(chi2_1 > LRT) + 0.55 P(chi2_2 > LRT)]
    from scipy.stats import chi2
    for lrt in (0.0, 1e-8, 0.3, 1.0, 2.7, 5.99, 12.0, 40.0):
        want = (2.0 / 3.0) * (0.45 * chi2.sf(lrt, 1) + 0.55 * chi2.sf(lrt, 2))
        assert abs(mr.p_value(lrt) - want) < 1e-14 * max(want, 1e-300) + 1e-300, lrt
    assert mr.p_value(0.0) == pytest.approx(2.0 / 3.0, abs=1e-16)


def test_ebf_conditionals_mix_to_the_site_likelihood():
    # the pattern likelihood is linear in one branch's matrix: conditioning
    # branch b on each class and mixing with (p1, 1 - p1) gives the site's
    # likelihood back, for every branch
    tree, aln, sim = sim_problem(7, 24, 95, gaps=0.1)
    prob = ho.Problem.from_refs(tree, aln)
    bl = branch_lengths(tree)
    code = ho.Code(0)
    for p in (0, 3, 8):
        m = mr.MemeCandidateOracle(prob, code, sim["theta"], aln.pos_freqs, aln.cf3x4, bl, p)
        a, w1, p1, bp = 0.9, 0.35, 0.7, 4.0
        lmix = m(a, w1, p1, bp)
        for n in range(prob.num_nodes):
            if n == prob.root:
                continue
            l1 = m(a, w1, p1, bp, cond_node=n, cond_class=0)
            lp = m(a, w1, p1, bp, cond_node=n, cond_class=1)
            mix = math.log(p1 * math.exp(l1 - lmix) + (1.0 - p1) * math.exp(lp - lmix)) + lmix
            assert abs(mix - lmix) < 1e-11 * abs(lmix), (p, n)


def synthetic_site(p):
    """A smooth per-pattern objective standing in for a site's lnL, with the
    conditional likelihoods linear in the conditioned class as a pattern's
    are."""
    rng = np.random.default_rng(100 + p)
    c = rng.uniform(0.2, 3.0, 4)
    ctr = (rng.uniform(0.1, 5.0), rng.uniform(0.0, 1.0), rng.uniform(0.05, 0.95), rng.uniform(0.5, 30.0))

    def f(a, w1, p1, bp, cond_node=-1, cond_class=0):
        v = -(c[0] * math.log(a / ctr[0]) ** 2 + c[1] * (w1 - ctr[1]) ** 2 + c[2] * (p1 - ctr[2]) ** 2
              + c[3] * 0.01 * math.log(bp / ctr[3]) ** 2) - 10.0 - p
        if cond_node >= 0:
            v += (0.1 * cond_node - 0.3) * (1.0 if cond_class else -1.0) + (1 + p) * math.log(bp / a) * cond_class
        return v
    return f


def fel_row(d):
    return [d["alpha"], d["beta"], d.get("beta_nuisance", 0.0), d["alpha_null"], d["lrt"], d["p_value"],
            d["total_branch_length"], d["log_l_alt"], d["log_l_null"], 0.0]


def compare_rows(got, ebf_got, want, ebf_want, exact=True):
    for k, name in enumerate(MEME_FIELD_NAMES):
        if name in ("total_branch_length", "evaluations"):
            continue
        if exact:
            assert got[k] == want[name], (name, got[k], want[name])
        else:
            assert abs(got[k] - want[name]) <= 1e-12 * max(1.0, abs(want[name])), (name, got[k], want[name])
    assert list(ebf_got) == list(ebf_want)


def test_native_driver_matches_oracle_sequence():
    # every pattern's candidate sequence and result from the lockstep native
    # driver equal the oracle's plain-loop analyze_pattern, for both start
    # regimes, an invariant pattern, and a site whose null beats the
    # alternative's first fit
    num_nodes, root = 7, 0
    fel_rows, inv, want, seqs = [], [], [], []
    cases = [(0.5, 2.0), (2.0, 0.5), (1.0, 1.0), (0.01, 30.0), (0.0, 0.0), (3.0, 3.5)]
    for p, (fa, fb) in enumerate(cases):
        d = dict(alpha=fa, beta=fb, alpha_null=0.5 * (fa + fb), lrt=0.0, p_value=1.0, total_branch_length=0.0,
                 log_l_alt=-12.0 - p, log_l_null=-13.0 - p)
        fel_rows.append(fel_row(d))
        inv.append(1 if p == 4 else 0)
        f = synthetic_site(p)
        seq = []

        def rec(a, w1, p1, bp, cond_node=-1, cond_class=0, f=f, seq=seq):
            seq.append((a, w1, p1, bp, cond_node, cond_class))
            return f(a, w1, p1, bp, cond_node=cond_node, cond_class=cond_class)
        r, ebf = mr.analyze_pattern(rec, bool(inv[-1]), d, num_nodes, root)
        want.append((r, ebf))
        seqs.append(seq)
    got_seq = [[] for _ in cases]

    def evaluate(pat, a, w1, p1, bp, cn, cc):
        out = np.zeros(len(pat))
        for k in range(len(pat)):
            p = int(pat[k])
            got_seq[p].append((a[k], w1[k], p1[k], bp[k], int(cn[k]), int(cc[k])))
            out[k] = synthetic_site(p)(a[k], w1[k], p1[k], bp[k], cond_node=int(cn[k]), cond_class=int(cc[k]))
        return out
    sites, ebf, calls, cands = meme_drive(inv, np.array(fel_rows), num_nodes, root, evaluate)
    for p in range(len(cases)):
        assert got_seq[p] == seqs[p], p
        compare_rows(sites[p], ebf[p], want[p][0], want[p][1])
        assert sites[p][MEME_FIELD_NAMES.index("evaluations")] == len(seqs[p])
    assert cands == sum(len(s) for s in seqs)
    assert calls < cands / 2  # lockstep: many candidates per call
    assert sites[4][MEME_FIELD_NAMES.index("p1")] == 1.0 and sites[4][MEME_FIELD_NAMES.index("p_value")] == 1.0
    assert sites[:, MEME_FIELD_NAMES.index("branches_under_selection")].max() > 0  # EBFs of 100 and more
    assert (sites[:, MEME_FIELD_NAMES.index("lrt")] > 0.0).sum() >= 2


def test_native_driver_reproduces_sequential_oracle():
    # the real candidate evaluator (the oracle's) under the native driver:
    # the oracle's sequential MEME bit for bit, EBFs included
    tree, aln, sim = sim_problem(6, 40, 96, omegas=(0.1, 3.0), weights=(0.7, 0.3))
    prob = ho.Problem.from_refs(tree, aln)
    bl = branch_lengths(tree)
    code = ho.Code(0)
    th = sim["theta"]
    pats = [p for p in range(prob.num_patterns) if not fr.is_invariant(prob.states[p])][:3] + \
        [p for p in range(prob.num_patterns) if fr.is_invariant(prob.states[p])][:1]
    ref = mr.meme(prob, code, th, aln.pos_freqs, aln.cf3x4, bl, 1.0, 0.5, patterns=pats, threads=4)
    scale = fr.mg94_scale(code, th, aln.pos_freqs, aln.cf3x4, 1.0, 0.5)
    blc, tested = fr.converted_lengths(bl, prob.root, scale)
    fel = []
    for p in pats:
        fel.append(fel_row(fr.analyze_pattern(fr.CandidateOracle(prob, code, th, aln.pos_freqs, aln.cf3x4, blc, p),
                                              prob.states[p], tested)))
    evs = [mr.MemeCandidateOracle(prob, code, th, aln.pos_freqs, aln.cf3x4, blc, p) for p in pats]

    def evaluate(pat, a, w1, p1, bp, cn, cc):
        return [evs[int(pat[k])](a[k], w1[k], p1[k], bp[k], cond_node=int(cn[k]), cond_class=int(cc[k]))
                for k in range(len(pat))]
    inv = [1 if fr.is_invariant(prob.states[p]) else 0 for p in pats]
    sites, ebf, calls, cands = meme_drive(inv, np.array(fel), prob.num_nodes, prob.root, evaluate)
    for i in range(len(pats)):
        compare_rows(sites[i], ebf[i], ref[i][0], ref[i][1])
        assert sites[i][MEME_FIELD_NAMES.index("evaluations")] == (0 if inv[i] else ref[i][0]["evaluations"])
    assert any(r[0]["lrt"] > 0.0 for r in ref) or any(r[0]["p_value"] < 1.0 for r in ref)


def test_driver_refusals():
    with pytest.raises(Exception):
        meme_drive([0], np.zeros((1, HY_FEL_FIELDS)), 5, 7, lambda *c: np.zeros(len(c[0])))  # root out of range


@pytest.mark.skipif(os.environ.get("HY_SLOW_ORACLE") != "1",
                    reason="about 12 CPU-minutes (8 processes, OMP_NUM_THREADS=1): set HY_SLOW_ORACLE=1 to run")
def test_oracle_meme_cd2_matches_hyphy_ground_truth():
    """The oracle's MEME on CD2 (the fitted tree and global MG94 rates of the
    reference's matched-tree FEL parity case) against HyPhy 2.5's per-site
    MEME LRT and p-values at the reference's MEME parity thresholds
    (tests/test_meme.cpp:66-103).  Measured: LRT r 0.992, rho 0.967;
    p-value r 0.994, rho 0.987."""
    import json
    from concurrent.futures import ProcessPoolExecutor

    from helpers import GOLDEN
    from test_fel_cpu import correlations, load_cd2_fel
    tree, aln, truth = load_cd2_fel()
    g = truth["global_mg94"]
    prob = ho.Problem.from_refs(tree, aln)
    chunks = [list(range(i, prob.num_patterns, 8)) for i in range(8)]
    with ProcessPoolExecutor(8) as ex:
        parts = list(ex.map(_meme_cd2_chunk, chunks))
    res = {}
    for ch, part in zip(chunks, parts):
        res.update(zip(ch, part))
    mt = json.load(open(os.path.join(GOLDEN, "cd2_meme_mle.json")))
    th = mt["thresholds"]
    lrt = [res[aln.site_to_pattern[s]]["lrt"] for s in range(187)]
    pv = [res[aln.site_to_pattern[s]]["p_value"] for s in range(187)]
    r_lrt, rho_lrt = correlations(lrt, [m[5] for m in mt["mle"]])
    r_p, rho_p = correlations(pv, [m[6] for m in mt["mle"]])
    assert r_lrt >= th["lrt_pearson"] and rho_lrt >= th["lrt_spearman"]
    assert r_p >= th["p_pearson"] and rho_p >= th["p_spearman"]


def _meme_cd2_chunk(patterns):
    from test_fel_cpu import load_cd2_fel
    tree, aln, truth = load_cd2_fel()
    g = truth["global_mg94"]
    prob = ho.Problem.from_refs(tree, aln)
    return [r for r, _ in mr.meme(prob, ho.Code(0), g["theta"], aln.pos_freqs, aln.cf3x4, branch_lengths(tree),
                                  g["alpha"], g["beta"], patterns=patterns)]
