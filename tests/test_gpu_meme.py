"""MEME on the GPU:

  * per-site candidates (hy_meme_eval, k_meme) against the oracle's
    restatement of eval_pattern_meme_lnl (meme.hpp:219-277): many candidates
    over many patterns in one call, all branches tested and with background
    branches, p1 at its bounds, repeated omegas (the spectrum pool), tips
    with gaps and ambiguity; the refusals;
  * the EBF conditionals (one branch held in one class) against the oracle
    and against the linearity identity p1 L_1 + (1 - p1) L_+ = L;
  * the whole MEME (hy_meme_run) on CD2 against HyPhy 2.5's per-site table
    at the reference's own parity thresholds, and against the oracle's
    sequential MEME on a few patterns."""
import numpy as np
import pytest

from helpers import branch_lengths, make_engine, rel, sim_problem
from oracle import hyoracle as ho
from oracle import meme_ref as mr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("background", [False, True])
def test_meme_candidates_match_oracle(background):
    tree, aln, sim = sim_problem(12, 60, 92, gaps=0.15)
    prob = ho.Problem.from_refs(tree, aln)
    bl = branch_lengths(tree)
    eng = make_engine(tree, aln)
    rng = np.random.default_rng(5)
    is_test = (rng.random(len(bl)) < 0.6).astype(np.uint8) if background else None
    P = prob.num_patterns
    n = 48
    pat = rng.integers(0, P, n).astype(np.int32)
    alpha = rng.uniform(0.05, 3.0, n)
    omega1 = rng.uniform(0.0, 1.0, n)
    p1 = rng.uniform(0.0, 1.0, n)
    p1[:3] = [0.0, 1.0, 1.0 - 1e-6]
    bplus = rng.uniform(0.2, 20.0, n)
    bplus[5:9] = bplus[4]  # repeated omegas from the pool
    alpha[5:9] = alpha[4]
    bbg = rng.uniform(0.1, 3.0, n) if background else None
    got = eng.meme_eval(pat, alpha, omega1, p1, bplus, sim["theta"], bl, beta_bg=bbg, is_test=is_test)
    got2 = eng.meme_eval(pat, alpha, omega1, p1, bplus, sim["theta"], bl, beta_bg=bbg, is_test=is_test)
    np.testing.assert_array_equal(got, got2)  # pooled spectra: the same values again
    for k in range(n):
        orc = mr.MemeCandidateOracle(prob, ho.Code(0), sim["theta"], aln.pos_freqs, aln.cf3x4, bl, int(pat[k]),
                                     is_test=is_test)
        ref = orc(alpha[k], omega1[k], p1[k], bplus[k], bbg[k] if background else 0.0)
        assert rel(got[k], ref) < 1e-9, (k, got[k], ref)
    eng.close()


def test_meme_refusals():
    from hyphy3.engine import HyError
    tree, aln, sim = sim_problem(6, 20, 93)
    bl = branch_lengths(tree)
    eng = make_engine(tree, aln)
    is_test = np.ones(len(bl), dtype=np.uint8)
    is_test[[n.id for n in tree.nodes if n.is_leaf][0]] = 0
    with pytest.raises(HyError):  # background branches without beta_bg
        eng.meme_eval([0], [1.0], [0.5], [0.5], [2.0], sim["theta"], bl, is_test=is_test)
    with pytest.raises(HyError):  # p1 outside [0, 1]
        eng.meme_eval([0], [1.0], [0.5], [1.5], [2.0], sim["theta"], bl)
    with pytest.raises(HyError):  # pattern out of range
        eng.meme_eval([10 ** 6], [1.0], [0.5], [0.5], [2.0], sim["theta"], bl)
    eng.close()


def test_meme_conditioned_candidates():
    # EBF conditionals (one branch held in one class) against the oracle, and
    # mixing the two conditionals of any branch with (p1, 1 - p1) gives the
    # site likelihood back (the likelihood is linear in one branch's matrix)
    tree, aln, sim = sim_problem(9, 40, 97, gaps=0.1)
    prob = ho.Problem.from_refs(tree, aln)
    bl = branch_lengths(tree)
    eng = make_engine(tree, aln)
    N = len(bl)
    nodes = [n for n in range(N) if n != prob.root]
    pats = [0, 4, 9]
    a, w1, p1, bp = 0.8, 0.3, 0.65, 5.0
    pat, cn, cc = [], [], []
    for p in pats:
        pat.append(p)
        cn.append(-1)
        cc.append(0)
        for n in nodes:
            for c in (0, 1):
                pat.append(p)
                cn.append(n)
                cc.append(c)
    got = eng.meme_eval(pat, a, w1, p1, bp, sim["theta"], bl, cond_node=cn, cond_class=cc)
    k = 0
    for p in pats:
        orc = mr.MemeCandidateOracle(prob, ho.Code(0), sim["theta"], aln.pos_freqs, aln.cf3x4, bl, p)
        lmix = got[k]
        assert rel(lmix, orc(a, w1, p1, bp)) < 1e-9
        k += 1
        for n in nodes:
            l1, lp = got[k], got[k + 1]
            assert rel(l1, orc(a, w1, p1, bp, cond_node=n, cond_class=0)) < 1e-9, (p, n)
            assert rel(lp, orc(a, w1, p1, bp, cond_node=n, cond_class=1)) < 1e-9, (p, n)
            mix = np.log(p1 * np.exp(l1 - lmix) + (1.0 - p1) * np.exp(lp - lmix)) + lmix
            assert abs(mix - lmix) < 1e-10 * abs(lmix), (p, n)
            k += 2
    from hyphy3.engine import HyError
    with pytest.raises(HyError):  # the root is no branch
        eng.meme_eval([0], a, w1, p1, bp, sim["theta"], bl, cond_node=[prob.root], cond_class=[0])
    with pytest.raises(HyError):  # classes are 0 and 1
        eng.meme_eval([0], a, w1, p1, bp, sim["theta"], bl, cond_node=[nodes[0]], cond_class=[2])
    eng.close()


def test_meme_cd2_against_hyphy_and_oracle():
    """MEME on CD2 (the fitted tree and global MG94 rates of the reference's
    matched-tree FEL parity case) against HyPhy 2.5's per-site MEME table at
    the thresholds of the reference's own MEME parity test
    (tests/test_meme.cpp:66-103, which fits its own global model first), and
    against the oracle's sequential MEME on a few patterns."""
    import json
    import os

    import hyphy3
    from helpers import GOLDEN
    from hyphy3 import meme as pm
    from oracle import fel_ref as fr
    from test_fel_cpu import correlations, load_cd2_fel
    tree, aln, truth = load_cd2_fel()
    g = truth["global_mg94"]
    mt = json.load(open(os.path.join(GOLDEN, "cd2_meme_mle.json")))
    h_aln = hyphy3.Alignment.from_fasta(os.path.join(GOLDEN, "cd2.fna"))
    h_tree = hyphy3.Tree.from_newick_file(os.path.join(GOLDEN, "cd2_fitted.nwk"))
    print("[meme] CD2 MEME on the GPU", flush=True)
    meme = pm.MEME(h_tree, h_aln, g["theta"], g["alpha"], g["beta"])
    try:
        res = meme.run()
        st = meme.stats
        print(f"[meme] {st['patterns']} patterns, {st['calls']} launches, {st['candidates']} candidates, "
              f"{st['seconds']:.2f} s", flush=True)
        assert len(res) == 187
        mle = mt["mle"]
        th = mt["thresholds"]
        lrt = np.array([r["lrt"] for r in res])
        pv = np.array([r["p_value"] for r in res])
        r_lrt, rho_lrt = correlations(lrt, [m[5] for m in mle])
        r_p, rho_p = correlations(pv, [m[6] for m in mle])
        print(f"[meme] vs HyPhy 2.5: LRT r {r_lrt:.5f} rho {rho_lrt:.5f}, p r {r_p:.5f} rho {rho_p:.5f}", flush=True)
        assert r_lrt >= th["lrt_pearson"] and rho_lrt >= th["lrt_spearman"]
        assert r_p >= th["p_pearson"] and rho_p >= th["p_spearman"]
        for r in res:
            assert 0.0 <= r["p_value"] <= 1.0 and r["lrt"] >= 0.0
            assert 1e-6 <= r["p1"] <= 1.0 and abs(r["p1"] + r["p_plus"] - 1.0) < 1e-12
            assert r["branch_ebf"][meme.root] == 1.0
            assert r["branches_under_selection"] == sum(e >= 100.0 for e in r["branch_ebf"])
        # against the oracle's sequential MEME on a few patterns: both routes
        # evaluate each pattern's candidates in the same order with values
        # equal to ~1e-12, so the fits end at the same points unless a near
        # tie in a simplex comparison sends one elsewhere
        prob = ho.Problem.from_refs(tree, aln)
        by_pat = {}
        for s, p in enumerate(aln.site_to_pattern):
            by_pat.setdefault(int(p), res[s])
        sel = [p for p in sorted(by_pat) if by_pat[p]["lrt"] > 0.0][:3]
        sel += [p for p in sorted(by_pat) if by_pat[p]["lrt"] == 0.0 and not fr.is_invariant(prob.states[p])][:1]
        ref = mr.meme(prob, ho.Code(0), g["theta"], aln.pos_freqs, aln.cf3x4, branch_lengths(tree), g["alpha"],
                      g["beta"], patterns=sel)
        same = 0
        for p, (want, ebf) in zip(sel, ref):
            got = by_pat[p]
            print(f"[meme] pattern {p}: LRT {got['lrt']:.6f} (oracle {want['lrt']:.6f}), p {got['p_value']:.6f}, "
                  f"evaluations {got['evaluations']} (oracle {want['evaluations']}), length "
                  f"{got['total_branch_length']:.6f} (oracle {want['total_branch_length']:.6f})", flush=True)
            # the same number of evaluations: the same simplex path, so the
            # same fitted point and values equal to rounding; otherwise a near
            # tie (vertices clamped onto one point of a bound) took another
            # path to a point within the simplex tolerance (1e-5 in f and in
            # every coordinate).  EBFs are only compared on the same path:
            # with omega1 near 0 the class-1 conditional of a branch with a
            # non-synonymous change moves by orders of magnitude within that
            # tolerance
            if got["evaluations"] == want["evaluations"]:
                same += 1
                tol_f, tol_x = 1e-9, 1e-8
                for n, e in enumerate(ebf):
                    assert abs(got["branch_ebf"][n] - e) <= tol_x * max(1.0, e), (p, n)
            else:
                tol_f, tol_x = 1e-4, 1e-3
            assert abs(got["lrt"] - want["lrt"]) < tol_f * max(1.0, want["lrt"])
            assert abs(got["log_l_meme"] - want["log_l_meme"]) < tol_f * max(1.0, abs(want["log_l_meme"]))
            assert abs(got["total_branch_length"] - want["total_branch_length"]) < \
                tol_x * max(1.0, want["total_branch_length"])
        print(f"[meme] {same} of {len(sel)} patterns took the oracle's simplex path", flush=True)
    finally:
        meme.engine.close()
