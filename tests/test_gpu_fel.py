"""FEL per-site refits on the GPU (hy_fel_eval and hyphy3.fel):

  * candidate log-likelihoods against the oracle's eval_pattern_lnl
    restatement: FP64, 1e-9 relative wherever rounding of the transition
    matrices' entries cannot move lnL by more than 1e-10, else within twice
    that rounding bound (very short or very long times, where the eigen route
    of the reference is itself that inexact); over the start grid, the box corners
    and random points, each pattern's candidates changing between calls the
    way an optimiser moves them (warm-started spectra), on CD2 and on a random
    tree with polytomies, gaps and ambiguous codons; the generator mean rates
    against the oracle's;
  * the spectrum pool: repeated candidates (in one call and across calls)
    give identical values, and values stay right after the pool grows and
    evicts;
  * argument checks (out of range, negative or non-finite rates);
  * the whole FEL on CD2 (the reference's matched-tree parity case) against
    HyPhy 2.5's per-site table at the reference test's thresholds
    (tests/test_fel_cd2_parity.cpp:67-120) and against the oracle's
    sequential FEL.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from helpers import GOLDEN, branch_lengths, make_engine
from oracle import data_ref as dr
from oracle import fel_ref as fr
from oracle import hyoracle as ho
from test_fel_cpu import correlations, load_cd2_fel
from test_gpu_sweep import random_alignment, random_newick
from hyphy3 import fel as pf
from hyphy3.engine import HyError

pytestmark = pytest.mark.gpu


def check_candidates(eng, prob, aln, theta, blc, calls, label):
    code = ho.Code(0)
    got = []  # (pattern, alpha, beta, lnL, mean rate) in call order
    for ps, a, b in calls:
        out, mu = eng.fel_eval(ps, a, b, theta, blc, scale=True)
        got.extend(zip(ps.tolist(), a.tolist(), b.tolist(), out.tolist(), mu.tolist()))
    by_pattern = {}
    for g in got:
        by_pattern.setdefault(g[0], []).append(g)

    def oracle(p):
        ev = fr.CandidateOracle(prob, code, theta, aln.pos_freqs, aln.cf3x4, blc, p)
        out = []
        for g in by_pattern[p]:
            ref = ev(g[1], g[2])
            out.append((g, ref, ev.rounding_bound()))
        return out

    with ThreadPoolExecutor(8) as ex:
        checked = [x for lst in ex.map(oracle, sorted(by_pattern)) for x in lst]
    worst, n_tight, n_wide = 0.0, 0, 0
    for (p, a, b, v, mu), ref, bound in checked:
        if ref == -1e20:
            assert v == -1e20, (label, p, a, b)
            continue
        err = abs(v - ref)
        # both routes round P's entries by at most the bound's d_b: where that
        # moves lnL by less than 1e-10 the two must agree to 1e-9 relative,
        # elsewhere to within twice the bound
        if bound < 1e-10 * max(1.0, abs(ref)):
            worst = max(worst, err / max(1.0, abs(ref)))
            assert err < 1e-9 * max(1.0, abs(ref)), (label, p, a, b, v, ref)
            n_tight += 1
        else:
            assert err <= 2.0 * bound + 1e-9 * max(1.0, abs(ref)), (label, p, a, b, v, ref, bound)
            n_wide += 1
        m = fr.mg94_scale(code, theta, aln.pos_freqs, aln.cf3x4, 1.0, b / max(a, 1e-12))
        assert abs(mu - m) < 1e-12 * m, (label, p, mu, m)
    print(f"[fel] {label}: {len(calls)} calls, {len(checked)} candidates ({n_tight} at 1e-9, {n_wide} within the "
          f"rounding bound), worst relative lnL error {worst:.2e}", flush=True)
    assert n_tight > 0.5 * len(checked)


def candidate_calls(rng, P, n_random):
    ps_all = np.arange(P, dtype=np.int32)
    calls = []
    for a, b in fr.START_GRID + ((1e-4, 1000.0), (1000.0, 1e-4), (0.01, 50.0), (1e-4, 1e-4)):
        calls.append((ps_all, np.full(P, a), np.full(P, b)))
    # optimiser-like walks: each pattern moves by small relative steps
    a = rng.uniform(0.05, 20.0, size=P)
    b = rng.uniform(0.05, 20.0, size=P)
    for _ in range(n_random):
        a = np.clip(a * np.exp(rng.normal(0.0, 0.2, size=P)), 1e-4, 1000.0)
        b = np.clip(b * np.exp(rng.normal(0.0, 0.2, size=P)), 1e-4, 1000.0)
        sub = np.sort(rng.choice(P, size=max(1, P // 2), replace=False)).astype(np.int32)
        calls.append((sub, a[sub].copy(), b[sub].copy()))
    return calls


def test_candidates_cd2():
    tree, aln, truth = load_cd2_fel()
    g = truth["global_mg94"]
    prob = ho.Problem.from_refs(tree, aln)
    scale = fr.mg94_scale(ho.Code(0), g["theta"], aln.pos_freqs, aln.cf3x4, g["alpha"], g["beta"])
    blc, _ = fr.converted_lengths(branch_lengths(tree), tree.root, scale)
    eng = make_engine(tree, aln)
    try:
        print("[fel] CD2 candidates", flush=True)
        check_candidates(eng, prob, aln, g["theta"], blc, candidate_calls(np.random.default_rng(31), prob.num_patterns, 6),
                         "cd2")
    finally:
        eng.close()


def test_candidates_polytomies_gaps():
    rng = np.random.default_rng(77)
    n = 24
    tree = dr.TreeRef.from_newick(random_newick(rng, n))
    aln = dr.AlignmentRef(["T%d" % i for i in range(n)], random_alignment(rng, n, 40, gaps=0.2))
    prob = ho.Problem.from_refs(tree, aln)
    theta = list(rng.uniform(0.3, 2.0, size=5))
    blc = branch_lengths(tree) * 4.0
    eng = make_engine(tree, aln)
    try:
        print("[fel] random tree with polytomies", flush=True)
        check_candidates(eng, prob, aln, theta, blc, candidate_calls(rng, prob.num_patterns, 4), "polytomies")
    finally:
        eng.close()


def test_argument_checks():
    tree, aln, truth = load_cd2_fel()
    g = truth["global_mg94"]
    bl = branch_lengths(tree)
    eng = make_engine(tree, aln)
    try:
        print("[fel] argument checks", flush=True)
        assert eng.fel_eval([], [], [], g["theta"], bl).size == 0
        with pytest.raises(HyError):
            eng.fel_eval([aln.num_patterns], [1.0], [1.0], g["theta"], bl)
        with pytest.raises(HyError):
            eng.fel_eval([0], [-1.0], [1.0], g["theta"], bl)
        with pytest.raises(HyError):
            eng.fel_eval([0], [1.0], [float("nan")], g["theta"], bl)
        # the handle still works after refused calls
        v = eng.fel_eval([0, 1], [1.0, 1.0], [1.0, 1.0], g["theta"], bl)
        assert np.all(np.isfinite(v))
    finally:
        eng.close()


def test_spectrum_pool_repeats_growth_eviction():
    tree, aln, truth = load_cd2_fel()
    g = truth["global_mg94"]
    prob = ho.Problem.from_refs(tree, aln)
    bl = branch_lengths(tree) * 3.1
    eng = make_engine(tree, aln)
    rng = np.random.default_rng(5)
    try:
        print("[fel] spectrum pool", flush=True)
        # the same candidate twice in one call, and again in the next call (pool hit)
        ps = np.array([4, 4, 9, 4], dtype=np.int32)
        a = np.array([0.7, 0.7, 0.7, 2.1])
        b = np.array([1.3, 1.3, 1.3, 3.9])  # the last one: same omega, other alpha
        v1 = eng.fel_eval(ps, a, b, g["theta"], bl)
        v2 = eng.fel_eval(ps, a, b, g["theta"], bl)
        assert v1[0] == v1[1] and np.array_equal(v1, v2)
        # three calls of 150 distinct omegas each: the pool grows past its
        # first size and then evicts; every value against the oracle
        checks = []
        for rep in range(3):
            n = 150
            pp = rng.integers(0, prob.num_patterns, size=n).astype(np.int32)
            aa = rng.uniform(0.2, 5.0, size=n)
            bb = aa * np.exp(rng.uniform(-3.0, 3.0, size=n))
            checks.append((pp, aa, bb, eng.fel_eval(pp, aa, bb, g["theta"], bl)))
        pp, aa, bb, _ = checks[0]  # the first call's omegas again, after growth and eviction
        checks.append((pp, aa, bb, eng.fel_eval(pp, aa, bb, g["theta"], bl)))
        worst = 0.0
        for pp, aa, bb, vv in checks:
            for i in range(len(pp)):
                ev = fr.CandidateOracle(prob, ho.Code(0), g["theta"], aln.pos_freqs, aln.cf3x4, bl, int(pp[i]))
                ref = ev(float(aa[i]), float(bb[i]))
                bound = ev.rounding_bound()
                err = abs(vv[i] - ref)
                assert err <= 2.0 * bound + 1e-9 * max(1.0, abs(ref)), (i, vv[i], ref, bound)
                if bound < 1e-10 * max(1.0, abs(ref)):
                    worst = max(worst, err / max(1.0, abs(ref)))
        print(f"[fel] pool: worst relative lnL error {worst:.2e}", flush=True)
        assert worst < 1e-9
        # a new theta empties the pool; values follow the new model
        th2 = list(np.asarray(g["theta"]) * 1.3)
        v3 = eng.fel_eval(ps[:1], a[:1], b[:1], th2, bl)
        ref = fr.CandidateOracle(prob, ho.Code(0), th2, aln.pos_freqs, aln.cf3x4, bl, 4)(0.7, 1.3)
        assert abs(v3[0] - ref) < 1e-9 * max(1.0, abs(ref))
    finally:
        eng.close()


def test_fel_cd2_against_hyphy_and_oracle():
    import hyphy3
    tree, aln, truth = load_cd2_fel()
    g = truth["global_mg94"]
    h_aln = hyphy3.Alignment.from_fasta(os.path.join(GOLDEN, "cd2.fna"))
    h_tree = hyphy3.Tree.from_newick_file(os.path.join(GOLDEN, "cd2_fitted.nwk"))
    print("[fel] CD2 FEL on the GPU", flush=True)
    fel = pf.FEL(h_tree, h_aln, g["theta"], g["alpha"], g["beta"])
    try:
        res = fel.run()
        st = fel.stats
        print(f"[fel] {st['patterns']} patterns, {st['calls']} launches, {st['candidates']} candidates, "
              f"{st['seconds'] * 1e3:.1f} ms", flush=True)
        assert len(res) == 187
        mle = truth["mle"]
        lrt = np.array([r["lrt"] for r in res])
        pv = np.array([r["p_value"] for r in res])
        th = truth["thresholds"]
        r_lrt, rho_lrt = correlations(lrt, [m[3] for m in mle])
        r_p, rho_p = correlations(pv, [m[4] for m in mle])
        print(f"[fel] vs HyPhy 2.5: LRT r {r_lrt:.5f} rho {rho_lrt:.5f}, p r {r_p:.5f} rho {rho_p:.5f}", flush=True)
        assert r_lrt > th["lrt_pearson"] and rho_lrt > th["lrt_spearman"]
        assert r_p > th["p_pearson"] and rho_p > th["p_spearman"]
        # against the oracle's sequential FEL: the same evaluation sequence per
        # pattern, candidates equal to ~1e-12, so most refits end identically;
        # a near-tie in a simplex comparison may still send one elsewhere
        prob = ho.Problem.from_refs(tree, aln)
        ref = fr.fel(prob, ho.Code(0), g["theta"], aln.pos_freqs, aln.cf3x4, branch_lengths(tree), g["alpha"],
                     g["beta"], threads=8)
        d_lrt = np.array([abs(res[s]["lrt"] - ref[aln.site_to_pattern[s]]["lrt"]) for s in range(187)])
        d_a = np.array([abs(res[s]["alpha"] - ref[aln.site_to_pattern[s]]["alpha"]) /
                        max(1.0, ref[aln.site_to_pattern[s]]["alpha"]) for s in range(187)])
        print(f"[fel] vs oracle: LRT |diff| median {np.median(d_lrt):.2e} max {d_lrt.max():.2e}; "
              f"alpha rel max {d_a.max():.2e}; sites within 1e-6: {(d_lrt < 1e-6).mean():.3f}", flush=True)
        assert (d_lrt < 1e-6).mean() > 0.9
        assert d_lrt.max() < 0.05
        r_o, _ = correlations(lrt, [ref[aln.site_to_pattern[s]]["lrt"] for s in range(187)])
        assert r_o > 0.99999
    finally:
        fel.engine.close()
