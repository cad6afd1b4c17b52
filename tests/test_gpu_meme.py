"""MEME's per-site candidates on the GPU (hy_meme_eval, k_meme) against the
oracle's restatement of eval_pattern_meme_lnl (meme.hpp:219-277): many
candidates over many patterns in one call, all branches tested and with
background branches, p1 at its bounds, repeated omegas (the spectrum pool),
tips with gaps and ambiguity; and the refusals."""
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
