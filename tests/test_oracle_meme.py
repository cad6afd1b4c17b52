"""Pins of the oracle's MEME per-site likelihood (oracle/meme_ref.py,
meme.hpp:219-277) against the FEL candidate oracle (fel.hpp:367-385, pinned
in test_fel_cpu.py), which builds its generators at alpha = 1 and scales
time instead: with p1 = 1 MEME is FEL at (a, a w1); with p1 = 0 it is FEL at
(a, b_plus); background branches at b_bg = b_plus change nothing when
p1 = 0."""
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
