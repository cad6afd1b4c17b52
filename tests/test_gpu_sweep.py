"""Randomised parity sweep of the MI355X engine against the FP64 CPU oracle
(GPU): random topologies (including polytomies and caterpillars), taxon and
codon counts, gap fractions, omega mixtures, time conventions, tile widths
and both precisions, each checked for logL and the full gradient."""
import numpy as np
import pytest

from helpers import branch_lengths, grad_err, make_engine, rel
from oracle import data_ref as dr
from oracle import hyoracle as ho

pytestmark = pytest.mark.gpu

CODONS = dr.Code().sense_codons


def random_newick(rng, n):
    """Random topology with occasional polytomies and random branch lengths."""
    nodes = ["T%d:%.4f" % (i, rng.exponential(0.15) + 1e-3) for i in range(n)]
    while len(nodes) > 3:
        k = 2 if rng.uniform() < 0.85 or len(nodes) < 5 else int(rng.integers(3, min(5, len(nodes) - 1) + 1))
        idx = sorted(rng.choice(len(nodes), size=k, replace=False), reverse=True)
        picked = [nodes.pop(i) for i in idx]
        nodes.append("(%s):%.4f" % (",".join(picked), rng.exponential(0.1) + 1e-3))
    return "(%s);" % ",".join(nodes)


def random_alignment(rng, n, L, gaps):
    seqs = []
    for _ in range(n):
        cods = list(rng.choice(CODONS, size=L))
        for j in range(L):
            u = rng.uniform()
            if u < gaps:
                cods[j] = "---"
            elif u < gaps * 1.3:
                cods[j] = cods[j][:2] + "N"
        seqs.append("".join(cods))
    return seqs


CASES = [(seed, int(n), int(L)) for seed, n, L in zip(range(24), [4, 5, 6, 7, 9, 11, 13, 16, 20, 24, 31, 40] * 2,
                                                      [1, 3, 17, 16, 33, 48, 65, 80, 7, 120, 31, 64] * 2)]


@pytest.mark.parametrize("seed,n,L", CASES)
def test_random_problem(seed, n, L):
    rng = np.random.default_rng(1000 + seed)
    tree = dr.TreeRef.from_newick(random_newick(rng, n))
    names = ["T%d" % i for i in range(n)]
    aln = dr.AlignmentRef(names, random_alignment(rng, n, L, gaps=0.1 * (seed % 3)))
    K = 1 + seed % 3
    omegas = list(rng.uniform(0.05, 3.0, size=K))
    w = rng.dirichlet(np.ones(K))
    theta = rng.uniform(0.5, 2.0, size=5)
    conv = seed % 2
    N = len(tree.nodes)
    bl = branch_lengths(tree)
    mix = ho.Mixture(omegas, np.zeros(N, dtype=np.int32), [[(k, float(w[k])) for k in range(K)]], conv)
    ref = ho.mixture_eval(ho.Problem.from_refs(tree, aln), ho.Code(0), ho.params(1.0, 1.0, theta), aln.pos_freqs,
                          aln.cf3x4, mix, bl)
    eng = make_engine(tree, aln)
    eng.set_mixture(K, np.zeros(N, dtype=np.int32), [list(range(K))], convention=conv)
    for T in (16, 32, 64):
        eng.set_tile(T)
        for prec, tl, tg in ((0, 1e-9, 1e-7), (1, 1e-5, 1e-4)):
            r = eng.eval(bl, omegas, w, theta, precision=prec)
            msg = (seed, n, L, K, conv, T, prec)
            assert rel(r["log_likelihood"], ref["log_likelihood"]) < tl, msg
            for k in ("grad_bl", "grad_omega", "grad_w", "grad_theta"):
                assert grad_err(r[k], ref[k]) < tg, msg + (k,)
