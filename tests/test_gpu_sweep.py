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


LARGE = [(201, 48, 900, 3, 0, 0.0), (202, 80, 1500, 2, 1, 0.1), (203, 128, 600, 1, 0, 0.05),
         (204, 180, 1200, 3, 1, 0.0), (205, 100, 2400, 2, 0, 0.1), (206, 150, 300, 3, 1, 0.2)]


@pytest.mark.parametrize("seed,n,L,K,conv,gaps", LARGE)
def test_larger_problems_split_trees(seed, n, L, K, conv, gaps):
    """Simulated alignments large enough for subtree parts, the lean FP32
    pruning variants (with pending messages spilled to HBM in deep parts) and
    the default variants on the top part; the engine's own tile choice, then
    the autotuner's, against the oracle; some with test/background classes."""
    from helpers import sim_problem
    rng = np.random.default_rng(seed)
    omegas = list(rng.uniform(0.05, 3.0, size=K))
    w = list(rng.dirichlet(np.ones(K)))
    tree, aln, sim = sim_problem(n, L, seed, omegas=omegas, weights=w, gaps=gaps)
    N = len(tree.nodes)
    two = seed % 2 == 0
    bc = np.array([0 if (not two or i % 3) else 1 for i in range(N)], dtype=np.int32)
    om_all = omegas + ([0.6] if two else [])
    classes_o = [[(k, w[k]) for k in range(K)]] + ([[(K, 1.0)]] if two else [])
    classes_e = [list(range(K))] + ([[K]] if two else [])
    cw = w + ([1.0] if two else [])
    bl = branch_lengths(tree)
    mix = ho.Mixture(om_all, bc, classes_o, conv)
    ref = ho.mixture_eval(ho.Problem.from_refs(tree, aln), ho.Code(0), ho.params(1.0, 1.0, sim["theta"]),
                          aln.pos_freqs, aln.cf3x4, mix, bl, threads=16)
    eng = make_engine(tree, aln)
    eng.set_mixture(len(om_all), bc, classes_e, convention=conv)
    for tuned in (False, True):
        for prec, tl, tg in ((0, 1e-9, 1e-7), (1, 1e-5, 1e-4)):
            if tuned:
                eng.autotune(bl, om_all, cw, sim["theta"], precision=prec, reps=1)
            r = eng.eval(bl, om_all, cw, sim["theta"], precision=prec)
            msg = (seed, n, L, K, conv, gaps, tuned, prec, eng.last_tile(), eng.last_parts())
            print(msg, r["log_likelihood"], ref["log_likelihood"])
            if not tuned:
                assert eng.last_parts() > 0, msg  # few tiles: split trees
            assert rel(r["log_likelihood"], ref["log_likelihood"]) < tl, msg
            for k in ("grad_bl", "grad_omega", "grad_w", "grad_theta"):
                assert grad_err(r[k], ref[k]) < tg, msg + (k,)
    eng.close()
